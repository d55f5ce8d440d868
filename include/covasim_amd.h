/*
 * covasim_amd.h -- C ABI of the MI355X (gfx950) daily-transmission-step library.
 *
 * Plain extern "C" entry points over plain device pointers and sizes; no C++
 * or torch types cross the boundary. All pointers are device pointers unless
 * a parameter name ends in _host. Every call is asynchronous on the caller's
 * HIP stream (hipStream_t passed as void*), returns 0 on success and a
 * negative CVK_E* code otherwise; cvk_last_error() gives the message of the
 * most recent failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the Covasim 3.1.7 checkout):
 *   cvk_viral_load     covasim/utils.py:39-79   compute_viral_load
 *   cvk_trans_sus      covasim/utils.py:82-90   compute_trans_sus
 *   cvk_infections     covasim/utils.py:93-128  compute_infections
 *   cvk_agent_pass     covasim/people.py:164-186 People.update_states_pre (+ check_* :211-312, :368-374)
 *                      covasim/immunity.py:303-350 check_immunity, utils.py:39-79 (viral-load phase)
 *   cvk_resample_layer covasim/base.py:1849-1876 Layer.update (frac = 1)
 *   cvk_resample_layer_frac covasim/base.py:1866-1875 Layer.update (frac < 1)
 *   cvk_choose         covasim/utils.py:430-443 choose / immunity.py:117-130 variant import choice
 *   cvk_edge_pass      covasim/sim.py:611-649   variant x layer loop of compute_trans_sus + compute_infections,
 *                      with People.infect's first-occurrence dedup (people.py:465) fused as a min-key
 *   cvk_infect_pass    covasim/people.py:435-586 People.infect (+ immunity.py:138-202 update_peak_nab)
 *   cvk_count_pass     covasim/sim.py:651-674   stock counts (base.py:1133-1145) + immunity.py:205-213 update_nab
 *   cvk_run_days       covasim/sim.py:558-685   Sim.step, repeated (the day loop of Sim.run, sim.py:733-755)
 *   cvk_pop_*          covasim/population.py:143-364 make_randpop / make_*_contacts,
 *                      people.py:139-160 set_prognoses (population generation, before the day loop)
 *   cvk_finalize       covasim/sim.py:764-1001 finalize, compute_states, compute_doubling, compute_r_eff
 *   cvk_reduce         covasim/run.py:220-303  MultiSim.reduce (median / quantiles or mean / std)
 *   cvk_move_edges     covasim/interventions.py:589-673 clip_edges (Layer.pop_inds / append, base.py:1739-1768)
 *   cvk_diag_apply     covasim/interventions.py:858-981 test_prob (People.test, people.py:589-617) and
 *                      :984-1151 contact_tracing (Layer.find_contacts, utils.py:131-147;
 *                      People.schedule_quarantine, people.py:620-641)
 *   cvk_states_post    covasim/people.py:189-196 update_states_post (check_diagnosed :315-332,
 *                      check_quar :335-358, check_enter_iso :361-365)
 */
#ifndef COVASIM_AMD_H
#define COVASIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVK_MAX_LAYERS 8
#define CVK_MAX_VARIANTS 8
#define CVK_N_DURS 9

/* status codes */
#define CVK_OK 0
#define CVK_E_INVALID (-1)   /* invalid argument (null pointer, bad size, bad enum) */
#define CVK_E_CAPACITY (-2)  /* an output buffer is too small */
#define CVK_E_HIP (-3)       /* a HIP runtime call failed */
#define CVK_E_ALIGN (-4)     /* a pointer is not aligned as required */

/* state bits of cvk_people.flags (defaults.py:54-71 order) */
#define CVK_SUSCEPTIBLE (1u << 0)
#define CVK_NAIVE (1u << 1)
#define CVK_EXPOSED (1u << 2)
#define CVK_INFECTIOUS (1u << 3)
#define CVK_SYMPTOMATIC (1u << 4)
#define CVK_SEVERE (1u << 5)
#define CVK_CRITICAL (1u << 6)
#define CVK_TESTED (1u << 7)
#define CVK_DIAGNOSED (1u << 8)
#define CVK_RECOVERED (1u << 9)
#define CVK_KNOWN_DEAD (1u << 10)
#define CVK_DEAD (1u << 11)
#define CVK_KNOWN_CONTACT (1u << 12)
#define CVK_QUARANTINED (1u << 13)
#define CVK_ISOLATED (1u << 14)
#define CVK_VACCINATED (1u << 15)
#define CVK_EARLY (1u << 16)             /* derived daily: in the high viral-load phase */
#define CVK_INFVAR_SHIFT 20              /* bits 20..23: infectious_variant, 0xF = undefined */
/* cvk_people.vstate: bits 0..3 exposed_variant, 4..7 recovered_variant (0xF = undefined),
 * 8..15 exposed_by_variant bitmask, 16..23 infectious_by_variant bitmask */
#define CVK_VAR_NONE 0xFu

/* duration distribution codes (utils.py:156-236 names) */
#define CVK_DIST_UNIFORM 0
#define CVK_DIST_NORMAL 1
#define CVK_DIST_NORMAL_POS 2
#define CVK_DIST_NORMAL_INT 3
#define CVK_DIST_LOGNORMAL 4
#define CVK_DIST_LOGNORMAL_INT 5

/* duration slots of cvk_params.dur (parameters.py:85-95) */
#define CVK_DUR_EXP2INF 0
#define CVK_DUR_INF2SYM 1
#define CVK_DUR_SYM2SEV 2
#define CVK_DUR_SEV2CRIT 3
#define CVK_DUR_ASYM2REC 4
#define CVK_DUR_MILD2REC 5
#define CVK_DUR_SEV2REC 6
#define CVK_DUR_CRIT2REC 7
#define CVK_DUR_CRIT2DIE 8

/* infection-log layer codes besides 0..n_layers-1 */
#define CVK_LAYER_SEED 14
#define CVK_LAYER_IMPORT 15

/* per-day result columns (counts, uint32) */
enum {
  CVK_R_NEW_INFECTIONS = 0, CVK_R_NEW_REINFECTIONS, CVK_R_NEW_INFECTIOUS, CVK_R_NEW_SYMPTOMATIC,
  CVK_R_NEW_SEVERE, CVK_R_NEW_CRITICAL, CVK_R_NEW_RECOVERIES, CVK_R_NEW_DEATHS, CVK_R_NEW_KNOWN_DEATHS,
  CVK_R_N_SUSCEPTIBLE, CVK_R_N_EXPOSED, CVK_R_N_INFECTIOUS, CVK_R_N_SYMPTOMATIC, CVK_R_N_SEVERE,
  CVK_R_N_CRITICAL, CVK_R_N_RECOVERED, CVK_R_N_DEAD, CVK_R_N_DIAGNOSED, CVK_R_N_KNOWN_DEAD,
  CVK_R_N_QUARANTINED, CVK_R_N_ISOLATED, CVK_R_N_VACCINATED, CVK_R_N_IMPORTS,
  CVK_R_MID_SEVERE, CVK_R_MID_CRITICAL,   /* counts after update_states_pre, for bed limits */
  CVK_R_N_ALIVE_NAB,                      /* alive agents (denominator of pop_nabs) */
  CVK_R_NEW_TESTS, CVK_R_NEW_DIAGNOSES, CVK_R_NEW_QUARANTINED, CVK_R_NEW_ISOLATED,  /* testing and tracing */
  CVK_R_BYVAR,                            /* start of 6 x CVK_MAX_VARIANTS by-variant columns */
  CVK_R_COLS = CVK_R_BYVAR + 6 * CVK_MAX_VARIANTS
};
/* by-variant column offsets: CVK_R_BYVAR + k*CVK_MAX_VARIANTS + v */
#define CVK_RV_NEW_INFECTIONS 0
#define CVK_RV_NEW_SYMPTOMATIC 1
#define CVK_RV_NEW_SEVERE 2
#define CVK_RV_NEW_INFECTIOUS 3
#define CVK_RV_N_EXPOSED 4
#define CVK_RV_N_INFECTIOUS 5
/* per-day float columns (double): pop_nabs numerator, sus_imm sum, symp_imm sum */
#define CVK_RF_NAB_SUM 0
#define CVK_RF_SUS_IMM_SUM 1
#define CVK_RF_SYMP_IMM_SUM 2
#define CVK_RF_COLS 4

typedef struct {
  int32_t dist;   /* CVK_DIST_* */
  int32_t pad;
  double par1, par2;   /* as given to sample() */
  double mu, sigma;    /* lognormal: parameters of the underlying normal (utils.py:223-227) */
} cvk_dur;

/* Scalar parameters of one sim (parameters.py:15-151). Plain old data, copied by value. */
typedef struct {
  int64_t n_agents;
  int32_t n_variants, n_layers;
  uint64_t seed;                 /* Philox key */
  float frac_time, load_ratio, high_cap;   /* viral_dist */
  float vl_hi, vl_lo;            /* viral load in the early / late phase (utils.py:78), float32 */
  float asymp_factor;
  float beta_v[CVK_MAX_VARIANTS];          /* default_float(beta*rel_beta*variant rel_beta), sim.py:627 */
  float beta_layer[CVK_MAX_LAYERS], iso_factor[CVK_MAX_LAYERS], quar_factor[CVK_MAX_LAYERS];
  float rel_symp_prob[CVK_MAX_VARIANTS], rel_severe_prob[CVK_MAX_VARIANTS];
  float rel_crit_prob[CVK_MAX_VARIANTS], rel_death_prob[CVK_MAX_VARIANTS];
  float no_hosp_factor, no_icu_factor, trans_redux, nab_boost;
  int64_t n_beds_hosp, n_beds_icu;         /* -1 = no constraint */
  cvk_dur dur[CVK_N_DURS];
  cvk_dur nab_init;
  int32_t use_waning, nab_kin_len;
  double rel_imm_symp[3];        /* asymptomatic, mild, severe */
  double nab_norm;               /* 1 + alpha_inf_diff */
  double ve_exp_alpha[3], ve_beta[3];   /* sus, symp, sev: exp(alpha), beta (immunity.py:216-247) */
  float immunity[CVK_MAX_VARIANTS][CVK_MAX_VARIANTS];
  const double* nab_kin;         /* device, length nab_kin_len (immunity.py:356-401) */
} cvk_params;

/* Agent state, structure of arrays of length n (device pointers). */
typedef struct {
  int64_t n;
  uint32_t* flags;
  uint32_t* vstate;
  float *date_exposed, *date_infectious, *date_symptomatic, *date_severe, *date_critical;
  float *date_recovered, *date_dead, *date_diagnosed, *date_end_isolation;
  float *dur_exp2inf, *dur_inf2sym, *dur_sym2sev, *dur_sev2crit, *dur_disease;
  float *rel_trans, *rel_sus, *symp_prob, *severe_prob, *crit_prob, *death_prob;
  int32_t *n_infections, *n_breakthroughs;
  float *nab, *peak_nab;
  int32_t* t_nab_event;
  float *sus_imm, *symp_imm, *sev_imm;   /* [n_variants][n] */
} cvk_people;

/* No reference counterpart: source lists of a static layer, both directions of every stored pair
 * grouped by source agent. The slots [indptr[i], indptr[i + 1]) of agent i hold target[k] and
 * edge_dir[k] = edge index << 1 | direction (0: agent i is p1, 1: it is p2). Self-loops and edges
 * with an endpoint outside [0, n) are left out (they cannot transmit). Built by cvk_sources_build;
 * the order of an agent's slots is unspecified (results do not depend on it). */
typedef struct {
  const int32_t* indptr;     /* [n + 1] */
  const int32_t* target;     /* [n_slots] */
  const uint32_t* edge_dir;  /* [n_slots] */
  int64_t n_slots;
} cvk_sources;

/* One contact layer: int32 endpoints and float32 per-edge beta (base.py:1610-1676). */
typedef struct {
  int32_t* p1;
  int32_t* p2;
  float* beta;
  int64_t n_edges;
  int32_t dynamic;     /* resampled each day (dynam_layer); the resample sets every beta to 1 */
  int32_t beta_ones;   /* nonzero when every beta[e] is 1.0f (the edge pass then skips the beta gather) */
  const cvk_sources* sources;  /* optional (static layers): the edge pass walks the lists of the day's
                                  infectious agents instead of streaming p1/p2; must describe p1/p2 */
} cvk_layer;

/* Device workspace (sizes from cvk_workspace_sizes). */
typedef struct {
  uint64_t* winner;      /* [n] min transmission key per target; all-ones = none */
  uint32_t* touched;     /* [(n + 31) / 32] bit i: agent i holds a key this day (cleared by the infect pass) */
  uint32_t* bitmap;      /* [bitmap_words] bit i: agent i infectious with rel_trans != 0 */
  uint32_t* ctl;         /* [64] control words: log count, choose state, flags */
  uint32_t* results;     /* [npts][CVK_R_COLS] */
  double* results_f;     /* [npts][CVK_RF_COLS] */
  int32_t* log;          /* [log_capacity][4] infection records: source, target, day, layer | variant<<8 */
  int64_t log_capacity;
  uint32_t* select_hist; /* [256] histogram scratch of cvk_choose */
  uint32_t* coarse;      /* [>= 4 * ceil(N / 2048)] bit j of word w: some agent of [512w + 16j, 512w + 16j + 16)
                            is infectious (OR of bitmap bits); the edge pass stages it in LDS when the
                            bitmap does not fit there (the caller may allocate bitmap_words) */
  uint64_t* rec;         /* [n] per-agent edge-pass record: low word float (rel_trans if infectious,
                            else rel_sus), high word the flags */
  uint32_t* partials;    /* [CVK_PARTIAL_COLS][CVK_MAX_REDUCE_BLOCKS] per-block counts (column-major) */
  double* partials_f;    /* [CVK_RF_COLS][CVK_MAX_REDUCE_BLOCKS] per-block sums (column-major) */
  uint16_t* bound;       /* [n] per-agent upper bound of the edge-pass factor: infectious agents
                            variant << 13 | code of rel_trans x own factors, susceptible agents the code
                            of rel_sus x quarantine factor, others 0; code = float bits >> 18 rounded up */
  int32_t npts;
  int32_t pad;
  uint64_t* day_clock;   /* optional [3 * npts + 1] (null: off): the device wall clock when day t's first pass
                            starts (entry t), and when the closing pass of a run's last day starts (entry
                            t1); a day's duration in a run is day_clock[t + 1] - day_clock[t] ticks of
                            cvk_wall_clock_rate Hz. Entries npts + 1 + 2t and npts + 2 + 2t: the earliest
                            workgroup start (atomic min) and the latest workgroup end (atomic max) of day
                            t's edge pass; the caller sets them to ~0 and 0 before the run. No reference
                            counterpart (measurement). */
  uint32_t* aux;         /* optional [bitmap_words] (null: not kept): bit i = agent i is susceptible with rel_sus != 0
                            and not infectious (it can be a target); written by the agent pass of the fused day
                            loop when its edge pass reads it (CVK_TUNE_TARGET_BITS). No reference counterpart. */
} cvk_work;

#define CVK_MAX_REDUCE_BLOCKS 65536
#define CVK_PARTIAL_COLS 64

/* control word indices of cvk_work.ctl */
#define CVK_CTL_UNUSED0 0          /* reserved */
#define CVK_CTL_LOG_COUNT 2       /* 64-bit counter in words 2 (lo) and 3 (hi) */
#define CVK_CTL_LOG_OVERFLOW 4
#define CVK_CTL_SEL_PREFIX 6      /* cvk_choose radix-select state: 64-bit prefix in words 6, 7 */
#define CVK_CTL_SEL_REMAIN 8
#define CVK_CTL_SEL_ALL 9
#define CVK_CTL_SEL_SHORT 10      /* set when fewer agents were eligible than requested */
#define CVK_CTL_SRC_SLOTS 12      /* 64-bit count (words 12, 13) of source-list slots walked by edge passes,
                                     when CVK_TUNE_COUNT_SOURCES is on */
#define CVK_CTL_SRC_AGENTS 14     /* 64-bit count (words 14, 15) of (agent, layer) lists walked */
#define CVK_CTL_SRC_STREAMED 16   /* 64-bit count (words 16, 17) of edge passes that streamed the layers with
                                     source lists (walk decision: too many infectious agents) */
#define CVK_CTL_WORDS 64

/* ---- library ---- */
const char* cvk_last_error(void);
/* ABI version: changes whenever a public struct, an entry point's signature or a tunable's meaning
 * changes. Bindings compare it with the value they were written for. */
#define CVK_ABI_VERSION 9
int cvk_version(void);
/* No reference counterpart: the frequency (Hz) of the device wall clock of cvk_work.day_clock. */
int cvk_wall_clock_rate(int64_t* hz_out);
/* No reference counterpart: a 64-bit hash of the HIP and header sources the library was built from
 * (set by the build script), so measurements taken on one kernel build can be told from another's. */
uint64_t cvk_source_hash(void);
/* No reference counterpart: ABI self-description for bindings. Writes sizeof/offsetof values of
 * the structs above into out[0..m) and returns m (see cvk_day.hip for the order). */
int cvk_abi_layout(int64_t* out, int32_t n);
/* No reference counterpart: launch tunables (process-wide). CVK_TUNE_EDGE_LDS = 0 keeps the edge
 * pass's infectious bitmap in global memory (L2) even when it fits in LDS; 1 (default) stages it. */
#define CVK_TUNE_EDGE_LDS 1
/* CVK_TUNE_EDGE_CHUNKS: 4-edge chunks per lane of an edge-pass wave tile, 2 (default) or 4. */
#define CVK_TUNE_EDGE_CHUNKS 2
/* CVK_TUNE_GENERATE_DYNAMIC: 1 = inside cvk_run_days / graphs, the edge and infect passes compute
 * each day's dynamic-layer endpoints from the resample stream, and the layer arrays are written
 * once, with the last day's contacts, when the run closes its last day; 0 = every day's resample is
 * written to the layer and read back; 2 (default) = 1 from 2^25 dynamic edges up, else 0 (never with
 * testing and tracing, which read each day's contacts). The results are identical. */
#define CVK_TUNE_GENERATE_DYNAMIC 4
/* CVK_TUNE_OVERLAP_RESAMPLE: 1 (default) = inside cvk_run_days / graphs, day t+1's resample of the
 * dynamic layers runs as extra workgroups of day t's infect pass, and that pass takes dynamic-layer
 * sources from the resample stream; 0 = each day's resample runs inside that day's agent pass. */
#define CVK_TUNE_OVERLAP_RESAMPLE 5
/* CVK_TUNE_SOURCE_LISTS: static layers with source lists are walked through the day's infectious agents'
 * lists: 2 (default) on days when at most CVK_TUNE_SOURCE_PERMILLE (default 30) agents per thousand are
 * infectious (streamed on the others), 1 always, 0 never. CVK_TUNE_COUNT_SOURCES: 1 = edge passes add the
 * slots and lists they walk to the CVK_CTL_SRC_* counters (0 default). The results are identical either way. */
#define CVK_TUNE_SOURCE_LISTS 8
#define CVK_TUNE_COUNT_SOURCES 9
#define CVK_TUNE_SOURCE_PERMILLE 10
/* CVK_TUNE_EDGE_COARSE: 1 (default) = when the infectious bitmap does not fit in LDS, the edge pass
 * stages the coarse map there and reads bitmap words from L2 only for endpoints whose 16-agent group
 * has an infectious agent; 0 = every bitmap word is read from L2. */
#define CVK_TUNE_EDGE_COARSE 6
/* CVK_TUNE_EDGE_PREFETCH: 1 = each edge-pass wave issues its next tile's endpoint loads after this
 * tile's bound loads, before its candidates are processed; 0 (default) = after. */
#define CVK_TUNE_EDGE_PREFETCH 7
/* CVK_TUNE_TARGET_BITS: inside cvk_run_days / graphs, with cvk_work.aux given and the infectious bitmap not
 * staged in LDS (large populations, whose codes miss L2), 1 (default) = a mixed edge of a wave tile reads its
 * target's aux bit (L2-resident, N/8 bytes) instead of the target's bound code, and its prefilter bound takes
 * the day's largest target code (the agent pass's maximum) with the source's own code; 0 = both endpoints'
 * codes are gathered. The results are identical. */
#define CVK_TUNE_TARGET_BITS 11
int cvk_set_tuning(int32_t key, int32_t value);
/* No reference counterpart: compute units and LDS bytes per workgroup of the current device. */
int cvk_device_info(int32_t* n_cu, int32_t* max_lds_per_block);
/* Sizes in bytes of the workspace arrays for n agents and npts days. */
int cvk_workspace_sizes(int64_t n_agents, int32_t npts, int64_t* winner_bytes, int64_t* touched_bytes,
                        int64_t* bitmap_bytes, int64_t* results_bytes, int64_t* results_f_bytes);
/* Bytes of the remaining workspace arrays: rec (8 per agent), partials, partials_f and bound (2 per agent). */
int cvk_workspace_sizes2(int64_t n_agents, int64_t* rec_bytes, int64_t* partials_bytes, int64_t* partials_f_bytes,
                         int64_t* bound_bytes);

/* ---- 1:1 mirrors of the reference's numba kernels (parity boundary) ---- */

/* utils.py:39-79. out[i] = viral load of agent i on day t. */
int cvk_viral_load(int32_t t, const float* time_start, const float* time_recovered, const float* time_dead,
                   int64_t n, float frac_time, float load_ratio, float high_cap, float* out, void* stream);

/* utils.py:82-90. Boolean inputs are uint8 (0/1) arrays of length n. */
int cvk_trans_sus(const float* rel_trans, const float* rel_sus, const uint8_t* inf, const uint8_t* sus,
                  float beta_layer, const float* viral_load, const uint8_t* symp, const uint8_t* iso,
                  const uint8_t* quar, float asymp_factor, float iso_factor, float quar_factor,
                  const float* immunity_factors, int64_t n, float* rel_trans_out, float* rel_sus_out,
                  void* stream);

/* utils.py:93-128. Writes the transmitting (source, target) pairs in the reference's order:
 * direction (p1->p2, then p2->p1, or only p1->p2 if legacy), then ascending edge index.
 * uniforms: NULL for counter-based Philox draws keyed by (seed, t, variant, layer), else a
 * device float64 array [2][n_edges] of per-(direction, edge) uniforms. prob_out (may be NULL)
 * receives the float32 probability of each emitted pair. *count_host receives the number of
 * pairs; returns CVK_E_CAPACITY (after writing count) if it exceeds capacity. Synchronises
 * the stream to read the count. */
int cvk_infections(float beta, const int32_t* p1, const int32_t* p2, const float* layer_betas,
                   int64_t n_edges, const float* rel_trans, const float* rel_sus, int64_t n_agents,
                   const double* uniforms, uint64_t seed, int32_t t, int32_t variant, int32_t layer,
                   int32_t legacy, int32_t* src_out, int32_t* tgt_out, float* prob_out, int64_t capacity,
                   int64_t* count_host, void* stream);

/* No reference counterpart: cvk_infections without any allocation, for callers that cannot call
 * hipMalloc (serving loops, harnesses that forbid allocation). Same arguments, results and errors
 * as cvk_infections; its per-block counts live in a fixed device array of the library (256 KiB per
 * GPU, held until the process ends) instead of a per-call allocation. Limit: at most
 * CVK_INFECTIONS_NOALLOC_MAX_SLOTS edge slots (2 * n_edges, or n_edges with legacy != 0); larger
 * calls return CVK_E_CAPACITY before any device work. Calls that overlap on one GPU (two host
 * threads, or two streams) share that array and race: serialise them. */
#define CVK_INFECTIONS_NOALLOC_MAX_BLOCKS 65536
#define CVK_INFECTIONS_NOALLOC_MAX_SLOTS (CVK_INFECTIONS_NOALLOC_MAX_BLOCKS * 2048LL)
int cvk_infections_noalloc(float beta, const int32_t* p1, const int32_t* p2, const float* layer_betas,
                           int64_t n_edges, const float* rel_trans, const float* rel_sus, int64_t n_agents,
                           const double* uniforms, uint64_t seed, int32_t t, int32_t variant, int32_t layer,
                           int32_t legacy, int32_t* src_out, int32_t* tgt_out, float* prob_out, int64_t capacity,
                           int64_t* count_host, void* stream);

/* ---- production path: one day is agent_pass -> resample -> choose(imports) -> edge_pass ->
 *      infect_pass -> count_pass; cvk_run_days chains them for a range of days ---- */

/* Seed infections at t = 0 (sim.py:505-532): choose k distinct agents (eligible = all, or
 * susceptible only when susceptible_only != 0) with Philox selection keys on stream
 * `stream_id`, and register them as infections of class `layer_code` / `variant` in
 * work->winner / work->touched (bitmap). */
int cvk_choose(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
               uint32_t stream_id, int64_t k, int32_t susceptible_only, int32_t layer_code,
               int32_t variant, void* stream);

int cvk_agent_pass(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                   void* stream);
int cvk_resample_layer(const cvk_layer* layer, int32_t layer_index, const cvk_params* params, int32_t t,
                       void* stream);
/* Layer.update with frac < 1 (covasim/base.py:1866-1875): k = round(n_edges * frac) edges, a uniform
 * k-subset (the k smallest keys (Philox word of (edge, t, layer) << 32) | edge), get new endpoints,
 * uniform over [0, n_agents) with replacement, and weight 1; the other edges keep theirs. A chosen
 * edge gets exactly what cvk_resample_layer gives it on day t, so k = n_edges is the full resample.
 * work: cvk_resample_frac_work_bytes() bytes of device memory, 256-byte aligned. */
int64_t cvk_resample_frac_work_bytes(void);
int cvk_resample_layer_frac(const cvk_layer* layer, int32_t layer_index, const cvk_params* params, int32_t t,
                            int64_t k, void* work, int64_t work_bytes, void* stream);
/* uniforms: NULL (Philox) or per layer a device float64 array [n_variants][2][n_edges].
 * Must follow cvk_agent_pass of the same day (it reads that pass's bitmaps, records and bound codes). */
int cvk_edge_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                  cvk_work* work, int32_t t, const double* const* uniforms, void* stream);
int cvk_infect_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                    cvk_work* work, int32_t t, int32_t count_flows, void* stream);
int cvk_count_pass(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                   void* stream);

/* ---- testing, contact tracing, quarantine and isolation ----
 * One cvk_diag_op is one test_prob (interventions.py:858-981) or contact_tracing
 * (interventions.py:984-1151) intervention; cvk_diag_apply runs it on day t (nothing outside its
 * days), after the day's imports and before transmission, in the sim's intervention order;
 * cvk_states_post then runs update_states_post (people.py:189-196) once per day.
 * Draws (Philox4x32-10 keyed by the seed): the test of agent i on day t uses
 * (i, t, 0x90000000 | slot, 0) -- words 0,1: tested (u < probability), words 2,3: true positive
 * (u < sensitivity) -- and (i, t, 0x90000000 | slot, 1) words 0,1: not lost (u < 1 - loss_prob);
 * the tracing of contact i found in layer l uses (i, t, 0xA0000000 | slot << 8 | l, 0), words 0,1
 * (u < trace_prob[l]). Uniforms are 53-bit doubles, compared with the double probabilities. The options
 * (subtarget, ILI, swab delay, tracing capacity) follow the fields below. */
#define CVK_OP_TEST 1
#define CVK_OP_TRACE 2
/* quarantine testing policies of get_quar_inds (interventions.py:682-710) */
#define CVK_QP_NONE 0    /* nobody is tested as quarantined */
#define CVK_QP_START 1   /* date_quarantined == t - 1 */
#define CVK_QP_END 2     /* date_end_quarantine == t + 1 */
#define CVK_QP_BOTH 3    /* either of the two */
#define CVK_QP_DAILY 4   /* quarantined */
#define CVK_QP_DAYS 5    /* date_quarantined == t - 1 - q for some q in quar_days */
#define CVK_MAX_QUAR_DAYS 8
#define CVK_PENDING_NONE INT32_MIN

typedef struct {
  int32_t kind;                 /* CVK_OP_TEST or CVK_OP_TRACE */
  int32_t slot;                 /* 0..255, the Philox stream slot: the intervention's index in the sim */
  int32_t start_day, end_day;   /* acts on days start_day <= t <= end_day; end_day < 0: no end */
  /* CVK_OP_TEST: test probabilities by symptom and quarantine-test state (diagnosed agents: 0) */
  double symp_prob, asymp_prob, symp_quar_prob, asymp_quar_prob;
  double sensitivity;
  double keep_prob;             /* 1 - loss_prob */
  double test_delay;            /* date_diagnosed = float32(t + test_delay) */
  int32_t quar_policy;          /* CVK_QP_* */
  int32_t n_quar_days;
  int32_t quar_days[CVK_MAX_QUAR_DAYS];
  /* subtarget (interventions.py:150-187, 958-960): the agents in the device bitmap sub_mask
   * [(n + 31) / 32] test with probability sub_vals[i] (device [n]), or sub_val when sub_vals is null,
   * in place of the probability above (diagnosed agents still 0); null sub_mask: no subtarget */
  const uint32_t* sub_mask;
  const double* sub_vals;
  double sub_val;
  /* ili_prev (interventions.py:936-942): on day start_day + r, r < n_ili_days, ili_counts[r] agents
   * (host array) chosen uniformly without replacement have influenza-like symptoms; those without
   * COVID symptoms test with symp_prob (quarantined or not). The choice is the ili_counts[r] smallest
   * keys (Philox (i, t, 0xC0000000 | slot, 0) word 0 << 32) | i. Null ili_counts: no ILI. */
  const int64_t* ili_counts;
  int32_t n_ili_days;
  /* swab_delay (interventions.py:924-933): n_swab_bins > 0 replaces symp_prob, for the symptomatic
   * agents outside quarantine testing, by swab_base[s] * (n_symp / count[s]), s = t - date_symptomatic,
   * count[s] = symptomatic agents at s days since onset, n_symp = all symptomatic agents;
   * swab_base (device [n_swab_bins]) = pdf(s) * (symp_prob / (1 - s * symp_prob) when s * symp_prob < 1,
   * else 1); n_swab_bins must exceed every s (the sim's days suffice) and fit cvk_diag.swab_bins. */
  int32_t n_swab_bins;
  const double* swab_base;
  /* CVK_OP_TRACE */
  double trace_prob[CVK_MAX_LAYERS];      /* 0: the layer is not traced */
  double trace_time[CVK_MAX_LAYERS];      /* >= 0; quarantine starts on day int(t + trace_time) */
  int32_t trace_period[CVK_MAX_LAYERS];   /* int(quar_period - trace_time): days of quarantine */
  int32_t presumptive;          /* trace the exposed agents tested today, not those diagnosed today */
  int32_t pad1;
  /* capacity (interventions.py:1078-1083): >= 0: when more agents are to be traced today, capacity of
   * them chosen uniformly without replacement are traced -- the smallest keys (Philox (i, t,
   * 0xD0000000 | slot, 0) word 0 << 32) | i among them; < 0: no limit */
  int64_t capacity;
} cvk_diag_op;

/* Testing and quarantine state beside cvk_people (device arrays of n agents). */
typedef struct {
  float *date_tested, *date_pos_test, *date_known_contact, *date_quarantined, *date_end_quarantine;
  int32_t* pending;    /* [ring][n]: row d % ring holds, per agent, the latest end day of the quarantines
                          scheduled to start on day d (CVK_PENDING_NONE: none); initialise to NONE */
  uint64_t* cases;     /* [(n + 63) / 64] bitmap of the cases a tracing op is tracing (scratch) */
  uint32_t* coarse;    /* [cvk_diag_coarse_words(n)] bit k: some case among agents [k << c, (k + 1) << c),
                          c = the smallest shift that fits the map in 8 KiB; initialise to 0 (each
                          tracing op clears it) */
  uint32_t* contacts;  /* [n] bit l: found in layer l by the running tracing op; initialise to 0 (the op
                          clears what it set) */
  uint32_t* part;      /* [4][CVK_DIAG_MAX_BLOCKS] per-block counts of the day (tests, diagnoses,
                          quarantined, isolated); initialise to 0 (cvk_states_post sums them into the
                          day's row and clears them) */
  int32_t ring;        /* > the largest int(trace_time) of every tracing op */
  int32_t swab_bins;   /* room of swab_hist / swab_prob: >= every test op's n_swab_bins (0: none) */
  uint64_t* select;    /* [130] scratch of the uniform choices (ILI agents, traced cases under a capacity) */
  uint32_t* swab_hist; /* [swab_bins + 1] scratch: symptomatic agents by days since onset, then all of them */
  double* swab_prob;   /* [swab_bins] scratch: the day's test probability by days since onset */
} cvk_diag;

#define CVK_DIAG_MAX_BLOCKS 4096
int64_t cvk_diag_coarse_words(int64_t n);

/* Runs one test_prob / contact_tracing op on day t (its tests are counted by the next
 * cvk_states_post of the day). */
int cvk_diag_apply(const cvk_people* people, const cvk_layer* layers, const cvk_params* params, cvk_work* work,
                   const cvk_diag* diag, const cvk_diag_op* op, int32_t t, void* stream);
/* update_states_post on day t: check_diagnosed, check_quar, check_enter_iso, and the edge records
 * of the agents whose quarantine or isolation changed (so it runs after cvk_agent_pass and before
 * cvk_edge_pass); then the day's counts (the ops' tests and its own diagnoses, quarantines and
 * isolations) go to CVK_R_NEW_TESTS .. CVK_R_NEW_ISOLATED of row t. */
int cvk_states_post(const cvk_people* people, const cvk_params* params, cvk_work* work, const cvk_diag* diag,
                    int32_t t, void* stream);

/* Import schedule for cvk_run_days: on day days[i], n_imports[i] susceptible agents are
 * infected with variant variants[i] (immunity.py:117-130). Host arrays. */
typedef struct {
  int32_t n;
  const int32_t* days;
  const int32_t* variants;
  const int64_t* n_imports;
  /* Per-day transmission rates, as interventions set them (interventions.py:530-586 change_beta):
   * day t uses row min(t, n_beta_days - 1) of beta_v_days [n_beta_days][CVK_MAX_VARIANTS] (the
   * float32 beta * rel_beta * variant rel_beta of sim.py:627) and of beta_layer_days
   * [n_beta_days][CVK_MAX_LAYERS]; n_beta_days = 0 (or NULL tables) keeps cvk_params' values. */
  int32_t n_beta_days;
  const float* beta_v_days;
  const float* beta_layer_days;
  /* Testing and tracing: with diag != NULL each day runs diag_ops[0..n_diag_ops) (host array, the
   * sim's order) after the imports, then cvk_states_post, before transmission. */
  int32_t n_diag_ops;
  const cvk_diag_op* diag_ops;
  const cvk_diag* diag;
} cvk_imports;

/* No reference counterpart: how cvk_run_days / cvk_graph_create would run n_days days of these layers
 * under the current tunables. out[0] = 1 if dynamic-layer endpoints are generated (CVK_TUNE_GENERATE_DYNAMIC),
 * out[1] = 1 if each next day's resample runs beside the infect pass (CVK_TUNE_OVERLAP_RESAMPLE).
 * Host-only; reads no device memory. */
int cvk_day_plan(const cvk_layer* layers, const cvk_params* params, const cvk_imports* imports, int32_t n_days,
                 int32_t* out);
/* Runs days [t0, t1) (t0 == 0 also seeds pop_infected infections first). use_graph != 0
 * captures the launches into one hipGraph and launches it once. */
int cvk_run_days(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                 int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports,
                 int32_t use_graph, void* stream);

/* Capture days [t0, t1) (the cvk_run_days sequence) into an instantiated hipGraph; replay it with
 * cvk_graph_launch (each replay re-runs the captured days on the same buffers, so the caller
 * resets the state in between); release with cvk_graph_destroy. The day loop of Sim.run
 * (sim.py:733-755) as one launch. */
int cvk_graph_create(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                     int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports, void* stream,
                     void** graph_exec_out);
int cvk_graph_launch(void* graph_exec, void* stream);
int cvk_graph_destroy(void* graph_exec);

/* ---- batched many-sim path ----
 * No reference counterpart: the reference runs independent sims one per process
 * (MultiSim.run -> run.py:1477 sc.parallelize over Sim.run, sim.py:688-761). Here S sims share
 * one launch per phase per day. Agents of all sims are concatenated (sim s owns storage agents
 * [agent_offsets[s], agent_offsets[s+1])); each layer's edge lists are concatenated over sims
 * (each sim's run starting on a 4-edge boundary) with endpoints given as indices within the sim. Counter-based draws and transmission keys use
 * indices within the sim, so every sim of a batch produces exactly what cvk_run_days produces
 * for it alone. All sims share n_layers, n_variants, use_waning and the dynamic-layer flags. */
#define CVK_BATCH_MAX_SIMS 4096

typedef struct {
  int32_t n_sims, n_layers, n_variants, use_waning;
  int64_t n_agents_total;
  const int64_t* agent_offsets_host;   /* [n_sims + 1] */
  const int64_t* edge_starts_host;     /* [n_layers][n_sims]: first edge of sim s in layer l (a multiple of 4) */
  const int64_t* edge_counts_host;     /* [n_layers][n_sims]: edges of sim s in layer l */
  const cvk_params* params_host;       /* [n_sims]; params[s].n_agents is sim s's agent count */
  const int32_t* dynamic_host;         /* [n_layers]: nonzero = resampled every day */
  cvk_people people;                   /* device SoA over all agents; [variant][agent] rows have stride n_agents_total */
  int32_t* p1[CVK_MAX_LAYERS];         /* device, 16-byte aligned, concatenated over sims */
  int32_t* p2[CVK_MAX_LAYERS];
  float* beta[CVK_MAX_LAYERS];
  int32_t beta_ones[CVK_MAX_LAYERS];   /* nonzero when every weight of the layer is 1.0f */
  /* Source lists (cvk_sources, local indices) of sim s's static layer l at sources[s * n_layers + l]
   * (host array; null: none). A static layer with lists for every sim is walked, for the sims with at
   * most CVK_TUNE_SOURCE_PERMILLE infectious agents per thousand (the same rule as cvk_edge_pass),
   * instead of streamed. The results are identical either way. */
  const cvk_sources* sources;
  /* Per-day rates of each sim (change_beta, interventions.py:530-586), as cvk_imports holds them for one sim:
   * day t of sim s uses row min(t, n_beta_days - 1) of beta_v_days [n_sims][n_beta_days][CVK_MAX_VARIANTS]
   * and beta_layer_days [n_sims][n_beta_days][CVK_MAX_LAYERS] (host arrays); n_beta_days = 0 keeps the
   * params' values. */
  int32_t n_beta_days;
  const float* beta_v_days;
  const float* beta_layer_days;
} cvk_batch_desc;

/* Import events of a batch: on day days[j], counts[j * n_sims + s] susceptible agents of sim s are
 * infected with variant variants[j] (immunity.py:117-130). Host arrays. */
typedef struct {
  int32_t n_events;
  const int32_t* days;
  const int32_t* variants;
  const int64_t* counts;
} cvk_batch_imports;

/* Launch shape of the batch kernels (chosen by an autotuner; see cvk_batch_set_config). */
typedef struct {
  int32_t edge_block;       /* 512 or 1024 threads per edge-pass workgroup */
  int32_t edge_chunks;      /* 2 or 4 four-edge chunks per lane of an edge-pass wave tile */
  int32_t edge_blocks_per_cu;  /* 1 or 2 */
  int32_t agents_per_thread;   /* 1, 2 or 4 in the agent pass */
} cvk_batch_config;

/* Copies the descriptor's tables to the device and allocates the batch workspace (device
 * buffers owned by the handle: bitmaps, records, keys, partial rows, results, log). */
int cvk_batch_create(const cvk_batch_desc* desc, int32_t npts, int64_t log_capacity, void** handle_out);
int cvk_batch_destroy(void* handle);
int cvk_batch_set_config(void* handle, const cvk_batch_config* config);
int cvk_batch_get_config(void* handle, cvk_batch_config* config_out);
/* Clears keys, counters, results and the log (graph-safe fill kernels on `stream`). */
int cvk_batch_reset(void* handle, void* stream);
/* Runs days [t0, t1) of every sim; t0 == 0 first seeds pop_infected[s] agents of sim s (host array). */
int cvk_batch_run_days(void* handle, int32_t t0, int32_t t1, const int64_t* pop_infected_host,
                       const cvk_batch_imports* imports, int32_t use_graph, void* stream);
/* Captures cvk_batch_run_days(t0, t1) into an instantiated hipGraph (launch/destroy with
 * cvk_graph_launch / cvk_graph_destroy). */
int cvk_batch_graph_create(void* handle, int32_t t0, int32_t t1, const int64_t* pop_infected_host,
                           const cvk_batch_imports* imports, void* stream, void** graph_exec_out);
/* No reference counterpart (measurement): enable != 0 makes eager cvk_batch_run_days calls record a HIP
 * event pair around each day's edge pass; cvk_batch_edge_times then writes the durations (ms) of days
 * [0, n) into ms_out (-1 where a day was not timed since the last enable) and returns how many it wrote.
 * Graph captures are never timed. */
int cvk_batch_time_edges(void* handle, int32_t enable);
int cvk_batch_edge_times(void* handle, float* ms_out, int32_t n);
/* Synchronous copies of the result tables: results [n_sims][npts][CVK_R_COLS] uint32 and
 * results_f [n_sims][npts][CVK_RF_COLS] (either may be NULL). */
int cvk_batch_copy_results(void* handle, uint32_t* results_host, double* results_f_host);
/* Synchronous copy of up to capacity infection records (source, target, day, layer | variant << 8
 * | sim << 12; source and target are indices within the sim); *count_out = records written. */
int cvk_batch_copy_log(void* handle, int32_t* log_host, int64_t capacity, int64_t* count_out);
/* Device pointers of the batch's result tables (valid while the handle lives): results
 * [n_sims][npts][CVK_R_COLS] uint32 and results_f [n_sims][npts][CVK_RF_COLS] double. */
int cvk_batch_results(void* handle, const uint32_t** results_out, const double** results_f_out);

/* ---- results: finalisation and reduction over sims ----
 * cvk_finalize replaces covasim/sim.py:764-850 (Sim.finalize: cumulative flows, compute_states),
 * :872-895 (compute_doubling) and :898-1001 (compute_r_eff, method 'daily') for one sim or a batch,
 * on the device. Output rows are doubles [n_sims][CVK_F_ROWS][npts] and, by variant,
 * [n_sims][CVK_FV_ROWS][n_variants][npts]. Flows follow the order new_ / cum_ + infections,
 * reinfections, infectious, symptomatic, severe, critical, recoveries, deaths, tests, diagnoses,
 * known_deaths, quarantined, isolated, doses, vaccinated (defaults.py result_flows; flows the day
 * loop does not produce are 0); stocks n_ + susceptible, exposed, infectious, symptomatic, severe,
 * critical, recovered, dead, diagnosed, known_dead, quarantined, isolated, vaccinated. */
#define CVK_F_N_FLOWS 15
#define CVK_F_N_STOCKS 13
enum {
  CVK_F_NEW = 0,                              /* + flow */
  CVK_F_CUM = CVK_F_NEW + CVK_F_N_FLOWS,      /* + flow */
  CVK_F_STOCK = CVK_F_CUM + CVK_F_N_FLOWS,    /* + stock; n_susceptible recomputed as compute_states does */
  CVK_F_N_ALIVE = CVK_F_STOCK + CVK_F_N_STOCKS,
  CVK_F_N_NAIVE, CVK_F_N_PREINFECTIOUS, CVK_F_N_REMOVED, CVK_F_PREVALENCE, CVK_F_INCIDENCE,
  CVK_F_FRAC_VACCINATED, CVK_F_N_IMPORTS, CVK_F_POP_NABS, CVK_F_POP_PROTECTION, CVK_F_POP_SYMP_PROTECTION,
  CVK_F_DOUBLING_TIME, CVK_F_R_EFF, CVK_F_TEST_YIELD, CVK_F_REL_TEST_YIELD,
  CVK_F_ROWS
};
enum {
  CVK_FV_NEW = 0,       /* + 0 infections, 1 symptomatic, 2 severe, 3 infectious (by variant) */
  CVK_FV_CUM = 4,       /* + the same four */
  CVK_FV_N_EXPOSED = 8, CVK_FV_N_INFECTIOUS, CVK_FV_INCIDENCE, CVK_FV_PREVALENCE,
  CVK_FV_ROWS
};

typedef struct {
  int32_t n_sims, npts, n_variants, use_waning;
  int32_t r_eff_smoothing;          /* sc.smooth repeats of compute_r_eff (default 2) */
  int32_t doubling_window;          /* compute_doubling window (default 3) */
  double max_doubling_time;         /* default 30 */
  const double* pop_infected_host;  /* [n_sims] */
  const int32_t* initial_period_host;  /* [n_sims]: int(min(npts, exp2inf par1 + asym2rec par1)), sim.py:934-935 */
  const int64_t* agent_offsets_host;   /* [n_sims + 1]: sim s owns agents [off[s], off[s+1]) */
} cvk_finalize_desc;

int64_t cvk_finalize_work_bytes(const cvk_finalize_desc* desc);
/* results / results_f: the day loop's tables ([n_sims][npts][...], as cvk_batch_results or a
 * cvk_work's). imports: [n_sims][npts] imported agents per day (NULL = none). Dates: the agents'
 * float32 date_infectious, date_recovered, date_dead (NaN = undefined) over all sims' agents.
 * Synchronous on `stream` (per-sim offsets are uploaded into `work`). */
int cvk_finalize(const cvk_finalize_desc* desc, const uint32_t* results, const double* results_f,
                 const double* imports, const float* date_infectious, const float* date_recovered,
                 const float* date_dead, double* out, double* out_variant, void* work, int64_t work_bytes,
                 void* stream);
/* MultiSim.reduce (covasim/run.py:220-303) over n_sims rows of n_cols doubles (values[s][c]):
 * use_mean = 0: center = median, low / high = quantiles q_low, q_high (numpy's 'linear' method; a
 * column holding NaN gives NaN); use_mean = 1: center = mean, low / high = mean -/+ bounds * std
 * (numpy's pairwise sums, ddof 0). n_sims <= CVK_BATCH_MAX_SIMS. */
int cvk_reduce(int32_t n_sims, int64_t n_cols, const double* values, int32_t use_mean, double q_low, double q_high,
               double bounds, double* center, double* low, double* high, void* stream);

/* ---- population generation ----
 * Replaces covasim/population.py:143-226 make_randpop (ages, sexes), :241-284 make_random_contacts,
 * :287-329 make_microstructured_contacts, :332-364 make_hybrid_contacts and the per-agent draws of
 * people.py:139-160 set_prognoses + utils.py:409-426 n_neg_binomial (rel_trans). The population is
 * a pure function of (seed, tables, sizes): every draw comes from Philox4x32-10 with the counter
 * (c0, c1) = an agent, household, member or edge index, c2 = a CVK_POP_* stream id (| layer slot),
 * so the result does not depend on launch shape and a CPU restatement reproduces it bit for bit.
 *
 * Draw rules (u01(a, b) = ((a >> 5) * 2^26 + (b >> 6)) / 2^53; words w = (x, y, z, w)):
 *  age of agent i    stream AGE (c0 = i): bin = #{k : age_cdf[k] < u01(x, y)} (capped at the last
 *                    row), age = (float)(age_lo[bin] + age_width[bin] * u01(z, w)) (double math).
 *  sort key of i     stream AGE_NOISE (c0 = i): s = (x + y + z + w) * 2^-32 (Irwin-Hall of 4, exact),
 *                    key = clamp(floor(age_d + (s - 2) * age_noise) + 64, 0, 255), age_d the double age.
 *  household h size  stream HH_SIZE (c0, c1 = h): entry = #{k : hh_cdf[k] < u01(x, y)} (capped),
 *                    size = hh_size_min + entry. Households h = 0, 1, ... are taken until they cover
 *                    n agents; the last one is cut to fit.
 *  household order   correlate = 0: household h holds agents [start_h, start_h + size_h) as drawn.
 *                    correlate = 1: agents are ordered by (key, agent index); household h takes the
 *                    next size_h agents of that order and is then moved to position pi(h), pi a
 *                    4-round Feistel permutation of [0, H) (round r: F = philox(R, r, HH_PERM, 0).x
 *                    & mask over 2b-bit words, b = max(1, ceil(bits(H - 1) / 2)), cycle-walked below
 *                    H); final agent indices follow the permuted household order.
 *  household edges   every pair (a, b), a < b, of each household, in (p1, p2) order.
 *  layer members     agents with lo <= age < hi, ascending (all agents when mapping is not used).
 *  contact counts    member m, stream COUNT | slot (c0 = m): k = the smallest count whose CDF reaches
 *                    u01(x, y) (Poisson: p0 = e^-lam, p_{k+1} = p_k * lam / (k+1); negative binomial:
 *                    p_{k+1} = p_k * ((k + r) / (k + 1)) * q); edges of m = round-half-even(k / 2).
 *  contact partners  edge e, stream EDGE | slot (c0, c1 = e >> 2): word e & 3 of the draw, mapped to
 *                    [0, n_members) by Lemire's multiply (rejecting low words below 2^32 mod n; retry
 *                    k uses word x of stream EDGE_RETRY | slot with c0, c1 = e, c3 = k, at most 255).
 *                    p1 = member of the edge (its own edges are consecutive), p2 = mapping[partner].
 *  sex, rel_trans    stream PERSON (c0 = i): sex = x & 1; rel_trans = (float)(trans_or[bin] *
 *                    (nb * nb_step)), nb the negative binomial count of u01(y, z) as above
 *                    (p0 = nb_p0, r = nb_r, q = nb_q); bin = #{k : prog_cutoff[k] <= age} - 1.
 */
#define CVK_POP_MAX_AGE_BINS 32
#define CVK_POP_MAX_HH_SIZES 64
#define CVK_POP_MAX_PROG_BINS 16
#define CVK_POP_COUNT_POISSON 0
#define CVK_POP_COUNT_NEGBIN 1

typedef struct {
  uint64_t seed;
  int32_t n_age_bins;          /* rows of the age table, 1..CVK_POP_MAX_AGE_BINS */
  int32_t n_hh_sizes;          /* entries of hh_cdf, 1..CVK_POP_MAX_HH_SIZES */
  int32_t hh_size_min;         /* size of hh_cdf entry 0 (1 for categorical sizes, 0 for Poisson) */
  int32_t correlate;           /* 1: households from the noisy-age order, shuffled; 0: consecutive agents */
  int32_t n_prog_bins;         /* 1..CVK_POP_MAX_PROG_BINS */
  int32_t pad;
  double age_cdf[CVK_POP_MAX_AGE_BINS];
  double age_lo[CVK_POP_MAX_AGE_BINS];
  double age_width[CVK_POP_MAX_AGE_BINS];
  double hh_cdf[CVK_POP_MAX_HH_SIZES];
  double age_noise;            /* 15 * sqrt(3) gives a sort-key noise of standard deviation 15 years */
  double prog_cutoff[CVK_POP_MAX_PROG_BINS];
  double trans_or[CVK_POP_MAX_PROG_BINS];
  float symp_prob[CVK_POP_MAX_PROG_BINS], severe_prob[CVK_POP_MAX_PROG_BINS];
  float crit_prob[CVK_POP_MAX_PROG_BINS], death_prob[CVK_POP_MAX_PROG_BINS];
  float rel_sus[CVK_POP_MAX_PROG_BINS];
  double nb_p0, nb_q, nb_r, nb_step;  /* rel_trans draw: P(0), 1 - p, dispersion, value of one count */
} cvk_popgen;

/* Contact-count distribution of a random layer. */
typedef struct {
  int32_t kind;                /* CVK_POP_COUNT_POISSON or CVK_POP_COUNT_NEGBIN */
  int32_t pad;
  double p0;                   /* P(count = 0): e^-lam, or (r / (r + mean))^r */
  double lam;                  /* Poisson mean */
  double r, q;                 /* negative binomial dispersion and 1 - p */
} cvk_pop_counts;

/* Bytes of the workspace the population calls take (`work`, 256-byte aligned device memory):
 * enough for n agents and h_cap household slots. */
int64_t cvk_pop_work_bytes(int64_t n, int64_t h_cap);
/* Ages (float32[n]) and, when g->correlate, sort keys (uint8[n]; may be NULL otherwise). */
int cvk_pop_ages(const cvk_popgen* g, int64_t n, float* age, uint8_t* key, void* stream);
/* Plans the households of n agents from household slots [0, h_cap): sizes, order and edge
 * offsets, kept in `work`. Writes totals[0] = H (households used, or -1 when h_cap slots do not
 * cover n agents: call again with a larger h_cap) and totals[1] = household edges. */
int cvk_pop_households_plan(const cvk_popgen* g, int64_t n, const uint8_t* key, int64_t h_cap, void* work,
                            int64_t work_bytes, int64_t* totals, void* stream);
/* After a plan with the same arguments: age_out[n] = age_in in household order (age_in itself
 * when !g->correlate), household edges p1/p2 (n_edges = totals[1]) and beta = 1 (beta may be NULL). */
int cvk_pop_households_fill(const cvk_popgen* g, int64_t n, const float* age_in, int64_t h_cap, void* work,
                            int64_t work_bytes, float* age_out, int32_t* p1, int32_t* p2, float* beta,
                            int64_t n_edges, void* stream);
/* Members of an age band: mapping[] = agents with lo <= age < hi, ascending; *total = count.
 * mapping must hold n entries. This call, cvk_pop_contacts_plan and cvk_pop_people reuse the
 * workspace: fill the households before calling them. */
int cvk_pop_members(int64_t n, const float* age, double lo, double hi, int32_t* mapping, int64_t* total,
                    void* work, int64_t work_bytes, void* stream);
/* Edge offsets of a random layer over n_members members: offsets[n_members + 1], the last entry
 * being the layer's edge count. `slot` is the layer's index in the population (0..255). */
int cvk_pop_contacts_plan(const cvk_popgen* g, int32_t slot, int64_t n_members, const cvk_pop_counts* counts,
                          int64_t* offsets, void* work, int64_t work_bytes, void* stream);
/* Fills a planned random layer: p1/p2 (mapping[member], or the member itself when mapping is
 * NULL), beta = 1 (beta may be NULL). p2 and beta must be 16-byte aligned. */
int cvk_pop_contacts_fill(const cvk_popgen* g, int32_t slot, int64_t n_members, const int32_t* mapping,
                          const int64_t* offsets, int32_t* p1, int32_t* p2, float* beta, int64_t n_edges,
                          void* stream);
/* Per-agent draws and prognoses of n agents with final ages `age`: sex (uint8), symp_prob,
 * severe_prob, crit_prob, death_prob, rel_sus, rel_trans (float32). */
int cvk_pop_people(const cvk_popgen* g, int64_t n, const float* age, uint8_t* sex, float* symp_prob,
                   float* severe_prob, float* crit_prob, float* death_prob, float* rel_sus, float* rel_trans,
                   void* work, int64_t work_bytes, void* stream);

/* ---- interventions ----
 * change_beta (interventions.py:530-586) needs no entry point of its own: its per-day rates go to
 * cvk_run_days / cvk_graph_create in cvk_imports.beta_v_days / beta_layer_days.
 * cvk_move_edges is the edge move of clip_edges (interventions.py:589-673, Layer.pop_inds / append,
 * base.py:1739-1768): k of a layer's n edges, chosen uniformly without replacement, move out. Edge e
 * has the key (philox(seed; e lo, t, 0x80000000 | slot, e hi).x << 32) | e and the k smallest keys
 * move. keep_* get the other n - k edges in their order, move_* the k moved edges in their order
 * (p1, p2, beta columns; point move_* at the end of the destination layer to append them). */
/* No reference counterpart: builds the source lists of a layer of n_agents agents (cvk_sources).
 * target / edge_dir need room for 2 * n_edges slots (capacity); *n_slots_out receives the count (a
 * synchronous read-back). work: cvk_sources_work_bytes(n_agents) bytes, 256-byte aligned. */
int64_t cvk_sources_work_bytes(int64_t n_agents);
int cvk_sources_build(const cvk_layer* layer, int64_t n_agents, int32_t* indptr, int32_t* target, uint32_t* edge_dir,
                      int64_t capacity, int64_t* n_slots_out, void* work, int64_t work_bytes, void* stream);
int64_t cvk_move_edges_work_bytes(int64_t n);
int cvk_move_edges(const int32_t* p1, const int32_t* p2, const float* beta, int64_t n, int64_t k, uint64_t seed,
                   int32_t t, int32_t slot, int32_t* keep_p1, int32_t* keep_p2, float* keep_beta, int32_t* move_p1,
                   int32_t* move_p2, float* move_beta, void* work, int64_t work_bytes, void* stream);

/* ---- differentiable relaxation ----
 * No reference counterpart: an expected-infection relaxation of the transmission step
 * (utils.py:88-123 with indicator states replaced by probabilities) for gradient-based
 * calibration. Per agent i and day t: susceptible probability s_i(t), infection probability
 * D_i(t) (D_i(0) = seed_frac), expected infectiousness tr_j(t) = rel_trans_j * sum_{d>=1}
 * D_j(t-d) w[bin_j][d], log-survival L_i(t) = sum over layers l and neighbours j of
 * log1p(-min(beta * beta_l * sigma[bin_i] * tr_j, 0.999999)), D_i(t) = s_i(t)(1 - e^{L_i}),
 * s_i(t+1) = s_i(t) e^{L_i}; E_t = sum_i s_i(t)(1 - e^{L_i}). Parameters theta (host doubles):
 * [beta, beta_l for each layer, sigma_b for the 10 age bins]. A layer resampled every day
 * (base.py:1849-1876) enters in expectation: with c_l = dyn_slots[l] = 2 E_l / n_agents contact slots
 * per agent, each with a uniformly drawn source, its term of L_i(t) is
 * c_l / n_agents * sum_{j != i} log1p(-min(beta * beta_l * sigma[bin_i] * tr_j, 0.999999)). */
typedef struct {
  int64_t n_agents;
  int32_t n_layers, n_days;
  int32_t kernel_len;          /* lags 0..kernel_len-1 of w_host per age bin */
  int32_t checkpoint_every;    /* days between stored states (the backward re-runs each segment) */
  int32_t precision;           /* 32 (float arithmetic) or 64 (double) */
  int32_t pad;
  double seed_frac;            /* pop_infected / n_agents */
  const int64_t* indptr[CVK_MAX_LAYERS];  /* device [n_agents + 1]: neighbour list offsets per layer */
  const int32_t* nbr[CVK_MAX_LAYERS];     /* device: both directions of every edge, no self-loops */
  const uint8_t* age_bin;      /* device [n_agents]: prognosis age bin 0..9 */
  const float* rel_trans;      /* device [n_agents] */
  const double* w_host;        /* host [10][kernel_len]: mean infectiousness d days after infection */
  double dyn_slots[CVK_MAX_LAYERS];  /* > 0: layer l is resampled every day and enters in expectation with
                                        this many contact slots per agent (indptr / nbr unused); 0: its lists */
} cvk_relax_desc;

int cvk_relax_create(const cvk_relax_desc* desc, void** handle_out);
int cvk_relax_destroy(void* handle);
/* expected_host[n_days] = E_t. Synchronous on `stream`. */
int cvk_relax_forward(void* handle, const double* theta_host, double* expected_host, void* stream);
/* grad_theta_host[1 + n_layers + 10] = sum_t grad_expected_host[t] * dE_t/dtheta (adjoint, after a
 * forward with the same theta). Synchronous on `stream`. */
int cvk_relax_backward(void* handle, const double* theta_host, const double* grad_expected_host,
                       double* grad_theta_host, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* COVASIM_AMD_H */
