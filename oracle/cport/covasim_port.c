/*
 * covasim_port.c -- compiled CPU restatement of Covasim 3.1.7's daily transmission step.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline and a second, independent restatement
 * that the numpy oracle is checked against. Never linked into the product.
 *
 * It follows the reference's structure (the numba kernels run one call per variant and
 * layer, each compacting the nonzero-probability edges per direction, utils.py:93-128;
 * People.infect keeps the first occurrence of each target, people.py:465) with the
 * counter-based Philox draws of oracle/philox.py, so its trajectories equal the numpy
 * oracle's. Compile with -ffp-contract=off (float32 products must not fuse).
 * With -fopenmp the per-edge probability loop is split over threads (numba 'safe'-style).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXV 8
#define MAXL 8

enum { DIST_UNIFORM = 0, DIST_NORMAL, DIST_NORMAL_POS, DIST_NORMAL_INT, DIST_LOGNORMAL, DIST_LOGNORMAL_INT };

typedef struct {
  int32_t dist;
  int32_t pad;
  double par1, par2, mu, sigma;
} port_dur;

typedef struct {
  int64_t n;
  int32_t n_variants, n_layers, use_waning, nab_kin_len;
  uint64_t seed;
  float frac_time, load_ratio, high_cap, asymp_factor;
  float beta_v[MAXV];
  float beta_layer[MAXL], iso_factor[MAXL], quar_factor[MAXL];
  int32_t dynamic[MAXL];
  float rel_symp[MAXV], rel_sev[MAXV], rel_crit[MAXV], rel_death[MAXV];
  float no_hosp_factor, no_icu_factor, trans_redux, nab_boost;
  int64_t n_beds_hosp, n_beds_icu;
  port_dur dur[9];
  port_dur nab_init;
  double rel_imm_symp[3], nab_norm, ve_exp_alpha[3], ve_beta[3];
  float immunity[MAXV][MAXV];
  const double* nab_kin;
} port_params;

/* agent state: uint8 booleans and float32 dates (NaN = undefined), as the reference's People */
typedef struct {
  uint8_t *susceptible, *naive, *exposed, *infectious, *symptomatic, *severe, *critical, *diagnosed, *recovered,
      *known_dead, *dead, *known_contact, *quarantined, *isolated;
  float *exposed_variant, *infectious_variant, *recovered_variant;
  uint8_t *exposed_by_variant, *infectious_by_variant; /* [V][n] */
  float *date_exposed, *date_infectious, *date_symptomatic, *date_severe, *date_critical, *date_recovered,
      *date_dead, *date_diagnosed, *date_end_isolation;
  float *dur_exp2inf, *dur_inf2sym, *dur_sym2sev, *dur_sev2crit, *dur_disease;
  float *rel_trans, *rel_sus, *symp_prob, *severe_prob, *crit_prob, *death_prob;
  int32_t *n_infections, *n_breakthroughs, *t_nab_event;
  float *nab, *peak_nab;
  float *sus_imm, *symp_imm, *sev_imm; /* [V][n] */
} port_people;

typedef struct {
  int32_t* p1;
  int32_t* p2;
  float* beta;
  int64_t n_edges;
} port_layer;

/* per-day output row: the same columns as the GPU results table, as doubles */
enum {
  R_NEW_INFECTIONS = 0, R_NEW_REINFECTIONS, R_NEW_INFECTIOUS, R_NEW_SYMPTOMATIC, R_NEW_SEVERE, R_NEW_CRITICAL,
  R_NEW_RECOVERIES, R_NEW_DEATHS, R_NEW_KNOWN_DEATHS, R_N_SUSCEPTIBLE, R_N_EXPOSED, R_N_INFECTIOUS,
  R_N_SYMPTOMATIC, R_N_SEVERE, R_N_CRITICAL, R_N_RECOVERED, R_N_DEAD, R_COLS
};

/* ---------------- Philox4x32-10 (Salmon et al., SC'11), same streams as oracle/philox.py ---------------- */
#define STREAM_TX 0x10000000u
#define STREAM_INFECT 0x20000000u
#define STREAM_CHOOSE 0x30000000u
#define STREAM_IMPORT 0x40000000u
#define STREAM_RESAMPLE 0x50000000u

static void philox(uint32_t c[4], uint64_t seed) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

static double u01(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

/* Resampled endpoint j (0: p1, 1: p2) of edge e with first word w0: Lemire's multiply-and-reject,
 * rejected words replaced from the retry stream (oracle/philox.py, resample_edges). */
#define STREAM_RESAMPLE_RETRY 0x60000000u
static uint32_t resample_endpoint(uint64_t seed, int32_t t, uint32_t l, int64_t e, int j, uint32_t w0, uint32_t n) {
  const uint32_t thr = (uint32_t)((0x100000000ull - n) % n);
  uint64_t m = (uint64_t)w0 * n;
  for (uint32_t k = 0; k < 255 && (uint32_t)m < thr; ++k) {
    uint32_t c[4] = {(uint32_t)e, (uint32_t)t, STREAM_RESAMPLE_RETRY | l, (uint32_t)(((uint64_t)e >> 32) << 8) | k};
    philox(c, seed);
    m = (uint64_t)c[j] * n;
  }
  return (uint32_t)(m >> 32);
}

static void slot_pair(uint64_t seed, uint32_t agent, int32_t t, uint32_t slot, double* u0, double* u1) {
  uint32_t c[4] = {agent, (uint32_t)t, STREAM_INFECT, slot};
  philox(c, seed);
  *u0 = u01(c[0], c[1]);
  *u1 = u01(c[2], c[3]);
}

/* sample() of utils.py:156-236 on a Box-Muller normal */
static double draw(const port_dur* d, uint64_t seed, uint32_t agent, int32_t t, uint32_t slot) {
  double u0, u1;
  slot_pair(seed, agent, t, slot, &u0, &u1);
  const double z = sqrt(-2.0 * log1p(-u0)) * cos(2.0 * 3.141592653589793 * u1);
  switch (d->dist) {
    case DIST_UNIFORM: return d->par1 + (d->par2 - d->par1) * u0;
    case DIST_NORMAL: return d->par1 + d->par2 * z;
    case DIST_NORMAL_POS: return fabs(d->par1 + d->par2 * z);
    case DIST_NORMAL_INT: return rint(fabs(d->par1 + d->par2 * z));
    case DIST_LOGNORMAL: return d->par1 > 0 ? exp(d->mu + d->sigma * z) : 0.0;
    case DIST_LOGNORMAL_INT: return d->par1 > 0 ? rint(exp(d->mu + d->sigma * z)) : 0.0;
  }
  return 0.0;
}

/* ---------------- k smallest selection keys (uniform k-subset) ---------------- */
static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static int64_t choose_k(uint64_t seed, int32_t t, uint32_t stream, const int64_t* eligible, int64_t n_el, int64_t k,
                        int32_t* out) {
  if (k >= n_el) {
    for (int64_t j = 0; j < n_el; ++j) out[j] = (int32_t)eligible[j];
    return n_el;
  }
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n_el);
  for (int64_t j = 0; j < n_el; ++j) {
    uint32_t c[4] = {(uint32_t)eligible[j], (uint32_t)t, stream, 0};
    philox(c, seed);
    keys[j] = ((uint64_t)c[0] << 32) | (uint32_t)eligible[j];
  }
  qsort(keys, (size_t)n_el, sizeof(uint64_t), cmp_u64);
  for (int64_t j = 0; j < k; ++j) out[j] = (int32_t)(keys[j] & 0xFFFFFFFFu);
  free(keys);
  return k;
}

/* ---------------- People.infect (people.py:435-586) for unique, susceptible, sorted targets ---------------- */
static void infect_one(const port_params* P, port_people* S, int32_t i, int32_t t, int v, int hosp_max, int icu_max,
                       double* row, int count) {
  const int64_t n = P->n;
  const uint64_t seed = P->seed;
  if (S->peak_nab[i] != 0.0f) { /* breakthrough, :486-491 */
    if (S->n_breakthroughs[i] == 0) S->rel_trans[i] = S->rel_trans[i] * P->trans_redux;
    S->n_breakthroughs[i] += 1;
  }
  S->susceptible[i] = 0;
  S->naive[i] = 0;
  S->recovered[i] = 0;
  S->diagnosed[i] = 0;
  S->exposed[i] = 1;
  S->n_infections[i] += 1;
  S->exposed_variant[i] = (float)v;
  S->exposed_by_variant[(int64_t)v * n + i] = 1;
  if (count) {
    row[R_NEW_INFECTIONS] += 1;
    if (!isnan(S->date_recovered[i])) row[R_NEW_REINFECTIONS] += 1;
  }
  const float d_e2i = (float)draw(&P->dur[0], seed, (uint32_t)i, t, 0);
  S->dur_exp2inf[i] = d_e2i;
  S->date_exposed[i] = (float)t;
  const float date_inf = d_e2i + (float)t;
  S->date_infectious[i] = date_inf;
  S->date_symptomatic[i] = NAN;
  S->date_severe[i] = NAN;
  S->date_critical[i] = NAN;
  S->date_diagnosed[i] = NAN;
  float date_rec = NAN;
  double u_symp, u_sev, u_crit, u_death;
  slot_pair(seed, (uint32_t)i, t, 1, &u_symp, &u_sev);
  const float sp = (P->rel_symp[v] * S->symp_prob[i]) * (1.0f - S->symp_imm[(int64_t)v * n + i]);
  int cls = 0;
  if (!(u_symp < (double)sp)) {
    const double dd = draw(&P->dur[4], seed, (uint32_t)i, t, 3);
    date_rec = (float)((double)date_inf + dd);
    S->dur_disease[i] = (float)((double)d_e2i + dd);
  } else {
    const float d_i2s = (float)draw(&P->dur[1], seed, (uint32_t)i, t, 3);
    S->dur_inf2sym[i] = d_i2s;
    const float date_sym = date_inf + d_i2s;
    S->date_symptomatic[i] = date_sym;
    const float sevp = (P->rel_sev[v] * S->severe_prob[i]) * (1.0f - S->sev_imm[(int64_t)v * n + i]);
    if (!(u_sev < (double)sevp)) {
      cls = 1;
      const double dd = draw(&P->dur[5], seed, (uint32_t)i, t, 4);
      date_rec = (float)((double)date_sym + dd);
      S->dur_disease[i] = (float)((double)(d_e2i + d_i2s) + dd);
    } else {
      cls = 2;
      const float d_s2s = (float)draw(&P->dur[2], seed, (uint32_t)i, t, 4);
      S->dur_sym2sev[i] = d_s2s;
      const float date_sev = date_sym + d_s2s;
      S->date_severe[i] = date_sev;
      slot_pair(seed, (uint32_t)i, t, 2, &u_crit, &u_death);
      const float critp = (P->rel_crit[v] * S->crit_prob[i]) * (hosp_max ? P->no_hosp_factor : 1.0f);
      if (!(u_crit < (double)critp)) {
        const double dd = draw(&P->dur[6], seed, (uint32_t)i, t, 5);
        date_rec = (float)((double)date_sev + dd);
        S->dur_disease[i] = (float)((double)((d_e2i + d_i2s) + d_s2s) + dd);
      } else {
        const float d_s2c = (float)draw(&P->dur[3], seed, (uint32_t)i, t, 5);
        S->dur_sev2crit[i] = d_s2c;
        const float date_crit = date_sev + d_s2c;
        S->date_critical[i] = date_crit;
        const float deathp = (P->rel_death[v] * S->death_prob[i]) * (icu_max ? P->no_icu_factor : 1.0f);
        const float dsum = ((d_e2i + d_i2s) + d_s2s) + d_s2c;
        if (!(u_death < (double)deathp)) {
          const double dd = draw(&P->dur[7], seed, (uint32_t)i, t, 6);
          date_rec = (float)((double)date_crit + dd);
          S->dur_disease[i] = (float)((double)dsum + dd);
        } else {
          const double dd = draw(&P->dur[8], seed, (uint32_t)i, t, 6);
          S->date_dead[i] = (float)((double)date_crit + dd);
          S->dur_disease[i] = (float)((double)dsum + dd);
        }
      }
    }
  }
  S->date_recovered[i] = date_rec;
  if (P->use_waning) { /* update_peak_nab, immunity.py:138-202 */
    if (S->nab[i] > 0.0f) {
      S->peak_nab[i] = S->peak_nab[i] * P->nab_boost;
    } else {
      const double init_nab = draw(&P->nab_init, seed, (uint32_t)i, t, 7);
      S->peak_nab[i] = (float)((pow(2.0, init_nab) * P->rel_imm_symp[cls]) * P->nab_norm);
    }
    S->t_nab_event[i] = t;
  }
}

typedef struct {
  int32_t target;
  int32_t order;
} tgt_rec;

static int cmp_tgt(const void* a, const void* b) {
  const tgt_rec* x = (const tgt_rec*)a;
  const tgt_rec* y = (const tgt_rec*)b;
  if (x->target != y->target) return x->target < y->target ? -1 : 1;
  return x->order < y->order ? -1 : (x->order > y->order);
}

/* unique targets, first occurrence (np.unique(return_index=True)), susceptible only */
static void infect_list(const port_params* P, port_people* S, const int32_t* targets, int64_t k, int32_t t, int v,
                        int hosp_max, int icu_max, double* row, int count) {
  if (k <= 0) return;
  tgt_rec* r = (tgt_rec*)malloc(sizeof(tgt_rec) * (size_t)k);
  for (int64_t j = 0; j < k; ++j) {
    r[j].target = targets[j];
    r[j].order = (int32_t)j;
  }
  qsort(r, (size_t)k, sizeof(tgt_rec), cmp_tgt);
  for (int64_t j = 0; j < k; ++j) {
    if (j > 0 && r[j].target == r[j - 1].target) continue;
    const int32_t i = r[j].target;
    if (!S->susceptible[i]) continue;
    infect_one(P, S, i, t, v, hosp_max, icu_max, row, count);
  }
  free(r);
}

/* ---------------- one day (sim.py:558-685) ---------------- */
int port_step(const port_params* P, port_people* S, port_layer* layers, int32_t t, int64_t pop_infected,
              const int32_t* imp_days, const int32_t* imp_vars, const int64_t* imp_counts, int32_t n_imp,
              double* row, float* scratch_f /* [3n] */, int32_t* scratch_i /* [n + 2 * max edges] */) {
  const int64_t n = P->n;
  const int nv = P->n_variants;
  const float tf = (float)t;
  if (t == 0 && pop_infected > 0) { /* init_infections, sim.py:505-532 */
    int64_t* all = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t j = 0; j < n; ++j) all[j] = j;
    const int64_t k = choose_k(P->seed, 0, STREAM_CHOOSE, all, n, pop_infected, scratch_i);
    free(all);
    double junk[R_COLS];
    infect_list(P, S, scratch_i, k, 0, 0, 0, 0, junk, 0);
  }
  /* update_states_pre, people.py:164-186 */
  double c_inf = 0, c_sym = 0, c_sev = 0, c_crit = 0, c_rec = 0, c_kd = 0, c_dead = 0;
#pragma omp parallel for reduction(+ : c_inf, c_sym, c_sev, c_crit, c_rec, c_kd, c_dead) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const int is_exp = S->exposed[i];
    if (is_exp) {
      if (!S->infectious[i] && !isnan(S->date_infectious[i]) && tf >= S->date_infectious[i]) {
        S->infectious[i] = 1;
        S->infectious_variant[i] = S->exposed_variant[i];
        const int ev = (int)S->exposed_variant[i];
        S->infectious_by_variant[(int64_t)ev * n + i] = 1;
        c_inf += 1;
      }
      if (!S->symptomatic[i] && !isnan(S->date_symptomatic[i]) && tf >= S->date_symptomatic[i]) {
        S->symptomatic[i] = 1;
        c_sym += 1;
      }
      if (!S->severe[i] && !isnan(S->date_severe[i]) && tf >= S->date_severe[i]) {
        S->severe[i] = 1;
        c_sev += 1;
      }
      if (!S->critical[i] && !isnan(S->date_critical[i]) && tf >= S->date_critical[i]) {
        S->critical[i] = 1;
        c_crit += 1;
      }
      if (!S->recovered[i] && !isnan(S->date_recovered[i]) && tf >= S->date_recovered[i]) {
        S->exposed[i] = S->infectious[i] = S->symptomatic[i] = S->severe[i] = S->critical[i] = 0;
        S->recovered[i] = 1;
        S->recovered_variant[i] = S->exposed_variant[i];
        S->infectious_variant[i] = NAN;
        S->exposed_variant[i] = NAN;
        for (int v = 0; v < nv; ++v) {
          S->exposed_by_variant[(int64_t)v * n + i] = 0;
          S->infectious_by_variant[(int64_t)v * n + i] = 0;
        }
        if (P->use_waning) {
          S->susceptible[i] = 1;
          S->diagnosed[i] = 0;
        }
        c_rec += 1;
      }
    }
    if (S->isolated[i] && !isnan(S->date_end_isolation[i]) && tf >= S->date_end_isolation[i]) S->isolated[i] = 0;
    if (is_exp && !S->dead[i] && !isnan(S->date_dead[i]) && tf >= S->date_dead[i]) {
      S->dead[i] = 1;
      if (S->diagnosed[i]) {
        S->known_dead[i] = 1;
        c_kd += 1;
      }
      S->susceptible[i] = S->exposed[i] = S->infectious[i] = S->symptomatic[i] = S->severe[i] = S->critical[i] = 0;
      S->known_contact[i] = S->quarantined[i] = S->recovered[i] = 0;
      S->infectious_variant[i] = S->exposed_variant[i] = S->recovered_variant[i] = NAN;
      c_dead += 1;
    }
    if (P->use_waning) { /* check_immunity, immunity.py:303-350 */
      const double nab = (double)S->nab[i];
      const int was_inf = tf >= S->date_recovered[i];
      for (int v = 0; v < nv; ++v) {
        double imm = 0.0;
        if (was_inf) imm = (double)P->immunity[v][(int)S->recovered_variant[i]];
        const double eff = nab * imm;
        const double e0 = P->ve_exp_alpha[0] * pow(eff, P->ve_beta[0]);
        const double e1 = P->ve_exp_alpha[1] * pow(eff, P->ve_beta[1]);
        const double e2 = P->ve_exp_alpha[2] * pow(eff, P->ve_beta[2]);
        S->sus_imm[(int64_t)v * n + i] = (float)(e0 / (1 + e0));
        S->symp_imm[(int64_t)v * n + i] = (float)(e1 / (1 + e1));
        S->sev_imm[(int64_t)v * n + i] = (float)(e2 / (1 + e2));
      }
    }
  }
  row[R_NEW_INFECTIOUS] += c_inf;
  row[R_NEW_SYMPTOMATIC] += c_sym;
  row[R_NEW_SEVERE] += c_sev;
  row[R_NEW_CRITICAL] += c_crit;
  row[R_NEW_RECOVERIES] += c_rec;
  row[R_NEW_KNOWN_DEATHS] += c_kd;
  row[R_NEW_DEATHS] += c_dead;
  /* update_contacts: dynamic layers, base.py:1849-1876 */
  for (int l = 0; l < P->n_layers; ++l) {
    if (!P->dynamic[l]) continue;
    port_layer* L = &layers[l];
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (L->n_edges + 1) / 2; ++q) {  /* one draw per pair of edges */
      uint32_t c[4] = {(uint32_t)q, (uint32_t)t, STREAM_RESAMPLE | (uint32_t)l, (uint32_t)((uint64_t)q >> 32)};
      philox(c, P->seed);
      for (int h = 0; h < 2 && 2 * q + h < L->n_edges; ++h) {
        const int64_t e = 2 * q + h;
        L->p1[e] = (int32_t)resample_endpoint(P->seed, t, (uint32_t)l, e, 0, c[2 * h], (uint32_t)n);
        L->p2[e] = (int32_t)resample_endpoint(P->seed, t, (uint32_t)l, e, 1, c[2 * h + 1], (uint32_t)n);
        L->beta[e] = 1.0f;
      }
    }
  }
  int64_t n_sev = 0, n_crit = 0;
  for (int64_t i = 0; i < n; ++i) {
    n_sev += S->severe[i];
    n_crit += S->critical[i];
  }
  const int hosp_max = P->n_beds_hosp >= 0 && n_sev > P->n_beds_hosp;
  const int icu_max = P->n_beds_icu >= 0 && n_crit > P->n_beds_icu;
  /* variant imports, immunity.py:117-130 */
  for (int j = 0; j < n_imp; ++j) {
    if (imp_days[j] != t) continue;
    int64_t* el = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i)
      if (S->susceptible[i]) el[m++] = i;
    const int64_t k = choose_k(P->seed, t, STREAM_IMPORT | (uint32_t)imp_vars[j], el, m, imp_counts[j], scratch_i);
    free(el);
    infect_list(P, S, scratch_i, k, t, imp_vars[j], 0, 0, row, 1);
  }
  /* viral load, utils.py:39-79 */
  float* vl = scratch_f;
  float* rt = scratch_f + n;
  float* rs = scratch_f + 2 * n;
  const float one = 1.0f;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const float start = S->date_infectious[i];
    const float stop = isnan(S->date_dead[i]) ? S->date_recovered[i] : S->date_dead[i];
    const float total = stop - start;
    const float trans_day = P->frac_time * total;
    const float tp = trans_day > P->high_cap ? P->high_cap / total : P->frac_time;
    const int early = (tf - start) / total < tp;
    vl[i] = (P->load_ratio * (early ? 1.0f : 0.0f) + one * (early ? 0.0f : 1.0f)) /
            (one + P->frac_time * (P->load_ratio - one));
  }
  /* transmission: variants x layers (sim.py:622-649) */
  int32_t* tlist = scratch_i;
  for (int v = 0; v < nv; ++v) {
    int any = 0;
    for (int64_t i = 0; i < n && !any; ++i) any = S->infectious[i] && S->infectious_variant[i] == (float)v;
    if (!any) continue;
    const float beta = P->beta_v[v];
    for (int l = 0; l < P->n_layers; ++l) {
      /* compute_trans_sus, utils.py:82-90 */
      const float bl = P->beta_layer[l], isf = P->iso_factor[l], qf = P->quar_factor[l], af = P->asymp_factor;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        const int inf = S->infectious[i] && S->infectious_variant[i] == (float)v;
        const float fsymp = S->symptomatic[i] ? 1.0f : 0.0f;
        const float fiso = S->isolated[i] ? 1.0f : 0.0f;
        const float fquar = S->quarantined[i] ? 1.0f : 0.0f;
        const float f_asymp = fsymp + (1.0f - fsymp) * af;
        const float f_iso = (1.0f - fiso) + fiso * isf;
        const float f_quar = (1.0f - fquar) + fquar * qf;
        float x = S->rel_trans[i] * (inf ? 1.0f : 0.0f);
        x = x * f_quar;
        x = x * f_asymp;
        x = x * f_iso;
        x = x * bl;
        x = x * vl[i];
        rt[i] = x;
        float y = S->rel_sus[i] * (S->susceptible[i] ? 1.0f : 0.0f);
        y = y * f_quar;
        y = y * (1.0f - S->sus_imm[(int64_t)v * n + i]);
        rs[i] = y;
      }
      /* compute_infections, utils.py:93-128: direction-major, edge-ascending */
      const port_layer* L = &layers[l];
      const int64_t E = L->n_edges;
      uint8_t* hit = (uint8_t*)malloc((size_t)(2 * E + 1));
      /* probabilities and uniform tests per (direction, edge); in parallel with OpenMP */
#pragma omp parallel for schedule(static)
      for (int64_t s2 = 0; s2 < 2 * E; ++s2) {
        const int dir = s2 >= E;
        const int64_t e = dir ? s2 - E : s2;
        const int32_t sv = dir ? L->p2[e] : L->p1[e];
        const int32_t tv = dir ? L->p1[e] : L->p2[e];
        hit[s2] = 0;
        const float st = rt[sv];
        if (st == 0.0f) continue;
        const float p = ((beta * L->beta[e]) * st) * rs[tv];
        if (p == 0.0f) continue;
        uint32_t c[4] = {(uint32_t)e, (uint32_t)t, STREAM_TX | ((uint32_t)v << 8) | (uint32_t)l,
                         (uint32_t)((uint64_t)e >> 32)};
        philox(c, P->seed);
        const double u = dir ? u01(c[2], c[3]) : u01(c[0], c[1]);
        hit[s2] = u < (double)p;
      }
      /* compaction in the reference's order: direction-major, edge-ascending */
      int64_t k = 0;
      for (int64_t s2 = 0; s2 < 2 * E; ++s2)
        if (hit[s2]) tlist[k++] = s2 >= E ? L->p1[s2 - E] : L->p2[s2];
      free(hit);
      infect_list(P, S, tlist, k, t, v, hosp_max, icu_max, row, 1);
    }
  }
  /* stocks (sim.py:651-656) and update_nab (sim.py:666-669) */
  double s0 = 0, s1 = 0, s2c = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0;
#pragma omp parallel for reduction(+ : s0, s1, s2c, s3, s4, s5, s6, s7) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    s0 += S->susceptible[i];
    s1 += S->exposed[i];
    s2c += S->infectious[i];
    s3 += S->symptomatic[i];
    s4 += S->severe[i];
    s5 += S->critical[i];
    s6 += S->recovered[i];
    s7 += S->dead[i];
    if (P->use_waning && S->peak_nab[i] != 0.0f) {
      int32_t ts = t - S->t_nab_event[i];
      if (ts < 0) ts = 0;
      if (ts >= P->nab_kin_len) ts = P->nab_kin_len - 1;
      float nab = (float)((double)S->nab[i] + P->nab_kin[ts] * (double)S->peak_nab[i]);
      nab = nab < 0.0f ? 0.0f : nab;
      nab = nab > S->peak_nab[i] ? S->peak_nab[i] : nab;
      S->nab[i] = nab;
    }
  }
  row[R_N_SUSCEPTIBLE] += s0;
  row[R_N_EXPOSED] += s1;
  row[R_N_INFECTIOUS] += s2c;
  row[R_N_SYMPTOMATIC] += s3;
  row[R_N_SEVERE] += s4;
  row[R_N_CRITICAL] += s5;
  row[R_N_RECOVERED] += s6;
  row[R_N_DEAD] += s7;
  return 0;
}

int port_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Exported for popgen_port.c (the population generator's restatement, same library). */
void port_philox(uint32_t c[4], uint64_t seed) { philox(c, seed); }
