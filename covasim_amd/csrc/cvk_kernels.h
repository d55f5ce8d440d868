// Device building blocks of the day loop, shared by the single-sim kernels (cvk_day.hip) and
// the batched many-sim kernels (cvk_batch.hip). Each block works on one sim at a time through a
// view whose pointers are already offset to that sim's slice, so everything below indexes agents
// and edges by their index within the sim: counter-based draws and transmission keys are
// therefore identical whether a sim runs alone or inside a batch.
#pragma once

#include "cvk_device.h"

namespace cvk {

constexpr uint32_t STATE_CLEAR_ON_RECOVERY = CVK_EXPOSED | CVK_INFECTIOUS | CVK_SYMPTOMATIC | CVK_SEVERE | CVK_CRITICAL;
constexpr uint32_t STATE_CLEAR_ON_DEATH = CVK_SUSCEPTIBLE | CVK_EXPOSED | CVK_INFECTIOUS | CVK_SYMPTOMATIC |
                                          CVK_SEVERE | CVK_CRITICAL | CVK_KNOWN_CONTACT | CVK_QUARANTINED |
                                          CVK_RECOVERED;
constexpr uint32_t INFVAR_MASK = 0xFu << CVK_INFVAR_SHIFT;
constexpr int MAXV = CVK_MAX_VARIANTS;
constexpr int AGENT_BLOCK = 256;

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }

// Pointers that came out of device memory (the batched views) have no known address space, so the
// compiler emits flat accesses, which count on the LDS wait counter too: every LDS wait would also wait
// for outstanding memory loads. These casts name the global address space.
#define CVK_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline CVK_GLOBAL T* gmem(T* p) {
  return (CVK_GLOBAL T*)p;
}

// A copy of a wave-uniform struct held in scalar registers (each word through readfirstlane), so that
// its fields are neither vector registers nor reloaded after the kernel's stores.
template <class T>
__device__ inline T uniform_copy(const T& in) {
  static_assert(sizeof(T) % 4 == 0, "uniform_copy needs whole words");
  T out;
  const uint32_t* src = (const uint32_t*)&in;
  uint32_t* dst = (uint32_t*)&out;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}


__device__ inline uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// calc_VE, immunity.py:216-247, in float64 then stored as float32.
__device__ inline float calc_ve(double nab, double exp_alpha, double beta) {
  const double exp_lo = exp_alpha * pow(nab, beta);
  return (float)(exp_lo / (1.0 + exp_lo));
}

// ---------------------------------------------------------------------------
// Reductions without device-scope fences: every block of a reducing kernel stores its partial
// row (column-major, `stride` entries per column); the rows are summed after the kernel
// boundary. Integer columns are order-free; float columns are summed in a fixed order.
// Sums blocks [b0, b0 + nb) of every column (one wave per column, 8 loads in flight per lane);
// totals land in s_tot[NC] / s_totf[ND]. Block-wide (contains a barrier).
// ---------------------------------------------------------------------------
template <int NC, int ND>
__device__ void sum_partials(const uint32_t* partials, const double* partials_f, int64_t stride, uint32_t b0,
                             uint32_t nb, uint32_t* s_tot, double* s_totf) {
  const int n_waves = (int)(blockDim.x >> 6);
  const int w = threadIdx.x >> 6;
  for (int k = w; k < NC + ND; k += n_waves) {
    if (k < NC) {
      const uint32_t* col = partials + (size_t)k * stride + b0;
      uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t c0 = 0; c0 < nb; c0 += 512) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const uint32_t b = c0 + j * 64 + lane_id();
          acc[j] += b < nb ? col[b] : 0u;
        }
      }
      uint32_t tot = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
      tot = wave_sum(tot);
      if (lane_id() == 0) s_tot[k] = tot;
    } else {
      const double* col = partials_f + (size_t)(k - NC) * stride + b0;
      double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t c0 = 0; c0 < nb; c0 += 512) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const uint32_t b = c0 + j * 64 + lane_id();
          acc[j] += b < nb ? col[b] : 0.0;
        }
      }
      double tot = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
      tot = wave_sum_d(tot);
      if (lane_id() == 0) s_totf[k - NC] = tot;
    }
  }
  __syncthreads();
}

// ===========================================================================
// Agent pass: [previous-day stocks + update_nab] + transitions + bitmaps + edge records
// ===========================================================================
// stock columns (0..29) and day columns (30..46) of the partial rows
constexpr int SC_ALIVE = 13, SC_EXPV = 14, SC_INFV = 14 + MAXV, N_STOCK_COLS = 14 + 2 * MAXV;
constexpr int DC_BASE = N_STOCK_COLS, N_DAY_COLS = 9 + MAXV;
// column of the agents whose infectious bit this pass set (the edge pass's choice between walking
// source lists and streaming reads it)
constexpr int DC_BITS = N_STOCK_COLS + N_DAY_COLS, N_SUM_COLS = DC_BITS + 1;
// column of the block's largest target code (not summed: the edge pass takes the maximum over the blocks),
// written with the target bits (AgentSpan.aux)
constexpr int DC_TMAX = N_SUM_COLS, N_AGENT_COLS = DC_TMAX + 1;
constexpr int N_AGENT_FCOLS = 3;
static_assert(N_AGENT_COLS <= CVK_PARTIAL_COLS, "partial row too small");

// Upper-bound codes of the edge-pass prefilter: the float's bits above bit 18 (8 exponent and 5
// mantissa bits), rounded up, so that bound_decode(bound_code(x)) >= x for every x >= 0.
constexpr uint32_t BOUND_CODE_MAX = 0x1FDFu;  // the largest code that decodes to a finite float
__device__ inline uint32_t bound_code(float x) {
  const uint32_t c = (__float_as_uint(x) + 0x3FFFFu) >> 18;
  return c < BOUND_CODE_MAX ? c : BOUND_CODE_MAX;
}
__device__ inline float bound_decode(uint32_t code) { return __uint_as_float((code & 0x1FFFu) << 18); }

// Edge-pass record of an agent with flags f: the flags, and rel_trans if infectious, else rel_sus.
__device__ inline uint64_t edge_record(uint32_t f, float rt, float rs) {
  const float val = (f & CVK_INFECTIOUS) ? rt : rs;
  return ((uint64_t)f << 32) | (uint64_t)__float_as_uint(val);
}

// Prefilter bound of the edge pass (wave_tile_pass): the source's factors for any layer (variant
// in bits 13..15), or the target's; 0 for agents that neither transmit nor acquire.
__device__ inline uint16_t edge_bound(uint32_t f, float rt, float rs, const cvk_params& prm, float quar_max,
                                      float iso_max) {
  uint32_t bc = 0;
  if ((f & CVK_INFECTIOUS) && rt != 0.0f) {
    float ub = rt;
    if (f & CVK_QUARANTINED) ub = ub * quar_max;
    if (!(f & CVK_SYMPTOMATIC)) ub = ub * prm.asymp_factor;
    if (f & CVK_ISOLATED) ub = ub * iso_max;
    ub = ub * ((f & CVK_EARLY) ? prm.vl_hi : prm.vl_lo);
    bc = (inf_var(f) << 13) | bound_code(ub);
  } else if ((f & CVK_SUSCEPTIBLE) && rs != 0.0f) {
    bc = bound_code((f & CVK_QUARANTINED) ? rs * quar_max : rs);
  }
  return (uint16_t)bc;
}

// The largest code a susceptible agent's edge_bound can take before the edge pass: quarantine at
// max(1, its largest factor), whether or not the agent is quarantined (testing and tracing change that
// flag, and rewrite the code, after the agent pass). 0 for other agents.
__device__ inline uint32_t target_bound_any(uint32_t f, float rs, float quar_any) {
  return ((f & CVK_SUSCEPTIBLE) && !(f & CVK_INFECTIOUS) && rs != 0.0f) ? bound_code(rs * quar_any) : 0u;
}

__device__ inline uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
  return v;
}

// Per-agent inputs, loaded in two round trips: the state words (and the NAb fields), then only the
// dates and factors the agent's state can use. Those of an agent that is not exposed are NaN / 0
// except rel_sus (susceptible), date_recovered (ever infected, with waning: check_immunity) and
// date_end_isolation (isolated): its transitions and viral-load phase then come out as before (none,
// and not early), and its edge record and bound are those of an agent that neither transmits nor
// acquires. A line of 32 agents none of whom is exposed is not fetched at all.
struct AgentIn {
  uint32_t f, vs;
  float d_inf, d_sym, d_sev, d_crit, d_rec, d_dead, d_iso, rt, rs, peak, nab;
  int32_t tne;
};

__device__ inline void load_agent_state(const cvk_people& P, int64_t g, bool waning, AgentIn& x) {
  x.f = P.flags[g];
  x.vs = P.vstate[g];
  if (waning) {
    x.peak = P.peak_nab[g];
    x.nab = P.nab[g];
    x.tne = P.t_nab_event[g];
  } else {
    x.peak = 0.0f;
    x.nab = 0.0f;
    x.tne = 0;
  }
}

// Every date and factor, independent of the state words (issued with them: one round trip).
__device__ inline void load_agent_all(const cvk_people& P, int64_t g, AgentIn& x) {
  x.d_inf = P.date_infectious[g];
  x.d_sym = P.date_symptomatic[g];
  x.d_sev = P.date_severe[g];
  x.d_crit = P.date_critical[g];
  x.d_rec = P.date_recovered[g];
  x.d_dead = P.date_dead[g];
  x.d_iso = P.date_end_isolation[g];
  x.rt = P.rel_trans[g];
  x.rs = P.rel_sus[g];
}

__device__ inline void load_agent_dates(const cvk_people& P, int64_t g, bool waning, AgentIn& x) {
  const float fnan = __builtin_nanf("");
  const bool exp = (x.f & CVK_EXPOSED) != 0;
  x.d_inf = exp ? P.date_infectious[g] : fnan;
  x.d_sym = exp ? P.date_symptomatic[g] : fnan;
  x.d_sev = exp ? P.date_severe[g] : fnan;
  x.d_crit = exp ? P.date_critical[g] : fnan;
  x.d_dead = exp ? P.date_dead[g] : fnan;
  x.d_rec = (exp || (waning && !(x.f & CVK_NAIVE))) ? P.date_recovered[g] : fnan;
  x.d_iso = (x.f & CVK_ISOLATED) ? P.date_end_isolation[g] : fnan;
  x.rt = exp ? P.rel_trans[g] : 0.0f;
  x.rs = (exp || (x.f & CVK_SUSCEPTIBLE)) ? P.rel_sus[g] : 0.0f;
}

// The agents one block of the agent pass handles: AGENT_BLOCK * APT consecutive agents of one sim.
struct AgentSpan {
  int64_t g0;          // storage index of the sim's agent 0 (the block's agents are g0 + li)
  int64_t li0;         // first local index of the block
  int64_t n_local;     // agents in the sim
  int64_t n_stride;    // row stride of the [variant][agent] arrays (all agents in storage)
  uint64_t* bitmap;    // the sim's infectious bitmap (64 agents per word, local indexing)
  uint32_t* coarse;    // one bit per 16 agents (OR of bitmap bits), one word per 512 agents; null: not
                       // written (needs AGENT_BLOCK * APT == 512 and li0 a multiple of 512)
  uint64_t* rec;       // the sim's edge records (local indexing)
  uint16_t* bound;     // the sim's upper-bound codes (local indexing)
  uint64_t* aux;       // the sim's target bits (susceptible with rel_sus != 0, not infectious) and, in the
                       // partial column DC_TMAX, the block's largest target code; null: neither is written
  uint32_t* partials;  // column-major partial rows; this block writes entry `part_col` of each column
  double* partials_f;
  int64_t part_stride;
  int64_t part_col;
};

// update_states_pre (people.py:164-312, 368-374) + check_immunity (immunity.py:303-350) + the
// viral-load phase (utils.py:58-78) for day t, preceded (closing != 0) by update_nab and the
// stock counts of day tc = transitions ? t - 1 : t (sim.py:651-674). Block-wide.
// LAZY: load the dates and factors in a second round trip, only where the state uses them (fewer
// bytes on quiet days); otherwise every field is loaded in one round trip.
template <int APT, bool LAZY>
__device__ void agent_block(const cvk_people& P, const cvk_params& prm, const AgentSpan& S, int32_t t,
                            int32_t prologue, int32_t transitions) {
  // per-wave counters: sums of per-agent booleans via ballot (wave-uniform, no shuffles)
  uint32_t c[N_AGENT_COLS];
#pragma unroll
  for (int k = 0; k < N_AGENT_COLS; ++k) c[k] = 0;
  double d[N_AGENT_FCOLS] = {0.0, 0.0, 0.0};
  const int nv = prm.n_variants;
  const bool waning = prm.use_waning != 0;
  const float tf = (float)t;
  const int64_t n_words = (S.n_local + 63) / 64;
  const bool closing = prologue || !transitions;  // stocks + update_nab of the previous (or this) day
  const int32_t tc = transitions ? t - 1 : t;
  uint32_t coarse_word = 0;  // this wave's bits of the block's coarse word (wave-uniform)
  float quar_max = 0.0f, iso_max = 0.0f;  // over layers: the bound covers every layer's factor
  for (int l = 0; l < prm.n_layers; ++l) {
    quar_max = fmaxf(quar_max, prm.quar_factor[l]);
    iso_max = fmaxf(iso_max, prm.iso_factor[l]);
  }
  const float quar_any = fmaxf(1.0f, quar_max);
  uint32_t tmax = 0;  // this lane's largest target code (target bits only)

  int64_t idx[APT];
  AgentIn in[APT];
#pragma unroll
  for (int r = 0; r < APT; ++r) {
    idx[r] = S.li0 + (int64_t)r * AGENT_BLOCK + threadIdx.x;
    if (idx[r] < S.n_local) {
      load_agent_state(P, S.g0 + idx[r], waning, in[r]);
      if (!LAZY && transitions) load_agent_all(P, S.g0 + idx[r], in[r]);
    }
  }
  if (LAZY && transitions) {  // a count-only pass reads the state words alone
#pragma unroll
    for (int r = 0; r < APT; ++r)
      if (idx[r] < S.n_local) load_agent_dates(P, S.g0 + idx[r], waning, in[r]);
  }
#pragma unroll
  for (int r = 0; r < APT; ++r) {
    const int64_t li = idx[r];
    const int64_t g = S.g0 + li;
    const bool valid = li < S.n_local;
    bool inf_bit = false;
    bool tgt_bit = false;
    uint32_t f_close = 0, vs_close = 0;  // state counted as the previous day's stocks
    uint32_t ev = 0;                     // this day's transition events (bits as the day columns)
    uint32_t ev_var = CVK_VAR_NONE;      // variant of a new infectious agent
    if (valid) {
      const AgentIn& x = in[r];
      uint32_t f = x.f;
      uint32_t vs = x.vs;
      f_close = f;
      vs_close = vs;
      if (closing && waning) {
        float nab = x.nab;
        if (x.peak != 0.0f) {  // has_nabs = true(peak_nab), sim.py:666-669; immunity.py:205-213
          int32_t ts = tc - x.tne;
          ts = ts < 0 ? 0 : (ts >= prm.nab_kin_len ? prm.nab_kin_len - 1 : ts);
          nab = (float)((double)nab + prm.nab_kin[ts] * (double)x.peak);
          nab = nab < 0.0f ? 0.0f : nab;
          nab = nab > x.peak ? x.peak : nab;
          P.nab[g] = nab;
          for (int v = 0; v < nv; ++v) {  // population means of sus/symp_imm (sim.py:673-674)
            d[1] += (double)P.sus_imm[(int64_t)v * S.n_stride + g];
            d[2] += (double)P.symp_imm[(int64_t)v * S.n_stride + g];
          }
          in[r].nab = nab;
        }
        if (!(f & CVK_DEAD)) d[0] += (double)nab;
      }
      if (transitions) {
        const uint32_t f0 = f, vs0 = vs;
        const bool is_exp = (f & CVK_EXPOSED) != 0;  // people.py:169 snapshot
        const float d_inf = x.d_inf, d_rec = x.d_rec, d_dead = x.d_dead;
        if (is_exp) {
          // check_infectious, people.py:222-232
          if (!(f & CVK_INFECTIOUS) && !isnan(d_inf) && tf >= d_inf) {
            const uint32_t evv = exp_var(vs);
            f = ((f | CVK_INFECTIOUS) & ~INFVAR_MASK) | (evv << CVK_INFVAR_SHIFT);
            if (evv < MAXV) vs |= 1u << (16 + evv);
            ev |= 1u << 0;
            ev_var = evv;
          }
          // check_symptomatic / severe / critical, people.py:235-253
          if (!(f & CVK_SYMPTOMATIC) && !isnan(x.d_sym) && tf >= x.d_sym) { f |= CVK_SYMPTOMATIC; ev |= 1u << 1; }
          if (!(f & CVK_SEVERE) && !isnan(x.d_sev) && tf >= x.d_sev) { f |= CVK_SEVERE; ev |= 1u << 2; }
          if (!(f & CVK_CRITICAL) && !isnan(x.d_crit) && tf >= x.d_crit) { f |= CVK_CRITICAL; ev |= 1u << 3; }
          // check_recovery, people.py:256-291
          if (!(f & CVK_RECOVERED) && !isnan(d_rec) && tf >= d_rec) {
            const uint32_t evv = exp_var(vs);
            f = ((f & ~STATE_CLEAR_ON_RECOVERY) | CVK_RECOVERED) | INFVAR_MASK;
            vs = (vs & ~0xFFFFFFu) | (evv << 4) | CVK_VAR_NONE;  // recovered_variant = exposed_variant
            if (waning) f = (f | CVK_SUSCEPTIBLE) & ~CVK_DIAGNOSED;
            ev |= 1u << 4;
          }
        }
        // check_exit_iso, people.py:368-374 (not filtered by exposure)
        if ((f & CVK_ISOLATED) && !isnan(x.d_iso) && tf >= x.d_iso) f &= ~CVK_ISOLATED;
        // check_death, people.py:294-312
        if (is_exp && !(f & CVK_DEAD) && !isnan(d_dead) && tf >= d_dead) {
          f |= CVK_DEAD;
          if (f & CVK_DIAGNOSED) { f |= CVK_KNOWN_DEAD; ev |= 1u << 6; }
          f = (f & ~STATE_CLEAR_ON_DEATH) | INFVAR_MASK;
          vs = (vs & ~0xFFu) | CVK_VAR_NONE | (CVK_VAR_NONE << 4);
          ev |= 1u << 5;
        }
        // check_immunity, immunity.py:303-350 (no vaccine sources on this path)
        if (waning && x.peak != 0.0f) {  // agents never infected keep sus/symp/sev_imm == 0
          const double nab = (double)in[r].nab;
          const bool was_inf = tf >= d_rec;
          const uint32_t rv = rec_var(vs);
          double prev_eff = -1.0;  // efficacies of the last distinct effective level (levels are >= 0)
          float ve_sus = 0.0f, ve_symp = 0.0f, ve_sev = 0.0f;
          for (int v = 0; v < nv; ++v) {
            const double imm = (was_inf && rv < MAXV) ? (double)prm.immunity[v][rv] : 0.0;
            const double eff = nab * imm;
            if (eff != prev_eff) {  // variants with the same cross-immunity share the three powers
              ve_sus = calc_ve(eff, prm.ve_exp_alpha[0], prm.ve_beta[0]);
              ve_symp = calc_ve(eff, prm.ve_exp_alpha[1], prm.ve_beta[1]);
              ve_sev = calc_ve(eff, prm.ve_exp_alpha[2], prm.ve_beta[2]);
              prev_eff = eff;
            }
            const int64_t off = (int64_t)v * S.n_stride + g;
            P.sus_imm[off] = ve_sus;
            P.symp_imm[off] = ve_symp;
            P.sev_imm[off] = ve_sev;
          }
        }
        // viral-load phase (utils.py:58-78): early iff (t - start)/total < switch point
        {
          const float stop = isnan(d_dead) ? d_rec : d_dead;
          const float total = stop - d_inf;
          const float trans_day = prm.frac_time * total;
          const float trans_point = trans_day > prm.high_cap ? prm.high_cap / total : prm.frac_time;
          const bool early = (tf - d_inf) / total < trans_point;
          f = early ? (f | CVK_EARLY) : (f & ~CVK_EARLY);
        }
        if (f != f0) P.flags[g] = f;
        if (vs != vs0) P.vstate[g] = vs;
        inf_bit = (f & CVK_INFECTIOUS) && x.rt != 0.0f;
        S.rec[li] = edge_record(f, x.rt, x.rs);
        S.bound[li] = edge_bound(f, x.rt, x.rs, prm, quar_max, iso_max);
        if (S.aux) {  // uniform
          const uint32_t tb = target_bound_any(f, x.rs, quar_any);
          tgt_bit = tb != 0u;
          tmax = max(tmax, tb);
        }
        if (f & CVK_SEVERE) ev |= 1u << 7;
        if (f & CVK_CRITICAL) ev |= 1u << 8;
      }
    }
    // wave counts (every lane reaches these ballots)
    if (closing) {
      const uint32_t stock_bits[13] = {CVK_SUSCEPTIBLE, CVK_EXPOSED, CVK_INFECTIOUS, CVK_SYMPTOMATIC, CVK_SEVERE,
                                       CVK_CRITICAL, CVK_RECOVERED, CVK_DEAD, CVK_DIAGNOSED, CVK_KNOWN_DEAD,
                                       CVK_QUARANTINED, CVK_ISOLATED, CVK_VACCINATED};
#pragma unroll
      for (int k = 0; k < 13; ++k) c[k] += (uint32_t)__popcll(__ballot(f_close & stock_bits[k]));
      c[SC_ALIVE] += (uint32_t)__popcll(__ballot(valid && !(f_close & CVK_DEAD)));
#pragma unroll
      for (int v = 0; v < MAXV; ++v) {
        if (v < nv) {  // nv is uniform: the ballots stay wave-wide
          c[SC_EXPV + v] += (uint32_t)__popcll(__ballot((vs_close >> (8 + v)) & 1u));
          c[SC_INFV + v] += (uint32_t)__popcll(__ballot((vs_close >> (16 + v)) & 1u));
        }
      }
    }
    if (transitions) {
#pragma unroll
      for (int k = 0; k < 9; ++k) c[DC_BASE + k] += (uint32_t)__popcll(__ballot((ev >> k) & 1u));
#pragma unroll
      for (int v = 0; v < MAXV; ++v)
        if (v < nv) c[DC_BASE + 9 + v] += (uint32_t)__popcll(__ballot(ev_var == (uint32_t)v));
      // one 64-bit ballot per wave covers 64 consecutive agents
      const uint64_t ib = __ballot(inf_bit);
      c[DC_BITS] += (uint32_t)__popcll(ib);
      const int64_t word = li >> 6;
      if (lane_id() == 0 && word < n_words) S.bitmap[word] = ib;
      if (S.aux) {  // uniform
        const uint64_t tw = __ballot(tgt_bit);
        if (lane_id() == 0 && word < n_words) S.aux[word] = tw;
      }
      // coarse bits of this ballot's four 16-agent groups, at their place in the block's 512 agents
      const uint32_t nib = (uint32_t)((ib & 0xFFFFull) != 0) | ((uint32_t)(((ib >> 16) & 0xFFFFull) != 0) << 1) |
                           ((uint32_t)(((ib >> 32) & 0xFFFFull) != 0) << 2) | ((uint32_t)((ib >> 48) != 0) << 3);
      coarse_word |= nib << (uint32_t)((((li - S.li0) & ~63ll) >> 4) & 31);
    }
  }

  c[DC_TMAX] = (S.aux && transitions) ? wave_max(tmax) : 0u;  // uniform

  // block partial row: per-wave counts through LDS (no atomics), float sums in wave order
  __shared__ uint32_t s_wc[AGENT_BLOCK / 64][N_AGENT_COLS];
  __shared__ double s_wd[AGENT_BLOCK / 64][N_AGENT_FCOLS];
  __shared__ uint32_t s_coarse[AGENT_BLOCK / 64];
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < N_AGENT_COLS; ++k) s_wc[w][k] = c[k];
    s_coarse[w] = coarse_word;
  }
  if (waning) {
#pragma unroll
    for (int k = 0; k < N_AGENT_FCOLS; ++k) {
      const double v = wave_sum_d(d[k]);
      if (lane_id() == 0) s_wd[w][k] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < N_SUM_COLS) {
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < AGENT_BLOCK / 64; ++j) tot += s_wc[j][threadIdx.x];
    S.partials[(size_t)threadIdx.x * S.part_stride + S.part_col] = tot;
  } else if (threadIdx.x == DC_TMAX) {
    if (S.aux) {
      uint32_t m = 0;
#pragma unroll
      for (int j = 0; j < AGENT_BLOCK / 64; ++j) m = max(m, s_wc[j][DC_TMAX]);
      S.partials[(size_t)DC_TMAX * S.part_stride + S.part_col] = m;
    }
  } else if (threadIdx.x >= 64 && threadIdx.x < 64 + N_AGENT_FCOLS) {
    const int k = threadIdx.x - 64;
    double tot = 0.0;
    if (waning)
      for (int j = 0; j < AGENT_BLOCK / 64; ++j) tot += s_wd[j][k];
    S.partials_f[(size_t)k * S.part_stride + S.part_col] = tot;
  } else if (threadIdx.x == 128 && S.coarse && transitions) {
    uint32_t cw = 0;
#pragma unroll
    for (int j = 0; j < AGENT_BLOCK / 64; ++j) cw |= s_coarse[j];
    S.coarse[S.li0 >> 9] = cw;
  }
}

// Writes one sim's day rows from the summed agent-pass partial rows (block-wide; s_tot/s_totf
// from sum_partials). results / results_f are the sim's [npts][cols] tables.
__device__ inline void write_agent_rows(uint32_t* results, double* results_f, const uint32_t* s_tot,
                                        const double* s_totf, int32_t t, int32_t prologue, int32_t transitions) {
  if (prologue || !transitions) {
    const int32_t tc = transitions ? t - 1 : t;
    uint32_t* row = results + (int64_t)tc * CVK_R_COLS;
    for (int k = threadIdx.x; k < N_STOCK_COLS; k += blockDim.x) {
      int col;
      if (k < 13) col = CVK_R_N_SUSCEPTIBLE + k;
      else if (k == SC_ALIVE) col = CVK_R_N_ALIVE_NAB;
      else if (k < SC_INFV) col = CVK_R_BYVAR + CVK_RV_N_EXPOSED * MAXV + (k - SC_EXPV);
      else col = CVK_R_BYVAR + CVK_RV_N_INFECTIOUS * MAXV + (k - SC_INFV);
      row[col] = s_tot[k];
    }
    if (threadIdx.x < N_AGENT_FCOLS) results_f[(int64_t)tc * CVK_RF_COLS + threadIdx.x] = s_totf[threadIdx.x];
  }
  if (transitions) {
    uint32_t* row = results + (int64_t)t * CVK_R_COLS;
    static constexpr int cols[9] = {CVK_R_NEW_INFECTIOUS, CVK_R_NEW_SYMPTOMATIC, CVK_R_NEW_SEVERE,
                                    CVK_R_NEW_CRITICAL, CVK_R_NEW_RECOVERIES, CVK_R_NEW_DEATHS,
                                    CVK_R_NEW_KNOWN_DEATHS, CVK_R_MID_SEVERE, CVK_R_MID_CRITICAL};
    for (int k = threadIdx.x; k < N_DAY_COLS; k += blockDim.x) {
      const int col = k < 9 ? cols[k] : CVK_R_BYVAR + CVK_RV_NEW_INFECTIOUS * MAXV + (k - 9);
      row[col] = s_tot[DC_BASE + k];
    }
  }
  __syncthreads();
}

// Dynamic-layer resample on day t: iid uniform endpoints (base.py:1849-1876), a pure function of
// (seed, t, layer, e), so a consumer may compute them instead of reading them back. Edges 2q and 2q+1
// share one draw (words 0, 1 and 2, 3 are their p1, p2). Each endpoint is an unbiased bounded
// integer by multiply-and-reject: w * n is accepted when its low word is at least
// (2^32 - n) mod n (resample_thr), else the same-position word of the edge's retry draws is used
// (at most RESAMPLE_MAX_RETRY; the last is accepted as is, with probability below 10^-600).
__device__ inline uint32_t resample_thr(uint32_t n) { return (0u - n) % n; }

__device__ inline int32_t resample_endpoint(uint64_t seed, int32_t t, uint32_t layer, uint32_t n, uint32_t thr,
                                            int64_t e, int j, uint32_t w) {
  uint64_t m = (uint64_t)w * n;
  if ((uint32_t)m < thr) {  // rare: < n / 2^32 of the words
    const uint64_t ue = (uint64_t)e;
    for (uint32_t k = 0; k < RESAMPLE_MAX_RETRY; ++k) {
      const u32x4 r = philox_seeded(seed, (uint32_t)ue, (uint32_t)t, STREAM_RESAMPLE_RETRY | layer,
                                    ((uint32_t)(ue >> 32) << 8) | k);
      m = (uint64_t)(j ? r.y : r.x) * n;
      if ((uint32_t)m >= thr) break;
    }
  }
  return (int32_t)(m >> 32);
}

struct EdgeEnds {
  int32_t p1, p2;
};

// The draw of pair q (edges 2q, 2q+1).
__device__ inline u32x4 resample_pair_draw(uint64_t seed, int32_t t, uint32_t layer, int64_t q) {
  const uint64_t uq = (uint64_t)q;
  return philox_seeded(seed, (uint32_t)uq, (uint32_t)t, STREAM_RESAMPLE | layer, (uint32_t)(uq >> 32));
}

__device__ inline EdgeEnds dynamic_endpoints(uint64_t seed, int32_t t, uint32_t layer, uint32_t n_agents, int64_t e) {
  const u32x4 w = resample_pair_draw(seed, t, layer, e >> 1);
  const uint32_t thr = resample_thr(n_agents);
  const bool odd = e & 1;
  return EdgeEnds{resample_endpoint(seed, t, layer, n_agents, thr, e, 0, odd ? w.z : w.x),
                  resample_endpoint(seed, t, layer, n_agents, thr, e, 1, odd ? w.w : w.y)};
}

// Writes the resample of pair q (edges 2q and, when it exists, 2q+1 of an n_edges layer), beta = 1.
// beta may be null when the layer's weights are already all ones (every resample after the first of
// a run). thr = resample_thr(n_agents).
__device__ inline void resample_pair(uint64_t seed, int32_t t, uint32_t layer, uint32_t n_agents, uint32_t thr,
                                     int64_t q, int64_t n_edges, int32_t* p1, int32_t* p2, float* beta) {
  const u32x4 w = resample_pair_draw(seed, t, layer, q);
  const int64_t e = 2 * q;
  const int32_t a1 = resample_endpoint(seed, t, layer, n_agents, thr, e, 0, w.x);
  const int32_t a2 = resample_endpoint(seed, t, layer, n_agents, thr, e, 1, w.y);
  if (e + 1 < n_edges) {
    const int32_t b1 = resample_endpoint(seed, t, layer, n_agents, thr, e + 1, 0, w.z);
    const int32_t b2 = resample_endpoint(seed, t, layer, n_agents, thr, e + 1, 1, w.w);
    typedef int32_t i32x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    *(CVK_GLOBAL i32x2*)gmem(p1 + e) = i32x2{a1, b1};
    *(CVK_GLOBAL i32x2*)gmem(p2 + e) = i32x2{a2, b2};
    if (beta) *(CVK_GLOBAL f32x2*)gmem(beta + e) = f32x2{1.0f, 1.0f};
  } else {
    gmem(p1)[e] = a1;
    gmem(p2)[e] = a2;
    if (beta) gmem(beta)[e] = 1.0f;
  }
}

// Selection key of cvk_choose: philox word << 32 | local agent index (distinct per agent).
__device__ inline uint64_t selection_key(uint64_t seed, int32_t t, uint32_t stream_id, uint32_t i) {
  const u32x4 w = philox_seeded(seed, i, (uint32_t)t, stream_id, 0u);
  return ((uint64_t)w.x << 32) | i;
}

// ===========================================================================
// Edge pass: compute_trans_sus + compute_infections + first-occurrence dedup
// ===========================================================================
// One sim as the edge pass sees it (pointers offset to the sim's slices; local indexing).
struct SimView {
  const uint64_t* rec;
  const uint16_t* bound;    // per-agent upper-bound codes (prefilter), see AgentSpan
  uint64_t* winner;
  const float* sus_imm;     // row v at sus_imm + v * sus_stride
  int64_t sus_stride;
  const uint32_t* bitmap;
  const uint32_t* coarse;   // one bit per 16 agents (single-sim day loop; see AgentSpan), or null
  const uint32_t* aux;      // target bits (see AgentSpan), or null
  uint32_t* touched;        // bitmap over storage indices (a_off + local): targets holding a key
  int64_t a_off;
  uint64_t seed;
  uint32_t n_agents;
  int32_t n_variants;
  int32_t use_sus_imm;
  float asymp_factor, vl_hi, vl_lo;
  float beta_v[MAXV];
};

// One layer of one sim.
struct SegView {
  const int32_t* p1;
  const int32_t* p2;
  const float* beta;
  const double* unif;  // [n_variants][2][n] supplied uniforms, or null
  int64_t n;
  float beta_layer, iso_factor, quar_factor;
  int32_t beta_ones;   // every per-edge beta is 1.0f: no gather needed
  int32_t l;           // layer index (key and counter field)
  int32_t gen;         // dynamic layer of a fused day loop: endpoints come from the day's resample
                       // stream (dynamic_endpoints), p1/p2 are not read
};

typedef int v4i __attribute__((ext_vector_type(4)));

// The float64 uniform of a candidate slot (utils.py:123): supplied, or drawn from the slot's
// counter stream (edge, day, variant, layer; the two directions use the two halves of one draw).
__device__ inline double edge_uniform(const SimView& S, const SegView& L, int32_t t, uint32_t v, int dir, int64_t e) {
  if (L.unif) return gmem(L.unif)[((int64_t)v * 2 + dir) * L.n + e];
  const uint64_t ue = (uint64_t)e;
  const u32x4 w = philox_seeded(S.seed, (uint32_t)ue, (uint32_t)t, STREAM_TX | (v << 8) | (uint32_t)L.l,
                                (uint32_t)(ue >> 32));
  return dir ? u01(w.z, w.w) : u01(w.x, w.y);
}

// Exact-rejection prefilter: p <= p_ub = beta_v * beta_layer * bound(src) * bound(tgt) (each bound at or
// above the factor it covers; the 1 + 2^-10 margin absorbs float rounding of both products), so
// u >= p_ub implies u >= p and the candidate is rejected without gathering the two 8-byte records.
constexpr float BOUND_MARGIN = 1.0f + 1.0f / 1024.0f;

__device__ inline v4i load4(const int32_t* p0, int64_t e, int64_t n) {
  const CVK_GLOBAL int32_t* p = gmem(p0);
  if (e + 3 < n) return __builtin_nontemporal_load((const CVK_GLOBAL v4i*)(p + e));
  v4i r = {0, 0, 0, 0};
  if (e < n) r[0] = p[e];
  if (e + 1 < n) r[1] = p[e + 1];
  if (e + 2 < n) r[2] = p[e + 2];
  return r;
}

// ---------------------------------------------------------------------------
// Wave-tile edge pass: a wave owns K x 256 consecutive edges of one layer and meets memory latency
// three times per tile at most (endpoints, bound codes, and, when the bitmap is in L2, the bits),
// because the per-candidate work in between touches no memory:
//   1. infectious bits of both endpoints; an edge is a candidate in at most one direction (the
//      infectious endpoint is the source; an infectious agent is never susceptible);
//   2. the 2-byte bound codes of both endpoints of every mixed edge, issued together;
//   3. candidates with a susceptible target go onto the wave's LDS stack with a rounded-up p_ub;
//      64 at a time they draw their uniform and test u < p_ub (no memory access);
//   4. the few that pass are queued (LDS) for the exact float32 test of the reference, which
//      gathers the two 8-byte records; the queue is drained 32 entries at a time, so that memory
//      round trip is shared by many tiles.
// The set of (target, key) pairs offered to atomicMin is exactly the reference's, in another order.
// ---------------------------------------------------------------------------
constexpr int EDGE_CQ = 320;  // candidate stack per wave: < 64 carried + 256 from one chunk
constexpr int EDGE_XQ = 32;   // deferred exact tests per wave
constexpr int EDGE_WAVE_WORDS = EDGE_CQ + 5 * EDGE_XQ;  // LDS words per wave

// p_ub kept in 18 bits (float bits 30..13, rounded up, so the decoded value is never below it);
// PCODE_INF decodes to +inf (always take the exact test).
constexpr uint32_t PCODE_INF = 0x7F800000u >> 13;
__device__ inline uint32_t pcode_up(float p) {
  const uint32_t c = (__float_as_uint(p) + 0x1FFFu) >> 13;
  return c < PCODE_INF ? c : PCODE_INF;
}
__device__ inline float pcode_decode(uint32_t c) { return __uint_as_float(c << 13); }

// The largest float <= u (u >= 0): u < p iff rz(u) < p for every float p, so the deferred test
// compares floats and stays exact.
__device__ inline float float_rz(double u) {
  const float f = (float)u;
  return (double)f > u ? __uint_as_float(__float_as_uint(f) - 1u) : f;
}

struct WaveQ {
  const float* beta;  // the sim's beta_v[MAXV] in LDS (a lane-varying index into the kernel argument would be
                      // a memory round trip per use, and a register select spills scalar registers)
  uint32_t* cand;  // [EDGE_CQ] lane | q << 6 | c << 8 | dir << 10 | v << 11 | pcode << 14
  uint32_t* xs;    // [EDGE_XQ] source
  uint32_t* xt;    // target
  uint32_t* xe;    // edge index, low 32 bits
  uint32_t* xm;    // dir | v << 1 | (edge >> 32) << 4
  uint32_t* xu;    // float_rz(u)
};
__device__ inline WaveQ wave_queues(uint32_t* base, const float* s_beta) {
  WaveQ Q;
  Q.beta = s_beta;
  Q.cand = base;
  Q.xs = base + EDGE_CQ;
  Q.xt = Q.xs + EDGE_XQ;
  Q.xe = Q.xt + EDGE_XQ;
  Q.xm = Q.xe + EDGE_XQ;
  Q.xu = Q.xm + EDGE_XQ;
  return Q;
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int K>
struct WaveTile {
  int64_t e0;       // first edge of the tile in its layer
  int32_t seg;      // segment (sim x layer) of the tile
  int32_t n_valid;  // edges of the tile that exist
  v4i x[K], y[K];   // lane's edges (c * 64 + lane) * 4 + q of the tile
};

// The tile's endpoints: loaded, or for a generated dynamic layer (GEN instantiations only) computed
// from the day's resample stream.
template <int K, bool GEN = false>
__device__ inline void load_wave_edges(const SimView& S, const SegView& L, int32_t t, WaveTile<K>& T) {
  constexpr int TILE = 64 * 4 * K;
  const int64_t rem = L.n - T.e0;
  T.n_valid = (int32_t)(rem < TILE ? rem : TILE);
  const uint32_t lane = lane_id();
  if (GEN && L.gen) {  // edges 4i .. 4i+3 of a lane are two resample pairs
    const uint32_t thr = resample_thr(S.n_agents);
#pragma unroll
    for (int c = 0; c < K; ++c) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t e = T.e0 + ((int64_t)c * 64 + lane) * 4 + 2 * h;
        u32x4 w = {0u, 0u, 0u, 0u};
        if (e < L.n) w = resample_pair_draw(S.seed, t, (uint32_t)L.l, e >> 1);
        const uint32_t l = (uint32_t)L.l;
        T.x[c][2 * h] = e < L.n ? resample_endpoint(S.seed, t, l, S.n_agents, thr, e, 0, w.x) : 0;
        T.y[c][2 * h] = e < L.n ? resample_endpoint(S.seed, t, l, S.n_agents, thr, e, 1, w.y) : 0;
        T.x[c][2 * h + 1] = e + 1 < L.n ? resample_endpoint(S.seed, t, l, S.n_agents, thr, e + 1, 0, w.z) : 0;
        T.y[c][2 * h + 1] = e + 1 < L.n ? resample_endpoint(S.seed, t, l, S.n_agents, thr, e + 1, 1, w.w) : 0;
      }
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < K; ++c) {
    const int64_t e = T.e0 + ((int64_t)c * 64 + lane) * 4;
    T.x[c] = load4(L.p1, e, L.n);
    T.y[c] = load4(L.p2, e, L.n);
  }
}

// The reference's float32 test for one candidate slot (utils.py:88-89, 116-123) and the
// first-occurrence dedup. uf = float_rz(u) of the slot's draw for variant vq.
__device__ inline void exact_test(const SimView& S, const SegView& L, int32_t t, int32_t src, int32_t tgt, int64_t e,
                                  uint32_t dir, uint32_t vq, float uf, const float* s_beta) {
  const uint64_t rsrc = gmem(S.rec)[src];
  const uint64_t rtgt = gmem(S.rec)[tgt];
  const uint32_t fs = (uint32_t)(rsrc >> 32);
  const uint32_t ft = (uint32_t)(rtgt >> 32);
  const uint32_t v = inf_var(fs);
  if (!(fs & CVK_INFECTIOUS) || v >= (uint32_t)S.n_variants || !(ft & CVK_SUSCEPTIBLE)) return;
  float rt = __uint_as_float((uint32_t)rsrc);                   // rel_trans * inf (== 1)
  if (fs & CVK_QUARANTINED) rt = rt * L.quar_factor;             // * f_quar
  if (!(fs & CVK_SYMPTOMATIC)) rt = rt * S.asymp_factor;         // * f_asymp
  if (fs & CVK_ISOLATED) rt = rt * L.iso_factor;                 // * f_iso
  rt = rt * L.beta_layer;                                        // * beta_layer
  rt = rt * ((fs & CVK_EARLY) ? S.vl_hi : S.vl_lo);              // * viral_load
  if (rt == 0.0f) return;                                        // utils.py:116
  float rs = __uint_as_float((uint32_t)rtgt);                   // rel_sus * sus (== 1)
  if (ft & CVK_QUARANTINED) rs = rs * L.quar_factor;
  if (S.use_sus_imm) rs = rs * (1.0f - gmem(S.sus_imm)[(int64_t)v * S.sus_stride + tgt]);
  const float eb = L.beta_ones ? 1.0f : gmem(L.beta)[e];
  const float p = ((s_beta[v] * eb) * rt) * rs;                  // utils.py:117
  if (p == 0.0f) return;                                         // utils.py:118
  const bool hit = v == vq ? uf < p : edge_uniform(S, L, t, v, (int)dir, e) < (double)p;  // utils.py:123
  if (!hit) return;
  const uint64_t key = make_key(1, v, (uint32_t)L.l, dir, (uint64_t)e);
  const uint64_t old = __hip_atomic_fetch_min(gmem(S.winner) + tgt, key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (old == KEY_NONE) {  // first key of this target today: mark it for the infect pass
    const int64_t g = S.a_off + tgt;
    __hip_atomic_fetch_or(gmem(S.touched) + (g >> 5), 1u << (g & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Exact tests of the n_x (<= 64) queued slots, one per lane; every slot is from segment L.
__device__ inline void flush_exact(const SimView& S, const SegView& L, int32_t t, const WaveQ& Q, uint32_t n_x) {
  wave_sync_lds();
  const uint32_t lane = lane_id();
  if (lane < n_x) {
    const uint32_t m = Q.xm[lane];
    const int64_t e = (int64_t)Q.xe[lane] | ((int64_t)(m >> 4) << 32);
    exact_test(S, L, t, (int32_t)Q.xs[lane], (int32_t)Q.xt[lane], e, m & 1u, (m >> 1) & 7u,
               __uint_as_float(Q.xu[lane]), Q.beta);
  }
  wave_sync_lds();
}

// The slots of this round that passed the prefilter (m = ballot of pass) go to the wave's exact queue,
// or are tested at once when more than the queue holds passed together. Wave-uniform call.
__device__ inline void queue_exact(const SimView& S, const SegView& L, int32_t t, const WaveQ& Q, uint64_t m, bool pass,
                                   int32_t src, int32_t tgt, int64_t e, uint32_t dir, uint32_t vq, double u,
                                   uint32_t& n_x) {
  const uint32_t k = (uint32_t)__popcll(m);
  if (k > (uint32_t)EDGE_XQ) {
    if (pass) exact_test(S, L, t, src, tgt, e, dir, vq, float_rz(u), Q.beta);
    return;
  }
  if (n_x + k > (uint32_t)EDGE_XQ) {
    flush_exact(S, L, t, Q, n_x);
    n_x = 0;
  }
  if (pass) {
    const uint32_t pos = n_x + lanes_below(m);
    Q.xs[pos] = (uint32_t)src;
    Q.xt[pos] = (uint32_t)tgt;
    Q.xe[pos] = (uint32_t)e;
    Q.xm[pos] = dir | (vq << 1) | ((uint32_t)((uint64_t)e >> 32) << 4);
    Q.xu[pos] = __float_as_uint(float_rz(u));
  }
  n_x += k;
}

// The prefilter bound code of a candidate slot (source code bs, target code bt != 0, variant vq = bs >> 13):
// p_ub rounded up, PCODE_INF for weighted layers and products that may have underflowed, 0 when the
// variant's or the layer's beta is 0.
__device__ inline uint32_t candidate_pcode(const SegView& L, const WaveQ& Q, uint32_t bs, uint32_t bt, uint32_t vq) {
  if (!L.beta_ones) return PCODE_INF;  // weighted layers: per-edge beta is not in the bound
  const float bv = Q.beta[vq];
  const float p_ub = (((bv * L.beta_layer) * bound_decode(bs)) * bound_decode(bt)) * BOUND_MARGIN;
  // p_ub == 0 with nonzero factors can only be underflow: leave such slots to the exact test
  return p_ub >= 0x1p-100f ? pcode_up(p_ub) : ((bv == 0.0f || L.beta_layer == 0.0f) ? 0u : PCODE_INF);
}

// One round: the top min(64, n_c) stacked candidates draw their uniform; those below p_ub are queued
// for the exact test (or tested at once when more than the queue holds pass together).
template <int K>
__device__ inline void candidate_round(const SimView& S, const SegView& L, int32_t t, const WaveQ& Q,
                                       const WaveTile<K>& T, uint32_t& n_c, uint32_t& n_x) {
  wave_sync_lds();
  const uint32_t lane = lane_id();
  const uint32_t take = n_c < 64u ? n_c : 64u;
  const uint32_t base = n_c - take;
  const bool act = lane < take;
  const uint32_t ent = act ? Q.cand[base + lane] : 0u;
  n_c = base;
  wave_sync_lds();  // the slots just read may be overwritten by the next pushes
  const uint32_t ol = ent & 63u, q = (ent >> 6) & 3u, c = (ent >> 8) & 3u, dir = (ent >> 10) & 1u;
  const uint32_t vq = (ent >> 11) & 7u;
  const int64_t e = T.e0 + ((int64_t)c * 64 + ol) * 4 + q;
  double u = 1.0;
  if (act) u = edge_uniform(S, L, t, vq, (int)dir, e);
  const bool pass = act && u < (double)pcode_decode(ent >> 14);
  const uint64_t m = __ballot(pass);
  if (!m) return;  // wave-uniform
  // endpoints of the passing slots, from the lanes that loaded them
  int32_t xv = 0, yv = 0;
#pragma unroll
  for (int cc = 0; cc < K; ++cc) {
    if (!__ballot(pass && c == (uint32_t)cc)) continue;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int32_t a = __shfl(T.x[cc][qq], (int)ol, 64);
      const int32_t b = __shfl(T.y[cc][qq], (int)ol, 64);
      if (c == (uint32_t)cc && q == (uint32_t)qq) {
        xv = a;
        yv = b;
      }
    }
  }
  const int32_t src = dir ? yv : xv, tgt = dir ? xv : yv;
  queue_exact(S, L, t, Q, m, pass, src, tgt, e, dir, vq, u, n_x);
}

// Where the edge pass finds the infectious bits: the fine bitmap staged in LDS; a coarse map in LDS
// (one bit per 16 agents) with fine words read from L2 only for endpoints whose group has an
// infectious agent; or the fine bitmap in L2 only.
constexpr int BITS_LDS = 0, BITS_COARSE = 1, BITS_L2 = 2;

// A tile between its two phases: which edges are mixed, and their endpoints' bound codes (the loads
// may still be in flight).
template <int K>
struct TileBounds {
  uint32_t mixed, from_x;  // bit c*4+q: exactly one endpoint infectious; it is x (p1)
  uint32_t bs[K * 4], bt[K * 4];
};

// Phase 1 of a wave tile: infectious bits of every endpoint, then the bound loads of the mixed edges
// (issued, not waited for). lds_bits: the LDS map of MODE (fine or coarse; unused for BITS_L2).
// TB: the targets' words of the aux bits (bt) instead of their codes (tile_candidates<K, true>).
template <int MODE, int K, bool TB = false>
__device__ inline void tile_bits_and_bounds(const SimView& S, const uint32_t* lds_bits, const WaveTile<K>& T,
                                            TileBounds<K>& B) {
  const CVK_GLOBAL uint16_t* bound = gmem(S.bound);
  const CVK_GLOBAL uint32_t* fine = gmem(S.bitmap);
  // The bitmap words of both endpoints of every edge, all requests in flight before the first use.
  // Branch-free: edges past the end of the layer were loaded as (0, 0), which is never mixed (an agent
  // cannot be infectious and not infectious), and endpoints outside [0, n) (the reference leaves them
  // undefined) are masked out after a clamped lookup.
  const uint32_t nm1 = S.n_agents - 1u;
  uint32_t okm = 0;  // bit c*4+q: both endpoints in [0, n)
  uint32_t wx[K * 4], wy[K * 4];
#pragma unroll
  for (int j = 0; j < K * 4; ++j) {
    const int c = j >> 2, q = j & 3;
    const uint32_t xi = (uint32_t)T.x[c][q], yi = (uint32_t)T.y[c][q];
    okm |= (uint32_t)(max(xi, yi) <= nm1) << j;
    const uint32_t xs = min(xi, nm1), ys = min(yi, nm1);
    if (MODE == BITS_LDS) {
      wx[j] = lds_bits[xs >> 5];
      wy[j] = lds_bits[ys >> 5];
    } else if (MODE == BITS_COARSE) {
      wx[j] = 0u;
      wy[j] = 0u;
      if ((lds_bits[xs >> 9] >> ((xs >> 4) & 31)) & 1u) wx[j] = fine[xs >> 5];
      if ((lds_bits[ys >> 9] >> ((ys >> 4) & 31)) & 1u) wy[j] = fine[ys >> 5];
    } else {
      wx[j] = fine[xs >> 5];
      wy[j] = fine[ys >> 5];
    }
  }
  uint32_t mixed = 0, from_x = 0;
#pragma unroll
  for (int j = 0; j < K * 4; ++j) {
    const int c = j >> 2, q = j & 3;
    const uint32_t ok = (okm >> j) & 1u;
    const uint32_t xi = (uint32_t)T.x[c][q], yi = (uint32_t)T.y[c][q];
    const uint32_t ix = (wx[j] >> (xi & 31)) & ok;
    const uint32_t iy = (wy[j] >> (yi & 31)) & ok;
    mixed |= ((ix ^ iy) & 1u) << j;
    from_x |= (ix & 1u) << j;
  }
  B.mixed = mixed;
  B.from_x = from_x;
  // bound codes of both endpoints of every mixed edge, all in flight together
#pragma unroll
  for (int j = 0; j < K * 4; ++j) {
    B.bs[j] = 0;
    B.bt[j] = 0;
    if ((mixed >> j) & 1u) {
      const uint32_t xi = (uint32_t)T.x[j >> 2][j & 3], yi = (uint32_t)T.y[j >> 2][j & 3];
      const bool sx = (from_x >> j) & 1u;
      B.bs[j] = bound[sx ? xi : yi];
      B.bt[j] = TB ? gmem(S.aux)[(sx ? yi : xi) >> 5] : (uint32_t)bound[sx ? yi : xi];
    }
  }
}

// Phase 2 of a wave tile: candidates onto the LDS stack, rounds of draws and prefilter tests, exact
// tests queued. n_x: the wave's exact-queue fill, whose slots all belong to segment L.
// TB: B.bt holds the targets' aux words; a target is a candidate if its bit is set, and its bound is tmax
// (nonzero), the day's largest target code (at or above the target's own code, so p_ub stays >= p).
template <int K, bool TB = false>
__device__ inline void tile_candidates(const SimView& S, const SegView& L, int32_t t, const WaveQ& Q,
                                       const WaveTile<K>& T, const TileBounds<K>& B, uint32_t& n_x,
                                       uint32_t tmax = 0u) {
  const uint32_t lane = lane_id();
  const uint32_t mixed = B.mixed, from_x = B.from_x;
  const uint32_t* bs = B.bs;
  const uint32_t* bt = B.bt;
  uint32_t n_c = 0;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    uint32_t ent[4];
    bool has[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = c * 4 + q;
      has[q] = false;
      ent[q] = 0;
      const uint32_t vq = bs[j] >> 13;
      // a non-infectious target's code is its susceptibility bound, or 0 if it cannot be infected; with
      // target bits, its bit says whether it can be, and tmax bounds it
      bool can;
      if (TB) {
        const uint32_t tgt = (uint32_t)(((from_x >> j) & 1u) ? T.y[c][q] : T.x[c][q]);
        can = ((bt[j] >> (tgt & 31u)) & 1u) != 0u;
      } else {
        can = bt[j] != 0u;
      }
      if (((mixed >> j) & 1u) && vq < (uint32_t)S.n_variants && can) {
        const uint32_t pc = candidate_pcode(L, Q, bs[j], TB ? tmax : bt[j], vq);
        if (pc) {
          has[q] = true;
          const uint32_t dir = ((from_x >> j) & 1u) ? 0u : 1u;
          ent[q] = lane | ((uint32_t)q << 6) | ((uint32_t)c << 8) | (dir << 10) | (vq << 11) | (pc << 14);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint64_t m = __ballot(has[q]);
      if (has[q]) Q.cand[n_c + lanes_below(m)] = ent[q];
      n_c += (uint32_t)__popcll(m);
    }
    while (n_c >= 64u) candidate_round<K>(S, L, t, Q, T, n_c, n_x);
  }
  if (n_c) candidate_round<K>(S, L, t, Q, T, n_c, n_x);
}

// One wave tile, both phases.
template <int MODE, int K>
__device__ inline void wave_tile_pass(const SimView& S, const SegView& L, int32_t t, const uint32_t* lds_bits,
                                      const WaveQ& Q, const WaveTile<K>& T, uint32_t& n_x) {
  TileBounds<K> B;
  tile_bits_and_bounds<MODE, K>(S, lds_bits, T, B);
  tile_candidates<K>(S, L, t, Q, T, B, n_x);
}

// ---------------------------------------------------------------------------
// Source lists of a static layer: the edge pass walks the lists of the day's infectious agents
// instead of streaming the layer. Slot k of agent i's list [indptr[i], indptr[i + 1]) is one
// direction of one edge: target[k], and edge_dir[k] = edge << 1 | dir (dir 0: i is p1). Every
// (source, target, edge, direction) slot with an infectious source is offered exactly as the
// streamed pass offers it (same prefilter, uniform and exact test), so the results are the same.
// ---------------------------------------------------------------------------
struct SrcView {
  const int32_t* indptr;
  const int32_t* target;
  const uint32_t* edge_dir;
};

constexpr int SRC_ITEM_WORDS = 64;                    // bitmap words per source work item (one per lane)
constexpr int SRC_ITEM_AGENTS = SRC_ITEM_WORDS * 32;  // agents per source work item
constexpr int SRC_UNROLL = 4;                         // slots per lane whose loads are in flight together

// The position of the j-th set bit (j < popcount) of x.
__device__ inline uint32_t nth_set_bit(uint32_t x, uint32_t j) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t half = 16; half; half >>= 1) {
    const uint32_t c = (uint32_t)__popc(x & ((1u << half) - 1u));
    if (j >= c) {
      j -= c;
      x >>= half;
      pos += half;
    }
  }
  return pos;
}

// Source work item: agents [a0, a0 + SRC_ITEM_AGENTS) of static layer L; lds_bits: the fine bitmap in LDS
// (MODE == BITS_LDS), otherwise it is read from global memory. Rounds of up to 64 infectious agents: lane k takes the (64 r + k)-th set bit of the
// item's words (k-th-bit selection over the words' prefix counts), loads its list bounds and bound
// code, then the round's slots are dealt over the lanes (each lane finds its slot's agent by a binary
// search over the round's prefix sums in LDS), SRC_UNROLL slots per lane with their loads in flight
// together. cnt: slots and lists walked (added to by lane 0).
template <int MODE>
__device__ inline void source_item_pass(const SimView& S, const SegView& L, const SrcView& R, int32_t t,
                                        const uint32_t* lds_bits, const WaveQ& Q, int64_t a0, uint32_t& n_x,
                                        uint64_t* cnt) {
  const uint32_t lane = lane_id();
  const int64_t n_words = ((int64_t)S.n_agents + 31) / 32;
  const int64_t w = (a0 >> 5) + lane;
  uint32_t bits = 0;
  if (lane < (uint32_t)SRC_ITEM_WORDS && w < n_words) bits = MODE == BITS_LDS ? lds_bits[w] : gmem(S.bitmap)[w];
  if (w == n_words - 1 && (S.n_agents & 31u)) bits &= (1u << (S.n_agents & 31u)) - 1u;  // agents < n only
  const uint32_t wc = (uint32_t)__popc(bits);
  uint32_t wincl = wc;  // inclusive scan of the words' counts
#pragma unroll
  for (int off = 1; off < SRC_ITEM_WORDS; off <<= 1) {
    const uint32_t y = __shfl_up(wincl, off, 64);
    if ((int)lane >= off) wincl += y;
  }
  const uint32_t n_agents = __shfl(wincl, SRC_ITEM_WORDS - 1, 64);
  if (n_agents == 0) return;  // wave-uniform
  uint32_t* s_off = Q.cand;         // [64] exclusive slot offsets of the round's agents
  uint32_t* s_first = Q.cand + 64;  // [64] first slot of each agent's list
  uint32_t* s_src = Q.cand + 128;   // [64] the agent
  uint32_t* s_bs = Q.cand + 192;    // [64] its bound code
  uint32_t* s_wex = Q.cand + 256;   // [SRC_ITEM_WORDS] exclusive prefix of the words' counts
  wave_sync_lds();
  if (lane < (uint32_t)SRC_ITEM_WORDS) s_wex[lane] = wincl - wc;
  wave_sync_lds();
  const CVK_GLOBAL int32_t* indptr = gmem(R.indptr);
  const CVK_GLOBAL int32_t* target = gmem(R.target);
  const CVK_GLOBAL uint32_t* edge_dir = gmem(R.edge_dir);
  const CVK_GLOBAL uint16_t* bound = gmem(S.bound);
  for (uint32_t r0 = 0; r0 < n_agents; r0 += 64) {  // wave-uniform
    const uint32_t g = r0 + lane;
    uint32_t wd = 0;  // the last word with s_wex <= g (zero-count words share their successor's prefix)
#pragma unroll
    for (uint32_t step = SRC_ITEM_WORDS / 2; step; step >>= 1)
      if (s_wex[wd + step] <= g) wd += step;
    const uint32_t word_bits = (uint32_t)__shfl(bits, (int)wd, 64);  // every lane active
    int32_t a = -1;
    if (g < n_agents) a = (int32_t)(a0 + (int64_t)wd * 32 + nth_set_bit(word_bits, g - s_wex[wd]));
    uint32_t first = 0, deg = 0, bs = 0;
    if (a >= 0) {
      first = (uint32_t)indptr[a];
      deg = (uint32_t)indptr[a + 1] - first;
      bs = bound[a];
    }
    uint32_t incl = deg;  // inclusive scan of the degrees over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, 64);
      if ((int)lane >= off) incl += y;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    if (lane == 0) {
      cnt[0] += total;
      cnt[1] += (uint32_t)__popcll(__ballot(a >= 0));
    }
    wave_sync_lds();
    s_off[lane] = incl - deg;
    s_first[lane] = first;
    s_src[lane] = (uint32_t)a;
    s_bs[lane] = bs;
    wave_sync_lds();
    for (uint32_t base = 0; base < total; base += 64 * SRC_UNROLL) {  // wave-uniform
      uint32_t slot[SRC_UNROLL], bsj[SRC_UNROLL];
      int32_t srcj[SRC_UNROLL], tgt[SRC_UNROLL];
      uint32_t ed[SRC_UNROLL], bt[SRC_UNROLL];
#pragma unroll
      for (int j = 0; j < SRC_UNROLL; ++j) {
        const uint32_t sidx = base + 64 * j + lane;
        uint32_t lo = 0;  // the last lane with s_off <= sidx (zero-degree lanes share offsets)
#pragma unroll
        for (uint32_t step = 32; step; step >>= 1)
          if (s_off[lo + step] <= sidx) lo += step;
        slot[j] = sidx < total ? s_first[lo] + (sidx - s_off[lo]) : 0xFFFFFFFFu;
        srcj[j] = (int32_t)s_src[lo];
        bsj[j] = s_bs[lo];
      }
#pragma unroll
      for (int j = 0; j < SRC_UNROLL; ++j) {
        tgt[j] = 0;
        ed[j] = 0;
        if (slot[j] != 0xFFFFFFFFu) {
          tgt[j] = target[slot[j]];
          ed[j] = edge_dir[slot[j]];
        }
      }
#pragma unroll
      for (int j = 0; j < SRC_UNROLL; ++j) bt[j] = slot[j] != 0xFFFFFFFFu ? bound[tgt[j]] : 0u;
#pragma unroll
      for (int j = 0; j < SRC_UNROLL; ++j) {
        if (base + 64 * j >= total) break;  // wave-uniform
        const int64_t e = (int64_t)(ed[j] >> 1);
        const uint32_t dir = ed[j] & 1u;
        const uint32_t vq = bsj[j] >> 13;
        bool pass = false;
        double u = 1.0;
        if (slot[j] != 0xFFFFFFFFu && vq < (uint32_t)S.n_variants && bt[j] != 0u) {
          const uint32_t pc = candidate_pcode(L, Q, bsj[j], bt[j], vq);
          if (pc) {
            u = edge_uniform(S, L, t, vq, (int)dir, e);
            pass = u < (double)pcode_decode(pc);
          }
        }
        const uint64_t m = __ballot(pass);
        if (m) queue_exact(S, L, t, Q, m, pass, srcj[j], tgt[j], e, dir, vq, u, n_x);
      }
    }
  }
}

// ===========================================================================
// People.infect for one target (people.py:470-579 + immunity.py:138-202)
// ===========================================================================
struct InfectOut {
  uint64_t key;  // the target's winning key (KEY_NONE: no infection)
  bool inf, reinf, symp, sev;
  uint32_t v, l;
};

// g: storage index; il: index within the sim (counter field of the draws); stride: row stride of
// the [variant][agent] arrays. Resets the target's winner key. The caller resolves the source.
// Every per-agent input is loaded before the first store (the stores may alias the inputs for the
// compiler, which would otherwise serialise one memory round trip per field).
__device__ inline InfectOut infect_target(const cvk_people& P, const cvk_params& prm, uint64_t* winner, int64_t g,
                                          uint32_t il, int64_t stride, int32_t t, bool hosp_max, bool icu_max) {
  InfectOut o = {KEY_NONE, false, false, false, false, 0u, 0u};
  const uint64_t key = winner[g];
  const uint32_t f0 = P.flags[g];
  o.key = key;
  if (key == KEY_NONE || !(f0 & CVK_SUSCEPTIBLE)) {  // people.py:470 keeps susceptibles only
    winner[g] = KEY_NONE;
    return o;
  }
  const uint32_t v = key_variant(key);
  // inputs (one round trip)
  const uint32_t vs0 = P.vstate[g];
  const float peak0 = P.peak_nab[g];
  const int32_t nb0 = P.n_breakthroughs[g];
  const float rt0 = P.rel_trans[g];
  const int32_t ninf0 = P.n_infections[g];
  const float drec0 = P.date_recovered[g];
  const float symp_p = P.symp_prob[g];
  const float sev_p = P.severe_prob[g];
  const float crit_p = P.crit_prob[g];
  const float death_p = P.death_prob[g];
  const float symp_imm = P.symp_imm[(int64_t)v * stride + g];
  const float sev_imm = P.sev_imm[(int64_t)v * stride + g];
  const float nab0 = prm.use_waning ? P.nab[g] : 0.0f;
  winner[g] = KEY_NONE;
  const uint64_t seed = prm.seed;
  o.v = v;
  o.l = key_layer(key);
  o.inf = true;
  // breakthrough, people.py:486-491
  if (peak0 != 0.0f) {
    if (nb0 == 0) P.rel_trans[g] = rt0 * prm.trans_redux;
    P.n_breakthroughs[g] = nb0 + 1;
  }
  // states, people.py:494-506
  P.flags[g] = (f0 & ~(CVK_SUSCEPTIBLE | CVK_NAIVE | CVK_RECOVERED | CVK_DIAGNOSED)) | CVK_EXPOSED;
  P.vstate[g] = (vs0 & ~0xFu) | v | (1u << (8 + v));
  P.n_infections[g] = ninf0 + 1;
  o.reinf = !isnan(drec0);
  // outcome tree, people.py:513-579
  const float d_e2i = (float)draw_dur(prm.dur[CVK_DUR_EXP2INF], seed, il, t, SLOT_EXP2INF);
  P.dur_exp2inf[g] = d_e2i;
  P.date_exposed[g] = (float)t;
  const float date_inf = d_e2i + (float)t;
  P.date_infectious[g] = date_inf;
  const float fnan = __builtin_nanf("");
  P.date_symptomatic[g] = fnan;
  P.date_severe[g] = fnan;
  P.date_critical[g] = fnan;
  P.date_diagnosed[g] = fnan;
  float date_rec = fnan;
  double u_symp, u_sev, u_crit, u_death;
  draw_pair(seed, il, t, SLOT_BERN_SYMP_SEV, u_symp, u_sev);
  const float sp = (prm.rel_symp_prob[v] * symp_p) * (1.0f - symp_imm);
  o.symp = u_symp < (double)sp;
  int imm_class = 0;  // 0 asymptomatic, 1 mild, 2 severe (rel_imm_symp)
  if (!o.symp) {
    const double dd = draw_dur(prm.dur[CVK_DUR_ASYM2REC], seed, il, t, SLOT_DUR_A);
    date_rec = (float)((double)date_inf + dd);
    P.dur_disease[g] = (float)((double)d_e2i + dd);
  } else {
    const float d_i2s = (float)draw_dur(prm.dur[CVK_DUR_INF2SYM], seed, il, t, SLOT_DUR_A);
    P.dur_inf2sym[g] = d_i2s;
    const float date_sym = date_inf + d_i2s;
    P.date_symptomatic[g] = date_sym;
    const float sevp = (prm.rel_severe_prob[v] * sev_p) * (1.0f - sev_imm);
    o.sev = u_sev < (double)sevp;
    if (!o.sev) {
      imm_class = 1;
      const double dd = draw_dur(prm.dur[CVK_DUR_MILD2REC], seed, il, t, SLOT_DUR_B);
      date_rec = (float)((double)date_sym + dd);
      P.dur_disease[g] = (float)((double)(d_e2i + d_i2s) + dd);
    } else {
      imm_class = 2;
      const float d_s2s = (float)draw_dur(prm.dur[CVK_DUR_SYM2SEV], seed, il, t, SLOT_DUR_B);
      P.dur_sym2sev[g] = d_s2s;
      const float date_sev = date_sym + d_s2s;
      P.date_severe[g] = date_sev;
      draw_pair(seed, il, t, SLOT_BERN_CRIT_DEATH, u_crit, u_death);
      const float critp = (prm.rel_crit_prob[v] * crit_p) * (hosp_max ? prm.no_hosp_factor : 1.0f);
      if (!(u_crit < (double)critp)) {
        const double dd = draw_dur(prm.dur[CVK_DUR_SEV2REC], seed, il, t, SLOT_DUR_C);
        date_rec = (float)((double)date_sev + dd);
        P.dur_disease[g] = (float)((double)((d_e2i + d_i2s) + d_s2s) + dd);
      } else {
        const float d_s2c = (float)draw_dur(prm.dur[CVK_DUR_SEV2CRIT], seed, il, t, SLOT_DUR_C);
        P.dur_sev2crit[g] = d_s2c;
        const float date_crit = date_sev + d_s2c;
        P.date_critical[g] = date_crit;
        const float deathp = (prm.rel_death_prob[v] * death_p) * (icu_max ? prm.no_icu_factor : 1.0f);
        const float dsum = ((d_e2i + d_i2s) + d_s2s) + d_s2c;
        if (!(u_death < (double)deathp)) {
          const double dd = draw_dur(prm.dur[CVK_DUR_CRIT2REC], seed, il, t, SLOT_DUR_D);
          date_rec = (float)((double)date_crit + dd);
          P.dur_disease[g] = (float)((double)dsum + dd);
        } else {
          const double dd = draw_dur(prm.dur[CVK_DUR_CRIT2DIE], seed, il, t, SLOT_DUR_D);
          P.date_dead[g] = (float)((double)date_crit + dd);
          P.dur_disease[g] = (float)((double)dsum + dd);
        }
      }
    }
  }
  P.date_recovered[g] = date_rec;
  // update_peak_nab, immunity.py:138-202
  if (prm.use_waning) {
    if (nab0 > 0.0f) {
      P.peak_nab[g] = peak0 * prm.nab_boost;
    } else {
      const double init_nab = draw_dur(prm.nab_init, seed, il, t, SLOT_NAB);
      P.peak_nab[g] = (float)((pow(2.0, init_nab) * prm.rel_imm_symp[imm_class]) * prm.nab_norm);
    }
    P.t_nab_event[g] = t;
  }
  return o;
}

// Source of an infection key (-1 for seeds and imports); p1/p2 are the sim's arrays of the key's layer.
// gen: the key's layer is a generated dynamic layer of day t (endpoints from the resample stream).
__device__ inline int32_t key_source(uint64_t key, const int32_t* p1, const int32_t* p2, bool gen = false,
                                     uint64_t seed = 0, int32_t t = 0, uint32_t n_agents = 0) {
  if (key_class(key) != 1) return -1;
  const int64_t e = (int64_t)key_index(key);
  if (gen) {
    const EdgeEnds d = dynamic_endpoints(seed, t, key_layer(key), n_agents, e);
    return key_dir(key) ? d.p2 : d.p1;
  }
  return key_dir(key) ? p2[e] : p1[e];
}

constexpr int INFECT_BLOCK = 256;
constexpr int INFECT_WAVE_LIST = 64 * 32;  // every bit of one wave's 64 words

// Compacts the set bits of this lane's touched word w (cleared in memory) into the wave's LDS list
// `list` (storage indices, ascending lane order); returns the wave's total. All lanes call it.
__device__ inline uint32_t compact_touched(uint32_t* touched, int64_t w, int64_t n_words, int64_t n, int32_t* list) {
  uint32_t bits = w < n_words ? touched[w] : 0u;
  if (bits) touched[w] = 0u;
  const uint32_t cnt = (uint32_t)__popc(bits);
  uint32_t incl = cnt;  // inclusive scan over the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if ((int)lane_id() >= off) incl += y;
  }
  uint32_t pos = incl - cnt;
  while (bits) {
    const int32_t b = __ffs(bits) - 1;
    bits &= bits - 1;
    const int64_t g = w * 32 + b;
    list[pos++] = g < n ? (int32_t)g : -1;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return total;
}
constexpr int N_INFECT_COLS = 2 + 3 * MAXV;

// Column of the day row for infect-pass counter k.
__device__ inline int infect_col(int k) {
  if (k == 0) return CVK_R_NEW_INFECTIONS;
  if (k == 1) return CVK_R_NEW_REINFECTIONS;
  return CVK_R_BYVAR + ((k - 2) / MAXV) * MAXV + (k - 2) % MAXV;  // infections, symptomatic, severe by variant
}

}  // namespace cvk
