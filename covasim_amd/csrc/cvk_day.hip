// Production day loop of the covasim_amd library (gfx950, wave64).
//
// One simulated day (covasim/sim.py:558-685) is three launches on one HIP stream
// (plus a radix-select choose on variant-import days):
//   day_begin    previous day's stocks + update_nab (sim.py:651-674), then this day's
//                update_states_pre transitions + check_immunity + viral-load phase bit
//                (people.py:164-312, immunity.py:303-350, utils.py:58-78), the two agent
//                bitmaps and edge records; extra blocks resample dynamic layers (base.py:1849-1876)
//   edge_pass    all layers x variants x both directions in one pass; infectious bitmap in LDS,
//                susceptible bitmap from L2, wave-local candidate queues, per-target atomicMin of
//                the (variant, layer, direction, edge) key (sim.py:611-649, people.py:465)
//   infect_pass  outcome tree, durations, NAb peak, infection log (people.py:435-586,
//                immunity.py:138-202)
// A count-only finish pass closes the last day of a run. Per-block counts are combined by the
// last block to finish (ticket in cvk_work.ctl), so no same-address atomics pile up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>

#include "cvk_device.h"
#include "cvk_host.h"

namespace cvk {

char* last_error_buffer() {
  static thread_local char buf[512] = {0};
  return buf;
}

constexpr uint32_t STATE_CLEAR_ON_RECOVERY = CVK_EXPOSED | CVK_INFECTIOUS | CVK_SYMPTOMATIC | CVK_SEVERE | CVK_CRITICAL;
constexpr uint32_t STATE_CLEAR_ON_DEATH = CVK_SUSCEPTIBLE | CVK_EXPOSED | CVK_INFECTIOUS | CVK_SYMPTOMATIC |
                                          CVK_SEVERE | CVK_CRITICAL | CVK_KNOWN_CONTACT | CVK_QUARANTINED |
                                          CVK_RECOVERED;
constexpr uint32_t INFVAR_MASK = 0xFu << CVK_INFVAR_SHIFT;
constexpr int MAXV = CVK_MAX_VARIANTS;
constexpr int AGENT_BLOCK = 256;
constexpr int AGENT_WAVES_MAX = 16;

__device__ inline uint32_t* day_row(const cvk_work& w, int32_t t) { return w.results + (int64_t)t * CVK_R_COLS; }
__device__ inline double* day_row_f(const cvk_work& w, int32_t t) { return w.results_f + (int64_t)t * CVK_RF_COLS; }

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }

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
// Reductions without device-scope fences: every block of a reducing kernel stores its
// partial row (no fence, no ticket); the rows are summed after the kernel boundary by
// finalize_partials (block 0 of the next kernel, or a one-block reduce kernel). Integer
// columns are order-free; float columns are summed in a fixed order (deterministic).
// ---------------------------------------------------------------------------
template <int NC, int ND>
__device__ void write_partials(const uint32_t (&c)[NC], const double (&d)[ND], uint32_t* s_red, double* s_redf,
                               uint32_t* partials, double* partials_f) {
  for (int k = threadIdx.x; k < NC; k += blockDim.x) s_red[k] = 0;
  for (int k = threadIdx.x; k < ND; k += blockDim.x) s_redf[k] = 0.0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const uint32_t v = wave_sum(c[k]);
    if (lane_id() == 0 && v) atomicAdd(&s_red[k], v);
  }
  // float sums: per-wave totals in LDS, combined by thread 0 in wave order
  __shared__ double s_wave[AGENT_WAVES_MAX][3];
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const double v = wave_sum_d(d[k]);
    if (lane_id() == 0) s_wave[threadIdx.x >> 6][k] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < ND; ++k) {
      double acc = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) acc += s_wave[w][k];
      s_redf[k] = acc;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NC; k += blockDim.x) partials[(size_t)k * CVK_MAX_REDUCE_BLOCKS + blockIdx.x] = s_red[k];
  for (int k = threadIdx.x; k < ND; k += blockDim.x) partials_f[(size_t)k * CVK_MAX_REDUCE_BLOCKS + blockIdx.x] = s_redf[k];
}

// Sums the n_blocks partial values of every column (column-major: one wave per column,
// coalesced loads); totals land in s_tot[NC] / s_totf[ND]. Float columns are summed in a
// fixed lane/shuffle order, so results are deterministic.
template <int NC, int ND>
__device__ void finalize_partials(const uint32_t* partials, const double* partials_f, uint32_t n_blocks,
                                  uint32_t* s_tot, double* s_totf) {
  const int n_waves = (int)(blockDim.x >> 6);
  const int w = threadIdx.x >> 6;
  for (int k = w; k < NC + ND; k += n_waves) {
    // 8 independent loads in flight per lane per step (the loop is latency-bound otherwise)
    if (k < NC) {
      const uint32_t* col = partials + (size_t)k * CVK_MAX_REDUCE_BLOCKS;
      uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t b0 = 0; b0 < n_blocks; b0 += 512) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const uint32_t b = b0 + j * 64 + lane_id();
          acc[j] += b < n_blocks ? col[b] : 0u;
        }
      }
      uint32_t tot = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
      tot = wave_sum(tot);
      if (lane_id() == 0) s_tot[k] = tot;
    } else {
      const double* col = partials_f + (size_t)(k - NC) * CVK_MAX_REDUCE_BLOCKS;
      double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t b0 = 0; b0 < n_blocks; b0 += 512) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const uint32_t b = b0 + j * 64 + lane_id();
          acc[j] += b < n_blocks ? col[b] : 0.0;
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
// day_begin: [previous-day stocks + update_nab] + transitions + bitmaps + records,
// with extra blocks resampling the dynamic layers.
// ===========================================================================
// stock columns (0..29) and day columns (30..46) of the partial rows
constexpr int SC_ALIVE = 13, SC_EXPV = 14, SC_INFV = 14 + MAXV, N_STOCK_COLS = 14 + 2 * MAXV;
constexpr int DC_BASE = N_STOCK_COLS, N_DAY_COLS = 9 + MAXV, N_AGENT_COLS = N_STOCK_COLS + N_DAY_COLS;
static_assert(N_AGENT_COLS <= CVK_PARTIAL_COLS, "partial row too small");

struct ResampleLayer {
  int32_t* p1;
  int32_t* p2;
  float* beta;
  int64_t n;
  uint32_t layer;
  uint32_t pad;
};

struct DayBeginArgs {
  int32_t t;
  int32_t prologue;     // count stocks and update NAbs of day t-1 first
  int32_t transitions;  // 0: count-only finish pass for day t
  int32_t n_resample;
  uint32_t agent_blocks;
  uint32_t pad;
  ResampleLayer rs[CVK_MAX_LAYERS];
};

__device__ inline void stock_counts(uint32_t f, uint32_t vs, int nv, uint32_t (&c)[N_AGENT_COLS]) {
  const uint32_t stock_bits[13] = {CVK_SUSCEPTIBLE, CVK_EXPOSED, CVK_INFECTIOUS, CVK_SYMPTOMATIC, CVK_SEVERE,
                                   CVK_CRITICAL, CVK_RECOVERED, CVK_DEAD, CVK_DIAGNOSED, CVK_KNOWN_DEAD,
                                   CVK_QUARANTINED, CVK_ISOLATED, CVK_VACCINATED};
#pragma unroll
  for (int k = 0; k < 13; ++k) c[k] += (f & stock_bits[k]) ? 1u : 0u;
  c[SC_ALIVE] += (f & CVK_DEAD) ? 0u : 1u;
#pragma unroll
  for (int v = 0; v < MAXV; ++v) {
    if (v < nv) {
      c[SC_EXPV + v] += (vs >> (8 + v)) & 1u;
      c[SC_INFV + v] += (vs >> (16 + v)) & 1u;
    }
  }
}

// Per-agent inputs of day_begin, all loaded before any is used (one memory round trip).
struct AgentIn {
  uint32_t f, vs;
  float d_inf, d_sym, d_sev, d_crit, d_rec, d_dead, d_iso, rt, rs, peak, nab;
  int32_t tne;
};

__device__ inline void load_agent(const cvk_people& P, int64_t i, bool waning, AgentIn& x) {
  x.f = P.flags[i];
  x.vs = P.vstate[i];
  x.d_inf = P.date_infectious[i];
  x.d_sym = P.date_symptomatic[i];
  x.d_sev = P.date_severe[i];
  x.d_crit = P.date_critical[i];
  x.d_rec = P.date_recovered[i];
  x.d_dead = P.date_dead[i];
  x.d_iso = P.date_end_isolation[i];
  x.rt = P.rel_trans[i];
  x.rs = P.rel_sus[i];
  if (waning) {
    x.peak = P.peak_nab[i];
    x.nab = P.nab[i];
    x.tne = P.t_nab_event[i];
  } else {
    x.peak = 0.0f;
    x.nab = 0.0f;
    x.tne = 0;
  }
}

constexpr int AGENTS_PER_THREAD = 2;

__global__ __launch_bounds__(AGENT_BLOCK) void day_begin_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                DayBeginArgs a) {
  if (blockIdx.x >= a.agent_blocks) {
    // ---- dynamic-layer resample: iid uniform endpoints per edge, beta = 1 (base.py:1849-1876)
    const int64_t stride = (int64_t)(gridDim.x - a.agent_blocks) * AGENT_BLOCK;
    const int64_t start = (int64_t)(blockIdx.x - a.agent_blocks) * AGENT_BLOCK + threadIdx.x;
    for (int r = 0; r < a.n_resample; ++r) {
      const ResampleLayer& R = a.rs[r];
      for (int64_t e = start; e < R.n; e += stride) {
        const uint64_t ue = (uint64_t)e;
        const u32x4 w = philox_seeded(prm.seed, (uint32_t)ue, (uint32_t)a.t, STREAM_RESAMPLE | R.layer,
                                      (uint32_t)(ue >> 32));
        R.p1[e] = (int32_t)below(w.x, w.y, (uint32_t)P.n);
        R.p2[e] = (int32_t)below(w.z, w.w, (uint32_t)P.n);
        R.beta[e] = 1.0f;
      }
    }
    return;
  }

  // per-wave counters: sums of per-agent booleans via ballot (wave-uniform, no shuffles)
  uint32_t c[N_AGENT_COLS];
#pragma unroll
  for (int k = 0; k < N_AGENT_COLS; ++k) c[k] = 0;
  double d[3] = {0.0, 0.0, 0.0};
  const int nv = prm.n_variants;
  const bool waning = prm.use_waning != 0;
  const int32_t t = a.t;
  const float tf = (float)t;
  const int64_t n_waves_bits = (P.n + 63) / 64;
  const bool closing = a.prologue || !a.transitions;  // stocks + update_nab of the previous (or this) day
  const int32_t tc = a.transitions ? t - 1 : t;

  int64_t idx[AGENTS_PER_THREAD];
  AgentIn in[AGENTS_PER_THREAD];
#pragma unroll
  for (int r = 0; r < AGENTS_PER_THREAD; ++r) {
    idx[r] = ((int64_t)blockIdx.x * AGENTS_PER_THREAD + r) * AGENT_BLOCK + threadIdx.x;
    if (idx[r] < P.n) load_agent(P, idx[r], waning, in[r]);
  }
#pragma unroll
  for (int r = 0; r < AGENTS_PER_THREAD; ++r) {
    const int64_t i = idx[r];
    const bool valid = i < P.n;
    bool inf_bit = false, sus_bit = false;
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
          P.nab[i] = nab;
          for (int v = 0; v < nv; ++v) {  // population means of sus/symp_imm (sim.py:673-674)
            d[1] += (double)P.sus_imm[(int64_t)v * P.n + i];
            d[2] += (double)P.symp_imm[(int64_t)v * P.n + i];
          }
          in[r].nab = nab;
        }
        if (!(f & CVK_DEAD)) d[0] += (double)nab;
      }
      if (a.transitions) {
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
          for (int v = 0; v < nv; ++v) {
            const double imm = (was_inf && rv < MAXV) ? (double)prm.immunity[v][rv] : 0.0;
            const double eff = nab * imm;
            const int64_t off = (int64_t)v * P.n + i;
            P.sus_imm[off] = calc_ve(eff, prm.ve_exp_alpha[0], prm.ve_beta[0]);
            P.symp_imm[off] = calc_ve(eff, prm.ve_exp_alpha[1], prm.ve_beta[1]);
            P.sev_imm[off] = calc_ve(eff, prm.ve_exp_alpha[2], prm.ve_beta[2]);
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
        if (f != f0) P.flags[i] = f;
        if (vs != vs0) P.vstate[i] = vs;
        inf_bit = (f & CVK_INFECTIOUS) && x.rt != 0.0f;
        sus_bit = (f & CVK_SUSCEPTIBLE) && x.rs != 0.0f;
        const float val = (f & CVK_INFECTIOUS) ? x.rt : x.rs;
        W.rec[i] = ((uint64_t)f << 32) | (uint64_t)__float_as_uint(val);
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
    if (a.transitions) {
#pragma unroll
      for (int k = 0; k < 9; ++k) c[DC_BASE + k] += (uint32_t)__popcll(__ballot((ev >> k) & 1u));
#pragma unroll
      for (int v = 0; v < MAXV; ++v)
        if (v < nv) c[DC_BASE + 9 + v] += (uint32_t)__popcll(__ballot(ev_var == (uint32_t)v));
      // one 64-bit ballot per wave covers 64 consecutive agents
      const uint64_t ib = __ballot(inf_bit);
      const uint64_t sb = __ballot(sus_bit);
      const int64_t wave = i >> 6;
      if (lane_id() == 0 && wave < n_waves_bits) {
        ((uint64_t*)W.bitmap)[wave] = ib;
        ((uint64_t*)W.susmap)[wave] = sb;
      }
    }
  }

  if (blockIdx.x == 0 && threadIdx.x == 0 && a.transitions) {
    W.ctl[CVK_CTL_TOUCHED] = 0;       // the new day's target list
    W.ctl[CVK_CTL_EDGE_TILES] = 0;    // the edge pass's tile counter
  }
  // block partial row: per-wave counts through LDS (no atomics), float sums in wave order
  __shared__ uint32_t s_wc[AGENT_BLOCK / 64][N_AGENT_COLS];
  __shared__ double s_wd[AGENT_BLOCK / 64][3];
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < N_AGENT_COLS; ++k) s_wc[w][k] = c[k];
  }
  if (waning) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double v = wave_sum_d(d[k]);
      if (lane_id() == 0) s_wd[w][k] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < N_AGENT_COLS) {
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < AGENT_BLOCK / 64; ++j) tot += s_wc[j][threadIdx.x];
    W.partials[(size_t)threadIdx.x * CVK_MAX_REDUCE_BLOCKS + blockIdx.x] = tot;
  } else if (threadIdx.x >= 64 && threadIdx.x < 67) {
    const int k = threadIdx.x - 64;
    double tot = 0.0;
    if (waning)
      for (int j = 0; j < AGENT_BLOCK / 64; ++j) tot += s_wd[j][k];
    W.partials_f[(size_t)k * CVK_MAX_REDUCE_BLOCKS + blockIdx.x] = tot;
  }
}

// Writes the day rows from the summed partial rows of day_begin (block-wide).
struct DayReduce {
  uint32_t n_blocks;  // 0: nothing to reduce
  int32_t t;
  int32_t prologue;
  int32_t transitions;
};

__device__ void day_reduce(const cvk_work& W, const DayReduce& r) {
  __shared__ uint32_t s_tot[N_AGENT_COLS];
  __shared__ double s_totf[3];
  finalize_partials<N_AGENT_COLS, 3>(W.partials, W.partials_f, r.n_blocks, s_tot, s_totf);
  if (r.prologue || !r.transitions) {
    const int32_t tc = r.transitions ? r.t - 1 : r.t;
    uint32_t* row = day_row(W, tc);
    for (int k = threadIdx.x; k < N_STOCK_COLS; k += blockDim.x) {
      int col;
      if (k < 13) col = CVK_R_N_SUSCEPTIBLE + k;
      else if (k == SC_ALIVE) col = CVK_R_N_ALIVE_NAB;
      else if (k < SC_INFV) col = CVK_R_BYVAR + CVK_RV_N_EXPOSED * MAXV + (k - SC_EXPV);
      else col = CVK_R_BYVAR + CVK_RV_N_INFECTIOUS * MAXV + (k - SC_INFV);
      row[col] = s_tot[k];
    }
    if (threadIdx.x < 3) day_row_f(W, tc)[threadIdx.x] = s_totf[threadIdx.x];
  }
  if (r.transitions) {
    uint32_t* row = day_row(W, r.t);
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

__global__ __launch_bounds__(1024) void day_reduce_kernel(cvk_work W, DayReduce r) { day_reduce(W, r); }

// ===========================================================================
// choose: the k smallest 64-bit selection keys (philox word << 32 | agent) among
// eligible agents, by an 8-pass byte-wise radix select; a uniform k-subset.
// ===========================================================================
__device__ inline uint64_t selection_key(uint64_t seed, int32_t t, uint32_t stream_id, uint32_t i) {
  const u32x4 w = philox_seeded(seed, i, (uint32_t)t, stream_id, 0u);
  return ((uint64_t)w.x << 32) | i;
}

__global__ void choose_init_kernel(cvk_work W, uint32_t k) {
  const int j = threadIdx.x;
  W.select_hist[j] = 0;
  if (j == 0) {
    W.ctl[CVK_CTL_SEL_PREFIX] = 0;
    W.ctl[CVK_CTL_SEL_PREFIX + 1] = 0;
    W.ctl[CVK_CTL_SEL_REMAIN] = k;
    W.ctl[CVK_CTL_SEL_ALL] = 0;
  }
}

__global__ __launch_bounds__(256) void choose_hist_kernel(cvk_people P, cvk_work W, uint64_t seed, int32_t t,
                                                          uint32_t stream_id, int32_t sus_only, int pass) {
  __shared__ uint32_t s_hist[256];
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t prefix = ((uint64_t)W.ctl[CVK_CTL_SEL_PREFIX + 1] << 32) | W.ctl[CVK_CTL_SEL_PREFIX];
  const bool all = W.ctl[CVK_CTL_SEL_ALL] != 0;
  const int shift = 56 - 8 * pass;
  if (!all) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * 256) {
      if (sus_only && !(P.flags[i] & CVK_SUSCEPTIBLE)) continue;
      const uint64_t key = selection_key(seed, t, stream_id, (uint32_t)i);
      if (pass > 0 && (key >> (shift + 8)) != (prefix >> (shift + 8))) continue;
      atomicAdd(&s_hist[(key >> shift) & 0xFF], 1u);
    }
  }
  __syncthreads();
  if (s_hist[threadIdx.x]) atomicAdd(&W.select_hist[threadIdx.x], s_hist[threadIdx.x]);
}

__global__ void choose_select_kernel(cvk_work W, int pass) {
  if (threadIdx.x != 0) return;
  if (W.ctl[CVK_CTL_SEL_ALL]) return;
  uint32_t remain = W.ctl[CVK_CTL_SEL_REMAIN];
  if (pass == 0) {
    uint64_t total = 0;
    for (int dd = 0; dd < 256; ++dd) total += W.select_hist[dd];
    if (remain >= total) {  // everyone eligible is chosen
      W.ctl[CVK_CTL_SEL_ALL] = 1;
      if (remain > total) W.ctl[CVK_CTL_SEL_SHORT] = 1;
      for (int dd = 0; dd < 256; ++dd) W.select_hist[dd] = 0;
      return;
    }
  }
  uint64_t prefix = ((uint64_t)W.ctl[CVK_CTL_SEL_PREFIX + 1] << 32) | W.ctl[CVK_CTL_SEL_PREFIX];
  const int shift = 56 - 8 * pass;
  uint32_t cum = 0;
  int digit = 255;
  for (int dd = 0; dd < 256; ++dd) {
    const uint32_t h = W.select_hist[dd];
    if (cum + h >= remain) { digit = dd; break; }
    cum += h;
  }
  remain -= cum;
  prefix |= (uint64_t)digit << shift;
  W.ctl[CVK_CTL_SEL_PREFIX] = (uint32_t)prefix;
  W.ctl[CVK_CTL_SEL_PREFIX + 1] = (uint32_t)(prefix >> 32);
  W.ctl[CVK_CTL_SEL_REMAIN] = remain;
  for (int dd = 0; dd < 256; ++dd) W.select_hist[dd] = 0;
}

__global__ __launch_bounds__(256) void choose_mark_kernel(cvk_people P, cvk_work W, uint64_t seed, int32_t t,
                                                          uint32_t stream_id, int32_t sus_only, uint32_t layer_code,
                                                          uint32_t variant) {
  const uint64_t thr = W.ctl[CVK_CTL_SEL_ALL] ? ~0ull
                                              : ((uint64_t)W.ctl[CVK_CTL_SEL_PREFIX + 1] << 32) | W.ctl[CVK_CTL_SEL_PREFIX];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * 256) {
    if (sus_only && !(P.flags[i] & CVK_SUSCEPTIBLE)) continue;
    const uint64_t key = selection_key(seed, t, stream_id, (uint32_t)i);
    if (key > thr) continue;
    const uint64_t wk = make_key(0, variant, layer_code, 0, (uint64_t)i);
    const uint64_t old = atomicMin((unsigned long long*)&W.winner[i], (unsigned long long)wk);
    if (old == KEY_NONE) W.touched[atomicAdd(&W.ctl[CVK_CTL_TOUCHED], 1u)] = (int32_t)i;
  }
}

// ===========================================================================
// edge_pass
// ===========================================================================
struct EdgeLayer {
  const int32_t* p1;
  const int32_t* p2;
  const float* beta;
  const double* unif;  // [n_variants][2][n] or null
  int64_t n;
  float beta_layer, iso_factor, quar_factor;
  int32_t beta_ones;   // every per-edge beta is 1.0f: no gather needed
};

struct EdgeArgs {
  EdgeLayer layer[CVK_MAX_LAYERS];
  int64_t tile_start[CVK_MAX_LAYERS + 1];
  int32_t n_layers, n_variants;
  float beta_v[MAXV];
  float asymp_factor, vl_hi, vl_lo;
  int32_t use_sus_imm;
  uint64_t seed;
  int32_t t;
  int32_t bitmap_words;
  const uint64_t* rec;
  const float* sus_imm;
  int64_t n_agents;
  const uint32_t* bitmap;
  const uint32_t* susmap;
  uint64_t* winner;
  int32_t* touched;
  uint32_t* touched_count;
  uint32_t* tile_counter;  // zeroed by the agent pass of the day
  cvk_work work;        // for the fused day reduction
  DayReduce reduce;     // day_begin partial rows to reduce in block 0 (n_blocks == 0: none)
};

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int EDGE_BLOCK = 1024;
constexpr int EDGE_CHUNKS = 2;                        // 4-edge chunks per thread per tile
constexpr int EDGE_TILE = EDGE_BLOCK * 4 * EDGE_CHUNKS;
constexpr int EDGE_QCAP = 64 * 8;                     // per-wave queue: 4 edges x 2 directions x 64 lanes

// A candidate slot (infectious source, susceptible target) of edge e of layer l: the float32
// products of compute_trans_sus (utils.py:88-89) and compute_infections (utils.py:117), the
// float64 uniform test (utils.py:123), and the first-occurrence dedup as an atomicMin.
__device__ inline void edge_candidate(const EdgeArgs& a, const EdgeLayer& L, int l, int dir, int64_t e, int32_t src,
                                      int32_t tgt) {
  const uint64_t rsrc = a.rec[src];
  const uint64_t rtgt = a.rec[tgt];
  const uint32_t fs = (uint32_t)(rsrc >> 32);
  const uint32_t ft = (uint32_t)(rtgt >> 32);
  const uint32_t v = inf_var(fs);
  if (!(fs & CVK_INFECTIOUS) || v >= (uint32_t)a.n_variants || !(ft & CVK_SUSCEPTIBLE)) return;
  float rt = __uint_as_float((uint32_t)rsrc);                   // rel_trans * inf (== 1)
  if (fs & CVK_QUARANTINED) rt = rt * L.quar_factor;             // * f_quar
  if (!(fs & CVK_SYMPTOMATIC)) rt = rt * a.asymp_factor;         // * f_asymp
  if (fs & CVK_ISOLATED) rt = rt * L.iso_factor;                 // * f_iso
  rt = rt * L.beta_layer;                                        // * beta_layer
  rt = rt * ((fs & CVK_EARLY) ? a.vl_hi : a.vl_lo);              // * viral_load
  if (rt == 0.0f) return;                                        // utils.py:116
  float rs = __uint_as_float((uint32_t)rtgt);                   // rel_sus * sus (== 1)
  if (ft & CVK_QUARANTINED) rs = rs * L.quar_factor;
  if (a.use_sus_imm) rs = rs * (1.0f - a.sus_imm[(int64_t)v * a.n_agents + tgt]);
  const float eb = L.beta_ones ? 1.0f : L.beta[e];
  const float p = ((a.beta_v[v] * eb) * rt) * rs;                // utils.py:117
  if (p == 0.0f) return;                                         // utils.py:118
  double u;
  if (L.unif) {
    u = L.unif[((int64_t)v * 2 + dir) * L.n + e];
  } else {
    const uint64_t ue = (uint64_t)e;
    const u32x4 w = philox_seeded(a.seed, (uint32_t)ue, (uint32_t)a.t, STREAM_TX | (v << 8) | (uint32_t)l,
                                  (uint32_t)(ue >> 32));
    u = dir ? u01(w.z, w.w) : u01(w.x, w.y);
  }
  if (!(u < (double)p)) return;                                  // utils.py:123
  const uint64_t key = make_key(1, v, (uint32_t)l, (uint32_t)dir, (uint64_t)e);
  const uint64_t old = atomicMin((unsigned long long*)&a.winner[tgt], (unsigned long long)key);
  if (old == KEY_NONE) a.touched[atomicAdd(a.touched_count, 1u)] = tgt;
}

template <bool LDS>
__device__ inline bool inf_bit(const EdgeArgs& a, const uint32_t* s_bits, uint32_t i) {
  const uint32_t w = LDS ? s_bits[i >> 5] : a.bitmap[i >> 5];
  return (w >> (i & 31)) & 1u;
}

__device__ inline bool sus_bit(const EdgeArgs& a, uint32_t i) { return (a.susmap[i >> 5] >> (i & 31)) & 1u; }

static_assert(EDGE_CHUNKS == 2, "EdgeTile holds two 4-edge chunks per thread");

// One tile's endpoints for this thread, in vector registers (two 4-edge chunks).
struct EdgeTile {
  int64_t tile;
  int64_t e0;       // first edge of the tile in its layer
  int32_t l;
  int32_t n_valid;  // edges of the tile that exist (the last tile of a layer is partial)
  v4i x0, y0, x1, y1;
};

__device__ inline v4i load4(const int32_t* p, int64_t e, int64_t n) {
  if (e + 3 < n) return __builtin_nontemporal_load((const v4i*)(p + e));
  v4i r = {0, 0, 0, 0};
  if (e < n) r[0] = p[e];
  if (e + 1 < n) r[1] = p[e + 1];
  if (e + 2 < n) r[2] = p[e + 2];
  return r;
}

// Locates a tile and issues its 16-byte endpoint loads (non-temporal: streamed once).
__device__ inline void load_tile(const EdgeArgs& a, int64_t tile, EdgeTile& T) {
  T.tile = tile;
  if (tile >= a.tile_start[a.n_layers]) return;
  int l = 0;
  while (l + 1 < a.n_layers && tile >= a.tile_start[l + 1]) ++l;
  T.l = l;
  const EdgeLayer& L = a.layer[l];
  T.e0 = (tile - a.tile_start[l]) * EDGE_TILE;
  const int64_t rem = L.n - T.e0;
  T.n_valid = (int32_t)(rem < EDGE_TILE ? rem : EDGE_TILE);
  const int64_t ea = T.e0 + (int64_t)threadIdx.x * 4;
  const int64_t eb = T.e0 + ((int64_t)EDGE_BLOCK + threadIdx.x) * 4;
  T.x0 = load4(L.p1, ea, L.n);
  T.y0 = load4(L.p2, ea, L.n);
  T.x1 = load4(L.p1, eb, L.n);
  T.y1 = load4(L.p2, eb, L.n);
}

template <bool LDS>
__device__ inline void process_chunk(const EdgeArgs& a, const uint32_t* s_bits, uint32_t* s_q, const EdgeLayer& L,
                                     const EdgeTile& T, int c, const v4i xs, const v4i ys) {
  const uint32_t lane = lane_id();
  const uint32_t n_agents = (uint32_t)a.n_agents;
  const uint32_t wave_base = threadIdx.x & ~63u;
  const int32_t local0 = (c * EDGE_BLOCK + (int32_t)threadIdx.x) * 4;  // tile-local index of edge q = 0
  // candidate bits (q*2 + dir): infectious source and susceptible target
  uint32_t cand = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t xi = (uint32_t)xs[q], yi = (uint32_t)ys[q];
    // endpoints outside [0, n) are ignored (the reference leaves them undefined)
    if (local0 + q < T.n_valid && xi < n_agents && yi < n_agents) {
      if (inf_bit<LDS>(a, s_bits, xi) && sus_bit(a, yi)) cand |= 1u << (2 * q);
      if (inf_bit<LDS>(a, s_bits, yi) && sus_bit(a, xi)) cand |= 2u << (2 * q);
    }
  }
  // wave-local compaction of the candidates into the LDS queue (ballot + mbcnt): code = lane << 3 | slot
  uint32_t total = 0;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool has = (cand >> b) & 1u;
    const uint64_t m = __ballot(has);
    if (has) s_q[total + lanes_below(m)] = (lane << 3) | (uint32_t)b;
    total += (uint32_t)__popcll(m);
  }
  if (total == 0) return;  // wave-uniform
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // every lane takes one queued slot per round; endpoints come from the owning lane's registers
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t j = base + lane;
    const uint32_t code = j < total ? s_q[j] : 0u;
    const int src_lane = (int)(code >> 3);
    const int q = (int)((code >> 1) & 3u);
    const int dir = (int)(code & 1u);
    const int x0 = __shfl(xs[0], src_lane, 64), x1 = __shfl(xs[1], src_lane, 64);
    const int x2 = __shfl(xs[2], src_lane, 64), x3 = __shfl(xs[3], src_lane, 64);
    const int y0 = __shfl(ys[0], src_lane, 64), y1 = __shfl(ys[1], src_lane, 64);
    const int y2 = __shfl(ys[2], src_lane, 64), y3 = __shfl(ys[3], src_lane, 64);
    if (j < total) {
      const int32_t xv = q == 0 ? x0 : (q == 1 ? x1 : (q == 2 ? x2 : x3));
      const int32_t yv = q == 0 ? y0 : (q == 1 ? y1 : (q == 2 ? y2 : y3));
      const int64_t e = T.e0 + ((int64_t)c * EDGE_BLOCK + wave_base + (uint32_t)src_lane) * 4 + q;
      edge_candidate(a, L, T.l, dir, e, dir ? yv : xv, dir ? xv : yv);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool LDS>
__device__ inline void process_tile(const EdgeArgs& a, const uint32_t* s_bits, uint32_t* s_q, const EdgeTile& T) {
  const EdgeLayer& L = a.layer[T.l];
  process_chunk<LDS>(a, s_bits, s_q, L, T, 0, T.x0, T.y0);
  process_chunk<LDS>(a, s_bits, s_q, L, T, 1, T.x1, T.y1);
}

template <bool LDS>
__global__ __launch_bounds__(EDGE_BLOCK) void edge_pass_kernel(EdgeArgs a) {
  extern __shared__ uint32_t s_mem[];
  __shared__ int64_t s_tiles[2];
  const int nbw = LDS ? a.bitmap_words : 0;
  uint32_t* s_bits = s_mem;
  uint32_t* s_q = s_mem + nbw + (threadIdx.x >> 6) * EDGE_QCAP;  // this wave's candidate queue
  if (blockIdx.x == 0 && a.reduce.n_blocks) day_reduce(a.work, a.reduce);  // the agent pass's counts
  if (threadIdx.x == 0) s_tiles[0] = atomicAdd(a.tile_counter, 1u);
  if (LDS) {
    const uint4* src = (const uint4*)a.bitmap;
    uint4* dst = (uint4*)s_bits;
    for (int j = threadIdx.x; j < (nbw >> 2); j += EDGE_BLOCK) dst[j] = src[j];
  }
  __syncthreads();
  const int64_t n_tiles = a.tile_start[a.n_layers];
  // tiles are handed out dynamically; the next tile's endpoints are in flight while this one is processed
  EdgeTile cur, nxt;
  load_tile(a, s_tiles[0], cur);
  int par = 1;
  while (cur.tile < n_tiles) {
    if (threadIdx.x == 0) s_tiles[par] = atomicAdd(a.tile_counter, 1u);
    __syncthreads();
    load_tile(a, s_tiles[par], nxt);
    par ^= 1;
    process_tile<LDS>(a, s_bits, s_q, cur);
    cur = nxt;
  }
}

// ===========================================================================
// infect_pass
// ===========================================================================
struct InfectArgs {
  const int32_t* p1[CVK_MAX_LAYERS];
  const int32_t* p2[CVK_MAX_LAYERS];
  int32_t t;
  int32_t count_flows;
  int32_t allow_beds;
  int32_t pad;
};

constexpr int INFECT_BLOCK = 256;
constexpr int N_INFECT_COLS = 2 + 3 * MAXV;

__global__ __launch_bounds__(INFECT_BLOCK) void infect_pass_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                   InfectArgs ia) {
  __shared__ uint32_t s_cnt[N_INFECT_COLS];
  if (threadIdx.x < N_INFECT_COLS) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int32_t t = ia.t;
  const uint32_t n_touched = W.ctl[CVK_CTL_TOUCHED];
  const uint32_t* row = day_row(W, t);
  const bool hosp_max = ia.allow_beds && prm.n_beds_hosp >= 0 && (int64_t)row[CVK_R_MID_SEVERE] > prm.n_beds_hosp;
  const bool icu_max = ia.allow_beds && prm.n_beds_icu >= 0 && (int64_t)row[CVK_R_MID_CRITICAL] > prm.n_beds_icu;
  const uint64_t seed = prm.seed;
  const uint32_t stride = gridDim.x * INFECT_BLOCK;
  const uint32_t n_rounds = (n_touched + stride - 1) / stride;
  for (uint32_t r = 0; r < n_rounds; ++r) {
    const uint32_t j = r * stride + blockIdx.x * INFECT_BLOCK + threadIdx.x;
    bool inf = false, reinf = false, symp = false, sev = false;
    uint32_t v = 0, l = 0;
    int32_t source = -1, i = 0;
    if (j < n_touched) {
      i = W.touched[j];
      const uint64_t key = W.winner[i];
      W.winner[i] = KEY_NONE;
      uint32_t f = P.flags[i];
      if (key != KEY_NONE && (f & CVK_SUSCEPTIBLE)) {  // people.py:470 keeps susceptibles only
        v = key_variant(key);
        l = key_layer(key);
        if (key_class(key) == 1) {
          const int64_t e = (int64_t)key_index(key);
          source = key_dir(key) ? ia.p2[l][e] : ia.p1[l][e];
        }
        inf = true;
        // breakthrough, people.py:486-491
        if (P.peak_nab[i] != 0.0f) {
          const int32_t nb = P.n_breakthroughs[i];
          if (nb == 0) P.rel_trans[i] = P.rel_trans[i] * prm.trans_redux;
          P.n_breakthroughs[i] = nb + 1;
        }
        // states, people.py:494-506
        f = (f & ~(CVK_SUSCEPTIBLE | CVK_NAIVE | CVK_RECOVERED | CVK_DIAGNOSED)) | CVK_EXPOSED;
        P.flags[i] = f;
        P.vstate[i] = (P.vstate[i] & ~0xFu) | v | (1u << (8 + v));
        P.n_infections[i] += 1;
        reinf = !isnan(P.date_recovered[i]);
        // outcome tree, people.py:513-579
        const float d_e2i = (float)draw_dur(prm.dur[CVK_DUR_EXP2INF], seed, (uint32_t)i, t, SLOT_EXP2INF);
        P.dur_exp2inf[i] = d_e2i;
        P.date_exposed[i] = (float)t;
        const float date_inf = d_e2i + (float)t;
        P.date_infectious[i] = date_inf;
        const float fnan = __builtin_nanf("");
        P.date_symptomatic[i] = fnan;
        P.date_severe[i] = fnan;
        P.date_critical[i] = fnan;
        P.date_diagnosed[i] = fnan;
        float date_rec = fnan;
        double u_symp, u_sev, u_crit, u_death;
        draw_pair(seed, (uint32_t)i, t, SLOT_BERN_SYMP_SEV, u_symp, u_sev);
        const float sp = (prm.rel_symp_prob[v] * P.symp_prob[i]) * (1.0f - P.symp_imm[(int64_t)v * P.n + i]);
        symp = u_symp < (double)sp;
        int imm_class = 0;  // 0 asymptomatic, 1 mild, 2 severe (rel_imm_symp)
        if (!symp) {
          const double dd = draw_dur(prm.dur[CVK_DUR_ASYM2REC], seed, (uint32_t)i, t, SLOT_DUR_A);
          date_rec = (float)((double)date_inf + dd);
          P.dur_disease[i] = (float)((double)d_e2i + dd);
        } else {
          const float d_i2s = (float)draw_dur(prm.dur[CVK_DUR_INF2SYM], seed, (uint32_t)i, t, SLOT_DUR_A);
          P.dur_inf2sym[i] = d_i2s;
          const float date_sym = date_inf + d_i2s;
          P.date_symptomatic[i] = date_sym;
          const float sevp = (prm.rel_severe_prob[v] * P.severe_prob[i]) * (1.0f - P.sev_imm[(int64_t)v * P.n + i]);
          sev = u_sev < (double)sevp;
          if (!sev) {
            imm_class = 1;
            const double dd = draw_dur(prm.dur[CVK_DUR_MILD2REC], seed, (uint32_t)i, t, SLOT_DUR_B);
            date_rec = (float)((double)date_sym + dd);
            P.dur_disease[i] = (float)((double)(d_e2i + d_i2s) + dd);
          } else {
            imm_class = 2;
            const float d_s2s = (float)draw_dur(prm.dur[CVK_DUR_SYM2SEV], seed, (uint32_t)i, t, SLOT_DUR_B);
            P.dur_sym2sev[i] = d_s2s;
            const float date_sev = date_sym + d_s2s;
            P.date_severe[i] = date_sev;
            draw_pair(seed, (uint32_t)i, t, SLOT_BERN_CRIT_DEATH, u_crit, u_death);
            const float critp = (prm.rel_crit_prob[v] * P.crit_prob[i]) * (hosp_max ? prm.no_hosp_factor : 1.0f);
            if (!(u_crit < (double)critp)) {
              const double dd = draw_dur(prm.dur[CVK_DUR_SEV2REC], seed, (uint32_t)i, t, SLOT_DUR_C);
              date_rec = (float)((double)date_sev + dd);
              P.dur_disease[i] = (float)((double)((d_e2i + d_i2s) + d_s2s) + dd);
            } else {
              const float d_s2c = (float)draw_dur(prm.dur[CVK_DUR_SEV2CRIT], seed, (uint32_t)i, t, SLOT_DUR_C);
              P.dur_sev2crit[i] = d_s2c;
              const float date_crit = date_sev + d_s2c;
              P.date_critical[i] = date_crit;
              const float deathp = (prm.rel_death_prob[v] * P.death_prob[i]) * (icu_max ? prm.no_icu_factor : 1.0f);
              const float dsum = ((d_e2i + d_i2s) + d_s2s) + d_s2c;
              if (!(u_death < (double)deathp)) {
                const double dd = draw_dur(prm.dur[CVK_DUR_CRIT2REC], seed, (uint32_t)i, t, SLOT_DUR_D);
                date_rec = (float)((double)date_crit + dd);
                P.dur_disease[i] = (float)((double)dsum + dd);
              } else {
                const double dd = draw_dur(prm.dur[CVK_DUR_CRIT2DIE], seed, (uint32_t)i, t, SLOT_DUR_D);
                P.date_dead[i] = (float)((double)date_crit + dd);
                P.dur_disease[i] = (float)((double)dsum + dd);
              }
            }
          }
        }
        P.date_recovered[i] = date_rec;
        // update_peak_nab, immunity.py:138-202
        if (prm.use_waning) {
          if (P.nab[i] > 0.0f) {
            P.peak_nab[i] = P.peak_nab[i] * prm.nab_boost;
          } else {
            const double init_nab = draw_dur(prm.nab_init, seed, (uint32_t)i, t, SLOT_NAB);
            P.peak_nab[i] = (float)((pow(2.0, init_nab) * prm.rel_imm_symp[imm_class]) * prm.nab_norm);
          }
          P.t_nab_event[i] = t;
        }
      }
    }
    // infection log, people.py:509-511: one wave-aggregated append per round
    const uint64_t m = __ballot(inf);
    if (m) {
      unsigned long long base = 0;
      if (lane_id() == 0) base = atomicAdd((unsigned long long*)&W.ctl[CVK_CTL_LOG_COUNT], (unsigned long long)__popcll(m));
      base = __shfl(base, 0, 64);
      if (inf) {
        const unsigned long long pos = base + lanes_below(m);
        if ((int64_t)pos < W.log_capacity) {
          int4 rec;
          rec.x = source;
          rec.y = i;
          rec.z = t;
          rec.w = (int32_t)(l | (v << 8));
          ((int4*)W.log)[pos] = rec;
        } else {
          W.ctl[CVK_CTL_LOG_OVERFLOW] = 1;
        }
      }
    }
    if (ia.count_flows && m) {
      const uint32_t n_inf = (uint32_t)__popcll(m);
      const uint32_t n_reinf = (uint32_t)__popcll(__ballot(inf && reinf));
      if (lane_id() == 0) {
        atomicAdd(&s_cnt[0], n_inf);
        if (n_reinf) atomicAdd(&s_cnt[1], n_reinf);
      }
      for (int vv = 0; vv < prm.n_variants; ++vv) {
        const uint32_t a0 = (uint32_t)__popcll(__ballot(inf && v == (uint32_t)vv));
        const uint32_t a1 = (uint32_t)__popcll(__ballot(inf && symp && v == (uint32_t)vv));
        const uint32_t a2 = (uint32_t)__popcll(__ballot(inf && sev && v == (uint32_t)vv));
        if (lane_id() == 0) {
          if (a0) atomicAdd(&s_cnt[2 + vv], a0);
          if (a1) atomicAdd(&s_cnt[2 + MAXV + vv], a1);
          if (a2) atomicAdd(&s_cnt[2 + 2 * MAXV + vv], a2);
        }
      }
    }
  }
  __syncthreads();
  if (ia.count_flows && threadIdx.x < N_INFECT_COLS && s_cnt[threadIdx.x]) {
    const int k = threadIdx.x;
    int col;
    if (k == 0) col = CVK_R_NEW_INFECTIONS;
    else if (k == 1) col = CVK_R_NEW_REINFECTIONS;
    else col = CVK_R_BYVAR + ((k - 2) / MAXV) * MAXV + (k - 2) % MAXV;  // infections, symptomatic, severe
    atomicAdd(&day_row(W, t)[col], s_cnt[k]);
  }
}

// ===========================================================================
// host side
// ===========================================================================
static int g_edge_lds = 1;  // tunable: stage the infectious bitmap in LDS when it fits

struct DeviceCaps {
  int device = -1;
  int n_cu = 0;
  int max_lds = 0;
};

static DeviceCaps& caps() {
  static thread_local DeviceCaps c;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return c;
  if (c.device != dev) {
    c.device = dev;
    if (hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) c.n_cu = 1;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
    c.max_lds = lds;
  }
  return c;
}

static int grid_for(int64_t n, int block) {
  const int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

static int64_t bitmap_words_for(int64_t n) { return ((n + 127) / 128) * 4; }  // multiple of 4 words

static bool people_ok(const cvk_people* P) {
  return P && P->n > 0 && P->flags && P->vstate && P->date_exposed && P->date_infectious && P->date_symptomatic &&
         P->date_severe && P->date_critical && P->date_recovered && P->date_dead && P->date_diagnosed &&
         P->date_end_isolation && P->dur_exp2inf && P->dur_inf2sym && P->dur_sym2sev && P->dur_sev2crit &&
         P->dur_disease && P->rel_trans && P->rel_sus && P->symp_prob && P->severe_prob && P->crit_prob &&
         P->death_prob && P->n_infections && P->n_breakthroughs && P->nab && P->peak_nab && P->t_nab_event &&
         P->sus_imm && P->symp_imm && P->sev_imm;
}

static bool work_ok(const cvk_work* W) {
  return W && W->winner && W->touched && W->bitmap && W->ctl && W->results && W->results_f && W->select_hist &&
         W->susmap && W->rec && W->partials && W->partials_f && W->npts > 0 && (W->log_capacity == 0 || W->log);
}

static int validate(const cvk_people* P, const cvk_params* prm, const cvk_work* W, int32_t t) {
  if (!people_ok(P)) return fail(CVK_E_INVALID, "people struct has a null pointer or n <= 0");
  if (!prm || prm->n_agents != P->n) return fail(CVK_E_INVALID, "params.n_agents does not match people.n");
  if (P->n >= (1ll << 31) - 1) return fail(CVK_E_INVALID, "more than 2^31-2 agents");
  if ((int64_t)grid_for(P->n, AGENT_BLOCK * AGENTS_PER_THREAD) > CVK_MAX_REDUCE_BLOCKS)
    return fail(CVK_E_CAPACITY, "too many agents for the reduction workspace (CVK_MAX_REDUCE_BLOCKS)");
  if (prm->n_variants < 1 || prm->n_variants > CVK_MAX_VARIANTS)
    return fail(CVK_E_INVALID, "n_variants must be in [1, 8]");
  if (prm->n_layers < 0 || prm->n_layers > CVK_MAX_LAYERS) return fail(CVK_E_INVALID, "n_layers must be in [0, 8]");
  if (prm->use_waning && (!prm->nab_kin || prm->nab_kin_len <= 0))
    return fail(CVK_E_INVALID, "use_waning requires a nab_kin table");
  if (!work_ok(W)) return fail(CVK_E_INVALID, "work struct has a null pointer");
  if (((uintptr_t)W->bitmap | (uintptr_t)W->susmap | (uintptr_t)W->rec) & 15)
    return fail(CVK_E_ALIGN, "bitmap, susmap and rec must be 16-byte aligned");
  if (t < 0 || t >= W->npts) return fail(CVK_E_INVALID, "day index outside the results table");
  return CVK_OK;
}

static uint32_t agent_blocks_for(int64_t n) { return (uint32_t)grid_for(n, AGENT_BLOCK * AGENTS_PER_THREAD); }

static int day_reduce_launch(cvk_work* work, const DayReduce& r, hipStream_t st) {
  if (!r.n_blocks) return CVK_OK;
  hipLaunchKernelGGL(day_reduce_kernel, dim3(1), dim3(1024), 0, st, *work, r);
  return check_launch("day_reduce");
}

static int day_begin_launch(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                            cvk_work* work, int32_t t, int prologue, int transitions, int resample, hipStream_t st,
                            DayReduce* reduce_out) {
  DayBeginArgs a;
  memset(&a, 0, sizeof(a));
  a.t = t;
  a.prologue = prologue;
  a.transitions = transitions;
  a.agent_blocks = agent_blocks_for(people->n);
  int64_t max_dyn = 0;
  if (resample) {
    for (int l = 0; l < params->n_layers; ++l) {
      const cvk_layer& L = layers[l];
      if (!L.dynamic || L.n_edges <= 0) continue;
      if (!L.p1 || !L.p2 || !L.beta) return fail(CVK_E_INVALID, "dynamic layer has a null pointer");
      ResampleLayer& R = a.rs[a.n_resample++];
      R.p1 = L.p1;
      R.p2 = L.p2;
      R.beta = L.beta;
      R.n = L.n_edges;
      R.layer = (uint32_t)l;
      max_dyn = std::max(max_dyn, L.n_edges);
    }
  }
  const uint32_t rs_blocks =
      a.n_resample ? (uint32_t)std::min<int64_t>(grid_for(max_dyn, AGENT_BLOCK), 4 * std::max(caps().n_cu, 1)) : 0u;
  hipLaunchKernelGGL(day_begin_kernel, dim3(a.agent_blocks + rs_blocks), dim3(AGENT_BLOCK), 0, st, *people, *params,
                     *work, a);
  reduce_out->n_blocks = a.agent_blocks;
  reduce_out->t = t;
  reduce_out->prologue = prologue;
  reduce_out->transitions = transitions;
  return check_launch("day_begin");
}

}  // namespace cvk

using namespace cvk;

extern "C" const char* cvk_last_error(void) { return last_error_buffer(); }

extern "C" int cvk_set_tuning(int32_t key, int32_t value) {
  switch (key) {
    case CVK_TUNE_EDGE_LDS:
      g_edge_lds = value ? 1 : 0;
      return CVK_OK;
    default:
      return fail(CVK_E_INVALID, "cvk_set_tuning: unknown key");
  }
}
extern "C" int cvk_version(void) { return 2; }

// Sizes and selected field offsets of the ABI structs, so bindings can verify their mirrors.
extern "C" int cvk_abi_layout(int64_t* out, int32_t n) {
  const int64_t vals[] = {
      (int64_t)sizeof(cvk_dur),    (int64_t)sizeof(cvk_params), (int64_t)sizeof(cvk_people),
      (int64_t)sizeof(cvk_layer),  (int64_t)sizeof(cvk_work),   (int64_t)sizeof(cvk_imports),
      (int64_t)offsetof(cvk_params, dur),        (int64_t)offsetof(cvk_params, nab_init),
      (int64_t)offsetof(cvk_params, rel_imm_symp), (int64_t)offsetof(cvk_params, immunity),
      (int64_t)offsetof(cvk_params, nab_kin),    (int64_t)offsetof(cvk_people, sev_imm),
      (int64_t)offsetof(cvk_work, log_capacity), (int64_t)offsetof(cvk_work, npts),
      (int64_t)CVK_R_COLS, (int64_t)CVK_RF_COLS};
  const int32_t m = (int32_t)(sizeof(vals) / sizeof(vals[0]));
  if (!out || n < m) return fail(CVK_E_INVALID, "cvk_abi_layout: output too small");
  for (int32_t j = 0; j < m; ++j) out[j] = vals[j];
  return m;
}

// Device properties the launch configuration depends on: CUs, LDS bytes per block.
extern "C" int cvk_device_info(int32_t* n_cu, int32_t* max_lds_per_block) {
  DeviceCaps& c = caps();
  if (c.device < 0) return fail(CVK_E_HIP, "cvk_device_info: no HIP device");
  if (n_cu) *n_cu = c.n_cu;
  if (max_lds_per_block) *max_lds_per_block = c.max_lds;
  return CVK_OK;
}

extern "C" int cvk_workspace_sizes(int64_t n_agents, int32_t npts, int64_t* winner_bytes, int64_t* touched_bytes,
                                   int64_t* bitmap_bytes, int64_t* results_bytes, int64_t* results_f_bytes) {
  if (n_agents <= 0 || npts <= 0) return fail(CVK_E_INVALID, "cvk_workspace_sizes: sizes must be positive");
  if (winner_bytes) *winner_bytes = n_agents * 8;
  if (touched_bytes) *touched_bytes = n_agents * 4;
  if (bitmap_bytes) *bitmap_bytes = bitmap_words_for(n_agents) * 4;
  if (results_bytes) *results_bytes = (int64_t)npts * CVK_R_COLS * 4;
  if (results_f_bytes) *results_f_bytes = (int64_t)npts * CVK_RF_COLS * 8;
  return CVK_OK;
}

extern "C" int cvk_workspace_sizes2(int64_t n_agents, int64_t* rec_bytes, int64_t* partials_bytes,
                                    int64_t* partials_f_bytes) {
  if (n_agents <= 0) return fail(CVK_E_INVALID, "cvk_workspace_sizes2: n_agents must be positive");
  if (rec_bytes) *rec_bytes = n_agents * 8;
  if (partials_bytes) *partials_bytes = (int64_t)CVK_MAX_REDUCE_BLOCKS * CVK_PARTIAL_COLS * 4;
  if (partials_f_bytes) *partials_f_bytes = (int64_t)CVK_MAX_REDUCE_BLOCKS * CVK_RF_COLS * 8;
  return CVK_OK;
}

extern "C" int cvk_agent_pass(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                              void* stream) {
  int rc = validate(people, params, work, t);
  if (rc) return rc;
  DayReduce r;
  rc = day_begin_launch(people, nullptr, params, work, t, 0, 1, 0, (hipStream_t)stream, &r);
  if (rc) return rc;
  return day_reduce_launch(work, r, (hipStream_t)stream);
}

extern "C" int cvk_count_pass(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                              void* stream) {
  int rc = validate(people, params, work, t);
  if (rc) return rc;
  DayReduce r;
  rc = day_begin_launch(people, nullptr, params, work, t, 0, 0, 0, (hipStream_t)stream, &r);
  if (rc) return rc;
  return day_reduce_launch(work, r, (hipStream_t)stream);
}

extern "C" int cvk_resample_layer(const cvk_layer* layer, int32_t layer_index, const cvk_params* params, int32_t t,
                                  void* stream) {
  if (!layer || !params || layer->n_edges < 0 || (layer->n_edges > 0 && (!layer->p1 || !layer->p2 || !layer->beta)))
    return fail(CVK_E_INVALID, "cvk_resample_layer: invalid layer");
  if (layer_index < 0 || layer_index >= CVK_MAX_LAYERS) return fail(CVK_E_INVALID, "cvk_resample_layer: bad index");
  if (params->n_agents <= 0 || params->n_agents >= (1ll << 31))
    return fail(CVK_E_INVALID, "cvk_resample_layer: bad n_agents");
  if (layer->n_edges == 0) return CVK_OK;
  // only resample blocks: agent_blocks = 0
  DayBeginArgs a;
  memset(&a, 0, sizeof(a));
  a.t = t;
  a.n_resample = 1;
  a.agent_blocks = 0;
  a.rs[0].p1 = layer->p1;
  a.rs[0].p2 = layer->p2;
  a.rs[0].beta = layer->beta;
  a.rs[0].n = layer->n_edges;
  a.rs[0].layer = (uint32_t)layer_index;
  cvk_people P;
  memset(&P, 0, sizeof(P));
  P.n = params->n_agents;
  cvk_work W;
  memset(&W, 0, sizeof(W));
  const int blocks = (int)std::min<int64_t>(grid_for(layer->n_edges, AGENT_BLOCK), 4 * std::max(caps().n_cu, 1));
  hipLaunchKernelGGL(day_begin_kernel, dim3(blocks), dim3(AGENT_BLOCK), 0, (hipStream_t)stream, P, *params, W, a);
  return check_launch("cvk_resample_layer");
}

static int choose_impl(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                       uint32_t stream_id, int64_t k, int32_t susceptible_only, int32_t layer_code, int32_t variant,
                       hipStream_t st) {
  if (k < 0 || k > people->n) return fail(CVK_E_INVALID, "cvk_choose: k outside [0, n]");
  if (layer_code < 0 || layer_code > 15 || variant < 0 || variant >= params->n_variants)
    return fail(CVK_E_INVALID, "cvk_choose: bad layer code or variant");
  if (k == 0) return CVK_OK;
  const int grid = (int)std::min<int64_t>(grid_for(people->n, 256), 8 * std::max(caps().n_cu, 1));
  hipLaunchKernelGGL(choose_init_kernel, dim3(1), dim3(256), 0, st, *work, (uint32_t)k);
  for (int pass = 0; pass < 8; ++pass) {
    hipLaunchKernelGGL(choose_hist_kernel, dim3(grid), dim3(256), 0, st, *people, *work, params->seed, t, stream_id,
                       susceptible_only, pass);
    hipLaunchKernelGGL(choose_select_kernel, dim3(1), dim3(64), 0, st, *work, pass);
  }
  hipLaunchKernelGGL(choose_mark_kernel, dim3(grid), dim3(256), 0, st, *people, *work, params->seed, t, stream_id,
                     susceptible_only, (uint32_t)layer_code, (uint32_t)variant);
  return check_launch("cvk_choose");
}

extern "C" int cvk_choose(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                          uint32_t stream_id, int64_t k, int32_t susceptible_only, int32_t layer_code,
                          int32_t variant, void* stream) {
  int rc = validate(people, params, work, t);
  if (rc) return rc;
  return choose_impl(people, params, work, t, stream_id, k, susceptible_only, layer_code, variant,
                     (hipStream_t)stream);
}

// Dynamic LDS budget of the edge pass: the device's per-workgroup LDS minus the kernel's static LDS.
static int edge_dynamic_lds(bool lds_variant, int64_t* budget) {
  DeviceCaps& c = caps();
  static thread_local int configured_device = -1;
  static thread_local int64_t budgets[2] = {0, 0};
  if (configured_device != c.device) {
    const void* fns[2] = {(const void*)edge_pass_kernel<false>, (const void*)edge_pass_kernel<true>};
    for (int j = 0; j < 2; ++j) {
      hipFuncAttributes attr;
      if (hipFuncGetAttributes(&attr, fns[j]) != hipSuccess)
        return fail(CVK_E_HIP, "cannot query the edge pass attributes");
      const int64_t b = (int64_t)c.max_lds - (int64_t)attr.sharedSizeBytes;
      if (b <= 0 || hipFuncSetAttribute(fns[j], hipFuncAttributeMaxDynamicSharedMemorySize, (int)b) != hipSuccess)
        return fail(CVK_E_HIP, "cannot raise the LDS limit of the edge pass");
      budgets[j] = b;
    }
    configured_device = c.device;
  }
  *budget = budgets[lds_variant ? 1 : 0];
  return CVK_OK;
}

static int edge_pass_impl(const cvk_people* people, const cvk_layer* layers, const cvk_params* params, cvk_work* work,
                          int32_t t, const double* const* uniforms, hipStream_t st, const DayReduce* reduce) {
  EdgeArgs a;
  memset(&a, 0, sizeof(a));
  a.work = *work;
  if (reduce) a.reduce = *reduce;
  a.n_layers = params->n_layers;
  a.n_variants = params->n_variants;
  for (int v = 0; v < MAXV; ++v) a.beta_v[v] = params->beta_v[v];
  a.asymp_factor = params->asymp_factor;
  a.vl_hi = params->vl_hi;
  a.vl_lo = params->vl_lo;
  a.use_sus_imm = params->use_waning;
  a.seed = params->seed;
  a.t = t;
  a.rec = work->rec;
  a.sus_imm = people->sus_imm;
  a.n_agents = people->n;
  a.bitmap = work->bitmap;
  a.susmap = work->susmap;
  a.bitmap_words = (int32_t)bitmap_words_for(people->n);
  a.winner = work->winner;
  a.touched = work->touched;
  a.touched_count = work->ctl + CVK_CTL_TOUCHED;
  a.tile_counter = work->ctl + CVK_CTL_EDGE_TILES;

  DeviceCaps& c = caps();
  const int64_t queue_bytes = (int64_t)(EDGE_BLOCK / 64) * EDGE_QCAP * 4;
  const int64_t lds_bytes = (int64_t)a.bitmap_words * 4 + queue_bytes;
  int64_t budget = 0;
  int rc0 = edge_dynamic_lds(true, &budget);
  if (rc0) return rc0;
  const bool use_lds = lds_bytes <= budget && g_edge_lds != 0;
  int64_t tiles = 0;
  for (int l = 0; l < params->n_layers; ++l) {
    const cvk_layer& L = layers[l];
    if (L.n_edges < 0 || (L.n_edges > 0 && (!L.p1 || !L.p2 || !L.beta)))
      return fail(CVK_E_INVALID, "cvk_edge_pass: invalid layer");
    if (((uintptr_t)L.p1 | (uintptr_t)L.p2) & 15) return fail(CVK_E_ALIGN, "cvk_edge_pass: p1/p2 must be 16-byte aligned");
    a.layer[l].p1 = L.p1;
    a.layer[l].p2 = L.p2;
    a.layer[l].beta = L.beta;
    a.layer[l].unif = uniforms ? uniforms[l] : nullptr;
    a.layer[l].n = L.n_edges;
    a.layer[l].beta_layer = params->beta_layer[l];
    a.layer[l].iso_factor = params->iso_factor[l];
    a.layer[l].quar_factor = params->quar_factor[l];
    a.layer[l].beta_ones = L.beta_ones;
    a.tile_start[l] = tiles;
    tiles += (L.n_edges + EDGE_TILE - 1) / EDGE_TILE;
  }
  a.tile_start[params->n_layers] = tiles;
  for (int l = params->n_layers + 1; l <= CVK_MAX_LAYERS; ++l) a.tile_start[l] = tiles;
  if (tiles == 0) return reduce ? day_reduce_launch(work, *reduce, st) : CVK_OK;
  const int grid = (int)std::min<int64_t>(tiles, (int64_t)std::max(c.n_cu, 1));
  if (use_lds) {
    hipLaunchKernelGGL((edge_pass_kernel<true>), dim3(grid), dim3(EDGE_BLOCK), (size_t)lds_bytes, st, a);
  } else {
    hipLaunchKernelGGL((edge_pass_kernel<false>), dim3(grid), dim3(EDGE_BLOCK), (size_t)queue_bytes, st, a);
  }
  return check_launch("cvk_edge_pass");
}

extern "C" int cvk_edge_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                             cvk_work* work, int32_t t, const double* const* uniforms, void* stream) {
  int rc = validate(people, params, work, t);
  if (rc) return rc;
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_edge_pass: layers is null");
  return edge_pass_impl(people, layers, params, work, t, uniforms, (hipStream_t)stream, nullptr);
}

static int infect_impl(const cvk_people* people, const cvk_layer* layers, const cvk_params* params, cvk_work* work,
                       int32_t t, int32_t count_flows, int32_t allow_beds, hipStream_t st) {
  InfectArgs ia;
  memset(&ia, 0, sizeof(ia));
  for (int l = 0; l < params->n_layers; ++l) {
    ia.p1[l] = layers[l].p1;
    ia.p2[l] = layers[l].p2;
  }
  ia.t = t;
  ia.count_flows = count_flows;
  ia.allow_beds = allow_beds;
  const int grid = std::max(caps().n_cu, 1);
  hipLaunchKernelGGL(infect_pass_kernel, dim3(grid), dim3(INFECT_BLOCK), 0, st, *people, *params, *work, ia);
  return check_launch("cvk_infect_pass");
}

extern "C" int cvk_infect_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                               cvk_work* work, int32_t t, int32_t count_flows, void* stream) {
  int rc = validate(people, params, work, t);
  if (rc) return rc;
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_infect_pass: layers is null");
  return infect_impl(people, layers, params, work, t, count_flows, count_flows, (hipStream_t)stream);
}

static int run_days_impl(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                         int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports, hipStream_t st) {
  int rc;
  for (int32_t t = t0; t < t1; ++t) {
    if (t == 0 && pop_infected > 0) {  // init_infections, sim.py:505-532 (flows not counted)
      rc = choose_impl(people, params, work, 0, STREAM_CHOOSE, pop_infected, 0, CVK_LAYER_SEED, 0, st);
      if (rc) return rc;
      rc = infect_impl(people, layers, params, work, 0, 0, 0, st);
      if (rc) return rc;
    }
    // day t begins with the stocks of day t-1 (already closed if t == t0)
    DayReduce r;
    rc = day_begin_launch(people, layers, params, work, t, t > t0 ? 1 : 0, 1, 1, st, &r);
    if (rc) return rc;
    if (imports) {
      for (int j = 0; j < imports->n; ++j) {
        if (imports->days[j] != t) continue;
        const int32_t v = imports->variants[j];
        rc = choose_impl(people, params, work, t, STREAM_IMPORT | (uint32_t)v, imports->n_imports[j], 1,
                         CVK_LAYER_IMPORT, v, st);
        if (rc) return rc;
      }
    }
    rc = edge_pass_impl(people, layers, params, work, t, nullptr, st, &r);  // block 0 reduces day_begin's rows
    if (rc) return rc;
    rc = infect_impl(people, layers, params, work, t, 1, 1, st);
    if (rc) return rc;
  }
  // close the last day: stocks and NAb update
  DayReduce r;
  rc = day_begin_launch(people, layers, params, work, t1 - 1, 0, 0, 0, st, &r);
  if (rc) return rc;
  return day_reduce_launch(work, r, st);
}

static int check_run_args(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                          int32_t t0, int32_t t1) {
  int rc = validate(people, params, work, t0);
  if (rc) return rc;
  if (t1 < t0 || t1 > work->npts) return fail(CVK_E_INVALID, "bad day range");
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "layers is null");
  return CVK_OK;
}

extern "C" int cvk_graph_create(const cvk_people* people, cvk_layer* layers, const cvk_params* params,
                                cvk_work* work, int32_t t0, int32_t t1, int64_t pop_infected,
                                const cvk_imports* imports, void* stream, void** graph_exec_out);
extern "C" int cvk_graph_launch(void* graph_exec, void* stream);
extern "C" int cvk_graph_destroy(void* graph_exec);

extern "C" int cvk_run_days(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                            int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports,
                            int32_t use_graph, void* stream) {
  int rc = check_run_args(people, layers, params, work, t0, t1);
  if (rc) return rc;
  if (t1 == t0) return CVK_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return run_days_impl(people, layers, params, work, t0, t1, pop_infected, imports, st);
  void* exec = nullptr;
  rc = cvk_graph_create(people, layers, params, work, t0, t1, pop_infected, imports, stream, &exec);
  if (rc) return rc;
  rc = cvk_graph_launch(exec, stream);
  (void)hipStreamSynchronize(st);
  (void)cvk_graph_destroy(exec);
  return rc;
}

// Two-phase graph API: capture days [t0, t1) once, replay many times (inputs must be reset
// between replays by the caller; kernel arguments are baked into the graph).
extern "C" int cvk_graph_create(const cvk_people* people, cvk_layer* layers, const cvk_params* params,
                                cvk_work* work, int32_t t0, int32_t t1, int64_t pop_infected,
                                const cvk_imports* imports, void* stream, void** graph_exec_out) {
  if (!graph_exec_out) return fail(CVK_E_INVALID, "cvk_graph_create: null output handle");
  *graph_exec_out = nullptr;
  int rc = check_run_args(people, layers, params, work, t0, t1);
  if (rc) return rc;
  if (t1 <= t0) return fail(CVK_E_INVALID, "cvk_graph_create: empty day range");
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) return fail(CVK_E_INVALID, "cvk_graph_create: graph capture needs a non-default stream");
  int64_t budget = 0;
  rc = edge_dynamic_lds(true, &budget);  // not a stream operation: done before capture
  if (rc) return rc;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_graph_create: hipStreamBeginCapture failed");
  rc = run_days_impl(people, layers, params, work, t0, t1, pop_infected, imports, st);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess || !graph) return fail(CVK_E_HIP, "cvk_graph_create: hipStreamEndCapture failed");
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(CVK_E_HIP, "cvk_graph_create: hipGraphInstantiate failed");
  *graph_exec_out = (void*)exec;
  return CVK_OK;
}

extern "C" int cvk_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return fail(CVK_E_INVALID, "cvk_graph_launch: null graph");
  if (hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_graph_launch: hipGraphLaunch failed");
  return CVK_OK;
}

extern "C" int cvk_graph_destroy(void* graph_exec) {
  if (!graph_exec) return CVK_OK;
  return hipGraphExecDestroy((hipGraphExec_t)graph_exec) == hipSuccess ? CVK_OK
                                                                        : fail(CVK_E_HIP, "cvk_graph_destroy failed");
}
