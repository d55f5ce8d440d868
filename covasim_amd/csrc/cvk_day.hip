// Production day loop of the covasim_amd library (gfx950, wave64).
//
// One simulated day (covasim/sim.py:558-685) is six stages on one HIP stream:
//   agent_pass   update_states_pre transitions + check_immunity + viral-load phase bit +
//                infectious bitmap + transition flows           (people.py:164-312, immunity.py:303-350)
//   resample     dynamic layers: iid uniform endpoints per edge   (base.py:1849-1876)
//   choose       variant imports among susceptibles (radix select of Philox keys; immunity.py:117-130)
//   edge_pass    all layers x variants x both directions, fused; LDS-staged infectious bitmap;
//                per-target atomicMin of the (variant, layer, direction, edge) key (sim.py:611-649)
//   infect_pass  outcome tree, durations, NAb peak, infection log  (people.py:435-586, immunity.py:138-202)
//   count_pass   stocks + update_nab + population immunity sums   (sim.py:651-674, immunity.py:205-213)
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

__device__ inline uint32_t* day_row(const cvk_work& w, int32_t t) { return w.results + (int64_t)t * CVK_R_COLS; }
__device__ inline double* day_row_f(const cvk_work& w, int32_t t) { return w.results_f + (int64_t)t * CVK_RF_COLS; }

// Wave-aggregated count: every lane of the wave must reach this call.
__device__ inline void wave_count(uint32_t* s_cnt, int col, bool pred) {
  const uint64_t m = __ballot(pred);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&s_cnt[col], (uint32_t)__popcll(m));
}

// calc_VE, immunity.py:216-247, in float64 then stored as float32.
__device__ inline float calc_ve(double nab, double exp_alpha, double beta) {
  const double exp_lo = exp_alpha * pow(nab, beta);
  return (float)(exp_lo / (1.0 + exp_lo));
}

// ===========================================================================
// agent_pass
// ===========================================================================
constexpr int AGENT_BLOCK = 256;
constexpr int AGENT_NCOUNT = 9 + CVK_MAX_VARIANTS;  // 9 scalar flows + infectious-by-variant

__global__ __launch_bounds__(AGENT_BLOCK) void agent_pass_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                 int32_t t) {
  __shared__ uint32_t s_cnt[AGENT_NCOUNT];
  if (threadIdx.x < AGENT_NCOUNT) s_cnt[threadIdx.x] = 0;
  __syncthreads();

  const int64_t i = (int64_t)blockIdx.x * AGENT_BLOCK + threadIdx.x;
  const bool valid = i < P.n;
  const float tf = (float)t;
  bool c_inf = false, c_sym = false, c_sev = false, c_crit = false, c_rec = false, c_death = false, c_kdeath = false;
  uint32_t inf_v = CVK_VAR_NONE;
  bool bit = false, mid_sev = false, mid_crit = false;

  if (valid) {
    uint32_t f = P.flags[i];
    uint32_t vs = P.vstate[i];
    const uint32_t f0 = f, vs0 = vs;
    const bool is_exp = (f & CVK_EXPOSED) != 0;  // people.py:169 snapshot
    const float d_inf = P.date_infectious[i];
    const float d_rec = P.date_recovered[i];
    const float d_dead = P.date_dead[i];
    if (is_exp) {
      // check_infectious, people.py:222-232
      if (!(f & CVK_INFECTIOUS) && !isnan(d_inf) && tf >= d_inf) {
        const uint32_t ev = exp_var(vs);
        f = (f | CVK_INFECTIOUS) & ~INFVAR_MASK;
        f |= ev << CVK_INFVAR_SHIFT;
        if (ev < CVK_MAX_VARIANTS) vs |= 1u << (16 + ev);
        c_inf = true;
        inf_v = ev;
      }
      // check_symptomatic / severe / critical, people.py:235-253
      if (!(f & CVK_SYMPTOMATIC)) {
        const float d = P.date_symptomatic[i];
        if (!isnan(d) && tf >= d) { f |= CVK_SYMPTOMATIC; c_sym = true; }
      }
      if (!(f & CVK_SEVERE)) {
        const float d = P.date_severe[i];
        if (!isnan(d) && tf >= d) { f |= CVK_SEVERE; c_sev = true; }
      }
      if (!(f & CVK_CRITICAL)) {
        const float d = P.date_critical[i];
        if (!isnan(d) && tf >= d) { f |= CVK_CRITICAL; c_crit = true; }
      }
      // check_recovery, people.py:256-291
      if (!(f & CVK_RECOVERED) && !isnan(d_rec) && tf >= d_rec) {
        const uint32_t ev = exp_var(vs);
        f = (f & ~STATE_CLEAR_ON_RECOVERY) | CVK_RECOVERED;
        f |= INFVAR_MASK;  // infectious_variant = nan
        vs = (vs & ~0xFFFFFFu) | (ev << 4) | CVK_VAR_NONE;  // recovered_variant = exposed_variant; clear by-variant
        if (prm.use_waning) f = (f | CVK_SUSCEPTIBLE) & ~CVK_DIAGNOSED;
        c_rec = true;
      }
    }
    // check_exit_iso, people.py:368-374 (not filtered by exposure)
    if (f & CVK_ISOLATED) {
      const float d = P.date_end_isolation[i];
      if (!isnan(d) && tf >= d) f &= ~CVK_ISOLATED;
    }
    // check_death, people.py:294-312
    if (is_exp && !(f & CVK_DEAD) && !isnan(d_dead) && tf >= d_dead) {
      f |= CVK_DEAD;
      if (f & CVK_DIAGNOSED) { f |= CVK_KNOWN_DEAD; c_kdeath = true; }
      f &= ~STATE_CLEAR_ON_DEATH;
      f |= INFVAR_MASK;
      vs = (vs & ~0xFFu) | CVK_VAR_NONE | (CVK_VAR_NONE << 4);
      c_death = true;
    }

    // check_immunity, immunity.py:303-350 (no vaccine sources on this path)
    if (prm.use_waning) {
      const float peak = P.peak_nab[i];
      if (peak != 0.0f) {  // agents never infected keep sus/symp/sev_imm == 0
        const double nab = (double)P.nab[i];
        const bool was_inf = tf >= d_rec;
        const uint32_t rv = rec_var(vs);
        for (int v = 0; v < prm.n_variants; ++v) {
          double imm = 0.0;
          if (was_inf && rv < CVK_MAX_VARIANTS) imm = (double)prm.immunity[v][rv];
          const double eff = nab * imm;
          const int64_t off = (int64_t)v * P.n + i;
          P.sus_imm[off] = calc_ve(eff, prm.ve_exp_alpha[0], prm.ve_beta[0]);
          P.symp_imm[off] = calc_ve(eff, prm.ve_exp_alpha[1], prm.ve_beta[1]);
          P.sev_imm[off] = calc_ve(eff, prm.ve_exp_alpha[2], prm.ve_beta[2]);
        }
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
    bit = (f & CVK_INFECTIOUS) && P.rel_trans[i] != 0.0f;
    mid_sev = (f & CVK_SEVERE) != 0;
    mid_crit = (f & CVK_CRITICAL) != 0;
  }

  // infectious bitmap: one 64-bit ballot per wave covers 64 consecutive agents
  const uint64_t bits = __ballot(bit);
  const int64_t wave = i >> 6;
  if ((threadIdx.x & 63) == 0 && wave < (P.n + 63) / 64) ((uint64_t*)W.bitmap)[wave] = bits;

  wave_count(s_cnt, 0, c_inf);
  wave_count(s_cnt, 1, c_sym);
  wave_count(s_cnt, 2, c_sev);
  wave_count(s_cnt, 3, c_crit);
  wave_count(s_cnt, 4, c_rec);
  wave_count(s_cnt, 5, c_death);
  wave_count(s_cnt, 6, c_kdeath);
  wave_count(s_cnt, 7, mid_sev);
  wave_count(s_cnt, 8, mid_crit);
  for (int v = 0; v < prm.n_variants; ++v) wave_count(s_cnt, 9 + v, c_inf && inf_v == (uint32_t)v);
  __syncthreads();
  if (threadIdx.x < AGENT_NCOUNT && s_cnt[threadIdx.x]) {
    static constexpr int cols[9] = {CVK_R_NEW_INFECTIOUS, CVK_R_NEW_SYMPTOMATIC, CVK_R_NEW_SEVERE,
                                    CVK_R_NEW_CRITICAL, CVK_R_NEW_RECOVERIES, CVK_R_NEW_DEATHS,
                                    CVK_R_NEW_KNOWN_DEATHS, CVK_R_MID_SEVERE, CVK_R_MID_CRITICAL};
    const int k = threadIdx.x;
    const int col = k < 9 ? cols[k] : CVK_R_BYVAR + CVK_RV_NEW_INFECTIOUS * CVK_MAX_VARIANTS + (k - 9);
    atomicAdd(&day_row(W, t)[col], s_cnt[k]);
  }
}

// ===========================================================================
// resample (Layer.update with frac = 1)
// ===========================================================================
__global__ void resample_kernel(int32_t* __restrict__ p1, int32_t* __restrict__ p2, float* __restrict__ beta,
                                int64_t n_edges, uint32_t n_agents, uint64_t seed, int32_t t, uint32_t layer) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_edges) return;
  const uint64_t ue = (uint64_t)e;
  const u32x4 w = philox_seeded(seed, (uint32_t)ue, (uint32_t)t, STREAM_RESAMPLE | layer, (uint32_t)(ue >> 32));
  p1[e] = (int32_t)below(w.x, w.y, n_agents);
  p2[e] = (int32_t)below(w.z, w.w, n_agents);
  beta[e] = 1.0f;
}

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
    for (int d = 0; d < 256; ++d) total += W.select_hist[d];
    if (remain >= total) {  // everyone eligible is chosen
      W.ctl[CVK_CTL_SEL_ALL] = 1;
      if (remain > total) W.ctl[CVK_CTL_SEL_SHORT] = 1;
      for (int d = 0; d < 256; ++d) W.select_hist[d] = 0;
      return;
    }
  }
  uint64_t prefix = ((uint64_t)W.ctl[CVK_CTL_SEL_PREFIX + 1] << 32) | W.ctl[CVK_CTL_SEL_PREFIX];
  const int shift = 56 - 8 * pass;
  uint32_t cum = 0;
  int digit = 255;
  for (int d = 0; d < 256; ++d) {
    const uint32_t h = W.select_hist[d];
    if (cum + h >= remain) { digit = d; break; }
    cum += h;
  }
  remain -= cum;
  prefix |= (uint64_t)digit << shift;
  W.ctl[CVK_CTL_SEL_PREFIX] = (uint32_t)prefix;
  W.ctl[CVK_CTL_SEL_PREFIX + 1] = (uint32_t)(prefix >> 32);
  W.ctl[CVK_CTL_SEL_REMAIN] = remain;
  for (int d = 0; d < 256; ++d) W.select_hist[d] = 0;
}

__global__ __launch_bounds__(256) void choose_mark_kernel(cvk_people P, cvk_work W, uint64_t seed, int32_t t,
                                                          uint32_t stream_id, int32_t sus_only, uint32_t layer_code,
                                                          uint32_t variant, int32_t count_imports, uint32_t k) {
  const uint64_t thr = W.ctl[CVK_CTL_SEL_ALL] ? ~0ull
                                              : ((uint64_t)W.ctl[CVK_CTL_SEL_PREFIX + 1] << 32) | W.ctl[CVK_CTL_SEL_PREFIX];
  if (count_imports && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&day_row(W, t)[CVK_R_N_IMPORTS], k);
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
  int32_t pad;
};

struct EdgeArgs {
  EdgeLayer layer[CVK_MAX_LAYERS];
  int64_t tile_start[CVK_MAX_LAYERS + 1];
  int32_t n_layers, n_variants;
  float beta_v[CVK_MAX_VARIANTS];
  float asymp_factor, vl_hi, vl_lo;
  int32_t use_sus_imm;
  uint64_t seed;
  int32_t t;
  int32_t bitmap_words;
  const uint32_t* flags;
  const float* rel_trans;
  const float* rel_sus;
  const float* sus_imm;
  int64_t n_agents;
  const uint32_t* bitmap;
  uint64_t* winner;
  int32_t* touched;
  uint32_t* touched_count;
};

constexpr int EDGE_EPT = 8;  // edges per thread per tile (two 16-byte loads per endpoint array)

// One directed slot (source -> target) of edge e of layer l: the float32 products of
// compute_trans_sus (utils.py:88-89) and compute_infections (utils.py:117), then the
// float64 uniform test (utils.py:123).
__device__ inline void edge_slot(const EdgeArgs& a, const EdgeLayer& L, int l, int dir, int64_t e, int32_t src,
                                 int32_t tgt) {
  const uint32_t fs = a.flags[src];
  const uint32_t v = inf_var(fs);
  if (!(fs & CVK_INFECTIOUS) || v >= (uint32_t)a.n_variants) return;
  float rt = a.rel_trans[src];                                   // * inf (== 1)
  if (fs & CVK_QUARANTINED) rt = rt * L.quar_factor;             // * f_quar
  if (!(fs & CVK_SYMPTOMATIC)) rt = rt * a.asymp_factor;         // * f_asymp
  if (fs & CVK_ISOLATED) rt = rt * L.iso_factor;                 // * f_iso
  rt = rt * L.beta_layer;                                        // * beta_layer
  rt = rt * ((fs & CVK_EARLY) ? a.vl_hi : a.vl_lo);              // * viral_load
  if (rt == 0.0f) return;                                        // utils.py:116
  const uint32_t ft = a.flags[tgt];
  if (!(ft & CVK_SUSCEPTIBLE)) return;                            // rel_sus * sus == 0
  float rs = a.rel_sus[tgt];
  if (ft & CVK_QUARANTINED) rs = rs * L.quar_factor;
  if (a.use_sus_imm) rs = rs * (1.0f - a.sus_imm[(int64_t)v * a.n_agents + tgt]);
  const float p = ((a.beta_v[v] * L.beta[e]) * rt) * rs;         // utils.py:117
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
__device__ inline bool infectious_bit(const EdgeArgs& a, const uint32_t* s_bits, int32_t i) {
  const uint32_t w = LDS ? s_bits[i >> 5] : a.bitmap[i >> 5];
  return (w >> (i & 31)) & 1u;
}

template <bool LDS>
__device__ inline void edge_quad(const EdgeArgs& a, const uint32_t* s_bits, const EdgeLayer& L, int l, int64_t e,
                                 const int32_t (&x)[4], const int32_t (&y)[4], int cnt) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // endpoints outside [0, n) are ignored (the reference leaves them undefined)
    if (q < cnt && (uint32_t)x[q] < (uint32_t)a.n_agents && (uint32_t)y[q] < (uint32_t)a.n_agents) {
      const bool bx = infectious_bit<LDS>(a, s_bits, x[q]);
      const bool by = infectious_bit<LDS>(a, s_bits, y[q]);
      if (bx) edge_slot(a, L, l, 0, e + q, x[q], y[q]);
      if (by) edge_slot(a, L, l, 1, e + q, y[q], x[q]);
    }
  }
}

template <bool LDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void edge_pass_kernel(EdgeArgs a) {
  extern __shared__ uint32_t s_bits[];
  if (LDS) {
    const int nw4 = a.bitmap_words >> 2;  // bitmap_words is a multiple of 4
    const uint4* src = (const uint4*)a.bitmap;
    uint4* dst = (uint4*)s_bits;
    for (int j = threadIdx.x; j < nw4; j += BLOCK) dst[j] = src[j];
    __syncthreads();
  }
  const int64_t tile_elems = (int64_t)BLOCK * EDGE_EPT;
  const int64_t n_tiles = a.tile_start[a.n_layers];
  for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    int l = 0;
    while (l + 1 < a.n_layers && tile >= a.tile_start[l + 1]) ++l;
    const EdgeLayer& L = a.layer[l];
    const int64_t e_tile = (tile - a.tile_start[l]) * tile_elems;
#pragma unroll
    for (int c = 0; c < EDGE_EPT / 4; ++c) {
      const int64_t e = e_tile + ((int64_t)c * BLOCK + threadIdx.x) * 4;
      int32_t x[4], y[4];
      int cnt = 0;
      if (e + 3 < L.n) {
        const int4 vx = *(const int4*)(L.p1 + e);
        const int4 vy = *(const int4*)(L.p2 + e);
        x[0] = vx.x; x[1] = vx.y; x[2] = vx.z; x[3] = vx.w;
        y[0] = vy.x; y[1] = vy.y; y[2] = vy.z; y[3] = vy.w;
        cnt = 4;
      } else if (e < L.n) {
        cnt = (int)(L.n - e);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          x[q] = q < cnt ? L.p1[e + q] : 0;
          y[q] = q < cnt ? L.p2[e + q] : 0;
        }
      }
      if (cnt) edge_quad<LDS>(a, s_bits, L, l, e, x, y, cnt);
    }
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
constexpr int INFECT_NCOUNT = 2 + 3 * CVK_MAX_VARIANTS;

__global__ __launch_bounds__(INFECT_BLOCK) void infect_pass_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                   InfectArgs ia) {
  __shared__ uint32_t s_cnt[INFECT_NCOUNT];
  if (threadIdx.x < INFECT_NCOUNT) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int32_t t = ia.t;
  const uint32_t n_touched = W.ctl[CVK_CTL_TOUCHED];
  const uint32_t* row = day_row(W, t);
  const bool hosp_max = ia.allow_beds && prm.n_beds_hosp >= 0 && (int64_t)row[CVK_R_MID_SEVERE] > prm.n_beds_hosp;
  const bool icu_max = ia.allow_beds && prm.n_beds_icu >= 0 && (int64_t)row[CVK_R_MID_CRITICAL] > prm.n_beds_icu;
  const uint64_t seed = prm.seed;
  const uint32_t n_rounds = (n_touched + gridDim.x * INFECT_BLOCK - 1) / (gridDim.x * INFECT_BLOCK);
  for (uint32_t r = 0; r < n_rounds; ++r) {
    const uint32_t j = (r * gridDim.x + blockIdx.x) * INFECT_BLOCK + threadIdx.x;
    bool inf = false, reinf = false, symp = false, sev = false;
    uint32_t v = 0;
    if (j < n_touched) {
      const int32_t i = W.touched[j];
      const uint64_t key = W.winner[i];
      W.winner[i] = KEY_NONE;
      uint32_t f = P.flags[i];
      if (key != KEY_NONE && (f & CVK_SUSCEPTIBLE)) {  // people.py:470 keeps susceptibles only
        v = key_variant(key);
        const uint32_t cls = key_class(key);
        const uint32_t l = key_layer(key);
        int32_t source = -1;
        if (cls == 1) {
          const int64_t e = (int64_t)key_index(key);
          source = key_dir(key) ? ia.p2[l][e] : ia.p1[l][e];
        }
        inf = true;
        // breakthrough, people.py:486-491
        const bool breakthrough = P.peak_nab[i] != 0.0f;
        if (breakthrough) {
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
        // infection log, people.py:509-511
        const unsigned long long pos = atomicAdd((unsigned long long*)&W.ctl[CVK_CTL_LOG_COUNT], 1ull);
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
          const double d = draw_dur(prm.dur[CVK_DUR_ASYM2REC], seed, (uint32_t)i, t, SLOT_DUR_A);
          date_rec = (float)((double)date_inf + d);
          P.dur_disease[i] = (float)((double)d_e2i + d);
        } else {
          const float d_i2s = (float)draw_dur(prm.dur[CVK_DUR_INF2SYM], seed, (uint32_t)i, t, SLOT_DUR_A);
          P.dur_inf2sym[i] = d_i2s;
          const float date_sym = date_inf + d_i2s;
          P.date_symptomatic[i] = date_sym;
          const float sevp = (prm.rel_severe_prob[v] * P.severe_prob[i]) * (1.0f - P.sev_imm[(int64_t)v * P.n + i]);
          sev = u_sev < (double)sevp;
          if (!sev) {
            imm_class = 1;
            const double d = draw_dur(prm.dur[CVK_DUR_MILD2REC], seed, (uint32_t)i, t, SLOT_DUR_B);
            date_rec = (float)((double)date_sym + d);
            P.dur_disease[i] = (float)((double)(d_e2i + d_i2s) + d);
          } else {
            imm_class = 2;
            const float d_s2s = (float)draw_dur(prm.dur[CVK_DUR_SYM2SEV], seed, (uint32_t)i, t, SLOT_DUR_B);
            P.dur_sym2sev[i] = d_s2s;
            const float date_sev = date_sym + d_s2s;
            P.date_severe[i] = date_sev;
            draw_pair(seed, (uint32_t)i, t, SLOT_BERN_CRIT_DEATH, u_crit, u_death);
            const float critp = (prm.rel_crit_prob[v] * P.crit_prob[i]) * (hosp_max ? prm.no_hosp_factor : 1.0f);
            if (!(u_crit < (double)critp)) {
              const double d = draw_dur(prm.dur[CVK_DUR_SEV2REC], seed, (uint32_t)i, t, SLOT_DUR_C);
              date_rec = (float)((double)date_sev + d);
              P.dur_disease[i] = (float)((double)((d_e2i + d_i2s) + d_s2s) + d);
            } else {
              const float d_s2c = (float)draw_dur(prm.dur[CVK_DUR_SEV2CRIT], seed, (uint32_t)i, t, SLOT_DUR_C);
              P.dur_sev2crit[i] = d_s2c;
              const float date_crit = date_sev + d_s2c;
              P.date_critical[i] = date_crit;
              const float deathp = (prm.rel_death_prob[v] * P.death_prob[i]) * (icu_max ? prm.no_icu_factor : 1.0f);
              const float dsum = ((d_e2i + d_i2s) + d_s2s) + d_s2c;
              if (!(u_death < (double)deathp)) {
                const double d = draw_dur(prm.dur[CVK_DUR_CRIT2REC], seed, (uint32_t)i, t, SLOT_DUR_D);
                date_rec = (float)((double)date_crit + d);
                P.dur_disease[i] = (float)((double)dsum + d);
              } else {
                const double d = draw_dur(prm.dur[CVK_DUR_CRIT2DIE], seed, (uint32_t)i, t, SLOT_DUR_D);
                P.date_dead[i] = (float)((double)date_crit + d);
                P.dur_disease[i] = (float)((double)dsum + d);
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
            const double peak = (pow(2.0, init_nab) * prm.rel_imm_symp[imm_class]) * prm.nab_norm;
            P.peak_nab[i] = (float)peak;
          }
          P.t_nab_event[i] = t;
        }
      }
    }
    if (ia.count_flows) {
      wave_count(s_cnt, 0, inf);
      wave_count(s_cnt, 1, inf && reinf);
      for (int vv = 0; vv < prm.n_variants; ++vv) {
        wave_count(s_cnt, 2 + vv, inf && v == (uint32_t)vv);
        wave_count(s_cnt, 2 + CVK_MAX_VARIANTS + vv, inf && symp && v == (uint32_t)vv);
        wave_count(s_cnt, 2 + 2 * CVK_MAX_VARIANTS + vv, inf && sev && v == (uint32_t)vv);
      }
    }
  }
  __syncthreads();
  if (ia.count_flows && threadIdx.x < INFECT_NCOUNT && s_cnt[threadIdx.x]) {
    const int k = threadIdx.x;
    int col;
    if (k == 0) col = CVK_R_NEW_INFECTIONS;
    else if (k == 1) col = CVK_R_NEW_REINFECTIONS;
    else {
      const int kk = k - 2;
      const int which = kk / CVK_MAX_VARIANTS;  // 0 infections, 1 symptomatic, 2 severe
      const int vv = kk % CVK_MAX_VARIANTS;
      col = CVK_R_BYVAR + which * CVK_MAX_VARIANTS + vv;
    }
    atomicAdd(&day_row(W, t)[col], s_cnt[k]);
  }
}

__global__ void reset_touched_kernel(cvk_work W) {
  if (threadIdx.x == 0 && blockIdx.x == 0) W.ctl[CVK_CTL_TOUCHED] = 0;
}

// ===========================================================================
// count_pass (end of day): stocks, update_nab, population immunity sums
// ===========================================================================
constexpr int COUNT_BLOCK = 256;
constexpr int COUNT_NSTOCK = 14;  // 13 stocks + alive
constexpr int COUNT_NCOUNT = COUNT_NSTOCK + 2 * CVK_MAX_VARIANTS;

__global__ __launch_bounds__(COUNT_BLOCK) void count_pass_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                 int32_t t) {
  __shared__ uint32_t s_cnt[COUNT_NCOUNT];
  __shared__ double s_sum[3];
  if (threadIdx.x < COUNT_NCOUNT) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x < 3) s_sum[threadIdx.x] = 0.0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * COUNT_BLOCK + threadIdx.x;
  const bool valid = i < P.n;
  uint32_t f = 0, vs = 0;
  double nab_sum = 0.0, sus_sum = 0.0, symp_sum = 0.0;
  if (valid) {
    f = P.flags[i];
    vs = P.vstate[i];
    if (prm.use_waning) {
      const float peak = P.peak_nab[i];
      float nab = P.nab[i];
      if (peak != 0.0f) {  // update_nab over has_nabs = true(peak_nab), sim.py:666-669
        int32_t ts = t - P.t_nab_event[i];
        ts = ts < 0 ? 0 : (ts >= prm.nab_kin_len ? prm.nab_kin_len - 1 : ts);
        nab = (float)((double)nab + prm.nab_kin[ts] * (double)peak);
        nab = nab < 0.0f ? 0.0f : nab;
        nab = nab > peak ? peak : nab;
        P.nab[i] = nab;
        for (int v = 0; v < prm.n_variants; ++v) {
          sus_sum += (double)P.sus_imm[(int64_t)v * P.n + i];
          symp_sum += (double)P.symp_imm[(int64_t)v * P.n + i];
        }
      }
      if (!(f & CVK_DEAD)) nab_sum = (double)nab;
    }
  }
  const uint32_t stock_bits[13] = {CVK_SUSCEPTIBLE, CVK_EXPOSED, CVK_INFECTIOUS, CVK_SYMPTOMATIC, CVK_SEVERE,
                                   CVK_CRITICAL, CVK_RECOVERED, CVK_DEAD, CVK_DIAGNOSED, CVK_KNOWN_DEAD,
                                   CVK_QUARANTINED, CVK_ISOLATED, CVK_VACCINATED};
#pragma unroll
  for (int k = 0; k < 13; ++k) wave_count(s_cnt, k, valid && (f & stock_bits[k]));
  wave_count(s_cnt, 13, valid && !(f & CVK_DEAD));
  for (int v = 0; v < prm.n_variants; ++v) {
    wave_count(s_cnt, COUNT_NSTOCK + v, valid && ((vs >> (8 + v)) & 1u));
    wave_count(s_cnt, COUNT_NSTOCK + CVK_MAX_VARIANTS + v, valid && ((vs >> (16 + v)) & 1u));
  }
  if (prm.use_waning) {
    // wave reduction of the three sums, then one LDS atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
      nab_sum += __shfl_down(nab_sum, off, 64);
      sus_sum += __shfl_down(sus_sum, off, 64);
      symp_sum += __shfl_down(symp_sum, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&s_sum[0], nab_sum);
      atomicAdd(&s_sum[1], sus_sum);
      atomicAdd(&s_sum[2], symp_sum);
    }
  }
  __syncthreads();
  uint32_t* row = day_row(W, t);
  if (threadIdx.x < COUNT_NCOUNT && s_cnt[threadIdx.x]) {
    const int k = threadIdx.x;
    int col;
    if (k < 13) col = CVK_R_N_SUSCEPTIBLE + k;
    else if (k == 13) col = CVK_R_N_ALIVE_NAB;
    else if (k < COUNT_NSTOCK + CVK_MAX_VARIANTS) col = CVK_R_BYVAR + CVK_RV_N_EXPOSED * CVK_MAX_VARIANTS + (k - COUNT_NSTOCK);
    else col = CVK_R_BYVAR + CVK_RV_N_INFECTIOUS * CVK_MAX_VARIANTS + (k - COUNT_NSTOCK - CVK_MAX_VARIANTS);
    atomicAdd(&row[col], s_cnt[k]);
  }
  if (prm.use_waning && threadIdx.x < 3 && s_sum[threadIdx.x] != 0.0) {
    atomicAdd(&day_row_f(W, t)[threadIdx.x], s_sum[threadIdx.x]);
  }
}

// ===========================================================================
// host side
// ===========================================================================
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
         W->npts > 0 && (W->log_capacity == 0 || W->log);
}

static int validate(const cvk_people* P, const cvk_params* prm, const cvk_work* W, int32_t t, const char* who) {
  if (!people_ok(P)) return fail(CVK_E_INVALID, "people struct has a null pointer or n <= 0");
  if (!prm || prm->n_agents != P->n) return fail(CVK_E_INVALID, "params.n_agents does not match people.n");
  if (P->n >= (1ll << 31) - 1) return fail(CVK_E_INVALID, "more than 2^31-2 agents");
  if (prm->n_variants < 1 || prm->n_variants > CVK_MAX_VARIANTS)
    return fail(CVK_E_INVALID, "n_variants must be in [1, 8]");
  if (prm->n_layers < 0 || prm->n_layers > CVK_MAX_LAYERS) return fail(CVK_E_INVALID, "n_layers must be in [0, 8]");
  if (prm->use_waning && (!prm->nab_kin || prm->nab_kin_len <= 0))
    return fail(CVK_E_INVALID, "use_waning requires a nab_kin table");
  if (!work_ok(W)) return fail(CVK_E_INVALID, "work struct has a null pointer");
  if (t < 0 || t >= W->npts) return fail(CVK_E_INVALID, "day index outside the results table");
  (void)who;
  return CVK_OK;
}

static int grid_for(int64_t n, int block) {
  const int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace cvk

using namespace cvk;

extern "C" const char* cvk_last_error(void) { return last_error_buffer(); }
extern "C" int cvk_version(void) { return 1; }

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

extern "C" int cvk_agent_pass(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                              void* stream) {
  int rc = validate(people, params, work, t, "cvk_agent_pass");
  if (rc) return rc;
  hipLaunchKernelGGL(agent_pass_kernel, dim3(grid_for(people->n, AGENT_BLOCK)), dim3(AGENT_BLOCK), 0,
                     (hipStream_t)stream, *people, *params, *work, t);
  return check_launch("cvk_agent_pass");
}

extern "C" int cvk_resample_layer(const cvk_layer* layer, int32_t layer_index, const cvk_params* params, int32_t t,
                                  void* stream) {
  if (!layer || !params || layer->n_edges < 0 || (layer->n_edges > 0 && (!layer->p1 || !layer->p2 || !layer->beta)))
    return fail(CVK_E_INVALID, "cvk_resample_layer: invalid layer");
  if (layer_index < 0 || layer_index >= CVK_MAX_LAYERS) return fail(CVK_E_INVALID, "cvk_resample_layer: bad index");
  if (params->n_agents <= 0 || params->n_agents >= (1ll << 31))
    return fail(CVK_E_INVALID, "cvk_resample_layer: bad n_agents");
  if (layer->n_edges == 0) return CVK_OK;
  hipLaunchKernelGGL(resample_kernel, dim3(grid_for(layer->n_edges, 256)), dim3(256), 0, (hipStream_t)stream,
                     layer->p1, layer->p2, layer->beta, layer->n_edges, (uint32_t)params->n_agents, params->seed, t,
                     (uint32_t)layer_index);
  return check_launch("cvk_resample_layer");
}

static int choose_impl(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                       uint32_t stream_id, int64_t k, int32_t susceptible_only, int32_t layer_code, int32_t variant,
                       int32_t count_imports, hipStream_t st) {
  if (k < 0 || k > people->n) return fail(CVK_E_INVALID, "cvk_choose: k outside [0, n]");
  if (layer_code < 0 || layer_code > 15 || variant < 0 || variant >= params->n_variants)
    return fail(CVK_E_INVALID, "cvk_choose: bad layer code or variant");
  if (k == 0) return CVK_OK;
  const int grid = (int)std::min<int64_t>(grid_for(people->n, 256), 4 * std::max(caps().n_cu, 1) * 8);
  hipLaunchKernelGGL(choose_init_kernel, dim3(1), dim3(256), 0, st, *work, (uint32_t)k);
  for (int pass = 0; pass < 8; ++pass) {
    hipLaunchKernelGGL(choose_hist_kernel, dim3(grid), dim3(256), 0, st, *people, *work, params->seed, t, stream_id,
                       susceptible_only, pass);
    hipLaunchKernelGGL(choose_select_kernel, dim3(1), dim3(64), 0, st, *work, pass);
  }
  hipLaunchKernelGGL(choose_mark_kernel, dim3(grid), dim3(256), 0, st, *people, *work, params->seed, t, stream_id,
                     susceptible_only, (uint32_t)layer_code, (uint32_t)variant, count_imports, (uint32_t)k);
  return check_launch("cvk_choose");
}

extern "C" int cvk_choose(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                          uint32_t stream_id, int64_t k, int32_t susceptible_only, int32_t layer_code,
                          int32_t variant, void* stream) {
  int rc = validate(people, params, work, t, "cvk_choose");
  if (rc) return rc;
  return choose_impl(people, params, work, t, stream_id, k, susceptible_only, layer_code, variant, 0,
                     (hipStream_t)stream);
}

static int edge_pass_impl(const cvk_people* people, const cvk_layer* layers, const cvk_params* params, cvk_work* work,
                          int32_t t, const double* const* uniforms, hipStream_t st) {
  EdgeArgs a;
  memset(&a, 0, sizeof(a));
  a.n_layers = params->n_layers;
  a.n_variants = params->n_variants;
  for (int v = 0; v < CVK_MAX_VARIANTS; ++v) a.beta_v[v] = params->beta_v[v];
  a.asymp_factor = params->asymp_factor;
  a.vl_hi = params->vl_hi;
  a.vl_lo = params->vl_lo;
  a.use_sus_imm = params->use_waning;
  a.seed = params->seed;
  a.t = t;
  a.flags = people->flags;
  a.rel_trans = people->rel_trans;
  a.rel_sus = people->rel_sus;
  a.sus_imm = people->sus_imm;
  a.n_agents = people->n;
  a.bitmap = work->bitmap;
  a.bitmap_words = (int32_t)bitmap_words_for(people->n);
  a.winner = work->winner;
  a.touched = work->touched;
  a.touched_count = work->ctl + CVK_CTL_TOUCHED;

  DeviceCaps& c = caps();
  const int64_t lds_bytes = (int64_t)a.bitmap_words * 4;
  const bool use_lds = lds_bytes <= (int64_t)c.max_lds - 1024;
  const int block = use_lds ? 1024 : 256;
  const int64_t tile_elems = (int64_t)block * EDGE_EPT;
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
    a.tile_start[l] = tiles;
    tiles += (L.n_edges + tile_elems - 1) / tile_elems;
  }
  a.tile_start[params->n_layers] = tiles;
  for (int l = params->n_layers + 1; l <= CVK_MAX_LAYERS; ++l) a.tile_start[l] = tiles;
  if (tiles == 0) return CVK_OK;
  if (use_lds) {
    const int grid = (int)std::min<int64_t>(tiles, (int64_t)std::max(c.n_cu, 1));
    static thread_local int configured_device = -1;
    if (lds_bytes > 65536 && configured_device != c.device) {
      if (hipFuncSetAttribute((const void*)edge_pass_kernel<true, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              c.max_lds) != hipSuccess)
        return fail(CVK_E_HIP, "cvk_edge_pass: cannot raise the LDS limit");
      configured_device = c.device;
    }
    hipLaunchKernelGGL((edge_pass_kernel<true, 1024>), dim3(grid), dim3(1024), (size_t)lds_bytes, st, a);
  } else {
    const int grid = (int)std::min<int64_t>(tiles, (int64_t)std::max(c.n_cu, 1) * 8);
    hipLaunchKernelGGL((edge_pass_kernel<false, 256>), dim3(grid), dim3(256), 0, st, a);
  }
  return check_launch("cvk_edge_pass");
}

extern "C" int cvk_edge_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                             cvk_work* work, int32_t t, const double* const* uniforms, void* stream) {
  int rc = validate(people, params, work, t, "cvk_edge_pass");
  if (rc) return rc;
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_edge_pass: layers is null");
  return edge_pass_impl(people, layers, params, work, t, uniforms, (hipStream_t)stream);
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
  const int grid = std::max(caps().n_cu, 1) * 4;
  hipLaunchKernelGGL(infect_pass_kernel, dim3(grid), dim3(INFECT_BLOCK), 0, st, *people, *params, *work, ia);
  hipLaunchKernelGGL(reset_touched_kernel, dim3(1), dim3(64), 0, st, *work);
  return check_launch("cvk_infect_pass");
}

extern "C" int cvk_infect_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                               cvk_work* work, int32_t t, int32_t count_flows, void* stream) {
  int rc = validate(people, params, work, t, "cvk_infect_pass");
  if (rc) return rc;
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_infect_pass: layers is null");
  return infect_impl(people, layers, params, work, t, count_flows, count_flows, (hipStream_t)stream);
}

extern "C" int cvk_count_pass(const cvk_people* people, const cvk_params* params, cvk_work* work, int32_t t,
                              void* stream) {
  int rc = validate(people, params, work, t, "cvk_count_pass");
  if (rc) return rc;
  hipLaunchKernelGGL(count_pass_kernel, dim3(grid_for(people->n, COUNT_BLOCK)), dim3(COUNT_BLOCK), 0,
                     (hipStream_t)stream, *people, *params, *work, t);
  return check_launch("cvk_count_pass");
}

static int run_days_impl(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                         int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports, hipStream_t st) {
  int rc;
  for (int32_t t = t0; t < t1; ++t) {
    if (t == 0 && pop_infected > 0) {  // init_infections, sim.py:505-532 (flows not counted)
      rc = choose_impl(people, params, work, 0, STREAM_CHOOSE, pop_infected, 0, CVK_LAYER_SEED, 0, 0, st);
      if (rc) return rc;
      rc = infect_impl(people, layers, params, work, 0, 0, 0, st);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(agent_pass_kernel, dim3(grid_for(people->n, AGENT_BLOCK)), dim3(AGENT_BLOCK), 0, st, *people,
                       *params, *work, t);
    for (int l = 0; l < params->n_layers; ++l) {
      if (layers[l].dynamic) {
        rc = cvk_resample_layer(&layers[l], l, params, t, st);
        if (rc) return rc;
      }
    }
    if (imports) {
      for (int j = 0; j < imports->n; ++j) {
        if (imports->days[j] != t) continue;
        const int32_t v = imports->variants[j];
        rc = choose_impl(people, params, work, t, STREAM_IMPORT | (uint32_t)v, imports->n_imports[j], 1,
                         CVK_LAYER_IMPORT, v, 0, st);
        if (rc) return rc;
      }
    }
    rc = edge_pass_impl(people, layers, params, work, t, nullptr, st);
    if (rc) return rc;
    rc = infect_impl(people, layers, params, work, t, 1, 1, st);
    if (rc) return rc;
    hipLaunchKernelGGL(count_pass_kernel, dim3(grid_for(people->n, COUNT_BLOCK)), dim3(COUNT_BLOCK), 0, st, *people,
                       *params, *work, t);
    rc = check_launch("cvk_run_days");
    if (rc) return rc;
  }
  return CVK_OK;
}

extern "C" int cvk_run_days(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                            int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports,
                            int32_t use_graph, void* stream) {
  int rc = validate(people, params, work, t0, "cvk_run_days");
  if (rc) return rc;
  if (t1 < t0 || t1 > work->npts) return fail(CVK_E_INVALID, "cvk_run_days: bad day range");
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_run_days: layers is null");
  if (t1 == t0) return CVK_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return run_days_impl(people, layers, params, work, t0, t1, pop_infected, imports, st);
  if (st == nullptr) return fail(CVK_E_INVALID, "cvk_run_days: graph capture needs a non-default stream");
  // Configure the LDS limit outside the capture (hipFuncSetAttribute is not a stream operation).
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_run_days: hipStreamBeginCapture failed");
  rc = run_days_impl(people, layers, params, work, t0, t1, pop_infected, imports, st);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess || !graph) return fail(CVK_E_HIP, "cvk_run_days: hipStreamEndCapture failed");
  hipGraphExec_t exec = nullptr;
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(CVK_E_HIP, "cvk_run_days: hipGraphInstantiate failed");
  }
  const hipError_t el = hipGraphLaunch(exec, st);
  (void)hipStreamSynchronize(st);
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  return el == hipSuccess ? CVK_OK : fail(CVK_E_HIP, "cvk_run_days: hipGraphLaunch failed");
}

// Two-phase graph API: capture days [t0, t1) once, replay many times (inputs must be reset
// between replays by the caller; kernel arguments are baked into the graph).
extern "C" int cvk_graph_create(const cvk_people* people, cvk_layer* layers, const cvk_params* params,
                                cvk_work* work, int32_t t0, int32_t t1, int64_t pop_infected,
                                const cvk_imports* imports, void* stream, void** graph_exec_out) {
  if (!graph_exec_out) return fail(CVK_E_INVALID, "cvk_graph_create: null output handle");
  *graph_exec_out = nullptr;
  int rc = validate(people, params, work, t0, "cvk_graph_create");
  if (rc) return rc;
  if (t1 <= t0 || t1 > work->npts) return fail(CVK_E_INVALID, "cvk_graph_create: bad day range");
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_graph_create: layers is null");
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) return fail(CVK_E_INVALID, "cvk_graph_create: graph capture needs a non-default stream");
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
