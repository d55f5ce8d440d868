// Differentiable relaxation of the transmission step: expected infections per agent and day.
//
// The stochastic step draws, for every infectious source j and susceptible target i sharing an
// edge of layer l, an infection with probability x = beta * beta_l * rel_trans_j * viral_load *
// rel_sus_i (utils.py:88-123). The relaxation replaces the indicator states by probabilities:
//   s_i(t)   probability that i is still susceptible at the start of day t
//   D_i(t)   probability that i is infected on day t            (D_i(0) = seed fraction f)
//   tr_j(t)  expected infectiousness = rel_trans_j * sum_{d>=1} D_j(t-d) w_{bin(j)}(d), where w_b(d)
//            is the mean (viral load x infectious) factor d days after infection for age bin b
//   L_i(t)   = sum over layers l and neighbours j of i of log1p(-min(beta*beta_l*sigma_{bin(i)}*tr_j, XMAX))
//   D_i(t)   = s_i(t) * (1 - exp(L_i(t)))          s_i(t+1) = s_i(t) * exp(L_i(t))
//   E_t      = sum_i (transmission part of D_i(t))   (the expected new infections of day t)
// Parameters theta = (beta, beta_l for each layer, sigma_b for the 10 age bins of rel_sus).
//
// Contacts are undirected, so each layer is a neighbour CSR (both directions of every stored pair,
// self-loops dropped). The forward gathers over in-neighbours; the adjoint gathers over the same
// lists in the source role: every sum is a gather in a fixed order (deterministic, no float
// atomics). Reverse mode, day t from last to first (bars are adjoints, e_t = dJ/dE_t):
//   Dbar_i(t) = e_t + sum_{d>=1} Ibar_i(t+d) w_{bin(i)}(d)
//   Lbar_i(t) = s_i(t) e^{L_i} (sbar_i(t+1) - Dbar_i(t))      sbar_i(t) = Dbar_i (1 - e^{L_i}) + sbar_i(t+1) e^{L_i}
//   with g'(x) = -1/(1-x) below XMAX (0 above):
//   C_{a,l} = sum_{j in N_l(a)} g'(beta beta_l sigma_a tr_j) tr_j             (a as target)
//   Ibar_a(t) = rel_trans_a * sum_l beta beta_l sum_{j in N_l(a)} Lbar_j g'(beta beta_l sigma_j tr_a) sigma_j
//   dJ/dbeta += Lbar_a sigma_a sum_l beta_l C_{a,l};  dJ/dbeta_l += beta Lbar_a sigma_a C_{a,l};
//   dJ/dsigma_{bin(a)} += beta Lbar_a sum_l beta_l C_{a,l}
// The forward keeps s at every checkpoint day; the backward re-runs each checkpoint segment's
// forward (storing L, tr, s for its days) and then walks the segment backwards.
//
// A layer resampled every day enters in expectation (c_l = 2 E_l / N slots per agent, uniform sources,
// self-pairs left out): with x_{l,b} = beta beta_l sigma_b, g(y) = log1p(-min(y, XMAX)) and
// S_{l,b}(t) = sum_j g(x_{l,b} tr_j), its term of L_i is c_l / N (S_{l,bin(i)} - g(x_{l,bin(i)} tr_i)). Its
// adjoint needs, per day, Lambda_b = sum_{bin(i)=b} Lbar_i, G_{l,b} = sum_j g'(x_{l,b} tr_j) tr_j and
// H_{l,b} = sum_{bin(i)=b} Lbar_i g'(x_{l,b} tr_i) tr_i (block partials summed in a fixed order):
//   Ibar_a += rel_trans_a c_l / N sum_b x_{l,b} g'(x_{l,b} tr_a) (Lambda_b - [bin(a) = b] Lbar_a)
//   xbar_{l,b} = c_l / N (Lambda_b G_{l,b} - H_{l,b}), which feeds dJ/dbeta, dJ/dbeta_l and dJ/dsigma_b.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "cvk_host.h"
#include "cvk_kernels.h"

namespace cvk {

constexpr int RELAX_BLOCK = 256;
constexpr int RELAX_BINS = 10;
constexpr int RELAX_MAX_PARAMS = 1 + CVK_MAX_LAYERS + RELAX_BINS;
constexpr double RELAX_XMAX = 0.999999;
constexpr int RELAX_UNROLL = 8;  // neighbour-list entries whose gathers are in flight together  // per-edge probabilities are capped below 1 (log1p(-x) stays finite)

template <class R>
struct RelaxTheta {
  R beta;
  R bl[CVK_MAX_LAYERS];
  R sig[RELAX_BINS];
};

template <class R>
struct RelaxDev {
  int64_t n;
  int32_t n_layers, n_days, kmax, pad;   // kmax: longest lag with w != 0 that is kept (<= n_days)
  const int64_t* indptr[CVK_MAX_LAYERS];  // [n + 1]
  const int32_t* nbr[CVK_MAX_LAYERS];
  const uint8_t* bin;    // [n] age bin 0..9
  const float* rt;       // [n] rel_trans
  const R* w;            // [10][kmax + 1]
  R* delta;              // [n_days][n]
  R* s;                  // [n] running susceptible probability
  R* ibar;               // [n_days][n]
  R* sbar;               // [n]
  R* lbar;               // [n]
  double* epart;         // [n_days][n_blocks] per-block sums of E_t
  double* gpart;         // [RELAX_MAX_PARAMS][n_blocks] per-block gradient sums
  R* pack;               // [n][4] the adjoint's per-neighbour inputs of the day: tr, lbar * sig(bin), sig(bin), 0
  int32_t n_blocks;
  R seed_frac;
  // layers resampled every day (in expectation)
  int32_t n_dyn;
  int32_t dyn_layer[CVK_MAX_LAYERS];  // their layer indices
  R dyn_slots[CVK_MAX_LAYERS];        // c_l per layer (0: the layer's neighbour lists)
  double* dyn_part;   // [(RELAX_BINS + 2 n_dyn RELAX_BINS)][n_dyn_blocks] per-workgroup sums
  int32_t n_dyn_blocks;  // workgroups of the per-bin sums (RELAX_DYN_SPAN agents each)
  double* dyn_fsum;   // [n_dyn][RELAX_BINS] the day's S_{l,b} (forward)
  double* dyn_red;    // [RELAX_BINS + 2 n_dyn RELAX_BINS]: Lambda_b, then G_{l,b}, then H_{l,b} (adjoint day)
  double* g_dyn;      // [RELAX_MAX_PARAMS] the dynamic layers' parameter partials, summed over days
};



__device__ inline float r_log1p(float x) { return log1pf(x); }
__device__ inline double r_log1p(double x) { return log1p(x); }
__device__ inline float r_expm1(float x) { return expm1f(x); }
__device__ inline double r_expm1(double x) { return expm1(x); }
__device__ inline float r_exp(float x) { return expf(x); }
__device__ inline double r_exp(double x) { return exp(x); }

// g(y) = log1p(-min(y, XMAX)) and g'(y) (0 above XMAX)
template <class R>
__device__ inline R relax_g(R y, R xmax) { return r_log1p(-(y < xmax ? y : xmax)); }
template <class R>
__device__ inline R relax_gp(R y, R xmax) { return y < xmax ? (R)1 / (y - (R)1) : (R)0; }

// Block sum of per-thread doubles into out (thread 0 of the block returns it). Fixed order.
__device__ inline double block_sum_d(double v, double* s_w) {
  v = wave_sum_d(v);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  double tot = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < RELAX_BLOCK / 64; ++k) tot += s_w[k];
  __syncthreads();
  return tot;
}

// tr_j(t) = rel_trans_j * sum_{d=1}^{min(kmax, t)} D_j(t-d) w_{bin(j)}(d)
template <class R>
__global__ __launch_bounds__(RELAX_BLOCK) void relax_trans_kernel(RelaxDev<R> D, int32_t t, R* tr) {
  const int64_t j = (int64_t)blockIdx.x * RELAX_BLOCK + threadIdx.x;
  if (j >= D.n) return;
  const int32_t dmax = t < D.kmax ? t : D.kmax;
  const R* wb = D.w + (int64_t)D.bin[j] * (D.kmax + 1);
  R acc = 0;
  for (int32_t d = 1; d <= dmax; ++d) acc += D.delta[(int64_t)(t - d) * D.n + j] * wb[d];
  tr[j] = (R)D.rt[j] * acc;
}

// L_i(t) over the neighbour lists, then D_i(t), s_i(t+1) and the block's share of E_t.
// rec_L / rec_s (may be null): L_i(t) and s_i(t) for the backward sweep.
template <class R>
__global__ __launch_bounds__(RELAX_BLOCK) void relax_forward_kernel(RelaxDev<R> D, RelaxTheta<R> th, int32_t t,
                                                                    const R* tr, R* rec_L, R* rec_s) {
  __shared__ double s_w[RELAX_BLOCK / 64];
  const int64_t i = (int64_t)blockIdx.x * RELAX_BLOCK + threadIdx.x;
  double e_part = 0.0;
  if (i < D.n) {
    const R sig = th.sig[D.bin[i]];
    const R xmax = (R)RELAX_XMAX;
    R L = 0;
    for (int l = 0; l < D.n_layers; ++l) {
      if (D.dyn_slots[l] != (R)0) continue;  // resampled daily: the expected term below
      const R bb = (th.beta * th.bl[l]) * sig;
      const int32_t* nbr = D.nbr[l];
      const int64_t k1 = D.indptr[l][i + 1];
      // RELAX_UNROLL neighbours at a time: their indices, then their tr values, all in flight together;
      // the sum runs in list order as before
      for (int64_t k = D.indptr[l][i]; k < k1; k += RELAX_UNROLL) {
        int32_t j[RELAX_UNROLL];
        R trj[RELAX_UNROLL];
#pragma unroll
        for (int u = 0; u < RELAX_UNROLL; ++u) j[u] = k + u < k1 ? nbr[k + u] : -1;
#pragma unroll
        for (int u = 0; u < RELAX_UNROLL; ++u) trj[u] = j[u] >= 0 ? tr[j[u]] : (R)0;
#pragma unroll
        for (int u = 0; u < RELAX_UNROLL; ++u) {
          if (j[u] < 0) break;
          R x = bb * trj[u];
          x = x < xmax ? x : xmax;
          L += r_log1p(-x);
        }
      }
    }
    for (int k = 0; k < D.n_dyn; ++k) {  // c_l / N (S_{l,b} - g(x_{l,b} tr_i))
      const int l = D.dyn_layer[k];
      const int b = D.bin[i];
      const R own = relax_g<R>(((th.beta * th.bl[l]) * sig) * tr[i], xmax);
      L += D.dyn_slots[l] * ((R)(D.dyn_fsum[k * RELAX_BINS + b] / (double)D.n) - own / (R)D.n);
    }
    const R s0 = D.s[i];
    const R inf = s0 * (-r_expm1(L));
    D.s[i] = s0 * r_exp(L);
    D.delta[(int64_t)t * D.n + i] = (t == 0 ? D.seed_frac : (R)0) + inf;
    if (rec_L) rec_L[i] = L;
    if (rec_s) rec_s[i] = s0;
    e_part = (double)inf;
  }
  const double tot = block_sum_d(e_part, s_w);
  if (threadIdx.x == 0) D.epart[(int64_t)t * D.n_blocks + blockIdx.x] = tot;
}

// Per-bin sums run over RELAX_DYN_ITEMS agents per thread (agent blockIdx.x * RELAX_DYN_SPAN + u * RELAX_BLOCK
// + threadIdx.x for u = 0, 1, ...): each thread sums its agents in registers, then one reduction per group of
// rows (wave sums, then the waves in order) gives the workgroup's partial of every row. Fixed order.
constexpr int RELAX_DYN_ITEMS = 4;
constexpr int RELAX_DYN_SPAN = RELAX_BLOCK * RELAX_DYN_ITEMS;

// part[(row0 + q) * stride + blockIdx.x] = the workgroup's sum of v[q], q < NV.
template <int NV>
__device__ inline void block_rows_store(const double (&v)[NV], double* s_w, double* part, int64_t row0,
                                        int64_t stride) {
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double w = wave_sum_d(v[q]);
    if (lane_id() == 0) s_w[(threadIdx.x >> 6) * NV + q] = w;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double tot = 0.0;
    for (int k = 0; k < RELAX_BLOCK / 64; ++k) tot += s_w[k * NV + threadIdx.x];
    part[(row0 + threadIdx.x) * stride + blockIdx.x] = tot;
  }
  __syncthreads();
}

// The day's S_{l,b} = sum_j g(x_{l,b} tr_j) of every daily-resampled layer: per-workgroup sums (row k * 10 + b).
template <class R>
__global__ __launch_bounds__(RELAX_BLOCK) void relax_dyn_sum_kernel(RelaxDev<R> D, RelaxTheta<R> th, const R* tr) {
  __shared__ double s_w[(RELAX_BLOCK / 64) * RELAX_BINS];
  const R xmax = (R)RELAX_XMAX;
  R trv[RELAX_DYN_ITEMS];
  bool ok[RELAX_DYN_ITEMS];
#pragma unroll
  for (int u = 0; u < RELAX_DYN_ITEMS; ++u) {
    const int64_t j = (int64_t)blockIdx.x * RELAX_DYN_SPAN + u * RELAX_BLOCK + threadIdx.x;
    ok[u] = j < D.n;
    trv[u] = ok[u] ? tr[j] : (R)0;
  }
  for (int k = 0; k < D.n_dyn; ++k) {  // uniform: every thread reaches the barriers
    const int l = D.dyn_layer[k];
    double acc[RELAX_BINS];
#pragma unroll
    for (int b = 0; b < RELAX_BINS; ++b) {
      const R x = (th.beta * th.bl[l]) * th.sig[b];
      double v = 0.0;
#pragma unroll
      for (int u = 0; u < RELAX_DYN_ITEMS; ++u)
        if (ok[u]) v += (double)relax_g<R>(x * trv[u], xmax);
      acc[b] = v;
    }
    block_rows_store<RELAX_BINS>(acc, s_w, D.dyn_part, (int64_t)k * RELAX_BINS, D.n_dyn_blocks);
  }
}

// The adjoint day's sums: Lambda_b (rows 0..9), G_{l,b} (rows 10 + k * 10 + b), H_{l,b} (rows 10 + 10 n_dyn + ...).
template <class R>
__global__ __launch_bounds__(RELAX_BLOCK) void relax_dyn_adj_sum_kernel(RelaxDev<R> D, RelaxTheta<R> th,
                                                                        const R* tr) {
  __shared__ double s_w[(RELAX_BLOCK / 64) * RELAX_BINS];
  const R xmax = (R)RELAX_XMAX;
  int ba[RELAX_DYN_ITEMS];
  R tra[RELAX_DYN_ITEMS], lb[RELAX_DYN_ITEMS];
#pragma unroll
  for (int u = 0; u < RELAX_DYN_ITEMS; ++u) {
    const int64_t a = (int64_t)blockIdx.x * RELAX_DYN_SPAN + u * RELAX_BLOCK + threadIdx.x;
    const bool ok = a < D.n;
    ba[u] = ok ? D.bin[a] : -1;
    tra[u] = ok ? tr[a] : (R)0;
    lb[u] = ok ? D.lbar[a] : (R)0;
  }
  const int64_t nb = D.n_dyn_blocks;
  double acc[RELAX_BINS];
#pragma unroll
  for (int b = 0; b < RELAX_BINS; ++b) {
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < RELAX_DYN_ITEMS; ++u)
      if (ba[u] == b) v += (double)lb[u];
    acc[b] = v;
  }
  block_rows_store<RELAX_BINS>(acc, s_w, D.dyn_part, 0, nb);
  for (int k = 0; k < D.n_dyn; ++k) {
    const int l = D.dyn_layer[k];
    double hsum[RELAX_BINS];
#pragma unroll
    for (int b = 0; b < RELAX_BINS; ++b) {
      const R x = (th.beta * th.bl[l]) * th.sig[b];
      double g = 0.0, h = 0.0;
#pragma unroll
      for (int u = 0; u < RELAX_DYN_ITEMS; ++u) {
        if (ba[u] < 0) continue;
        const R gt = relax_gp<R>(x * tra[u], xmax) * tra[u];
        g += (double)gt;
        if (ba[u] == b) h += (double)(lb[u] * gt);
      }
      acc[b] = g;
      hsum[b] = h;
    }
    block_rows_store<RELAX_BINS>(acc, s_w, D.dyn_part, (int64_t)RELAX_BINS + k * RELAX_BINS, nb);
    block_rows_store<RELAX_BINS>(hsum, s_w, D.dyn_part, (int64_t)RELAX_BINS + (D.n_dyn + k) * RELAX_BINS, nb);
  }
}

// xbar_{l,b} = c_l / N (Lambda_b G_{l,b} - H_{l,b}) into the parameter partials (one thread, fixed order).
template <class R>
__global__ void relax_dyn_grad_kernel(RelaxDev<R> D, RelaxTheta<R> th) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nl = D.n_layers;
  for (int k = 0; k < D.n_dyn; ++k) {
    const int l = D.dyn_layer[k];
    for (int b = 0; b < RELAX_BINS; ++b) {
      const double lam = D.dyn_red[b];
      const double G = D.dyn_red[RELAX_BINS + k * RELAX_BINS + b];
      const double H = D.dyn_red[RELAX_BINS + (D.n_dyn + k) * RELAX_BINS + b];
      const double xbar = (double)D.dyn_slots[l] / (double)D.n * (lam * G - H);
      D.g_dyn[0] += xbar * (double)th.bl[l] * (double)th.sig[b];
      D.g_dyn[1 + l] += xbar * (double)th.beta * (double)th.sig[b];
      D.g_dyn[1 + nl + b] += xbar * (double)th.beta * (double)th.bl[l];
    }
  }
}

// Phase 1 of the adjoint of day t: Dbar, Lbar_i(t), sbar_i(t).
template <class R>
__global__ __launch_bounds__(RELAX_BLOCK) void relax_adj_agent_kernel(RelaxDev<R> D, int32_t t, R ebar, const R* L_t,
                                                                      const R* s_t) {
  const int64_t a = (int64_t)blockIdx.x * RELAX_BLOCK + threadIdx.x;
  if (a >= D.n) return;
  const R* wb = D.w + (int64_t)D.bin[a] * (D.kmax + 1);
  R dbar = ebar;
  const int32_t dmax = std::min(D.kmax, D.n_days - 1 - t);
  for (int32_t d = 1; d <= dmax; ++d) dbar += D.ibar[(int64_t)(t + d) * D.n + a] * wb[d];
  const R L = L_t[a];
  const R e = r_exp(L);
  const R s0 = s_t[a];
  const R sb_next = D.sbar[a];
  D.lbar[a] = (s0 * e) * (sb_next - dbar);
  D.sbar[a] = dbar * (-r_expm1(L)) + sb_next * e;
}

// Between the two phases: what phase 2 reads of every neighbour j, packed in one 4-wide record (one
// gather per neighbour slot instead of three: tr_j, lbar_j and j's age bin).
template <class R>
__global__ __launch_bounds__(RELAX_BLOCK) void relax_pack_kernel(RelaxDev<R> D, RelaxTheta<R> th, const R* tr) {
  typedef R r4 __attribute__((ext_vector_type(4)));
  const int64_t j = (int64_t)blockIdx.x * RELAX_BLOCK + threadIdx.x;
  if (j >= D.n) return;
  const R sig = th.sig[D.bin[j]];
  ((r4*)D.pack)[j] = r4{tr[j], D.lbar[j] * sig, sig, (R)0};
}

// Phase 2 of the adjoint of day t: Ibar_a(t) (source role) and the parameter partials (target role).
template <class R>
__global__ __launch_bounds__(RELAX_BLOCK) void relax_adj_edge_kernel(RelaxDev<R> D, RelaxTheta<R> th, int32_t t,
                                                                     const R* tr) {
  __shared__ double s_w[RELAX_BLOCK / 64];
  const int64_t a = (int64_t)blockIdx.x * RELAX_BLOCK + threadIdx.x;
  const R xmax = (R)RELAX_XMAX;
  double g_beta = 0.0, g_bl[CVK_MAX_LAYERS], g_sig = 0.0;
#pragma unroll
  for (int l = 0; l < CVK_MAX_LAYERS; ++l) g_bl[l] = 0.0;
  int b_a = 0;
  if (a < D.n) {
    b_a = D.bin[a];
    const R sig_a = th.sig[b_a];
    const R tr_a = tr[a];
    const R lbar_a = D.lbar[a];
    R ibar = 0;
    R sum_blc = 0;  // sum_l beta_l C_{a,l}
#pragma unroll
    for (int l = 0; l < CVK_MAX_LAYERS; ++l) {
      if (l >= D.n_layers) break;
      if (D.dyn_slots[l] != (R)0) continue;  // resampled daily: below
      const R bb = th.beta * th.bl[l];
      const R bbs = bb * sig_a;
      R c = 0, src = 0;
      const int32_t* nbr = D.nbr[l];
      const int64_t k1 = D.indptr[l][a + 1];
      for (int64_t k = D.indptr[l][a]; k < k1; k += RELAX_UNROLL) {  // gathers in flight together, sums in order
        typedef R r4 __attribute__((ext_vector_type(4)));
        int32_t jj[RELAX_UNROLL];
        r4 pj[RELAX_UNROLL];
#pragma unroll
        for (int u = 0; u < RELAX_UNROLL; ++u) jj[u] = k + u < k1 ? nbr[k + u] : -1;
#pragma unroll
        for (int u = 0; u < RELAX_UNROLL; ++u) pj[u] = jj[u] >= 0 ? ((const r4*)D.pack)[jj[u]] : r4{0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < RELAX_UNROLL; ++u) {
          if (jj[u] < 0) break;
          const R tr_j = pj[u].x;
          const R x_in = bbs * tr_j;  // j -> a
          if (x_in < xmax) c += tr_j / (x_in - (R)1);
          const R sig_j = pj[u].z;
          const R x_out = (bb * sig_j) * tr_a;  // a -> j
          if (x_out < xmax) src += pj[u].y / (x_out - (R)1);  // (lbar_j * sig_j) / (x_out - 1)
        }
      }
      ibar += bb * src;
      sum_blc += th.bl[l] * c;
      g_bl[l] = (double)(th.beta * lbar_a * sig_a * c);
    }
    for (int k = 0; k < D.n_dyn; ++k) {  // the daily-resampled layers' source role
      const int l = D.dyn_layer[k];
      R acc = 0;
      for (int b = 0; b < RELAX_BINS; ++b) {
        const R x = (th.beta * th.bl[l]) * th.sig[b];
        const R lam = (R)D.dyn_red[b] - (b == b_a ? lbar_a : (R)0);
        acc += x * relax_gp<R>(x * tr_a, xmax) * lam;
      }
      ibar += D.dyn_slots[l] / (R)D.n * acc;
    }
    D.ibar[(int64_t)t * D.n + a] = (R)D.rt[a] * ibar;
    g_beta = (double)(lbar_a * sig_a * sum_blc);
    g_sig = (double)(th.beta * lbar_a * sum_blc);
  }
  // block partials, accumulated over days in this block's column (fixed order, no atomics)
  const int64_t nb = D.n_blocks;
  double tot = block_sum_d(g_beta, s_w);
  if (threadIdx.x == 0) D.gpart[blockIdx.x] += tot;
#pragma unroll
  for (int l = 0; l < CVK_MAX_LAYERS; ++l) {
    if (l < D.n_layers) {  // uniform: every thread reaches the barriers
      tot = block_sum_d(g_bl[l], s_w);
      if (threadIdx.x == 0) D.gpart[(1 + l) * nb + blockIdx.x] += tot;
    }
  }
  for (int b = 0; b < RELAX_BINS; ++b) {
    tot = block_sum_d(b_a == b ? g_sig : 0.0, s_w);
    if (threadIdx.x == 0) D.gpart[(1 + D.n_layers + b) * nb + blockIdx.x] += tot;
  }
}

// Sums `rows` columns of per-block partials (row-major [rows][n_blocks]) into out[rows] (fixed order).
__global__ __launch_bounds__(256) void relax_reduce_kernel(const double* part, int64_t n_blocks, int32_t rows,
                                                           double* out) {
  __shared__ double s_w[4];
  for (int32_t r = blockIdx.x; r < rows; r += gridDim.x) {
    double acc = 0.0;
    for (int64_t b = threadIdx.x; b < n_blocks; b += 256) acc += part[(int64_t)r * n_blocks + b];
    acc = wave_sum_d(acc);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[r] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
    __syncthreads();
  }
}

template <class R>
__global__ void relax_fill_kernel(R* p, int64_t n, R v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct RelaxBase {
  virtual ~RelaxBase() {}
  virtual int forward(const double* theta, double* e_host, hipStream_t st) = 0;
  virtual int backward(const double* theta, const double* ebar_host, double* grad_host, hipStream_t st) = 0;
  int device = -1;
};

template <class R>
struct Relax : RelaxBase {
  RelaxDev<R> D;
  int32_t ckpt = 10;
  std::vector<R*> ck_s;        // s at days 0, ckpt, 2 ckpt, ...
  R* seg_L = nullptr;          // [ckpt][n]
  R* seg_s = nullptr;          // [ckpt][n]
  R* seg_tr = nullptr;         // [ckpt][n]
  R* w = nullptr;
  double* e_dev = nullptr;     // [n_days]
  double* g_dev = nullptr;     // [RELAX_MAX_PARAMS]
  std::vector<void*> owned;
  bool have_forward = false;
  std::vector<double> theta_fwd;

  ~Relax() override {
    for (void* p : owned)
      if (p) (void)hipFree(p);
  }

  template <class T>
  int alloc(T** p, int64_t count) {
    *p = nullptr;
    if (hipMalloc((void**)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T)) != hipSuccess)
      return fail(CVK_E_HIP, "cvk_relax: device allocation failed");
    owned.push_back((void*)*p);
    return CVK_OK;
  }

  RelaxTheta<R> theta_of(const double* th) const {
    RelaxTheta<R> t;
    memset(&t, 0, sizeof(t));
    t.beta = (R)th[0];
    for (int l = 0; l < D.n_layers; ++l) t.bl[l] = (R)th[1 + l];
    for (int b = 0; b < RELAX_BINS; ++b) t.sig[b] = (R)th[1 + D.n_layers + b];
    return t;
  }

  int grid() const { return (int)((D.n + RELAX_BLOCK - 1) / RELAX_BLOCK); }

  int fill(R* p, int64_t n, R v, hipStream_t st) {
    hipLaunchKernelGGL(relax_fill_kernel<R>, dim3(1024), dim3(256), 0, st, p, n, v);
    return check_launch("cvk_relax fill");
  }

  // Runs days [t0, t1) from the current D.s; records L/s/tr of each day into the segment buffers if rec.
  int run_days(const RelaxTheta<R>& th, int32_t t0, int32_t t1, bool rec, R* tr_scratch, hipStream_t st) {
    for (int32_t t = t0; t < t1; ++t) {
      R* tr = rec ? seg_tr + (int64_t)(t - t0) * D.n : tr_scratch;
      hipLaunchKernelGGL(relax_trans_kernel<R>, dim3(grid()), dim3(RELAX_BLOCK), 0, st, D, t, tr);
      if (D.n_dyn) {
        hipLaunchKernelGGL(relax_dyn_sum_kernel<R>, dim3(D.n_dyn_blocks), dim3(RELAX_BLOCK), 0, st, D, th,
                           (const R*)tr);
        hipLaunchKernelGGL(relax_reduce_kernel, dim3(D.n_dyn * RELAX_BINS), dim3(256), 0, st, (const double*)D.dyn_part,
                           (int64_t)D.n_dyn_blocks, D.n_dyn * RELAX_BINS, D.dyn_fsum);
      }
      hipLaunchKernelGGL(relax_forward_kernel<R>, dim3(grid()), dim3(RELAX_BLOCK), 0, st, D, th, t, (const R*)tr,
                         rec ? seg_L + (int64_t)(t - t0) * D.n : (R*)nullptr,
                         rec ? seg_s + (int64_t)(t - t0) * D.n : (R*)nullptr);
    }
    return check_launch("cvk_relax forward");
  }

  int forward(const double* theta, double* e_host, hipStream_t st) override {
    const RelaxTheta<R> th = theta_of(theta);
    int rc = fill(D.s, D.n, (R)1 - D.seed_frac, st);
    if (rc) return rc;
    for (int32_t k = 0; k * ckpt < D.n_days; ++k) {
      const int32_t t0 = k * ckpt, t1 = std::min(D.n_days, t0 + ckpt);
      if (hipMemcpyAsync(ck_s[k], D.s, (size_t)D.n * sizeof(R), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(CVK_E_HIP, "cvk_relax: checkpoint copy failed");
      rc = run_days(th, t0, t1, false, seg_tr, st);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(relax_reduce_kernel, dim3(std::min(D.n_days, 1024)), dim3(256), 0, st, D.epart,
                       (int64_t)D.n_blocks, D.n_days, e_dev);
    rc = check_launch("cvk_relax reduce");
    if (rc) return rc;
    if (hipMemcpyAsync(e_host, e_dev, (size_t)D.n_days * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(CVK_E_HIP, "cvk_relax: copy of E failed");
    have_forward = true;
    theta_fwd.assign(theta, theta + 1 + D.n_layers + RELAX_BINS);
    return CVK_OK;
  }

  int backward(const double* theta, const double* ebar_host, double* grad_host, hipStream_t st) override {
    const int np = 1 + D.n_layers + RELAX_BINS;
    if (!have_forward || !std::equal(theta_fwd.begin(), theta_fwd.end(), theta))
      return fail(CVK_E_INVALID, "cvk_relax_backward: needs a forward with the same parameters first");
    const RelaxTheta<R> th = theta_of(theta);
    int rc = fill(D.sbar, D.n, (R)0, st);
    if (!rc) rc = fill(D.ibar, (int64_t)D.n_days * D.n, (R)0, st);
    if (!rc) {
      hipLaunchKernelGGL(relax_fill_kernel<double>, dim3(256), dim3(256), 0, st, D.gpart,
                         (int64_t)RELAX_MAX_PARAMS * D.n_blocks, 0.0);
      hipLaunchKernelGGL(relax_fill_kernel<double>, dim3(1), dim3(256), 0, st, D.g_dyn, (int64_t)RELAX_MAX_PARAMS, 0.0);
      rc = check_launch("cvk_relax fill");
    }
    if (rc) return rc;
    const int32_t n_seg = (D.n_days + ckpt - 1) / ckpt;
    for (int32_t k = n_seg - 1; k >= 0; --k) {
      const int32_t t0 = k * ckpt, t1 = std::min(D.n_days, t0 + ckpt);
      if (hipMemcpyAsync(D.s, ck_s[k], (size_t)D.n * sizeof(R), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(CVK_E_HIP, "cvk_relax: checkpoint restore failed");
      rc = run_days(th, t0, t1, true, nullptr, st);  // recompute the segment, recording L, s, tr
      if (rc) return rc;
      for (int32_t t = t1 - 1; t >= t0; --t) {
        const int64_t off = (int64_t)(t - t0) * D.n;
        hipLaunchKernelGGL(relax_adj_agent_kernel<R>, dim3(grid()), dim3(RELAX_BLOCK), 0, st, D, t, (R)ebar_host[t],
                           (const R*)(seg_L + off), (const R*)(seg_s + off));
        if (D.n_dyn) {
          const int rows = RELAX_BINS * (1 + 2 * D.n_dyn);
          hipLaunchKernelGGL(relax_dyn_adj_sum_kernel<R>, dim3(D.n_dyn_blocks), dim3(RELAX_BLOCK), 0, st, D, th,
                             (const R*)(seg_tr + off));
          hipLaunchKernelGGL(relax_reduce_kernel, dim3(rows), dim3(256), 0, st, (const double*)D.dyn_part,
                             (int64_t)D.n_dyn_blocks, rows, D.dyn_red);
          hipLaunchKernelGGL(relax_dyn_grad_kernel<R>, dim3(1), dim3(64), 0, st, D, th);
        }
        hipLaunchKernelGGL(relax_pack_kernel<R>, dim3(grid()), dim3(RELAX_BLOCK), 0, st, D, th,
                           (const R*)(seg_tr + off));
        hipLaunchKernelGGL(relax_adj_edge_kernel<R>, dim3(grid()), dim3(RELAX_BLOCK), 0, st, D, th, t,
                           (const R*)(seg_tr + off));
      }
      rc = check_launch("cvk_relax backward");
      if (rc) return rc;
    }
    hipLaunchKernelGGL(relax_reduce_kernel, dim3(np), dim3(256), 0, st, D.gpart, (int64_t)D.n_blocks, np, g_dev);
    rc = check_launch("cvk_relax reduce");
    if (rc) return rc;
    std::vector<double> gd(RELAX_MAX_PARAMS, 0.0);
    if (hipMemcpyAsync(grad_host, g_dev, (size_t)np * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(gd.data(), D.g_dyn, (size_t)np * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(CVK_E_HIP, "cvk_relax: copy of the gradient failed");
    for (int k = 0; k < np; ++k) grad_host[k] += gd[k];  // the daily-resampled layers' partials
    return CVK_OK;
  }
};

template <class R>
static int relax_create(const cvk_relax_desc* d, RelaxBase** out) {
  std::unique_ptr<Relax<R>> r(new Relax<R>());
  RelaxDev<R>& D = r->D;
  memset(&D, 0, sizeof(D));
  D.n = d->n_agents;
  D.n_layers = d->n_layers;
  D.n_days = d->n_days;
  D.kmax = std::min(d->kernel_len - 1, d->n_days);
  for (int l = 0; l < d->n_layers; ++l) {
    D.indptr[l] = d->indptr[l];
    D.nbr[l] = d->nbr[l];
  }
  D.bin = d->age_bin;
  D.rt = d->rel_trans;
  D.seed_frac = (R)d->seed_frac;
  D.n_blocks = (int32_t)((D.n + RELAX_BLOCK - 1) / RELAX_BLOCK);
  D.n_dyn_blocks = (int32_t)((D.n + RELAX_DYN_SPAN - 1) / RELAX_DYN_SPAN);
  D.n_dyn = 0;
  for (int l = 0; l < d->n_layers; ++l) {
    D.dyn_slots[l] = (R)d->dyn_slots[l];
    if (d->dyn_slots[l] > 0.0) D.dyn_layer[D.n_dyn++] = l;
  }
  r->ckpt = d->checkpoint_every;
  int rc = CVK_OK;
  rc = rc ? rc : r->alloc(&r->w, (int64_t)RELAX_BINS * (D.kmax + 1));
  rc = rc ? rc : r->alloc(&D.delta, (int64_t)D.n_days * D.n);
  rc = rc ? rc : r->alloc(&D.s, D.n);
  rc = rc ? rc : r->alloc(&D.ibar, (int64_t)D.n_days * D.n);
  rc = rc ? rc : r->alloc(&D.sbar, D.n);
  rc = rc ? rc : r->alloc(&D.lbar, D.n);
  rc = rc ? rc : r->alloc(&D.epart, (int64_t)D.n_days * D.n_blocks);
  rc = rc ? rc : r->alloc(&D.gpart, (int64_t)RELAX_MAX_PARAMS * D.n_blocks);
  rc = rc ? rc : r->alloc(&D.pack, 4 * D.n);
  rc = rc ? rc : r->alloc(&D.dyn_part, (int64_t)RELAX_BINS * (1 + 2 * std::max(D.n_dyn, 1)) * D.n_blocks);
  rc = rc ? rc : r->alloc(&D.dyn_fsum, (int64_t)RELAX_BINS * std::max(D.n_dyn, 1));
  rc = rc ? rc : r->alloc(&D.dyn_red, (int64_t)RELAX_BINS * (1 + 2 * std::max(D.n_dyn, 1)));
  rc = rc ? rc : r->alloc(&D.g_dyn, RELAX_MAX_PARAMS);
  rc = rc ? rc : r->alloc(&r->seg_L, (int64_t)r->ckpt * D.n);
  rc = rc ? rc : r->alloc(&r->seg_s, (int64_t)r->ckpt * D.n);
  rc = rc ? rc : r->alloc(&r->seg_tr, (int64_t)r->ckpt * D.n);
  rc = rc ? rc : r->alloc(&r->e_dev, D.n_days);
  rc = rc ? rc : r->alloc(&r->g_dev, RELAX_MAX_PARAMS);
  const int32_t n_seg = (D.n_days + r->ckpt - 1) / r->ckpt;
  for (int32_t k = 0; k < n_seg && !rc; ++k) {
    R* p = nullptr;
    rc = r->alloc(&p, D.n);
    r->ck_s.push_back(p);
  }
  if (rc) return rc;
  // infectiousness table w[b][d], d = 0..kmax, converted to R
  std::vector<R> wh((size_t)RELAX_BINS * (D.kmax + 1));
  for (int b = 0; b < RELAX_BINS; ++b)
    for (int dd = 0; dd <= D.kmax; ++dd) wh[(size_t)b * (D.kmax + 1) + dd] = (R)d->w_host[(size_t)b * d->kernel_len + dd];
  if (hipMemcpy(r->w, wh.data(), wh.size() * sizeof(R), hipMemcpyHostToDevice) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_relax_create: upload failed");
  D.w = r->w;
  if (hipGetDevice(&r->device) != hipSuccess) return fail(CVK_E_HIP, "cvk_relax_create: no device");
  *out = r.release();
  return CVK_OK;
}

}  // namespace cvk

using namespace cvk;

extern "C" int cvk_relax_create(const cvk_relax_desc* d, void** handle_out) {
  if (!handle_out) return fail(CVK_E_INVALID, "cvk_relax_create: null handle");
  *handle_out = nullptr;
  if (!d || d->n_agents <= 0 || d->n_agents >= (1ll << 31) || d->n_layers < 0 || d->n_layers > CVK_MAX_LAYERS ||
      d->n_days < 1 || d->kernel_len < 2 || d->checkpoint_every < 1 || !d->w_host || !d->age_bin || !d->rel_trans ||
      d->seed_frac < 0 || d->seed_frac > 1 || (d->precision != 32 && d->precision != 64))
    return fail(CVK_E_INVALID, "cvk_relax_create: invalid descriptor");
  for (int l = 0; l < d->n_layers; ++l) {
    if (!(d->dyn_slots[l] >= 0.0)) return fail(CVK_E_INVALID, "cvk_relax_create: dyn_slots must be >= 0");
    if (d->dyn_slots[l] == 0.0 && (!d->indptr[l] || !d->nbr[l]))
      return fail(CVK_E_INVALID, "cvk_relax_create: null neighbour list");
  }
  RelaxBase* r = nullptr;
  const int rc = d->precision == 64 ? relax_create<double>(d, &r) : relax_create<float>(d, &r);
  if (rc) return rc;
  *handle_out = (void*)r;
  return CVK_OK;
}

extern "C" int cvk_relax_destroy(void* handle) {
  if (!handle) return fail(CVK_E_INVALID, "cvk_relax_destroy: null handle");
  delete (RelaxBase*)handle;
  return CVK_OK;
}

extern "C" int cvk_relax_forward(void* handle, const double* theta_host, double* expected_host, void* stream) {
  if (!handle || !theta_host || !expected_host) return fail(CVK_E_INVALID, "cvk_relax_forward: null argument");
  return ((RelaxBase*)handle)->forward(theta_host, expected_host, (hipStream_t)stream);
}

extern "C" int cvk_relax_backward(void* handle, const double* theta_host, const double* grad_expected_host,
                                  double* grad_theta_host, void* stream) {
  if (!handle || !theta_host || !grad_expected_host || !grad_theta_host)
    return fail(CVK_E_INVALID, "cvk_relax_backward: null argument");
  return ((RelaxBase*)handle)->backward(theta_host, grad_expected_host, grad_theta_host, (hipStream_t)stream);
}
