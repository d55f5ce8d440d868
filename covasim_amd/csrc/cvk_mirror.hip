// 1:1 mirrors of the reference's three numba kernels (covasim/utils.py:39-128),
// used as the parity boundary: same inputs, same float32 operation order, same
// output order. The production day loop (cvk_day.hip) fuses these steps.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cvk_device.h"
#include "cvk_host.h"

namespace cvk {

// utils.py:39-79
__global__ void viral_load_kernel(int32_t t, const float* __restrict__ time_start,
                                  const float* __restrict__ time_recovered, const float* __restrict__ time_dead,
                                  int64_t n, float frac_time, float load_ratio, float high_cap,
                                  float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float dead = time_dead[i];
  const float start = time_start[i];
  const float stop = isnan(dead) ? time_recovered[i] : dead;        // :61-63
  const float total = stop - start;                                   // :66
  const float trans_day = frac_time * total;                          // :67
  const float trans_point = trans_day > high_cap ? high_cap / total : frac_time;  // :68-73
  const bool early = ((float)t - start) / total < trans_point;        // :77
  const float one = 1.0f;
  const float num = load_ratio * (early ? 1.0f : 0.0f) + one * (early ? 0.0f : 1.0f);
  out[i] = num / (one + frac_time * (load_ratio - one));              // :78
}

// utils.py:82-90
__global__ void trans_sus_kernel(const float* __restrict__ rel_trans, const float* __restrict__ rel_sus,
                                 const uint8_t* __restrict__ inf, const uint8_t* __restrict__ sus, float beta_layer,
                                 const float* __restrict__ viral_load, const uint8_t* __restrict__ symp,
                                 const uint8_t* __restrict__ iso, const uint8_t* __restrict__ quar,
                                 float asymp_factor, float iso_factor, float quar_factor,
                                 const float* __restrict__ immunity_factors, int64_t n, float* __restrict__ rt_out,
                                 float* __restrict__ rs_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float fsymp = symp[i] ? 1.0f : 0.0f;
  const float fiso = iso[i] ? 1.0f : 0.0f;
  const float fquar = quar[i] ? 1.0f : 0.0f;
  const float f_asymp = fsymp + (1.0f - fsymp) * asymp_factor;  // :85 (bool + bool*f32)
  const float f_iso = (1.0f - fiso) + fiso * iso_factor;          // :86
  const float f_quar = (1.0f - fquar) + fquar * quar_factor;      // :87
  float rt = rel_trans[i] * (inf[i] ? 1.0f : 0.0f);               // :88, strictly left to right
  rt = rt * f_quar;
  rt = rt * f_asymp;
  rt = rt * f_iso;
  rt = rt * beta_layer;
  rt = rt * viral_load[i];
  float rs = rel_sus[i] * (sus[i] ? 1.0f : 0.0f);                 // :89
  rs = rs * f_quar;
  rs = rs * (1.0f - immunity_factors[i]);
  rt_out[i] = rt;
  rs_out[i] = rs;
}

// utils.py:114-127: slot s in [0, n_dirs*E) is (direction s / E, edge s % E).
struct InfArgs {
  float beta;
  const int32_t* p1;
  const int32_t* p2;
  const float* layer_betas;
  int64_t n_edges;
  const float* rel_trans;
  const float* rel_sus;
  const double* uniforms;
  uint64_t seed;
  int32_t t, variant, layer;
  int64_t n_slots;
};

__device__ inline bool slot_transmits(const InfArgs& a, int64_t s, int32_t& src, int32_t& tgt, float& p) {
  const int dir = s >= a.n_edges ? 1 : 0;
  const int64_t e = dir ? s - a.n_edges : s;
  const int32_t x = a.p1[e], y = a.p2[e];
  src = dir ? y : x;
  tgt = dir ? x : y;
  const float st = a.rel_trans[src];
  if (st == 0.0f) return false;                                        // :116
  p = ((a.beta * a.layer_betas[e]) * st) * a.rel_sus[tgt];             // :117
  if (p == 0.0f) return false;                                         // :118
  double u;
  if (a.uniforms) {
    u = a.uniforms[dir * a.n_edges + e];
  } else {
    const uint64_t ue = (uint64_t)e;
    const u32x4 w = philox_seeded(a.seed, (uint32_t)ue, (uint32_t)a.t,
                                  STREAM_TX | ((uint32_t)a.variant << 8) | (uint32_t)a.layer, (uint32_t)(ue >> 32));
    u = dir ? u01(w.z, w.w) : u01(w.x, w.y);
  }
  return u < (double)p;                                                // :123
}

constexpr int INF_BLOCK = 256;
constexpr int INF_ITEMS = 8;
constexpr int64_t INF_CHUNK = (int64_t)INF_BLOCK * INF_ITEMS;

__global__ __launch_bounds__(INF_BLOCK) void infections_count_kernel(InfArgs a, uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t s_count;
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * INF_CHUNK;
  uint32_t c = 0;
  for (int j = 0; j < INF_ITEMS; ++j) {
    const int64_t s = base + (int64_t)j * INF_BLOCK + threadIdx.x;
    if (s < a.n_slots) {
      int32_t src, tgt;
      float p;
      c += slot_transmits(a, s, src, tgt, p) ? 1u : 0u;
    }
  }
  if (c) atomicAdd(&s_count, c);
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = s_count;
}

// Exclusive scan of block counts in one workgroup (sequential chunks of 1024).
__global__ __launch_bounds__(1024) void scan_counts_kernel(uint32_t* __restrict__ counts, int64_t n_blocks,
                                                           uint64_t* __restrict__ total) {
  __shared__ uint64_t s_vals[1024];
  __shared__ uint64_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n_blocks; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const uint64_t v = i < n_blocks ? counts[i] : 0;
    s_vals[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint64_t add = threadIdx.x >= (unsigned)off ? s_vals[threadIdx.x - off] : 0;
      __syncthreads();
      s_vals[threadIdx.x] += add;
      __syncthreads();
    }
    const uint64_t incl = s_vals[threadIdx.x];
    if (i < n_blocks) counts[i] = (uint32_t)(s_carry + incl - v);  // offsets fit: capacity checked on host
    __syncthreads();
    if (threadIdx.x == 1023) s_carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

__global__ __launch_bounds__(INF_BLOCK) void infections_write_kernel(InfArgs a, const uint32_t* __restrict__ offsets,
                                                                     int64_t capacity, int32_t* __restrict__ src_out,
                                                                     int32_t* __restrict__ tgt_out,
                                                                     float* __restrict__ prob_out) {
  __shared__ uint32_t s_scan[INF_BLOCK];
  __shared__ uint32_t s_base;
  if (threadIdx.x == 0) s_base = offsets[blockIdx.x];
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * INF_CHUNK;
  // Items are visited in slot order: round j covers slots base + j*BLOCK + [0, BLOCK).
  for (int j = 0; j < INF_ITEMS; ++j) {
    const int64_t s = base + (int64_t)j * INF_BLOCK + threadIdx.x;
    int32_t src = 0, tgt = 0;
    float p = 0.0f;
    const bool hit = s < a.n_slots && slot_transmits(a, s, src, tgt, p);
    s_scan[threadIdx.x] = hit ? 1u : 0u;
    __syncthreads();
    for (int off = 1; off < INF_BLOCK; off <<= 1) {
      const uint32_t add = threadIdx.x >= (unsigned)off ? s_scan[threadIdx.x - off] : 0u;
      __syncthreads();
      s_scan[threadIdx.x] += add;
      __syncthreads();
    }
    const uint32_t incl = s_scan[threadIdx.x];
    const uint32_t round_total = s_scan[INF_BLOCK - 1];
    if (hit) {
      const int64_t pos = (int64_t)s_base + incl - 1;
      if (pos < capacity) {
        src_out[pos] = src;
        tgt_out[pos] = tgt;
        if (prob_out) prob_out[pos] = p;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base += round_total;
    __syncthreads();
  }
}

}  // namespace cvk

using namespace cvk;

extern "C" int cvk_viral_load(int32_t t, const float* time_start, const float* time_recovered,
                              const float* time_dead, int64_t n, float frac_time, float load_ratio, float high_cap,
                              float* out, void* stream) {
  if (n < 0 || (n > 0 && (!time_start || !time_recovered || !time_dead || !out)))
    return fail(CVK_E_INVALID, "cvk_viral_load: null pointer or negative n");
  if (n == 0) return CVK_OK;
  const int block = 256;
  const int64_t grid = (n + block - 1) / block;
  hipLaunchKernelGGL(viral_load_kernel, dim3((unsigned)grid), dim3(block), 0, (hipStream_t)stream, t, time_start,
                     time_recovered, time_dead, n, frac_time, load_ratio, high_cap, out);
  return check_launch("cvk_viral_load");
}

extern "C" int cvk_trans_sus(const float* rel_trans, const float* rel_sus, const uint8_t* inf, const uint8_t* sus,
                             float beta_layer, const float* viral_load, const uint8_t* symp, const uint8_t* iso,
                             const uint8_t* quar, float asymp_factor, float iso_factor, float quar_factor,
                             const float* immunity_factors, int64_t n, float* rel_trans_out, float* rel_sus_out,
                             void* stream) {
  if (n < 0 || (n > 0 && (!rel_trans || !rel_sus || !inf || !sus || !viral_load || !symp || !iso || !quar ||
                          !immunity_factors || !rel_trans_out || !rel_sus_out)))
    return fail(CVK_E_INVALID, "cvk_trans_sus: null pointer or negative n");
  if (n == 0) return CVK_OK;
  const int block = 256;
  const int64_t grid = (n + block - 1) / block;
  hipLaunchKernelGGL(trans_sus_kernel, dim3((unsigned)grid), dim3(block), 0, (hipStream_t)stream, rel_trans, rel_sus,
                     inf, sus, beta_layer, viral_load, symp, iso, quar, asymp_factor, iso_factor, quar_factor,
                     immunity_factors, n, rel_trans_out, rel_sus_out);
  return check_launch("cvk_trans_sus");
}

// Arguments of cvk_infections / cvk_infections_noalloc, validated; fills a. Returns CVK_OK, an
// error, or 1 when there is nothing to do (no edges).
static int infections_args(InfArgs& a, float beta, const int32_t* p1, const int32_t* p2, const float* layer_betas,
                           int64_t n_edges, const float* rel_trans, const float* rel_sus, int64_t n_agents,
                           const double* uniforms, uint64_t seed, int32_t t, int32_t variant, int32_t layer,
                           int32_t legacy, const int32_t* src_out, const int32_t* tgt_out, int64_t capacity,
                           int64_t* count_host, const char* name) {
  if (!count_host) return failf(CVK_E_INVALID, "%s: count_host is null", name);
  *count_host = 0;
  if (n_edges < 0 || n_agents < 0 || capacity < 0) return failf(CVK_E_INVALID, "%s: negative size", name);
  if (n_edges == 0) return 1;
  if (!p1 || !p2 || !layer_betas || !rel_trans || !rel_sus || (capacity > 0 && (!src_out || !tgt_out)))
    return failf(CVK_E_INVALID, "%s: null pointer", name);
  if (n_edges >= (1ll << 31)) return failf(CVK_E_INVALID, "%s: more than 2^31-1 edges", name);
  a.beta = beta;
  a.p1 = p1;
  a.p2 = p2;
  a.layer_betas = layer_betas;
  a.n_edges = n_edges;
  a.rel_trans = rel_trans;
  a.rel_sus = rel_sus;
  a.uniforms = uniforms;
  a.seed = seed;
  a.t = t;
  a.variant = variant;
  a.layer = layer;
  a.n_slots = legacy ? n_edges : 2 * n_edges;
  return CVK_OK;
}

// Count, scan, read the total back, write the ordered output. d_counts holds n_blocks block counts
// (overwritten, so it need not start at zero), d_total one 64-bit word.
static int infections_run(const InfArgs& a, int64_t n_blocks, uint32_t* d_counts, uint64_t* d_total,
                          int32_t* src_out, int32_t* tgt_out, float* prob_out, int64_t capacity, int64_t* count_host,
                          hipStream_t st, const char* name) {
  hipLaunchKernelGGL(infections_count_kernel, dim3((unsigned)n_blocks), dim3(INF_BLOCK), 0, st, a, d_counts);
  hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, st, d_counts, n_blocks, d_total);
  uint64_t total = 0;
  if (hipMemcpyAsync(&total, d_total, sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return failf(CVK_E_HIP, "%s: count readback failed", name);
  *count_host = (int64_t)total;
  if ((int64_t)total > capacity) return failf(CVK_E_CAPACITY, "%s: output capacity too small", name);
  if (total > 0)
    hipLaunchKernelGGL(infections_write_kernel, dim3((unsigned)n_blocks), dim3(INF_BLOCK), 0, st, a, d_counts,
                       capacity, src_out, tgt_out, prob_out);
  const int rc = check_launch(name);
  const hipError_t se = hipStreamSynchronize(st);
  if (rc != CVK_OK) return rc;
  return se == hipSuccess ? CVK_OK : failf(CVK_E_HIP, "%s: stream synchronize failed", name);
}

extern "C" int cvk_infections(float beta, const int32_t* p1, const int32_t* p2, const float* layer_betas,
                              int64_t n_edges, const float* rel_trans, const float* rel_sus, int64_t n_agents,
                              const double* uniforms, uint64_t seed, int32_t t, int32_t variant, int32_t layer,
                              int32_t legacy, int32_t* src_out, int32_t* tgt_out, float* prob_out, int64_t capacity,
                              int64_t* count_host, void* stream) {
  InfArgs a;
  int rc = infections_args(a, beta, p1, p2, layer_betas, n_edges, rel_trans, rel_sus, n_agents, uniforms, seed, t,
                           variant, layer, legacy, src_out, tgt_out, capacity, count_host, "cvk_infections");
  if (rc != CVK_OK) return rc == 1 ? CVK_OK : rc;
  const int64_t n_blocks = (a.n_slots + INF_CHUNK - 1) / INF_CHUNK;
  uint32_t* d_counts = nullptr;
  if (hipMalloc((void**)&d_counts, (size_t)n_blocks * sizeof(uint32_t) + 64) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_infections: hipMalloc failed");
  uint64_t* d_total = (uint64_t*)((char*)d_counts + ((n_blocks * sizeof(uint32_t) + 7) / 8) * 8);
  rc = infections_run(a, n_blocks, d_counts, d_total, src_out, tgt_out, prob_out, capacity, count_host,
                      (hipStream_t)stream, "cvk_infections");
  (void)hipFree(d_counts);
  return rc;
}

// ---- the same computation without allocation ----
// Scratch of cvk_infections_noalloc: the per-block counts (CVK_INFECTIONS_NOALLOC_MAX_BLOCKS words),
// then the 64-bit total. Every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` only names the symbol; its device address comes from hipGetSymbolAddress,
// once per device, and only that address is passed to kernels.
static_assert(CVK_INFECTIONS_NOALLOC_MAX_SLOTS == CVK_INFECTIONS_NOALLOC_MAX_BLOCKS * INF_CHUNK,
              "header limit and scratch size disagree");
static __device__ __align__(256) uint32_t scratch[CVK_INFECTIONS_NOALLOC_MAX_BLOCKS + 64];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The current device's address of `scratch` (resolved on the first call on that device).
static int scratch_for_current_device(uint32_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(CVK_E_HIP, "cvk_infections_noalloc: no current device");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return failf(CVK_E_HIP, "cvk_infections_noalloc: hipGetSymbolAddress failed (%s)", hipGetErrorString(e));
    g_scratch_ptr[dev] = (uint32_t*)p;
  }
  *out = g_scratch_ptr[dev];
  return CVK_OK;
}

extern "C" int cvk_infections_noalloc(float beta, const int32_t* p1, const int32_t* p2, const float* layer_betas,
                                      int64_t n_edges, const float* rel_trans, const float* rel_sus,
                                      int64_t n_agents, const double* uniforms, uint64_t seed, int32_t t,
                                      int32_t variant, int32_t layer, int32_t legacy, int32_t* src_out,
                                      int32_t* tgt_out, float* prob_out, int64_t capacity, int64_t* count_host,
                                      void* stream) {
  InfArgs a;
  int rc = infections_args(a, beta, p1, p2, layer_betas, n_edges, rel_trans, rel_sus, n_agents, uniforms, seed, t,
                           variant, layer, legacy, src_out, tgt_out, capacity, count_host, "cvk_infections_noalloc");
  if (rc != CVK_OK) return rc == 1 ? CVK_OK : rc;
  if (a.n_slots > CVK_INFECTIONS_NOALLOC_MAX_SLOTS)
    return failf(CVK_E_CAPACITY, "cvk_infections_noalloc: %lld edge slots exceed the fixed scratch's limit of %lld "
                "(use cvk_infections)",
                (long long)a.n_slots, (long long)CVK_INFECTIONS_NOALLOC_MAX_SLOTS);
  uint32_t* base = nullptr;
  rc = scratch_for_current_device(&base);
  if (rc != CVK_OK) return rc;
  const int64_t n_blocks = (a.n_slots + INF_CHUNK - 1) / INF_CHUNK;
  uint64_t* d_total = (uint64_t*)(base + CVK_INFECTIONS_NOALLOC_MAX_BLOCKS);  // 8-byte aligned: the block is a multiple of 64 words
  return infections_run(a, n_blocks, base, d_total, src_out, tgt_out, prob_out, capacity, count_host,
                        (hipStream_t)stream, "cvk_infections_noalloc");
}
