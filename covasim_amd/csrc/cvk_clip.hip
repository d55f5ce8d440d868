// Edge moves of clip_edges (covasim/interventions.py:589-673): k of a layer's n edges, chosen
// uniformly without replacement, leave it for another layer (the intervention's store, or back).
// The same choice drives the partial resample of a dynamic layer (Layer.update with frac < 1,
// covasim/base.py:1866-1875): the chosen edges get new iid uniform endpoints and weight 1.
//
// Choice: edge e gets the 64-bit key (Philox word << 32) | e, distinct for every edge; the moved
// edges are those whose key is at most the k-th smallest (radix select, cvk_select.h). The edges then
// split by one exclusive scan of the selection flags: kept edges in their order into keep_*, moved
// edges in their order into move_*.
#include <algorithm>
#include <string.h>

#include "cvk_device.h"
#include "cvk_host.h"
#include "cvk_kernels.h"
#include "cvk_scan.h"
#include "cvk_select.h"

namespace cvk {
namespace {

constexpr int CLIP_BLOCK = 256;

// stream_id: STREAM_CLIP | slot (clip_edges) or STREAM_PARTIAL | layer (partial resample)
__device__ inline uint64_t edge_key(uint64_t seed, int32_t t, uint32_t stream_id, int64_t e) {
  return item_key(seed, t, stream_id, e);
}

// The k-th smallest key of edges [0, n) into state[0] (k >= 1).
inline void select_edges(uint64_t seed, int32_t t, uint32_t stream_id, int64_t n, int64_t k, uint64_t* state,
                         uint32_t* hist, hipStream_t st) {
  select_kth(AllItemsKey{seed, t, stream_id}, n, k, state, hist, st);
}

struct Chosen {
  uint64_t seed;
  int32_t t;
  uint32_t stream_id;
  const uint64_t* kth;  // the k-th smallest key (all keys <= it are chosen); null: nothing chosen
  __device__ int64_t operator()(int64_t e) const { return kth && edge_key(seed, t, stream_id, e) <= *kth ? 1 : 0; }
};

// Partial resample: every chosen edge (key <= *kth) gets the endpoints the full resample of day t
// gives it (dynamic_endpoints) and weight 1; the other edges are left as they are.
__global__ __launch_bounds__(CLIP_BLOCK) void partial_resample_kernel(uint64_t seed, int32_t t, uint32_t layer,
                                                                      uint32_t n_agents, int64_t n,
                                                                      const uint64_t* kth, int32_t* p1, int32_t* p2,
                                                                      float* beta) {
  const uint64_t thr = *kth;
  for (int64_t e = (int64_t)blockIdx.x * CLIP_BLOCK + threadIdx.x; e < n; e += (int64_t)gridDim.x * CLIP_BLOCK) {
    if (edge_key(seed, t, STREAM_PARTIAL | layer, e) > thr) continue;
    const EdgeEnds d = dynamic_endpoints(seed, t, layer, n_agents, e);
    p1[e] = d.p1;
    p2[e] = d.p2;
    beta[e] = 1.0f;
  }
}

struct SplitEdges {
  const int32_t *p1, *p2;
  const float* beta;
  int32_t *kp1, *kp2, *mp1, *mp2;
  float *kbeta, *mbeta;
  __device__ void operator()(int64_t e, int64_t moved_before, int64_t chosen) const {
    if (chosen) {
      mp1[moved_before] = p1[e];
      mp2[moved_before] = p2[e];
      mbeta[moved_before] = beta[e];
    } else {
      const int64_t j = e - moved_before;
      kp1[j] = p1[e];
      kp2[j] = p2[e];
      kbeta[j] = beta[e];
    }
  }
};

inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

}  // namespace
}  // namespace cvk

using namespace cvk;

extern "C" int64_t cvk_move_edges_work_bytes(int64_t n) {
  if (n < 0) return 0;
  return align256(16) + align256(256 * 4) + align256(8 * (scan_tiles(n) + 1));
}

extern "C" int cvk_move_edges(const int32_t* p1, const int32_t* p2, const float* beta, int64_t n, int64_t k,
                              uint64_t seed, int32_t t, int32_t slot, int32_t* keep_p1, int32_t* keep_p2,
                              float* keep_beta, int32_t* move_p1, int32_t* move_p2, float* move_beta, void* work,
                              int64_t work_bytes, void* stream) {
  if (n < 0 || k < 0 || k > n || n > (1ll << 32) || slot < 0 || slot > 255)
    return fail(CVK_E_INVALID, "cvk_move_edges: need 0 <= k <= n <= 2^32 and a slot in 0..255");
  if (n > 0 && (!p1 || !p2 || !beta)) return fail(CVK_E_INVALID, "cvk_move_edges: null layer column");
  if (n - k > 0 && (!keep_p1 || !keep_p2 || !keep_beta)) return fail(CVK_E_INVALID, "cvk_move_edges: null keep column");
  if (k > 0 && (!move_p1 || !move_p2 || !move_beta)) return fail(CVK_E_INVALID, "cvk_move_edges: null move column");
  if (!work || ((uintptr_t)work & 255)) return fail(CVK_E_ALIGN, "cvk_move_edges: workspace must be 256-byte aligned");
  if (work_bytes < cvk_move_edges_work_bytes(n)) return fail(CVK_E_CAPACITY, "cvk_move_edges: workspace too small");
  if (n == 0) return CVK_OK;
  hipStream_t st = (hipStream_t)stream;
  char* base = (char*)work;
  uint64_t* state = (uint64_t*)base;
  uint32_t* hist = (uint32_t*)(base + align256(16));
  int64_t* tiles = (int64_t*)(base + align256(16) + align256(256 * 4));
  const uint32_t sid = STREAM_CLIP | (uint32_t)slot;
  if (k > 0) select_edges(seed, t, sid, n, k, state, hist, st);
  SplitEdges split{p1, p2, beta, keep_p1, keep_p2, move_p1, move_p2, keep_beta, move_beta};
  return scan(Chosen{seed, t, sid, k > 0 ? (const uint64_t*)state : nullptr}, split, n, tiles, nullptr, st);
}

extern "C" int64_t cvk_resample_frac_work_bytes(void) { return align256(16) + align256(256 * 4); }

extern "C" int cvk_resample_layer_frac(const cvk_layer* layer, int32_t layer_index, const cvk_params* params,
                                       int32_t t, int64_t k, void* work, int64_t work_bytes, void* stream) {
  if (!layer || !params || layer->n_edges < 0 || (layer->n_edges > 0 && (!layer->p1 || !layer->p2 || !layer->beta)))
    return fail(CVK_E_INVALID, "cvk_resample_layer_frac: invalid layer");
  if (layer_index < 0 || layer_index >= CVK_MAX_LAYERS)
    return fail(CVK_E_INVALID, "cvk_resample_layer_frac: bad index");
  if (params->n_agents <= 0 || params->n_agents >= (1ll << 31))
    return fail(CVK_E_INVALID, "cvk_resample_layer_frac: bad n_agents");
  if (k < 0 || k > layer->n_edges || layer->n_edges > (1ll << 32))
    return fail(CVK_E_INVALID, "cvk_resample_layer_frac: need 0 <= k <= n_edges <= 2^32");
  if (k == 0) return CVK_OK;
  if (!work || ((uintptr_t)work & 255))
    return fail(CVK_E_ALIGN, "cvk_resample_layer_frac: workspace must be 256-byte aligned");
  if (work_bytes < cvk_resample_frac_work_bytes())
    return fail(CVK_E_CAPACITY, "cvk_resample_layer_frac: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  uint64_t* state = (uint64_t*)work;
  uint32_t* hist = (uint32_t*)((char*)work + align256(16));
  const uint32_t l = (uint32_t)layer_index;
  select_edges(params->seed, t, STREAM_PARTIAL | l, layer->n_edges, k, state, hist, st);
  const uint32_t grid =
      (uint32_t)std::max<int64_t>(1, std::min<int64_t>((layer->n_edges + CLIP_BLOCK - 1) / CLIP_BLOCK, 4096));
  hipLaunchKernelGGL(partial_resample_kernel, dim3(grid), dim3(CLIP_BLOCK), 0, st, params->seed, t, l,
                     (uint32_t)params->n_agents, layer->n_edges, (const uint64_t*)state, layer->p1, layer->p2,
                     layer->beta);
  return check_launch("cvk_resample_layer_frac");
}
