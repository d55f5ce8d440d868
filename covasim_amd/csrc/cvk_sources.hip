// Source lists of a static contact layer (cvk_sources): both directions of every stored pair,
// grouped by source agent, so that the edge pass can walk the contacts of the day's infectious agents
// instead of streaming the whole layer. Built in three steps: per-agent degrees (atomic counts), an
// exclusive scan into the offsets, and a scatter of every slot to its source's next free position.
// Self-loops and edges with an endpoint outside [0, n) are left out: an infectious agent is never
// susceptible, and such an endpoint indexes no agent.
#include <algorithm>

#include "cvk_device.h"
#include "cvk_host.h"
#include "cvk_scan.h"

namespace cvk {
namespace {

constexpr int SRCB_BLOCK = 256;

__global__ __launch_bounds__(SRCB_BLOCK) void zero_i32_kernel(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * SRCB_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * SRCB_BLOCK) p[i] = 0;
}

__device__ inline bool slot_edge(const int32_t* p1, const int32_t* p2, int64_t e, int64_t n, int32_t& x, int32_t& y) {
  x = p1[e];
  y = p2[e];
  return x != y && x >= 0 && y >= 0 && x < n && y < n;
}

__global__ __launch_bounds__(SRCB_BLOCK) void degree_kernel(const int32_t* p1, const int32_t* p2, int64_t n_edges,
                                                            int64_t n, int32_t* deg) {
  for (int64_t e = (int64_t)blockIdx.x * SRCB_BLOCK + threadIdx.x; e < n_edges; e += (int64_t)gridDim.x * SRCB_BLOCK) {
    int32_t x, y;
    if (!slot_edge(p1, p2, e, n, x, y)) continue;
    atomicAdd(deg + x, 1);
    atomicAdd(deg + y, 1);
  }
}

struct DegSrc {
  const int32_t* deg;
  __device__ int64_t operator()(int64_t i) const { return deg[i]; }
};

// offsets into indptr, and the same offsets into deg, which becomes each agent's cursor
struct OffsetSink {
  int32_t* indptr;
  int32_t* cursor;
  __device__ void operator()(int64_t i, int64_t prefix, int64_t) const {
    indptr[i] = (int32_t)prefix;
    cursor[i] = (int32_t)prefix;
  }
};

__global__ void finish_kernel(const int64_t* total, int32_t* indptr, int64_t n) { indptr[n] = (int32_t)*total; }

__global__ __launch_bounds__(SRCB_BLOCK) void scatter_kernel(const int32_t* p1, const int32_t* p2, int64_t n_edges,
                                                             int64_t n, int32_t* cursor, int32_t* target,
                                                             uint32_t* edge_dir) {
  for (int64_t e = (int64_t)blockIdx.x * SRCB_BLOCK + threadIdx.x; e < n_edges; e += (int64_t)gridDim.x * SRCB_BLOCK) {
    int32_t x, y;
    if (!slot_edge(p1, p2, e, n, x, y)) continue;
    const int32_t i = atomicAdd(cursor + x, 1);
    target[i] = y;
    edge_dir[i] = (uint32_t)e << 1;
    const int32_t j = atomicAdd(cursor + y, 1);
    target[j] = x;
    edge_dir[j] = ((uint32_t)e << 1) | 1u;
  }
}

inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

}  // namespace
}  // namespace cvk

using namespace cvk;

extern "C" int64_t cvk_sources_work_bytes(int64_t n_agents) {
  if (n_agents < 0) return 0;
  return align256(4 * n_agents) + align256(8 * (scan_tiles(n_agents) + 1)) + align256(8);
}

extern "C" int cvk_sources_build(const cvk_layer* layer, int64_t n_agents, int32_t* indptr, int32_t* target,
                                 uint32_t* edge_dir, int64_t capacity, int64_t* n_slots_out, void* work,
                                 int64_t work_bytes, void* stream) {
  if (!n_slots_out) return fail(CVK_E_INVALID, "cvk_sources_build: n_slots_out is null");
  *n_slots_out = 0;
  if (!layer || layer->n_edges < 0 || (layer->n_edges > 0 && (!layer->p1 || !layer->p2)))
    return fail(CVK_E_INVALID, "cvk_sources_build: invalid layer");
  if (n_agents <= 0 || n_agents >= (1ll << 31) || !indptr)
    return fail(CVK_E_INVALID, "cvk_sources_build: need 0 < n_agents < 2^31 and an indptr array");
  if (layer->n_edges >= (1ll << 30))
    return fail(CVK_E_INVALID, "cvk_sources_build: more than 2^30 - 1 edges (slot offsets are int32)");
  if (capacity < 2 * layer->n_edges || (layer->n_edges > 0 && (!target || !edge_dir)))
    return fail(CVK_E_CAPACITY, "cvk_sources_build: target / edge_dir need room for 2 * n_edges slots");
  if (!work || ((uintptr_t)work & 255)) return fail(CVK_E_ALIGN, "cvk_sources_build: workspace must be 256-byte aligned");
  if (work_bytes < cvk_sources_work_bytes(n_agents)) return fail(CVK_E_CAPACITY, "cvk_sources_build: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  char* base = (char*)work;
  int32_t* deg = (int32_t*)base;
  int64_t* tiles = (int64_t*)(base + align256(4 * n_agents));
  int64_t* total = (int64_t*)(base + align256(4 * n_agents) + align256(8 * (scan_tiles(n_agents) + 1)));
  const int64_t n = n_agents, E = layer->n_edges;
  const uint32_t ga = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((n + SRCB_BLOCK - 1) / SRCB_BLOCK, 4096));
  const uint32_t ge = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((E + SRCB_BLOCK - 1) / SRCB_BLOCK, 4096));
  hipLaunchKernelGGL(zero_i32_kernel, dim3(ga), dim3(SRCB_BLOCK), 0, st, deg, n);
  if (E > 0) hipLaunchKernelGGL(degree_kernel, dim3(ge), dim3(SRCB_BLOCK), 0, st, layer->p1, layer->p2, E, n, deg);
  int rc = scan(DegSrc{deg}, OffsetSink{indptr, deg}, n, tiles, total, st);
  if (rc) return rc;
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(1), 0, st, (const int64_t*)total, indptr, n);
  if (E > 0)
    hipLaunchKernelGGL(scatter_kernel, dim3(ge), dim3(SRCB_BLOCK), 0, st, layer->p1, layer->p2, E, n, deg, target,
                       edge_dir);
  rc = check_launch("cvk_sources_build");
  if (rc) return rc;
  int64_t slots = 0;
  if (hipMemcpyAsync(&slots, total, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_sources_build: slot count read-back failed");
  *n_slots_out = slots;
  return CVK_OK;
}
