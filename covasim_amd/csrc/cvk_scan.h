// Three-phase exclusive scan shared by the set-up kernels (population generation, edge moves):
// phase A sums each 2048-element tile, phase B (one workgroup) scans the tile sums, phase C scans
// each tile again and hands every element its exclusive prefix through a sink functor
// sink(i, prefix, value). Sources and sinks are functors, so a scan can fuse the values it counts
// (flags, counts drawn on the fly) and what it writes (offsets, compacted arrays).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cvk_host.h"

namespace cvk {
namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;  // 2048 elements per scan workgroup

__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t* s_wave, int64_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  int64_t base = 0, all = 0;
  for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
    if (w < wave) base += s_wave[w];
    all += s_wave[w];
  }
  __syncthreads();
  *total = all;
  return base + x - v;
}

template <class Src>
__global__ __launch_bounds__(SCAN_BLOCK) void scan_tile_sum_kernel(Src src, int64_t n, int64_t* tile_sum) {
  __shared__ int64_t s_wave[SCAN_BLOCK / 64];
  const int64_t i0 = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t acc = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (i0 + k < n) acc += src(i0 + k);
  int64_t total;
  block_exclusive_scan(acc, s_wave, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// One workgroup: tile_sum[0..n_tiles) -> exclusive offsets in place; *total = the sum.
__global__ __launch_bounds__(SCAN_BLOCK) void scan_tile_offsets_kernel(int64_t* tile_sum, int64_t n_tiles,
                                                                      int64_t* total) {
  __shared__ int64_t s_wave[SCAN_BLOCK / 64];
  int64_t carry = 0;
  for (int64_t c = 0; c < n_tiles; c += SCAN_BLOCK) {
    const int64_t i = c + threadIdx.x;
    const int64_t v = i < n_tiles ? tile_sum[i] : 0;
    int64_t chunk;
    const int64_t ex = block_exclusive_scan(v, s_wave, &chunk);
    if (i < n_tiles) tile_sum[i] = carry + ex;
    carry += chunk;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class Src, class Sink>
__global__ __launch_bounds__(SCAN_BLOCK) void scan_apply_kernel(Src src, Sink sink, int64_t n,
                                                               const int64_t* tile_off) {
  __shared__ int64_t s_wave[SCAN_BLOCK / 64];
  const int64_t i0 = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t v[SCAN_ITEMS];
  int64_t acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = i0 + k < n ? src(i0 + k) : 0;
    acc += v[k];
  }
  int64_t total;
  int64_t run = tile_off[blockIdx.x] + block_exclusive_scan(acc, s_wave, &total);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (i0 + k < n) sink(i0 + k, run, v[k]);
    run += v[k];
  }
}

inline int64_t scan_tiles(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

template <class Src, class Sink>
int scan(Src src, Sink sink, int64_t n, int64_t* tile_scratch, int64_t* total, hipStream_t st) {
  const int64_t tiles = scan_tiles(n);
  if (tiles == 0) {
    if (total && hipMemsetAsync(total, 0, sizeof(int64_t), st) != hipSuccess)
      return fail(CVK_E_HIP, "population scan: hipMemsetAsync failed");
    return CVK_OK;
  }
  hipLaunchKernelGGL((scan_tile_sum_kernel<Src>), dim3((uint32_t)tiles), dim3(SCAN_BLOCK), 0, st, src, n, tile_scratch);
  hipLaunchKernelGGL(scan_tile_offsets_kernel, dim3(1), dim3(SCAN_BLOCK), 0, st, tile_scratch, tiles, total);
  hipLaunchKernelGGL((scan_apply_kernel<Src, Sink>), dim3((uint32_t)tiles), dim3(SCAN_BLOCK), 0, st, src, sink, n,
                     (const int64_t*)tile_scratch);
  return check_launch("scan");
}

}  // namespace
}  // namespace cvk
