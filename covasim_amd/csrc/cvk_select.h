// Uniform choice without replacement by keys: item i gets the 64-bit key (Philox word << 32) | i,
// distinct for every item, and the chosen items are the eligible ones whose key is at most the k-th
// smallest eligible key. The k-th smallest key comes from an 8-pass radix select (one 8-bit digit per
// pass: a 256-bucket histogram of the eligible keys that share the digits found so far, then one
// workgroup picks the digit where the count reaches k). With fewer than k eligible items the
// threshold ends as all ones and every eligible item is chosen.
//
// KeyFn: __device__ bool operator()(int64_t i, uint64_t& key) const -- false: item i is not eligible.
// Scratch: state = 2 words (digits found so far, keys still to take below them), hist = 256 words.
#pragma once

#include <algorithm>

#include "cvk_device.h"

namespace cvk {
namespace {  // each translation unit that includes this has its own kernels

constexpr int SELECT_BLOCK = 256;

// The 64-bit key of item i in stream stream_id on day t (items up to 2^32: counter word 3 holds i's high bits).
__device__ inline uint64_t item_key(uint64_t seed, int32_t t, uint32_t stream_id, int64_t i) {
  const u32x4 w = philox_seeded(seed, (uint32_t)i, (uint32_t)t, stream_id, (uint32_t)((uint64_t)i >> 32));
  return ((uint64_t)w.x << 32) | (uint32_t)i;
}

struct AllItemsKey {  // every item eligible
  uint64_t seed;
  int32_t t;
  uint32_t stream_id;
  __device__ bool operator()(int64_t i, uint64_t& key) const {
    key = item_key(seed, t, stream_id, i);
    return true;
  }
};

template <class KeyFn>
__global__ __launch_bounds__(SELECT_BLOCK) void select_hist_kernel(KeyFn f, int64_t n, int pass, const uint64_t* state,
                                                                   uint32_t* hist) {
  __shared__ uint32_t s_hist[256];
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const int shift = 56 - 8 * pass;
  const uint64_t prefix = state[0];
  const uint64_t high_mask = pass == 0 ? 0ull : ~0ull << (64 - 8 * pass);
  for (int64_t i = (int64_t)blockIdx.x * SELECT_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * SELECT_BLOCK) {
    uint64_t key;
    if (f(i, key) && (key & high_mask) == prefix) atomicAdd(&s_hist[(key >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (s_hist[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_hist[threadIdx.x]);
}

// One workgroup: the digit whose bucket holds the remaining k-th key; clears the histogram.
__global__ __launch_bounds__(SELECT_BLOCK) void select_digit_kernel(int pass, uint64_t* state, uint32_t* hist) {
  if (threadIdx.x == 0) {
    uint64_t remain = state[1];
    uint32_t digit = 255;
    for (uint32_t b = 0; b < 256; ++b) {
      if (remain <= hist[b]) {
        digit = b;
        break;
      }
      remain -= hist[b];
    }
    state[0] |= (uint64_t)digit << (56 - 8 * pass);
    state[1] = remain;
  }
  __syncthreads();
  hist[threadIdx.x] = 0;
}

__global__ void select_init_kernel(uint64_t* state, uint32_t* hist, uint64_t k) {
  if (threadIdx.x == 0) {
    state[0] = 0;
    state[1] = k;
  }
  hist[threadIdx.x] = 0;
}

// The k-th smallest eligible key of items [0, n) into state[0] (k >= 1); 17 launches on stream st.
template <class KeyFn>
inline void select_kth(KeyFn f, int64_t n, int64_t k, uint64_t* state, uint32_t* hist, hipStream_t st) {
  hipLaunchKernelGGL(select_init_kernel, dim3(1), dim3(SELECT_BLOCK), 0, st, state, hist, (uint64_t)k);
  const uint32_t grid = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((n + SELECT_BLOCK - 1) / SELECT_BLOCK, 4096));
  for (int pass = 0; pass < 8; ++pass) {
    hipLaunchKernelGGL(select_hist_kernel<KeyFn>, dim3(grid), dim3(SELECT_BLOCK), 0, st, f, n, pass,
                       (const uint64_t*)state, hist);
    hipLaunchKernelGGL(select_digit_kernel, dim3(1), dim3(SELECT_BLOCK), 0, st, pass, state, hist);
  }
}

}  // namespace
}  // namespace cvk
