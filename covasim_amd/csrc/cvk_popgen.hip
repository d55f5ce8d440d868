// Population generation on the device: ages, households, random contact layers and the
// per-agent prognosis draws (covasim/population.py:143-364, people.py:139-160,
// utils.py:409-426). Every draw is a Philox4x32-10 word addressed by (index, stream), so the
// population does not depend on launch shapes and oracle/cport/popgen_port.c reproduces it
// bit for bit. The draw rules are stated next to cvk_popgen in include/covasim_amd.h.
//
// Building blocks: a three-phase exclusive scan (tile sums, one-workgroup scan of the tile sums,
// tile-local scan into a caller-chosen sink), an 8-bit stable counting sort (LDS histograms per
// 4096-agent tile, scan in bucket-major order, wave-ballot ranking for the scatter) and
// thread-per-household / thread-per-member / thread-per-four-edges fill kernels.
// Population building runs once before the day loop; none of these kernels is on the per-day path.
#include <algorithm>
#include <string.h>
#include <vector>

#include "cvk_device.h"
#include "cvk_host.h"
#include "cvk_scan.h"

namespace cvk {
namespace {

constexpr int POP_BLOCK = 256;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_TILE = POP_BLOCK * SORT_ROUNDS;  // 4096 agents per sort workgroup
constexpr int N_BUCKETS = 256;
constexpr int64_t COUNT_KMAX = 1 << 16;
constexpr int GUIDE = 4096;  // buckets of the guide table over a count CDF
constexpr int64_t POP_MAX_N = (1ll << 31) - 1;

// ---------------------------------------------------------------------------------------------
// Draws (host + device, identical arithmetic; compiled with -ffp-contract=off)
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline int table_index(const double* cdf, int n, double u) {
  int k = 0;
  for (int j = 0; j < n; ++j) k += cdf[j] < u ? 1 : 0;
  return k < n ? k : n - 1;
}

struct AgeDraw {
  float age;
  uint8_t key;
};

__host__ __device__ inline AgeDraw draw_age(const cvk_popgen& g, uint32_t i) {
  const u32x4 w = philox_seeded(g.seed, i, 0u, STREAM_POP_AGE, 0u);
  const int bin = table_index(g.age_cdf, g.n_age_bins, u01(w.x, w.y));
  const double age = g.age_lo[bin] + g.age_width[bin] * u01(w.z, w.w);
  const u32x4 z = philox_seeded(g.seed, i, 0u, STREAM_POP_AGE_NOISE, 0u);
  const double s = ((double)z.x + (double)z.y + (double)z.z + (double)z.w) * (1.0 / 4294967296.0);
  double k = floor(age + (s - 2.0) * g.age_noise) + 64.0;
  k = k < 0.0 ? 0.0 : (k > 255.0 ? 255.0 : k);
  return AgeDraw{(float)age, (uint8_t)k};
}

__host__ __device__ inline int32_t draw_hh_size(const cvk_popgen& g, int64_t h) {
  const u32x4 w = philox_seeded(g.seed, (uint32_t)h, (uint32_t)((uint64_t)h >> 32), STREAM_POP_HH_SIZE, 0u);
  return g.hh_size_min + table_index(g.hh_cdf, g.n_hh_sizes, u01(w.x, w.y));
}

// Half-width in bits of the Feistel domain that covers [0, H).
__host__ __device__ inline uint32_t feistel_bits(int64_t H) {
  uint32_t nb = 0;
  for (uint64_t x = H > 1 ? (uint64_t)(H - 1) : 0; x; x >>= 1) ++nb;
  const uint32_t b = (nb + 1) >> 1;
  return b < 1 ? 1 : b;
}

__host__ __device__ inline uint32_t feistel_once(uint64_t seed, uint32_t x, uint32_t b) {
  const uint32_t mask = (1u << b) - 1u;
  uint32_t L = x >> b, R = x & mask;
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t F = philox_seeded(seed, R, r, STREAM_POP_HH_PERM, 0u).x & mask;
    const uint32_t nl = R;
    R = L ^ F;
    L = nl;
  }
  return (L << b) | R;
}

// pi(h): cycle-walk the 2b-bit Feistel permutation until it lands below H (always terminates:
// the cycle through h returns to h).
__host__ __device__ inline int64_t hh_position(uint64_t seed, int64_t h, int64_t H, uint32_t b) {
  uint32_t y = (uint32_t)h;
  do {
    y = feistel_once(seed, y, b);
  } while ((int64_t)y >= H);
  return (int64_t)y;
}

// round(k / 2) with ties to even, as numpy's round of k / 2.0.
__host__ __device__ inline int64_t half_round_even(int64_t k) {
  const int64_t m = k >> 1;
  return m + ((k & 1) & (m & 1));
}

// Lemire's multiply-and-reject onto [0, n); thr = 2^32 mod n.
__host__ __device__ inline uint32_t partner(uint64_t seed, uint32_t slot, int64_t e, uint32_t word, uint32_t n,
                                            uint32_t thr) {
  uint64_t m = (uint64_t)word * n;
  if ((uint32_t)m < thr) {
    for (uint32_t k = 0; k < RESAMPLE_MAX_RETRY; ++k) {
      const u32x4 r = philox_seeded(seed, (uint32_t)e, (uint32_t)((uint64_t)e >> 32), STREAM_POP_EDGE_RETRY | slot, k);
      m = (uint64_t)r.x * n;
      if ((uint32_t)m >= thr) break;
    }
  }
  return (uint32_t)(m >> 32);
}

struct ArrayI32 {
  const int32_t* a;
  __device__ int64_t operator()(int64_t i) const { return a[i]; }
};
struct ArrayI64 {
  const int64_t* a;
  __device__ int64_t operator()(int64_t i) const { return a[i]; }
};
struct WriteI64 {
  int64_t* out;
  __device__ void operator()(int64_t i, int64_t prefix, int64_t) const { out[i] = prefix; }
};
struct BandFlag {
  const float* age;
  double lo, hi;
  __device__ int64_t operator()(int64_t i) const {
    const double a = (double)age[i];
    return (a >= lo && a < hi) ? 1 : 0;
  }
};
struct WriteMember {
  int32_t* mapping;
  __device__ void operator()(int64_t i, int64_t prefix, int64_t v) const {
    if (v) mapping[prefix] = (int32_t)i;
  }
};
// Count drawn by inversion: the smallest k whose CDF value F[k] reaches u, as a walk k = 0, 1, ..
// along the pmf recurrence would find it (capped at COUNT_KMAX). F[0..K] is the table of
// count_table; the search is a binary search between the guide entries of u's bucket, and u
// above F[K] gives COUNT_KMAX (F is non-decreasing, and past K it no longer changes).
__device__ inline int64_t count_lookup(const double* F, const int32_t* G, int64_t K, double u) {
  const int b = (int)(u * (double)GUIDE);
  int64_t lo = G[b], hi = G[b + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (F[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  return lo > K ? COUNT_KMAX : lo;
}

struct MemberEdges {
  uint64_t seed;
  uint32_t slot;
  const double* F;
  const int32_t* G;
  const int64_t* K;
  __device__ int64_t operator()(int64_t m) const {
    const u32x4 w = philox_seeded(seed, (uint32_t)m, (uint32_t)((uint64_t)m >> 32), STREAM_POP_COUNT | slot, 0u);
    return half_round_even(count_lookup(F, G, *K, u01(w.x, w.y)));
  }
};

// ---------------------------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------------------------
struct PopWork {
  int32_t* order;      // [n] agent at each sorted position (correlate)
  int64_t* hist;       // [256][sort tiles]
  int32_t* size;       // [h_cap] drawn household sizes
  int64_t* start;      // [h_cap] first sorted position of household h
  int32_t* perm;       // [h_cap] pi(h)
  int32_t* at;         // [h_cap] household at position j (inverse of perm)
  int64_t* new_start;  // [h_cap + 1] size of the household at position j, scanned in place to its first agent
  int64_t* pair_start; // [h_cap + 1] edges of the household at position j, scanned in place to its first edge
  double* cdf;         // [COUNT_KMAX + 1] CDF table of a count distribution
  int32_t* guide;      // [GUIDE + 1] first table entry reaching j / GUIDE
  int64_t* tiles;      // scan tile sums
  int64_t* ctl;        // [8]: 0 = H, 1 = agents covered by h_cap slots, 2 = household edges, 3 = CDF table end
};

inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }
inline int64_t sort_tiles(int64_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }

int64_t work_bytes(int64_t n, int64_t h_cap, PopWork* w, char* base) {
  const int64_t longest = std::max<int64_t>(std::max<int64_t>(n + 1, h_cap + 1), N_BUCKETS * sort_tiles(n));
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  char* order = take(4 * std::max<int64_t>(n, 1));
  char* hist = take(8 * std::max<int64_t>(N_BUCKETS * sort_tiles(n), 1));
  char* size = take(4 * std::max<int64_t>(h_cap, 1));
  char* start = take(8 * std::max<int64_t>(h_cap, 1));
  char* perm = take(4 * std::max<int64_t>(h_cap, 1));
  char* at = take(4 * std::max<int64_t>(h_cap, 1));
  char* new_start = take(8 * (h_cap + 1));
  char* pair_start = take(8 * (h_cap + 1));
  char* cdf = take(8 * (COUNT_KMAX + 1));
  char* guide = take(4 * (GUIDE + 1));
  char* tiles = take(8 * (scan_tiles(longest) + 1));
  char* ctl = take(8 * 8);
  if (w) {
    w->order = (int32_t*)order;
    w->hist = (int64_t*)hist;
    w->size = (int32_t*)size;
    w->start = (int64_t*)start;
    w->perm = (int32_t*)perm;
    w->at = (int32_t*)at;
    w->cdf = (double*)cdf;
    w->guide = (int32_t*)guide;
    w->new_start = (int64_t*)new_start;
    w->pair_start = (int64_t*)pair_start;
    w->tiles = (int64_t*)tiles;
    w->ctl = (int64_t*)ctl;
  }
  return off;
}

int bind_work(int64_t n, int64_t h_cap, void* work, int64_t bytes, PopWork* w, const char* where) {
  if (!work || ((uintptr_t)work & 255)) return fail(CVK_E_ALIGN, where);
  if (bytes < work_bytes(n, h_cap, nullptr, nullptr)) return fail(CVK_E_CAPACITY, where);
  work_bytes(n, h_cap, w, (char*)work);
  return CVK_OK;
}

int check_tables(const cvk_popgen* g, const char* where) {
  if (!g || g->n_age_bins < 1 || g->n_age_bins > CVK_POP_MAX_AGE_BINS || g->n_hh_sizes < 1 ||
      g->n_hh_sizes > CVK_POP_MAX_HH_SIZES || g->hh_size_min < 0 || g->n_prog_bins < 1 ||
      g->n_prog_bins > CVK_POP_MAX_PROG_BINS)
    return fail(CVK_E_INVALID, where);
  return CVK_OK;
}

inline uint32_t grid_of(int64_t n) { return (uint32_t)std::max<int64_t>(1, std::min<int64_t>((n + POP_BLOCK - 1) / POP_BLOCK, 1 << 16)); }

#define POP_LOOP(i, n) for (int64_t i = (int64_t)blockIdx.x * POP_BLOCK + threadIdx.x; i < (n); i += (int64_t)gridDim.x * POP_BLOCK)

// ---------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------
// CDF table of a count distribution, built on the host: p_0 = p0, Poisson p_{k+1} = p_k * lam / (k+1),
// negative binomial p_{k+1} = p_k * ((k + r) / (k + 1)) * q, F[k] = F[k-1] + p_k for k = 0..K, K the
// first step where the pmf no longer rises and an added term leaves F unchanged (later terms are no
// larger, so F cannot change again), or COUNT_KMAX. The recurrence is serial, a few thousand steps:
// one host thread (IEEE double, no contraction) gives the same values a device lane would, without a
// one-lane kernel. The guide G[j] = first k with F[k] >= j / GUIDE (K + 1 if none) and K are uploaded
// beside it; the call waits for the copies, so the host buffers can go.
int count_table(const cvk_pop_counts& c, const PopWork& w, hipStream_t st) {
  std::vector<double> F;
  F.reserve(1024);
  int64_t k = 0;
  double p = c.p0, f = p;
  F.push_back(f);
  while (k < COUNT_KMAX) {
    double pn;
    if (c.kind == CVK_POP_COUNT_POISSON) {
      pn = p * c.lam / (double)(k + 1);
    } else {
      const double ratio = ((double)k + c.r) / ((double)k + 1.0);
      pn = p * ratio * c.q;
    }
    const double fn = f + pn;
    k += 1;
    F.push_back(fn);
    if (fn == f && pn <= p) break;
    p = pn;
    f = fn;
  }
  const int64_t K = k;
  std::vector<int32_t> G(GUIDE + 1);
  for (int j = 0; j <= GUIDE; ++j) {
    const double u = (double)j * (1.0 / (double)GUIDE);
    G[j] = (int32_t)(std::lower_bound(F.begin(), F.end(), u) - F.begin());
  }
  if (hipMemcpyAsync(w.cdf, F.data(), sizeof(double) * F.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(w.guide, G.data(), sizeof(int32_t) * G.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(w.ctl + 3, &K, sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(CVK_E_HIP, "population count table: upload failed");
  return CVK_OK;
}

__global__ __launch_bounds__(POP_BLOCK) void pop_age_kernel(cvk_popgen g, int64_t n, float* age, uint8_t* key) {
  POP_LOOP(i, n) {
    const AgeDraw d = draw_age(g, (uint32_t)i);
    age[i] = d.age;
    if (key) key[i] = d.key;
  }
}

__global__ __launch_bounds__(POP_BLOCK) void pop_hist_kernel(const uint8_t* key, int64_t n, int64_t n_tiles,
                                                             int64_t* hist) {
  __shared__ uint32_t s_hist[N_BUCKETS];
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int64_t i = t0 + r * POP_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&s_hist[key[i]], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = s_hist[threadIdx.x];
}

// Stable scatter: position = scanned bucket offset of this tile + agents of the same bucket
// earlier in the tile (earlier rounds via s_base, earlier waves via s_wcnt, earlier lanes by ballot).
__global__ __launch_bounds__(POP_BLOCK) void pop_scatter_kernel(const uint8_t* key, int64_t n, int64_t n_tiles,
                                                                const int64_t* hist_off, int32_t* order) {
  __shared__ int64_t s_base[N_BUCKETS];
  __shared__ uint32_t s_wcnt[POP_BLOCK / 64][N_BUCKETS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  s_base[threadIdx.x] = hist_off[(int64_t)threadIdx.x * n_tiles + blockIdx.x];
  const int64_t t0 = (int64_t)blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    for (int w = 0; w < POP_BLOCK / 64; ++w) s_wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = t0 + r * POP_BLOCK + threadIdx.x;
    const bool valid = i < n;
    const uint32_t k = valid ? key[i] : 0u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool on = (k >> bit) & 1u;
      const uint64_t b = __ballot(on);
      peers &= on ? b : ~b;
    }
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
    if (valid && rank == 0) s_wcnt[wave][k] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      int64_t pos = s_base[k] + rank;
      for (int w = 0; w < wave; ++w) pos += s_wcnt[w][k];
      order[pos] = (int32_t)i;
    }
    __syncthreads();
    int64_t add = 0;
    for (int w = 0; w < POP_BLOCK / 64; ++w) add += s_wcnt[w][threadIdx.x];
    s_base[threadIdx.x] += add;
    __syncthreads();
  }
  (void)lane;
}

__global__ __launch_bounds__(POP_BLOCK) void pop_hh_size_kernel(cvk_popgen g, int64_t h_cap, int32_t* size) {
  POP_LOOP(h, h_cap) size[h] = draw_hh_size(g, h);
}

__global__ __launch_bounds__(POP_BLOCK) void pop_hh_init_kernel(int64_t* ctl) {
  if (threadIdx.x == 0) ctl[0] = -1;
}

// The household holding agent position n - 1 is the last one: H = h + 1.
__global__ __launch_bounds__(POP_BLOCK) void pop_hh_count_kernel(int64_t n, int64_t h_cap, const int32_t* size,
                                                                 const int64_t* start, int64_t* ctl) {
  POP_LOOP(h, h_cap) {
    if (start[h] < n && start[h] + size[h] >= n) ctl[0] = h + 1;
  }
}

__device__ inline int64_t hh_members(int64_t h, int64_t H, int64_t n, const int32_t* size, const int64_t* start) {
  return h == H - 1 ? n - start[h] : (int64_t)size[h];
}

__global__ __launch_bounds__(POP_BLOCK) void pop_hh_perm_kernel(cvk_popgen g, int64_t n, int64_t h_cap,
                                                                const int32_t* size, const int64_t* start,
                                                                const int64_t* ctl, int32_t* perm, int32_t* at,
                                                                int64_t* new_size, int64_t* new_pairs) {
  const int64_t H = ctl[0];
  const uint32_t b = feistel_bits(H);
  POP_LOOP(h, h_cap + 1) {
    if (h >= H) {  // positions H..h_cap are never a pi target; H < 0 clears everything
      new_size[h] = 0;
      new_pairs[h] = 0;
      continue;
    }
    const int64_t m = hh_members(h, H, n, size, start);
    const int64_t j = g.correlate ? hh_position(g.seed, h, H, b) : h;
    perm[h] = (int32_t)j;
    at[j] = (int32_t)h;
    new_size[j] = m;
    new_pairs[j] = m * (m - 1) / 2;
  }
}

__global__ __launch_bounds__(POP_BLOCK) void pop_hh_totals_kernel(const int64_t* ctl, const int64_t* pairs_total,
                                                                  int64_t* totals) {
  if (threadIdx.x == 0) {
    totals[0] = ctl[0];
    totals[1] = ctl[0] < 0 ? 0 : *pairs_total;
  }
}

// Segment of x in a staged, non-decreasing offset table s[0..cnt] with s[0] <= x < s[cnt]: the j
// with s[j] <= x < s[j + 1] (empty segments are skipped).
__device__ inline int segment_of(const int64_t* s, int cnt, int64_t x) {
  int lo = 0, hi = cnt;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Workgroups over tiles of 256 household positions. Each stages its positions' first agents
// (ages pass) or first edges (edges pass) in LDS and then writes its contiguous output range with
// coalesced stores, every element finding its position by a binary search over the stage.
__global__ __launch_bounds__(POP_BLOCK) void pop_hh_ages_kernel(int64_t h_cap, const float* age_in, PopWork w,
                                                                float* age_out) {
  __shared__ int64_t s_first[POP_BLOCK + 1];
  __shared__ int64_t s_src[POP_BLOCK];
  const int64_t H = w.ctl[0];
  for (int64_t j0 = (int64_t)blockIdx.x * POP_BLOCK; j0 < H; j0 += (int64_t)gridDim.x * POP_BLOCK) {
    const int cnt = (int)std::min<int64_t>(POP_BLOCK, H - j0);
    __syncthreads();
    if (threadIdx.x < cnt) {
      s_first[threadIdx.x] = w.new_start[j0 + threadIdx.x];
      s_src[threadIdx.x] = w.start[w.at[j0 + threadIdx.x]];
    }
    if (threadIdx.x == 0) s_first[cnt] = w.new_start[j0 + cnt];
    __syncthreads();
    for (int64_t i = s_first[0] + threadIdx.x; i < s_first[cnt]; i += POP_BLOCK) {
      const int j = segment_of(s_first, cnt, i);
      age_out[i] = age_in[w.order[s_src[j] + (i - s_first[j])]];
    }
  }
}

__global__ __launch_bounds__(POP_BLOCK) void pop_hh_edges_kernel(int64_t h_cap, PopWork w, int32_t* p1, int32_t* p2,
                                                                 float* beta, int64_t n_edges) {
  __shared__ int64_t s_first[POP_BLOCK + 1];
  __shared__ int64_t s_agent[POP_BLOCK];
  __shared__ int32_t s_size[POP_BLOCK];
  const int64_t H = w.ctl[0];
  for (int64_t j0 = (int64_t)blockIdx.x * POP_BLOCK; j0 < H; j0 += (int64_t)gridDim.x * POP_BLOCK) {
    const int cnt = (int)std::min<int64_t>(POP_BLOCK, H - j0);
    __syncthreads();
    if (threadIdx.x < cnt) {
      const int64_t j = j0 + threadIdx.x;
      s_first[threadIdx.x] = w.pair_start[j];
      s_agent[threadIdx.x] = w.new_start[j];
      s_size[threadIdx.x] = (int32_t)(w.new_start[j + 1] - w.new_start[j]);
    }
    if (threadIdx.x == 0) s_first[cnt] = w.pair_start[j0 + cnt];
    __syncthreads();
    const int64_t e1 = std::min<int64_t>(s_first[cnt], n_edges);
    for (int64_t e = s_first[0] + threadIdx.x; e < e1; e += POP_BLOCK) {
      const int j = segment_of(s_first, cnt, e);
      // pair t of a size-m household, in (a, b) order: (0,1), (0,2), .., (0,m-1), (1,2), ..
      int64_t t = e - s_first[j];
      const int m = s_size[j];
      int a = 0;
      while (t >= m - 1 - a) {
        t -= m - 1 - a;
        ++a;
      }
      p1[e] = (int32_t)(s_agent[j] + a);
      p2[e] = (int32_t)(s_agent[j] + a + 1 + t);
      if (beta) beta[e] = 1.0f;
    }
  }
}

// Workgroups over tiles of 256 members: the tile's edge offsets and member ids are staged in LDS
// and the tile's edge range is written with coalesced stores.
__global__ __launch_bounds__(POP_BLOCK) void pop_p1_kernel(int64_t n_members, const int32_t* mapping,
                                                           const int64_t* offsets, int32_t* p1, int64_t n_edges) {
  __shared__ int64_t s_off[POP_BLOCK + 1];
  __shared__ int32_t s_who[POP_BLOCK];
  for (int64_t m0 = (int64_t)blockIdx.x * POP_BLOCK; m0 < n_members; m0 += (int64_t)gridDim.x * POP_BLOCK) {
    const int cnt = (int)std::min<int64_t>(POP_BLOCK, n_members - m0);
    __syncthreads();
    if (threadIdx.x < cnt) {
      s_off[threadIdx.x] = offsets[m0 + threadIdx.x];
      s_who[threadIdx.x] = mapping ? mapping[m0 + threadIdx.x] : (int32_t)(m0 + threadIdx.x);
    }
    if (threadIdx.x == 0) s_off[cnt] = offsets[m0 + cnt];
    __syncthreads();
    const int64_t e1 = std::min<int64_t>(s_off[cnt], n_edges);
    for (int64_t e = s_off[0] + threadIdx.x; e < e1; e += POP_BLOCK) p1[e] = s_who[segment_of(s_off, cnt, e)];
  }
}

// One Philox draw per four edges; full quads store 16 bytes per lane (p2 and beta 16-byte aligned).
__global__ __launch_bounds__(POP_BLOCK) void pop_p2_kernel(uint64_t seed, uint32_t slot, int64_t n_members,
                                                           const int32_t* mapping, int32_t* p2, float* beta,
                                                           int64_t n_edges) {
  typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const int64_t quads = (n_edges + 3) >> 2;
  const uint32_t n = (uint32_t)n_members, thr = (0u - n) % n;
  POP_LOOP(q, quads) {
    const u32x4 w = philox_seeded(seed, (uint32_t)q, (uint32_t)((uint64_t)q >> 32), STREAM_POP_EDGE | slot, 0u);
    const int64_t e = 4 * q;
    if (e + 3 < n_edges) {
      int32_t r0 = (int32_t)partner(seed, slot, e, w.x, n, thr), r1 = (int32_t)partner(seed, slot, e + 1, w.y, n, thr);
      int32_t r2 = (int32_t)partner(seed, slot, e + 2, w.z, n, thr), r3 = (int32_t)partner(seed, slot, e + 3, w.w, n, thr);
      if (mapping) {
        r0 = mapping[r0];
        r1 = mapping[r1];
        r2 = mapping[r2];
        r3 = mapping[r3];
      }
      *(i32x4*)(p2 + e) = i32x4{r0, r1, r2, r3};
      if (beta) *(f32x4*)(beta + e) = f32x4{1.0f, 1.0f, 1.0f, 1.0f};
    } else {
      const uint32_t words[4] = {w.x, w.y, w.z, w.w};
      for (int j = 0; e + j < n_edges; ++j) {
        const uint32_t r = partner(seed, slot, e + j, words[j], n, thr);
        p2[e + j] = mapping ? mapping[r] : (int32_t)r;
        if (beta) beta[e + j] = 1.0f;
      }
    }
  }
}

__global__ __launch_bounds__(POP_BLOCK) void pop_people_kernel(cvk_popgen g, int64_t n, const float* age,
                                                               const double* F, const int32_t* G, const int64_t* ctl,
                                                               uint8_t* sex, float* symp, float* severe, float* crit,
                                                               float* death, float* rel_sus, float* rel_trans) {
  const int64_t K = ctl[3];
  POP_LOOP(i, n) {
    const u32x4 w = philox_seeded(g.seed, (uint32_t)i, 0u, STREAM_POP_PERSON, 0u);
    const double a = (double)age[i];
    int bin = -1;
    for (int k = 0; k < g.n_prog_bins; ++k) bin += g.prog_cutoff[k] <= a ? 1 : 0;
    bin = bin < 0 ? 0 : bin;
    const int64_t draws = count_lookup(F, G, K, u01(w.y, w.z));
    sex[i] = (uint8_t)(w.x & 1u);
    symp[i] = g.symp_prob[bin];
    severe[i] = g.severe_prob[bin];
    crit[i] = g.crit_prob[bin];
    death[i] = g.death_prob[bin];
    rel_sus[i] = g.rel_sus[bin];
    rel_trans[i] = (float)(g.trans_or[bin] * ((double)draws * g.nb_step));
  }
}

}  // namespace
}  // namespace cvk

using namespace cvk;

extern "C" int64_t cvk_pop_work_bytes(int64_t n, int64_t h_cap) {
  if (n < 0 || h_cap < 0) return 0;
  return work_bytes(n, h_cap, nullptr, nullptr);
}

extern "C" int cvk_pop_ages(const cvk_popgen* g, int64_t n, float* age, uint8_t* key, void* stream) {
  if (check_tables(g, "cvk_pop_ages: invalid tables")) return CVK_E_INVALID;
  if (n < 0 || n > POP_MAX_N || (n > 0 && !age) || (n > 0 && g->correlate && !key))
    return fail(CVK_E_INVALID, "cvk_pop_ages: invalid n or null output");
  if (n == 0) return CVK_OK;
  hipLaunchKernelGGL(pop_age_kernel, dim3(grid_of(n)), dim3(POP_BLOCK), 0, (hipStream_t)stream, *g, n, age, key);
  return check_launch("cvk_pop_ages");
}

extern "C" int cvk_pop_households_plan(const cvk_popgen* g, int64_t n, const uint8_t* key, int64_t h_cap, void* work,
                                       int64_t work_bytes_, int64_t* totals, void* stream) {
  if (check_tables(g, "cvk_pop_households_plan: invalid tables")) return CVK_E_INVALID;
  if (n < 0 || n > POP_MAX_N || h_cap < 1 || h_cap > POP_MAX_N || !totals || (n > 0 && g->correlate && !key))
    return fail(CVK_E_INVALID, "cvk_pop_households_plan: invalid sizes or null pointer");
  PopWork w;
  int rc = bind_work(n, h_cap, work, work_bytes_, &w, "cvk_pop_households_plan: workspace too small or misaligned");
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const uint32_t gh = grid_of(h_cap + 1);
  hipLaunchKernelGGL(pop_hh_size_kernel, dim3(gh), dim3(POP_BLOCK), 0, st, *g, h_cap, w.size);
  if ((rc = scan(ArrayI32{w.size}, WriteI64{w.start}, h_cap, w.tiles, w.ctl + 1, st))) return rc;
  hipLaunchKernelGGL(pop_hh_init_kernel, dim3(1), dim3(POP_BLOCK), 0, st, w.ctl);
  if (n > 0) hipLaunchKernelGGL(pop_hh_count_kernel, dim3(gh), dim3(POP_BLOCK), 0, st, n, h_cap, w.size, w.start, w.ctl);
  else if (hipMemsetAsync(w.ctl, 0, sizeof(int64_t), st) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_pop_households_plan: hipMemsetAsync failed");
  if (g->correlate && n > 0) {
    const int64_t tiles = sort_tiles(n);
    hipLaunchKernelGGL(pop_hist_kernel, dim3((uint32_t)tiles), dim3(POP_BLOCK), 0, st, key, n, tiles, w.hist);
    if ((rc = scan(ArrayI64{w.hist}, WriteI64{w.hist}, N_BUCKETS * tiles, w.tiles, nullptr, st))) return rc;
    hipLaunchKernelGGL(pop_scatter_kernel, dim3((uint32_t)tiles), dim3(POP_BLOCK), 0, st, key, n, tiles,
                       (const int64_t*)w.hist, w.order);
  }
  hipLaunchKernelGGL(pop_hh_perm_kernel, dim3(gh), dim3(POP_BLOCK), 0, st, *g, n, h_cap, (const int32_t*)w.size,
                     (const int64_t*)w.start, (const int64_t*)w.ctl, w.perm, w.at, w.new_start, w.pair_start);
  if ((rc = scan(ArrayI64{w.new_start}, WriteI64{w.new_start}, h_cap + 1, w.tiles, nullptr, st))) return rc;
  if ((rc = scan(ArrayI64{w.pair_start}, WriteI64{w.pair_start}, h_cap + 1, w.tiles, w.ctl + 2, st))) return rc;
  hipLaunchKernelGGL(pop_hh_totals_kernel, dim3(1), dim3(POP_BLOCK), 0, st, (const int64_t*)w.ctl,
                     (const int64_t*)(w.ctl + 2), totals);
  return check_launch("cvk_pop_households_plan");
}

extern "C" int cvk_pop_households_fill(const cvk_popgen* g, int64_t n, const float* age_in, int64_t h_cap, void* work,
                                       int64_t work_bytes_, float* age_out, int32_t* p1, int32_t* p2, float* beta,
                                       int64_t n_edges, void* stream) {
  if (check_tables(g, "cvk_pop_households_fill: invalid tables")) return CVK_E_INVALID;
  if (n < 0 || n > POP_MAX_N || h_cap < 1 || n_edges < 0 || (n > 0 && (!age_in || !age_out)) ||
      (n_edges > 0 && (!p1 || !p2)))
    return fail(CVK_E_INVALID, "cvk_pop_households_fill: invalid sizes or null pointer");
  if (g->correlate && age_in == age_out && n > 0)
    return fail(CVK_E_INVALID, "cvk_pop_households_fill: age_out must not alias age_in when households are shuffled");
  PopWork w;
  int rc = bind_work(n, h_cap, work, work_bytes_, &w, "cvk_pop_households_fill: workspace too small or misaligned");
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (!g->correlate && n > 0 && age_out != age_in &&
      hipMemcpyAsync(age_out, age_in, sizeof(float) * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_pop_households_fill: hipMemcpyAsync failed");
  const uint32_t gt = grid_of(h_cap);
  if (g->correlate && n > 0)
    hipLaunchKernelGGL(pop_hh_ages_kernel, dim3(gt), dim3(POP_BLOCK), 0, st, h_cap, age_in, w, age_out);
  if (n_edges > 0)
    hipLaunchKernelGGL(pop_hh_edges_kernel, dim3(gt), dim3(POP_BLOCK), 0, st, h_cap, w, p1, p2, beta, n_edges);
  return check_launch("cvk_pop_households_fill");
}

extern "C" int cvk_pop_members(int64_t n, const float* age, double lo, double hi, int32_t* mapping, int64_t* total,
                               void* work, int64_t work_bytes_, void* stream) {
  if (n < 0 || n > POP_MAX_N || !total || (n > 0 && (!age || !mapping)))
    return fail(CVK_E_INVALID, "cvk_pop_members: invalid n or null pointer");
  PopWork w;
  int rc = bind_work(n, 1, work, work_bytes_, &w, "cvk_pop_members: workspace too small or misaligned");
  if (rc) return rc;
  return scan(BandFlag{age, lo, hi}, WriteMember{mapping}, n, w.tiles, total, (hipStream_t)stream);
}

static int check_counts(const cvk_pop_counts* c, const char* where) {
  if (!c || (c->kind != CVK_POP_COUNT_POISSON && c->kind != CVK_POP_COUNT_NEGBIN) || !(c->p0 > 0.0) ||
      !(c->p0 <= 1.0))
    return fail(CVK_E_INVALID, where);
  return CVK_OK;
}

extern "C" int cvk_pop_contacts_plan(const cvk_popgen* g, int32_t slot, int64_t n_members, const cvk_pop_counts* counts,
                                     int64_t* offsets, void* work, int64_t work_bytes_, void* stream) {
  if (!g || slot < 0 || slot > 255 || n_members < 0 || n_members > POP_MAX_N || !offsets)
    return fail(CVK_E_INVALID, "cvk_pop_contacts_plan: invalid argument");
  if (check_counts(counts, "cvk_pop_contacts_plan: count distribution needs a kind and 0 < p0 <= 1"))
    return CVK_E_INVALID;
  PopWork w;
  int rc = bind_work(n_members, 1, work, work_bytes_, &w, "cvk_pop_contacts_plan: workspace too small or misaligned");
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if ((rc = count_table(*counts, w, st))) return rc;
  return scan(MemberEdges{g->seed, (uint32_t)slot, w.cdf, w.guide, w.ctl + 3}, WriteI64{offsets}, n_members, w.tiles,
              offsets + n_members, st);
}

extern "C" int cvk_pop_contacts_fill(const cvk_popgen* g, int32_t slot, int64_t n_members, const int32_t* mapping,
                                     const int64_t* offsets, int32_t* p1, int32_t* p2, float* beta, int64_t n_edges,
                                     void* stream) {
  if (!g || slot < 0 || slot > 255 || n_members < 0 || n_members > POP_MAX_N || n_edges < 0 || !offsets ||
      (n_edges > 0 && (!p1 || !p2 || n_members == 0)))
    return fail(CVK_E_INVALID, "cvk_pop_contacts_fill: invalid argument");
  if (((uintptr_t)p2 | (uintptr_t)beta) & 15) return fail(CVK_E_ALIGN, "cvk_pop_contacts_fill: p2/beta need 16-byte alignment");
  if (n_edges == 0) return CVK_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(pop_p1_kernel, dim3(grid_of(n_members)), dim3(POP_BLOCK), 0, st, n_members, mapping, offsets, p1,
                     n_edges);
  hipLaunchKernelGGL(pop_p2_kernel, dim3(grid_of((n_edges + 3) >> 2)), dim3(POP_BLOCK), 0, st, g->seed,
                     (uint32_t)slot, n_members, mapping, p2, beta, n_edges);
  return check_launch("cvk_pop_contacts_fill");
}

extern "C" int cvk_pop_people(const cvk_popgen* g, int64_t n, const float* age, uint8_t* sex, float* symp_prob,
                              float* severe_prob, float* crit_prob, float* death_prob, float* rel_sus, float* rel_trans,
                              void* work, int64_t work_bytes_, void* stream) {
  if (check_tables(g, "cvk_pop_people: invalid tables")) return CVK_E_INVALID;
  if (n < 0 || n > POP_MAX_N ||
      (n > 0 && (!age || !sex || !symp_prob || !severe_prob || !crit_prob || !death_prob || !rel_sus || !rel_trans)))
    return fail(CVK_E_INVALID, "cvk_pop_people: invalid n or null pointer");
  if (!(g->nb_p0 > 0.0) || !(g->nb_p0 <= 1.0)) return fail(CVK_E_INVALID, "cvk_pop_people: nb_p0 must be in (0, 1]");
  PopWork w;
  int rc = bind_work(n, 1, work, work_bytes_, &w, "cvk_pop_people: workspace too small or misaligned");
  if (rc) return rc;
  if (n == 0) return CVK_OK;
  hipStream_t st = (hipStream_t)stream;
  cvk_pop_counts nb;
  memset(&nb, 0, sizeof(nb));
  nb.kind = CVK_POP_COUNT_NEGBIN;
  nb.p0 = g->nb_p0;
  nb.r = g->nb_r;
  nb.q = g->nb_q;
  if ((rc = count_table(nb, w, st))) return rc;
  hipLaunchKernelGGL(pop_people_kernel, dim3(grid_of(n)), dim3(POP_BLOCK), 0, st, *g, n, age, (const double*)w.cdf,
                     (const int32_t*)w.guide, (const int64_t*)w.ctl, sex, symp_prob, severe_prob, crit_prob, death_prob,
                     rel_sus, rel_trans);
  return check_launch("cvk_pop_people");
}
