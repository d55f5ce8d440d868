// Production day loop of the covasim_amd library (gfx950, wave64), one sim per call.
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
// A count-only finish pass closes the last day of a run. Per-block counts are stored as partial
// rows and summed after the next kernel boundary (no device-scope fences, no same-address atomics).
// The per-agent, per-edge and per-target logic lives in cvk_kernels.h (shared with the batch path).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>

#include "cvk_host.h"
#include "cvk_kernels.h"

namespace cvk {

char* last_error_buffer() {
  static thread_local char buf[512] = {0};
  return buf;
}

__device__ inline uint32_t* day_row(const cvk_work& w, int32_t t) { return w.results + (int64_t)t * CVK_R_COLS; }

// ===========================================================================
// day_begin: [previous-day stocks + update_nab] + transitions + bitmaps + records,
// with extra blocks resampling the dynamic layers.
// ===========================================================================
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
  int32_t clock_slot;   // >= 0: the launch's start time goes to work.day_clock[clock_slot]
  ResampleLayer rs[CVK_MAX_LAYERS];
};

constexpr int AGENTS_PER_THREAD = 2;
static_assert(AGENT_BLOCK * AGENTS_PER_THREAD == 512, "one coarse-map word per agent-pass workgroup");

__global__ __launch_bounds__(AGENT_BLOCK) void day_begin_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                DayBeginArgs a) {
  // the day's start on the device's constant-rate wall clock (work.day_clock, optional)
  if (blockIdx.x == 0 && threadIdx.x == 0 && W.day_clock && a.clock_slot >= 0)
    W.day_clock[a.clock_slot] = (uint64_t)wall_clock64();
  if (blockIdx.x >= a.agent_blocks) {
    const int64_t stride = (int64_t)(gridDim.x - a.agent_blocks) * AGENT_BLOCK;
    const int64_t start = (int64_t)(blockIdx.x - a.agent_blocks) * AGENT_BLOCK + threadIdx.x;
    const uint32_t thr = resample_thr((uint32_t)P.n);
    for (int r = 0; r < a.n_resample; ++r) {  // one thread per pair of edges
      const ResampleLayer& R = a.rs[r];
      for (int64_t q = start; q < (R.n + 1) / 2; q += stride)
        resample_pair(prm.seed, a.t, R.layer, (uint32_t)P.n, thr, q, R.n, R.p1, R.p2, R.beta);
    }
    return;
  }
  AgentSpan S;
  S.g0 = 0;
  S.li0 = (int64_t)blockIdx.x * (AGENT_BLOCK * AGENTS_PER_THREAD);
  S.n_local = P.n;
  S.n_stride = P.n;
  S.bitmap = (uint64_t*)W.bitmap;
  S.coarse = W.coarse;  // one word per block: AGENT_BLOCK * AGENTS_PER_THREAD == 512
  S.rec = W.rec;
  S.bound = W.bound;
  S.aux = (uint64_t*)W.aux;  // null unless the day's edge pass reads target bits
  S.partials = W.partials;
  S.partials_f = W.partials_f;
  S.part_stride = CVK_MAX_REDUCE_BLOCKS;
  S.part_col = blockIdx.x;
  agent_block<AGENTS_PER_THREAD, true>(P, prm, S, a.t, a.prologue, a.transitions);
}

// Writes the day rows from the summed partial rows of day_begin (block-wide).
// The infect pass of the fused day loop leaves its per-block flow counts in the last
// INFECT_PART_BLOCKS entries of the partial columns (the agent pass uses the first ones), so no
// result row is the target of one atomic per block; the next reduction adds them to their day's row.
constexpr int INFECT_PART_BLOCKS = 1024;
constexpr int INFECT_PART_BASE = CVK_MAX_REDUCE_BLOCKS - INFECT_PART_BLOCKS;

struct DayReduce {
  uint32_t n_blocks;  // agent-pass partial rows to reduce (0: none)
  int32_t t;
  int32_t prologue;
  int32_t transitions;
  uint32_t infect_blocks;  // infect-pass partial rows of day infect_t to add (0: none)
  int32_t infect_t;
};

static_assert(N_INFECT_COLS <= N_SUM_COLS, "day_reduce sums the infect columns into s_tot");
__device__ void day_reduce(const cvk_work& W, const DayReduce& r) {
  __shared__ uint32_t s_tot[N_SUM_COLS];
  __shared__ double s_totf[N_AGENT_FCOLS];
  if (r.n_blocks) {
    sum_partials<N_SUM_COLS, N_AGENT_FCOLS>(W.partials, W.partials_f, CVK_MAX_REDUCE_BLOCKS, 0, r.n_blocks, s_tot,
                                              s_totf);
    write_agent_rows(W.results, W.results_f, s_tot, s_totf, r.t, r.prologue, r.transitions);
  }
  if (r.infect_blocks) {
    sum_partials<N_INFECT_COLS, 0>(W.partials, W.partials_f, CVK_MAX_REDUCE_BLOCKS, INFECT_PART_BASE, r.infect_blocks,
                                   s_tot, s_totf);
    uint32_t* row = day_row(W, r.infect_t);
    if (threadIdx.x < N_INFECT_COLS) row[infect_col(threadIdx.x)] += s_tot[threadIdx.x];
  }
}

__global__ __launch_bounds__(1024) void day_reduce_kernel(cvk_work W, DayReduce r) { day_reduce(W, r); }

// ===========================================================================
// choose: the k smallest 64-bit selection keys (philox word << 32 | agent) among
// eligible agents, by an 8-pass byte-wise radix select; a uniform k-subset.
// ===========================================================================
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
    if (old == KEY_NONE) atomicOr(&W.touched[i >> 5], 1u << (i & 31));
  }
}

// ===========================================================================
// edge_pass
// ===========================================================================
constexpr int EDGE_BLOCK = 1024;
constexpr int EDGE_WAVES = EDGE_BLOCK / 64;

// Walk decision of the listed (source-list) layers, block-wide: walk (mask) unless `automatic` and the agent
// pass set more than max_permille infectious bits per thousand agents; every workgroup sums the same
// per-block counts (partial column DC_BITS of agent_blocks blocks), so all of them decide alike.
template <int BLOCK>
__device__ inline uint32_t walk_decision(const cvk_work& W, uint32_t mask, int32_t automatic, int32_t max_permille,
                                         uint32_t agent_blocks, uint32_t n_agents) {
  __shared__ uint32_t s_inf;
  if (!mask || !automatic) return mask;  // uniform over the grid
  if (threadIdx.x == 0) s_inf = 0;
  __syncthreads();
  uint32_t cnt = 0;
  for (uint32_t b = threadIdx.x; b < agent_blocks; b += BLOCK) cnt += W.partials[(int64_t)DC_BITS * CVK_MAX_REDUCE_BLOCKS + b];
  cnt = wave_sum(cnt);
  if (lane_id() == 0 && cnt) atomicAdd(&s_inf, cnt);
  __syncthreads();
  return (uint64_t)s_inf * 1000u > (uint64_t)max_permille * n_agents ? 0u : mask;
}

// The day's largest target code: the maximum of the agent pass's DC_TMAX column over its agent_blocks blocks
// (block-wide; s_max zeroed before a barrier that precedes this call, read after the next one).
template <int BLOCK>
__device__ inline void day_target_max(const cvk_work& W, uint32_t agent_blocks, uint32_t* s_max) {
  const uint32_t* col = W.partials + (int64_t)DC_TMAX * CVK_MAX_REDUCE_BLOCKS;
  uint32_t m = 0;
  for (uint32_t b0 = threadIdx.x; b0 < agent_blocks; b0 += 4 * BLOCK) {  // 4 loads in flight
#pragma unroll
    for (int u = 0; u < 4; ++u) m = max(m, b0 + u * BLOCK < agent_blocks ? col[b0 + u * BLOCK] : 0u);
  }
  m = wave_max(m);
  if (lane_id() == 0 && m) atomicMax(s_max, m);
}

struct EdgeArgs {
  SimView sim;
  SegView seg[CVK_MAX_LAYERS];
  // wave tiles (64 x 4 x K edges) in tile order: the listed layers (src_mask) first, then the others, each
  // group in layer order; position p of that order is layer pos_layer[p], its tiles start at tile_start[p]
  int64_t tile_start[CVK_MAX_LAYERS + 1];
  int32_t pos_layer[CVK_MAX_LAYERS];
  int64_t listed_tiles;  // tiles of the listed layers (tile_start[n_listed])
  SrcView src[CVK_MAX_LAYERS];             // source lists of the layers in src_mask
  int64_t item_start[CVK_MAX_LAYERS + 1];  // source work items (SRC_ITEM_AGENTS agents) per layer, prefix
  int32_t count_src;                       // add the slots and lists walked to the ctl counters
  uint32_t src_mask;     // bit l: layer l has source lists; on days they are walked its tiles are skipped
  int32_t src_auto;      // walk_decision's rule (0: the lists are always walked)
  int32_t src_max_permille;
  uint32_t agent_blocks;
  int32_t n_layers;
  int32_t lds_words;  // words of the LDS map (fine bitmap or coarse map, a multiple of 4)
  int32_t t;
  int32_t target_bits;  // 1: mixed edges read the targets' aux bits, bounded by the day's largest target code
                        // (the agent pass of day t has just written both)
  cvk_work work;     // for the fused day reduction
  DayReduce reduce;  // day_begin partial rows to reduce in block 0 (n_blocks == 0: none)
};

// On days the source lists are walked (walk_decision), every wave first walks its source items (item i
// goes to wave i mod the grid's working waves) and the listed layers' tiles are skipped. Then workgroup b
// takes its share of the (other) layers' tiles (G workgroups: [T*b/G, T*(b+1)/G) with the bitmap in LDS,
// else every G-th tile from b); its waves take the next
// tile from an LDS counter, so no wave waits for another and the global memory system sees no tile
// atomics. A wave's exact-test queue is drained when its work changes layer and at the end.
// MODE: BITS_LDS / BITS_COARSE / BITS_L2 (the LDS map is a.lds_words words of the fine bitmap or of
// the coarse map). GEN: some layer's endpoints are generated (CVK_TUNE_GENERATE_DYNAMIC); the other
// instantiations do not carry that code. PF: the next tile's endpoint loads are issued before this
// tile's candidates are processed. TB: target bits (a.target_bits; instantiated only where the bitmap is
// not in LDS).
template <int MODE, int K, bool GEN, bool PF, bool TB>
__global__ __launch_bounds__(EDGE_BLOCK) void edge_pass_kernel(EdgeArgs a) {
  extern __shared__ uint32_t s_mem[];
  __shared__ uint32_t s_next;
  __shared__ float s_beta[MAXV];
  __shared__ uint32_t s_tmax;  // target bits: the day's largest target code
  if (threadIdx.x == 0) s_tmax = 0u;
  // the launch's span on the device wall clock (work.day_clock, optional)
  unsigned long long* span = a.work.day_clock ? (unsigned long long*)a.work.day_clock + a.work.npts + 1 + 2 * a.t : nullptr;
  if (span && threadIdx.x == 0) atomicMin(span, (unsigned long long)wall_clock64());
  const int nbw = MODE == BITS_L2 ? 0 : a.lds_words;
  const WaveQ Q = wave_queues(s_mem + nbw + (threadIdx.x >> 6) * EDGE_WAVE_WORDS, s_beta);
  if (threadIdx.x < MAXV) s_beta[threadIdx.x] = a.sim.beta_v[threadIdx.x];
  // block 0 sums the previous passes' partial rows and, when other workgroups exist, takes no tiles
  const bool reducing = a.reduce.n_blocks || a.reduce.infect_blocks;
  if (blockIdx.x == 0 && reducing) day_reduce(a.work, a.reduce);
  if (threadIdx.x == 0) s_next = 0;
  if (MODE != BITS_L2) {
    const uint4* src = (const uint4*)(MODE == BITS_LDS ? a.sim.bitmap : a.sim.coarse);
    uint4* dst = (uint4*)s_mem;
    for (int j = threadIdx.x; j < (nbw >> 2); j += EDGE_BLOCK) dst[j] = src[j];
  }
  const bool walked = walk_decision<EDGE_BLOCK>(a.work, a.src_mask, a.src_auto, a.src_max_permille, a.agent_blocks,
                                                a.sim.n_agents) != 0;
  if (a.count_src && a.src_mask && !walked && blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd((unsigned long long*)(a.work.ctl + CVK_CTL_SRC_STREAMED), 1ull);
  constexpr bool tbits = TB && MODE != BITS_LDS;
  if (tbits) {  // uniform over the grid
    __syncthreads();  // s_tmax zeroed
    day_target_max<EDGE_BLOCK>(a.work, a.agent_blocks, &s_tmax);
  }
  __syncthreads();
  const uint32_t tmax = s_tmax;  // 0 with target bits: no agent can be infected today
  const uint32_t* bits = s_mem;
  const int64_t n_wt = a.tile_start[a.n_layers];
  const uint32_t skip = reducing && gridDim.x > 1 ? 1u : 0u;
  const int64_t workers = (int64_t)(gridDim.x - skip);
  const int64_t wb = (int64_t)blockIdx.x - skip;
  // This workgroup's tiles. With the bitmap in LDS (up to ~1.06M agents): its listed share [l0, l1) (empty on
  // walked days) and its other share [o0, o1), contiguous. Above that: every workers-th tile from base + wb,
  // base = 0, or on walked days the first tile after the listed layers' (whose tiles are skipped), so each
  // workgroup meets every part of every layer (measured: the 10M edge pass 856 -> 842 us; at 1M the
  // contiguous shares are faster, 40.8 against 42.9 us).
  constexpr bool STRIDED = MODE != BITS_LDS;
  const int64_t tl = a.listed_tiles, to = n_wt - tl;
  const int64_t l0 = wb < 0 || walked ? 0 : tl * wb / workers, l1 = wb < 0 || walked ? 0 : tl * (wb + 1) / workers;
  const int64_t o0 = wb < 0 ? 0 : tl + to * wb / workers, o1 = wb < 0 ? 0 : tl + to * (wb + 1) / workers;
  const int64_t base = walked ? tl : 0;
  const int64_t n_mine = !STRIDED ? (l1 - l0) + (o1 - o0)
                                  : (wb < 0 || n_wt - base <= wb ? 0 : (n_wt - base - wb + workers - 1) / workers);
  uint32_t n_x = 0;
  int32_t q_seg = a.pos_layer[0];
  if (walked && wb >= 0) {  // source items, one wave at a time over the grid's working waves
    uint64_t cnt[2] = {0, 0};
    const int64_t n_items = a.item_start[a.n_layers];
    const int64_t stride = workers * EDGE_WAVES;
    for (int64_t it = wb * EDGE_WAVES + (threadIdx.x >> 6); it < n_items; it += stride) {
      int l = 0;
      while (l + 1 < a.n_layers && it >= a.item_start[l + 1]) ++l;
      if (l != q_seg) {
        if (n_x) flush_exact(a.sim, a.seg[q_seg], a.t, Q, n_x);
        n_x = 0;
        q_seg = l;
      }
      source_item_pass<MODE>(a.sim, a.seg[l], a.src[l], a.t, bits, Q, (it - a.item_start[l]) * (int64_t)SRC_ITEM_AGENTS,
                             n_x, cnt);
    }
    if (a.count_src && lane_id() == 0 && (cnt[0] | cnt[1])) {
      atomicAdd((unsigned long long*)(a.work.ctl + CVK_CTL_SRC_SLOTS), (unsigned long long)cnt[0]);
      atomicAdd((unsigned long long*)(a.work.ctl + CVK_CTL_SRC_AGENTS), (unsigned long long)cnt[1]);
    }
  }
  // the wave's next tile from the block's counter: its layer and first edge (seg -1: none left)
  auto grab = [&](WaveTile<K>& T) {
    uint32_t k = 0;
    if (lane_id() == 0) k = atomicAdd(&s_next, 1u);
    k = __builtin_amdgcn_readfirstlane(__shfl(k, 0, 64));
    T.seg = -1;
    if ((int64_t)k >= n_mine) return;
    const int64_t tile = STRIDED ? base + wb + (int64_t)k * workers
                                 : ((int64_t)k < l1 - l0 ? l0 + k : o0 + (k - (l1 - l0)));
    int p = 0;
    while (p + 1 < a.n_layers && tile >= a.tile_start[p + 1]) ++p;
    T.seg = a.pos_layer[p];
    T.e0 = (tile - a.tile_start[p]) * (int64_t)(64 * 4 * K);
  };
  WaveTile<K> cur, nxt;
  grab(cur);
  if (cur.seg >= 0) load_wave_edges<K, GEN>(a.sim, a.seg[cur.seg], a.t, cur);
  while (cur.seg >= 0) {
    const int l = cur.seg;
    if (l != q_seg) {
      if (n_x) flush_exact(a.sim, a.seg[q_seg], a.t, Q, n_x);
      n_x = 0;
      q_seg = l;
    }
    TileBounds<K> B;
    tile_bits_and_bounds<MODE, K, tbits>(a.sim, bits, cur, B);
    // the next tile's endpoints are issued after this tile's bound loads (loads complete in order, so
    // waiting for the bounds does not wait for them) and arrive while the candidates are processed
    grab(nxt);
    if (PF && nxt.seg >= 0) load_wave_edges<K, GEN>(a.sim, a.seg[nxt.seg], a.t, nxt);
    if (!tbits) tile_candidates<K>(a.sim, a.seg[l], a.t, Q, cur, B, n_x);
    else if (tmax) tile_candidates<K, true>(a.sim, a.seg[l], a.t, Q, cur, B, n_x, tmax);
    if (!PF && nxt.seg >= 0) load_wave_edges<K, GEN>(a.sim, a.seg[nxt.seg], a.t, nxt);
    cur = nxt;
  }
  if (n_x) flush_exact(a.sim, a.seg[q_seg], a.t, Q, n_x);
  if (span) {  // uniform over the grid
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(span + 1, (unsigned long long)wall_clock64());
  }
}

// Launchers of every instantiation, indexed [mode][K == 4][GEN][PF][TB] (TB only where the bitmap is not in
// LDS: the LDS rows hold the TB = false kernel twice).
template <int MODE, int K, bool GEN, bool PF, bool TB>
static void edge_launch(dim3 g, dim3 b, size_t shmem, hipStream_t st, const EdgeArgs& a) {
  hipLaunchKernelGGL((edge_pass_kernel<MODE, K, GEN, PF, TB>), g, b, shmem, st, a);
}
typedef void (*EdgeLaunchFn)(dim3, dim3, size_t, hipStream_t, const EdgeArgs&);
#define CVK_EDGE_T(M, K, G, P) {edge_launch<M, K, G, P, false>, edge_launch<M, K, G, P, (M != BITS_LDS)>}
#define CVK_EDGE_K(M, K) \
  {{CVK_EDGE_T(M, K, false, false), CVK_EDGE_T(M, K, false, true)}, \
   {CVK_EDGE_T(M, K, true, false), CVK_EDGE_T(M, K, true, true)}}
#define CVK_EDGE_ROW(M) {CVK_EDGE_K(M, 2), CVK_EDGE_K(M, 4)}
static const EdgeLaunchFn kEdgeLaunch[3][2][2][2][2] = {CVK_EDGE_ROW(BITS_LDS), CVK_EDGE_ROW(BITS_COARSE),
                                                        CVK_EDGE_ROW(BITS_L2)};
#undef CVK_EDGE_T
#undef CVK_EDGE_K
#undef CVK_EDGE_ROW
#define CVK_EDGE_T(M, K, G, P) \
  {(const void*)edge_pass_kernel<M, K, G, P, false>, (const void*)edge_pass_kernel<M, K, G, P, (M != BITS_LDS)>}
#define CVK_EDGE_K(M, K) \
  {{CVK_EDGE_T(M, K, false, false), CVK_EDGE_T(M, K, false, true)}, \
   {CVK_EDGE_T(M, K, true, false), CVK_EDGE_T(M, K, true, true)}}
#define CVK_EDGE_ROW(M) {CVK_EDGE_K(M, 2), CVK_EDGE_K(M, 4)}
static const void* const kEdgeFns[3][2][2][2][2] = {CVK_EDGE_ROW(BITS_LDS), CVK_EDGE_ROW(BITS_COARSE),
                                                    CVK_EDGE_ROW(BITS_L2)};
#undef CVK_EDGE_T
#undef CVK_EDGE_K
#undef CVK_EDGE_ROW

// ===========================================================================
// infect_pass
// ===========================================================================
struct InfectArgs {
  const int32_t* p1[CVK_MAX_LAYERS];
  const int32_t* p2[CVK_MAX_LAYERS];
  int32_t t;
  int32_t count_flows;
  int32_t allow_beds;
  int32_t n_layers;
  uint32_t gen_mask;  // bit l: layer l's endpoints of day t come from the resample stream
  int32_t partial;    // 1: flow counts go to the infect partial rows (reduced later), 0: atomics on the row
  uint32_t infect_blocks;  // workgroups [0, infect_blocks) infect; the rest resample day rs_t
  int32_t rs_t;
  int32_t n_resample;
  ResampleLayer rs[CVK_MAX_LAYERS];
};

// Occupancy hint 3 waves/SIMD (168 VGPRs, a few spills in the latency-bound outcome tree): the store-bound
// resample workgroups sharing this kernel need the occupancy more than the infect lanes need registers.
__global__ __launch_bounds__(INFECT_BLOCK, 3) void infect_pass_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                      InfectArgs ia) {
  if (blockIdx.x >= ia.infect_blocks) {  // the next day's resample, one thread per pair of edges
    const int64_t stride = (int64_t)(gridDim.x - ia.infect_blocks) * INFECT_BLOCK;
    const int64_t start = (int64_t)(blockIdx.x - ia.infect_blocks) * INFECT_BLOCK + threadIdx.x;
    const uint32_t thr = resample_thr((uint32_t)P.n);
    for (int r = 0; r < ia.n_resample; ++r) {
      const ResampleLayer& R = ia.rs[r];
      for (int64_t q = start; q < (R.n + 1) / 2; q += stride)
        resample_pair(prm.seed, ia.rs_t, R.layer, (uint32_t)P.n, thr, q, R.n, R.p1, R.p2, R.beta);
    }
    return;
  }
  __shared__ uint32_t s_cnt[N_INFECT_COLS];
  if (threadIdx.x < N_INFECT_COLS) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int32_t t = ia.t;
  const uint32_t* row = day_row(W, t);
  const bool hosp_max = ia.allow_beds && prm.n_beds_hosp >= 0 && (int64_t)row[CVK_R_MID_SEVERE] > prm.n_beds_hosp;
  const bool icu_max = ia.allow_beds && prm.n_beds_icu >= 0 && (int64_t)row[CVK_R_MID_CRITICAL] > prm.n_beds_icu;
  // one touched-bitmap word (32 agents) per thread per sweep, compacted per wave into an LDS list;
  // each round gives every lane one target
  __shared__ int32_t s_list[INFECT_BLOCK / 64][INFECT_WAVE_LIST];
  int32_t* list = s_list[threadIdx.x >> 6];
  const int64_t n_words = (P.n + 31) / 32;
  const int64_t stride = (int64_t)ia.infect_blocks * INFECT_BLOCK;
  const int64_t n_sweeps = (n_words + stride - 1) / stride;
  for (int64_t sw = 0; sw < n_sweeps; ++sw) {
    const int64_t w = sw * stride + (int64_t)blockIdx.x * INFECT_BLOCK + threadIdx.x;
    const uint32_t total = compact_touched(W.touched, w, n_words, P.n, list);
    for (uint32_t base = 0; base < total; base += 64) {  // wave-uniform
      InfectOut o = {KEY_NONE, false, false, false, false, 0u, 0u};
      int32_t source = -1, i = 0;
      const uint32_t j = base + lane_id();
      const int32_t g = j < total ? list[j] : -1;
      if (g >= 0) {
        i = g;
        o = infect_target(P, prm, W.winner, i, (uint32_t)i, P.n, t, hosp_max, icu_max);
        if (o.inf && (int32_t)o.l < ia.n_layers)
          source = key_source(o.key, ia.p1[o.l], ia.p2[o.l], (ia.gen_mask >> o.l) & 1u, prm.seed, t, (uint32_t)P.n);
      }
      // infection log, people.py:509-511: one wave-aggregated append per round
      const uint64_t m = __ballot(o.inf);
      if (m) {
        unsigned long long base = 0;
        if (lane_id() == 0)
          base = atomicAdd((unsigned long long*)&W.ctl[CVK_CTL_LOG_COUNT], (unsigned long long)__popcll(m));
        base = __shfl(base, 0, 64);
        if (o.inf) {
          const unsigned long long pos = base + lanes_below(m);
          if ((int64_t)pos < W.log_capacity) {
            int4 rec;
            rec.x = source;
            rec.y = i;
            rec.z = t;
            rec.w = (int32_t)(o.l | (o.v << 8));
            ((int4*)W.log)[pos] = rec;
          } else {
            W.ctl[CVK_CTL_LOG_OVERFLOW] = 1;
          }
        }
      }
      if (ia.count_flows && m) {
        const uint32_t n_inf = (uint32_t)__popcll(m);
        const uint32_t n_reinf = (uint32_t)__popcll(__ballot(o.inf && o.reinf));
        if (lane_id() == 0) {
          atomicAdd(&s_cnt[0], n_inf);
          if (n_reinf) atomicAdd(&s_cnt[1], n_reinf);
        }
        for (int vv = 0; vv < prm.n_variants; ++vv) {
          const uint32_t a0 = (uint32_t)__popcll(__ballot(o.inf && o.v == (uint32_t)vv));
          const uint32_t a1 = (uint32_t)__popcll(__ballot(o.inf && o.symp && o.v == (uint32_t)vv));
          const uint32_t a2 = (uint32_t)__popcll(__ballot(o.inf && o.sev && o.v == (uint32_t)vv));
          if (lane_id() == 0) {
            if (a0) atomicAdd(&s_cnt[2 + vv], a0);
            if (a1) atomicAdd(&s_cnt[2 + MAXV + vv], a1);
            if (a2) atomicAdd(&s_cnt[2 + 2 * MAXV + vv], a2);
          }
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the list is rewritten by the next sweep
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  if (ia.count_flows && threadIdx.x < N_INFECT_COLS) {
    if (ia.partial)
      W.partials[(int64_t)threadIdx.x * CVK_MAX_REDUCE_BLOCKS + INFECT_PART_BASE + blockIdx.x] = s_cnt[threadIdx.x];
    else if (s_cnt[threadIdx.x])
      atomicAdd(&day_row(W, t)[infect_col(threadIdx.x)], s_cnt[threadIdx.x]);
  }
}

// ===========================================================================
// host side
// ===========================================================================
static int g_edge_lds = 1;     // tunable: stage the infectious bitmap in LDS when it fits
static int g_edge_chunks = 2;  // tunable: 4-edge chunks per lane per wave tile (2 or 4)
static int g_gen_dynamic = 2;  // tunable: the fused day loop computes dynamic layers' endpoints in place
                               // (1), writes each day's resample (0), or chooses by size (2)
// Auto choice: writing a day's resample of E dynamic edges costs 8E bytes of stores inside the
// latency-bound infect pass, generating them costs E/2 Philox blocks in the edge and infect passes.
// Measured: 10M dynamic edges (hybrid 1M) 0.092 ms/day writing against 0.097 generating; 100M
// (hybrid 10M) 1.70 against 1.48. Generation is chosen from 2^25 dynamic edges up.
constexpr int64_t GEN_AUTO_MIN_EDGES = (int64_t)1 << 25;
static int g_source_lists = 2;   // tunable: walk static layers through their source lists when given:
                                 // 0 never, 1 always, 2 on days with few infectious agents (g_source_permille)
static int g_source_permille = 30;  // tunable: the most infectious agents per thousand for which lists are walked
static int g_count_sources = 0;  // tunable: count the slots and lists the edge pass walks (ctl words)
static int g_overlap_resample = 1;  // tunable: the fused day loop resamples inside the previous infect pass
static int g_edge_coarse = 1;       // tunable: stage the coarse map in LDS when the bitmap does not fit
static int g_edge_prefetch = 0;     // tunable: issue the next wave tile's endpoint loads before the candidates
static int g_target_bits = 1;       // tunable: the fused day loop's edge pass reads target bits instead of target
                                    // codes when the bitmap is not staged in LDS (1) or never (0)
int overlap_resample_enabled() { return g_overlap_resample; }
int source_lists_mode() { return g_source_lists; }
int source_permille() { return g_source_permille; }
bool generate_dynamic_for(int64_t dyn_edges, bool diag) {
  return !diag && dyn_edges > 0 && (g_gen_dynamic == 1 || (g_gen_dynamic == 2 && dyn_edges >= GEN_AUTO_MIN_EDGES));
}

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
static int64_t coarse_words_for(int64_t n) { return ((n + 2047) / 2048) * 4; }  // 512 agents a word, multiple of 4

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
         W->coarse && W->rec && W->partials && W->partials_f && W->bound && W->npts > 0 &&
         (W->log_capacity == 0 || W->log);
}

static int validate(const cvk_people* P, const cvk_params* prm, const cvk_work* W, int32_t t) {
  if (!people_ok(P)) return fail(CVK_E_INVALID, "people struct has a null pointer or n <= 0");
  if (!prm || prm->n_agents != P->n) return fail(CVK_E_INVALID, "params.n_agents does not match people.n");
  if (P->n >= (1ll << 31) - 1) return fail(CVK_E_INVALID, "more than 2^31-2 agents");
  if ((int64_t)grid_for(P->n, AGENT_BLOCK * AGENTS_PER_THREAD) > INFECT_PART_BASE)
    return fail(CVK_E_CAPACITY, "too many agents for the reduction workspace (CVK_MAX_REDUCE_BLOCKS)");
  if (prm->n_variants < 1 || prm->n_variants > CVK_MAX_VARIANTS)
    return fail(CVK_E_INVALID, "n_variants must be in [1, 8]");
  if (prm->n_layers < 0 || prm->n_layers > CVK_MAX_LAYERS) return fail(CVK_E_INVALID, "n_layers must be in [0, 8]");
  if (prm->use_waning && (!prm->nab_kin || prm->nab_kin_len <= 0))
    return fail(CVK_E_INVALID, "use_waning requires a nab_kin table");
  if (!work_ok(W)) return fail(CVK_E_INVALID, "work struct has a null pointer");
  if (((uintptr_t)W->bitmap | (uintptr_t)W->coarse | (uintptr_t)W->rec) & 15)
    return fail(CVK_E_ALIGN, "bitmap, coarse and rec must be 16-byte aligned");
  if (t < 0 || t >= W->npts) return fail(CVK_E_INVALID, "day index outside the results table");
  return CVK_OK;
}

static uint32_t agent_blocks_for(int64_t n) { return (uint32_t)grid_for(n, AGENT_BLOCK * AGENTS_PER_THREAD); }

static int day_reduce_launch(cvk_work* work, const DayReduce& r, hipStream_t st) {
  if (!r.n_blocks && !r.infect_blocks) return CVK_OK;
  hipLaunchKernelGGL(day_reduce_kernel, dim3(1), dim3(1024), 0, st, *work, r);
  return check_launch("day_reduce");
}

// The dynamic layers' resample arguments of a day_begin launch; returns the blocks they need.
// resample: 0 none, 1 endpoints only (weights already all ones), 2 endpoints and weights.
static int fill_resample(DayBeginArgs& a, const cvk_layer* layers, const cvk_params* params, int resample,
                         uint32_t* rs_blocks) {
  int64_t max_dyn = 0;
  *rs_blocks = 0;
  if (resample) {
    for (int l = 0; l < params->n_layers; ++l) {
      const cvk_layer& L = layers[l];
      if (!L.dynamic || L.n_edges <= 0) continue;
      if (!L.p1 || !L.p2 || !L.beta) return fail(CVK_E_INVALID, "dynamic layer has a null pointer");
      if (((uintptr_t)L.p1 | (uintptr_t)L.p2 | (uintptr_t)L.beta) & 7)
        return fail(CVK_E_ALIGN, "dynamic layer: p1/p2/beta must be 8-byte aligned");
      ResampleLayer& R = a.rs[a.n_resample++];
      R.p1 = L.p1;
      R.p2 = L.p2;
      R.beta = resample == 2 ? L.beta : nullptr;
      R.n = L.n_edges;
      R.layer = (uint32_t)l;
      max_dyn = std::max(max_dyn, L.n_edges);
    }
  }
  *rs_blocks = a.n_resample ? (uint32_t)std::min<int64_t>(grid_for((max_dyn + 1) / 2, AGENT_BLOCK),
                                                          4 * std::max(caps().n_cu, 1))
                            : 0u;
  return CVK_OK;
}

static bool target_bits_for(int64_t n, const cvk_work* work);

static int day_begin_launch(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                            cvk_work* work, int32_t t, int prologue, int transitions, int resample, hipStream_t st,
                            DayReduce* reduce_out) {
  DayBeginArgs a;
  memset(&a, 0, sizeof(a));
  a.t = t;
  a.prologue = prologue;
  a.transitions = transitions;
  a.agent_blocks = agent_blocks_for(people->n);
  a.clock_slot = transitions ? t : t + 1;  // a closing pass marks the end of day t
  uint32_t rs_blocks = 0;
  int rc = fill_resample(a, layers, params, resample, &rs_blocks);
  if (rc) return rc;
  cvk_work w = *work;
  if (!(transitions && target_bits_for(people->n, work))) w.aux = nullptr;  // the target bits are for the edge pass
  hipLaunchKernelGGL(day_begin_kernel, dim3(a.agent_blocks + rs_blocks), dim3(AGENT_BLOCK), 0, st, *people, *params,
                     w, a);
  reduce_out->n_blocks = a.agent_blocks;
  reduce_out->t = t;
  reduce_out->prologue = prologue;
  reduce_out->transitions = transitions;
  reduce_out->infect_blocks = 0;
  reduce_out->infect_t = 0;
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
    case CVK_TUNE_EDGE_CHUNKS:
      if (value != 2 && value != 4) return fail(CVK_E_INVALID, "cvk_set_tuning: edge chunks must be 2 or 4");
      g_edge_chunks = value;
      return CVK_OK;
    case CVK_TUNE_GENERATE_DYNAMIC:
      if (value < 0 || value > 2) return fail(CVK_E_INVALID, "cvk_set_tuning: generate-dynamic must be 0, 1 or 2");
      g_gen_dynamic = value;
      return CVK_OK;
    case CVK_TUNE_OVERLAP_RESAMPLE:
      g_overlap_resample = value ? 1 : 0;
      return CVK_OK;
    case CVK_TUNE_EDGE_COARSE:
      g_edge_coarse = value ? 1 : 0;
      return CVK_OK;
    case CVK_TUNE_EDGE_PREFETCH:
      g_edge_prefetch = value ? 1 : 0;
      return CVK_OK;
    case CVK_TUNE_SOURCE_LISTS:
      if (value < 0 || value > 2) return fail(CVK_E_INVALID, "cvk_set_tuning: source lists must be 0, 1 or 2");
      g_source_lists = value;
      return CVK_OK;
    case CVK_TUNE_SOURCE_PERMILLE:
      if (value < 0 || value > 1000) return fail(CVK_E_INVALID, "cvk_set_tuning: permille must be in 0..1000");
      g_source_permille = value;
      return CVK_OK;
    case CVK_TUNE_COUNT_SOURCES:
      g_count_sources = value ? 1 : 0;
      return CVK_OK;
    case CVK_TUNE_TARGET_BITS:
      g_target_bits = value ? 1 : 0;
      return CVK_OK;
    default:
      return fail(CVK_E_INVALID, "cvk_set_tuning: unknown key");
  }
}
extern "C" int cvk_version(void) { return CVK_ABI_VERSION; }
#ifndef CVK_SOURCE_HASH
#define CVK_SOURCE_HASH 0ull
#endif
extern "C" uint64_t cvk_source_hash(void) { return CVK_SOURCE_HASH; }

// Sizes and selected field offsets of the ABI structs, so bindings can verify their mirrors.
extern "C" int cvk_abi_layout(int64_t* out, int32_t n) {
  const int64_t vals[] = {
      (int64_t)sizeof(cvk_dur),    (int64_t)sizeof(cvk_params), (int64_t)sizeof(cvk_people),
      (int64_t)sizeof(cvk_layer),  (int64_t)sizeof(cvk_work),   (int64_t)sizeof(cvk_imports),
      (int64_t)offsetof(cvk_params, dur),        (int64_t)offsetof(cvk_params, nab_init),
      (int64_t)offsetof(cvk_params, rel_imm_symp), (int64_t)offsetof(cvk_params, immunity),
      (int64_t)offsetof(cvk_params, nab_kin),    (int64_t)offsetof(cvk_people, sev_imm),
      (int64_t)offsetof(cvk_work, log_capacity), (int64_t)offsetof(cvk_work, npts),
      (int64_t)CVK_R_COLS, (int64_t)CVK_RF_COLS, (int64_t)sizeof(cvk_batch_desc),
      (int64_t)offsetof(cvk_batch_desc, people), (int64_t)sizeof(cvk_batch_imports), (int64_t)sizeof(cvk_batch_config),
      (int64_t)sizeof(cvk_relax_desc), (int64_t)sizeof(cvk_popgen), (int64_t)offsetof(cvk_popgen, nb_p0),
      (int64_t)sizeof(cvk_pop_counts), (int64_t)sizeof(cvk_diag_op), (int64_t)offsetof(cvk_diag_op, trace_prob),
      (int64_t)sizeof(cvk_diag), (int64_t)offsetof(cvk_imports, diag)};
  const int32_t m = (int32_t)(sizeof(vals) / sizeof(vals[0]));
  if (!out || n < m) return fail(CVK_E_INVALID, "cvk_abi_layout: output too small");
  for (int32_t j = 0; j < m; ++j) out[j] = vals[j];
  return m;
}

// Frequency of the device wall clock that stamps work.day_clock.
extern "C" int cvk_wall_clock_rate(int64_t* hz_out) {
  if (!hz_out) return fail(CVK_E_INVALID, "cvk_wall_clock_rate: null output");
  DeviceCaps& c = caps();
  if (c.device < 0) return fail(CVK_E_HIP, "cvk_wall_clock_rate: no HIP device");
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device) != hipSuccess || khz <= 0)
    return fail(CVK_E_HIP, "cvk_wall_clock_rate: attribute query failed");
  *hz_out = (int64_t)khz * 1000;
  return CVK_OK;
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
  if (touched_bytes) *touched_bytes = ((n_agents + 31) / 32) * 4;
  if (bitmap_bytes) *bitmap_bytes = bitmap_words_for(n_agents) * 4;
  if (results_bytes) *results_bytes = (int64_t)npts * CVK_R_COLS * 4;
  if (results_f_bytes) *results_f_bytes = (int64_t)npts * CVK_RF_COLS * 8;
  return CVK_OK;
}

extern "C" int cvk_workspace_sizes2(int64_t n_agents, int64_t* rec_bytes, int64_t* partials_bytes,
                                    int64_t* partials_f_bytes, int64_t* bound_bytes) {
  if (n_agents <= 0) return fail(CVK_E_INVALID, "cvk_workspace_sizes2: n_agents must be positive");
  if (rec_bytes) *rec_bytes = n_agents * 8;
  if (partials_bytes) *partials_bytes = (int64_t)CVK_MAX_REDUCE_BLOCKS * CVK_PARTIAL_COLS * 4;
  if (partials_f_bytes) *partials_f_bytes = (int64_t)CVK_MAX_REDUCE_BLOCKS * CVK_RF_COLS * 8;
  if (bound_bytes) *bound_bytes = n_agents * 2;
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
  if (((uintptr_t)layer->p1 | (uintptr_t)layer->p2 | (uintptr_t)layer->beta) & 7)
    return fail(CVK_E_ALIGN, "cvk_resample_layer: p1/p2/beta must be 8-byte aligned");
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

// Dynamic LDS budget of the edge pass: the device's per-workgroup LDS minus the kernel's static LDS
// (the same for every instantiation; each gets the raised limit).
static int edge_dynamic_lds(int64_t* budget) {
  DeviceCaps& c = caps();
  static thread_local int configured_device = -1;
  static thread_local int64_t value = 0;
  if (configured_device != c.device) {
    const void* const* fns = &kEdgeFns[0][0][0][0][0];
    int64_t b_min = -1;
    for (int j = 0; j < 48; ++j) {
      hipFuncAttributes attr;
      if (hipFuncGetAttributes(&attr, fns[j]) != hipSuccess)
        return fail(CVK_E_HIP, "cannot query the edge pass attributes");
      const int64_t b = (int64_t)c.max_lds - (int64_t)attr.sharedSizeBytes;
      if (b <= 0 || hipFuncSetAttribute(fns[j], hipFuncAttributeMaxDynamicSharedMemorySize, (int)b) != hipSuccess)
        return fail(CVK_E_HIP, "cannot raise the LDS limit of the edge pass");
      b_min = b_min < 0 ? b : std::min(b_min, b);
    }
    value = b_min;
    configured_device = c.device;
  }
  *budget = value;
  return CVK_OK;
}

// Whether the fused day loop's edge pass over n agents reads target bits (CVK_TUNE_TARGET_BITS), and so
// whether its agent pass writes them: with cvk_work.aux, when the infectious bitmap is not staged in LDS
// (the same rule as edge_pass_impl's choice of the bitmap's place; with it in LDS the codes fit in L2).
namespace cvk {
static bool target_bits_for(int64_t n, const cvk_work* work) {
  if (!work->aux || !g_target_bits) return false;
  int64_t budget = 0;
  if (edge_dynamic_lds(&budget)) return false;
  const int64_t queue_bytes = (int64_t)EDGE_WAVES * EDGE_WAVE_WORDS * 4;
  return !(g_edge_lds && bitmap_words_for(n) * 4 + queue_bytes <= budget);
}
}  // namespace cvk

// generate: dynamic layers' endpoints of day t are computed from the resample stream (fused day loop,
// which writes them to the layer only when it closes its last day).
static int edge_pass_impl(const cvk_people* people, const cvk_layer* layers, const cvk_params* params, cvk_work* work,
                          int32_t t, const double* const* uniforms, hipStream_t st, const DayReduce* reduce,
                          bool generate) {
  EdgeArgs a;
  memset(&a, 0, sizeof(a));
  a.work = *work;
  if (reduce) a.reduce = *reduce;
  a.n_layers = params->n_layers;
  SimView& S = a.sim;
  S.rec = work->rec;
  S.bound = work->bound;
  S.winner = work->winner;
  S.sus_imm = people->sus_imm;
  S.sus_stride = people->n;
  S.bitmap = work->bitmap;
  S.touched = work->touched;
  S.a_off = 0;
  S.seed = params->seed;
  S.n_agents = (uint32_t)people->n;
  S.n_variants = params->n_variants;
  S.use_sus_imm = params->use_waning;
  S.asymp_factor = params->asymp_factor;
  S.vl_hi = params->vl_hi;
  S.vl_lo = params->vl_lo;
  for (int v = 0; v < MAXV; ++v) S.beta_v[v] = params->beta_v[v];
  S.coarse = work->coarse;
  S.aux = work->aux;
  a.t = t;

  DeviceCaps& c = caps();
  const int64_t queue_bytes = (int64_t)EDGE_WAVES * EDGE_WAVE_WORDS * 4;
  int64_t budget = 0;
  int rc0 = edge_dynamic_lds(&budget);
  if (rc0) return rc0;
  const int64_t fine_words = bitmap_words_for(people->n), coarse_words = coarse_words_for(people->n);
  int mode = BITS_L2;
  if (g_edge_lds && fine_words * 4 + queue_bytes <= budget) mode = BITS_LDS;
  else if (g_edge_coarse && work->coarse && coarse_words * 4 + queue_bytes <= budget) mode = BITS_COARSE;
  a.lds_words = (int32_t)(mode == BITS_LDS ? fine_words : mode == BITS_COARSE ? coarse_words : 0);
  const int64_t wave_tile = 64 * 4 * (int64_t)g_edge_chunks;
  int64_t src_items = 0;
  for (int l = 0; l < params->n_layers; ++l) {
    const cvk_layer& L = layers[l];
    if (L.n_edges < 0 || (L.n_edges > 0 && (!L.p1 || !L.p2 || !L.beta)))
      return fail(CVK_E_INVALID, "cvk_edge_pass: invalid layer");
    if (((uintptr_t)L.p1 | (uintptr_t)L.p2) & 15) return fail(CVK_E_ALIGN, "cvk_edge_pass: p1/p2 must be 16-byte aligned");
    const bool listed = g_source_lists && L.sources && !L.dynamic && L.n_edges > 0;
    if (listed && (!L.sources->indptr || (L.sources->n_slots > 0 && (!L.sources->target || !L.sources->edge_dir))))
      return fail(CVK_E_INVALID, "cvk_edge_pass: invalid source lists");
    a.item_start[l] = src_items;
    if (listed) {
      a.src_mask |= 1u << l;
      a.src[l].indptr = L.sources->indptr;
      a.src[l].target = L.sources->target;
      a.src[l].edge_dir = L.sources->edge_dir;
      src_items += (people->n + SRC_ITEM_AGENTS - 1) / SRC_ITEM_AGENTS;
    }
    SegView& G = a.seg[l];
    G.p1 = L.p1;
    G.p2 = L.p2;
    G.beta = L.beta;
    G.unif = uniforms ? uniforms[l] : nullptr;
    G.n = L.n_edges;
    G.beta_layer = params->beta_layer[l];
    G.iso_factor = params->iso_factor[l];
    G.quar_factor = params->quar_factor[l];
    G.gen = generate && L.dynamic ? 1 : 0;
    G.beta_ones = L.beta_ones || G.gen;  // a resampled layer's weights are all 1
    G.l = l;
  }
  for (int l = params->n_layers; l <= CVK_MAX_LAYERS; ++l) a.item_start[l] = src_items;
  a.count_src = g_count_sources;
  // tile order: the listed layers (always walked: no tiles), then the others
  const bool always = g_source_lists == 1;
  int64_t wave_tiles = 0;
  int p = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int l = 0; l < params->n_layers; ++l) {
      const bool listed = (a.src_mask >> l) & 1u;
      if (listed != (pass == 0)) continue;
      a.pos_layer[p] = l;
      a.tile_start[p++] = wave_tiles;
      if (!(listed && always)) wave_tiles += (layers[l].n_edges + wave_tile - 1) / wave_tile;
    }
  for (int q = p; q <= CVK_MAX_LAYERS; ++q) a.tile_start[q] = wave_tiles;
  a.listed_tiles = 0;
  for (int q = 0; q < p; ++q)
    if ((a.src_mask >> a.pos_layer[q]) & 1u) a.listed_tiles = a.tile_start[q + 1];
  // the fused day loop (reduce given) runs the agent pass of day t right before, which then writes the target
  // bits and the largest target code (day_begin_launch makes the same choice); a standalone edge pass may
  // follow anything
  a.target_bits = reduce && reduce->n_blocks && reduce->transitions && target_bits_for(people->n, work) ? 1 : 0;
  a.src_auto = g_source_lists == 2 ? 1 : 0;
  a.src_max_permille = g_source_permille;
  a.agent_blocks = (uint32_t)agent_blocks_for(people->n);
  const int64_t work_items = std::max(wave_tiles, src_items);
  if (work_items == 0) return reduce ? day_reduce_launch(work, *reduce, st) : CVK_OK;
  // one workgroup per CU (the LDS bitmap takes most of a CU's LDS), at least one work item each
  const int grid = (int)std::min<int64_t>((work_items + EDGE_WAVES - 1) / EDGE_WAVES, (int64_t)std::max(c.n_cu, 1));
  const size_t shmem = (size_t)((int64_t)a.lds_words * 4 + queue_bytes);
  bool any_gen = false;
  for (int l = 0; l < params->n_layers; ++l) any_gen = any_gen || a.seg[l].gen;
  kEdgeLaunch[mode][g_edge_chunks == 4 ? 1 : 0][any_gen ? 1 : 0][g_edge_prefetch ? 1 : 0][a.target_bits ? 1 : 0](
      dim3(grid), dim3(EDGE_BLOCK),
                                                                                       shmem, st, a);
  return check_launch("cvk_edge_pass");
}

extern "C" int cvk_edge_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                             cvk_work* work, int32_t t, const double* const* uniforms, void* stream) {
  int rc = validate(people, params, work, t);
  if (rc) return rc;
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_edge_pass: layers is null");
  return edge_pass_impl(people, layers, params, work, t, uniforms, (hipStream_t)stream, nullptr, false);
}

// partial_blocks (fused day loop): non-null = the flow counts go to the infect partial rows, whose
// count of workgroups is stored there for the next reduction; null = atomics on the day row.
// resample_t >= 0: the launch also resamples the dynamic layers for day resample_t (extra workgroups);
// the caller sets generate so that no source lookup reads a layer being rewritten.
static int infect_impl(const cvk_people* people, const cvk_layer* layers, const cvk_params* params, cvk_work* work,
                       int32_t t, int32_t count_flows, int32_t allow_beds, hipStream_t st, bool generate = false,
                       uint32_t* partial_blocks = nullptr, int32_t resample_t = -1) {
  InfectArgs ia;
  memset(&ia, 0, sizeof(ia));
  for (int l = 0; l < params->n_layers; ++l) {
    ia.p1[l] = layers[l].p1;
    ia.p2[l] = layers[l].p2;
    if (generate && layers[l].dynamic) ia.gen_mask |= 1u << l;
  }
  ia.t = t;
  ia.n_layers = params->n_layers;
  ia.count_flows = count_flows;
  ia.allow_beds = allow_beds;
  // one touched-bitmap word per thread and sweep: enough workgroups for one sweep, at least one per CU and at
  // most three (the occupancy the kernel's register budget allows), so large populations keep more waves
  // in flight through the latency-bound outcome tree
  const int n_cu = std::max(caps().n_cu, 1);
  const int64_t want = (people->n / 32 + INFECT_BLOCK) / INFECT_BLOCK;
  const int grid = (int)std::min<int64_t>(std::min<int64_t>(std::max<int64_t>(want, n_cu), 3 * (int64_t)n_cu),
                                          INFECT_PART_BLOCKS);
  ia.partial = partial_blocks && count_flows ? 1 : 0;
  if (partial_blocks) *partial_blocks = ia.partial ? (uint32_t)grid : 0u;
  ia.infect_blocks = (uint32_t)grid;
  uint32_t rs_blocks = 0;
  if (resample_t >= 0) {
    DayBeginArgs ra;
    memset(&ra, 0, sizeof(ra));
    int rc = fill_resample(ra, layers, params, 1, &rs_blocks);
    if (rc) return rc;
    ia.rs_t = resample_t;
    ia.n_resample = ra.n_resample;
    for (int r = 0; r < ra.n_resample; ++r) ia.rs[r] = ra.rs[r];
  }
  hipLaunchKernelGGL(infect_pass_kernel, dim3(grid + rs_blocks), dim3(INFECT_BLOCK), 0, st, *people, *params, *work,
                     ia);
  return check_launch("cvk_infect_pass");
}

extern "C" int cvk_infect_pass(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                               cvk_work* work, int32_t t, int32_t count_flows, void* stream) {
  int rc = validate(people, params, work, t);
  if (rc) return rc;
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "cvk_infect_pass: layers is null");
  return infect_impl(people, layers, params, work, t, count_flows, count_flows, (hipStream_t)stream);
}

static bool has_dynamic(const cvk_layer* layers, const cvk_params* params) {
  for (int l = 0; l < params->n_layers; ++l)
    if (layers[l].dynamic && layers[l].n_edges > 0) return true;
  return false;
}

// How a run of days [t0, t1) treats dynamic layers. gen: no day's resample is written until the last
// day is closed; the edge and infect passes compute the endpoints from the resample stream (never
// with contact tracing, which reads each day's contacts). overlap: day t+1's resample runs as extra
// workgroups of day t's infect pass (store-bound work beside latency-bound work), which then takes
// dynamic-layer sources from the resample stream.
static void day_plan(const cvk_layer* layers, const cvk_params* params, const cvk_imports* imports, int32_t n_days,
                     bool* gen, bool* overlap) {
  const bool diag = imports && imports->diag;
  int64_t dyn_edges = 0;
  for (int l = 0; l < params->n_layers; ++l) dyn_edges += layers[l].dynamic ? layers[l].n_edges : 0;
  *gen = generate_dynamic_for(dyn_edges, diag) && has_dynamic(layers, params);
  *overlap = !*gen && g_overlap_resample && n_days > 1 && has_dynamic(layers, params);
}

extern "C" int cvk_day_plan(const cvk_layer* layers, const cvk_params* params, const cvk_imports* imports,
                            int32_t n_days, int32_t* out) {
  if (!params || !out || (params->n_layers > 0 && !layers)) return fail(CVK_E_INVALID, "cvk_day_plan: null argument");
  if (params->n_layers < 0 || params->n_layers > CVK_MAX_LAYERS) return fail(CVK_E_INVALID, "cvk_day_plan: n_layers");
  bool gen, overlap;
  day_plan(layers, params, imports, n_days, &gen, &overlap);
  out[0] = gen ? 1 : 0;
  out[1] = overlap ? 1 : 0;
  return CVK_OK;
}

static int run_days_impl(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                         int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports, hipStream_t st) {
  int rc;
  const bool diag = imports && imports->diag;
  bool gen, overlap;
  day_plan(layers, params, imports, t1 - t0, &gen, &overlap);
  uint32_t infect_blocks = 0;  // infect partial rows not yet added to their day's row
  const bool sched = imports && imports->n_beta_days > 0 && imports->beta_v_days && imports->beta_layer_days;
  cvk_params day_params = *params;
  for (int32_t t = t0; t < t1; ++t) {
    if (sched) {  // this day's rates; kernel arguments are copied at launch (and into a captured graph)
      const int32_t row = std::min<int32_t>(t, imports->n_beta_days - 1);
      memcpy(day_params.beta_v, imports->beta_v_days + (int64_t)row * CVK_MAX_VARIANTS, sizeof(day_params.beta_v));
      memcpy(day_params.beta_layer, imports->beta_layer_days + (int64_t)row * CVK_MAX_LAYERS,
             sizeof(day_params.beta_layer));
      params = &day_params;
    }
    if (t == 0 && pop_infected > 0) {  // init_infections, sim.py:505-532 (flows not counted)
      rc = choose_impl(people, params, work, 0, STREAM_CHOOSE, pop_infected, 0, CVK_LAYER_SEED, 0, st);
      if (rc) return rc;
      rc = infect_impl(people, layers, params, work, 0, 0, 0, st);
      if (rc) return rc;
    }
    // day t begins with the stocks of day t-1 (already closed if t == t0)
    DayReduce r;
    // the first resample of the run also sets every weight to 1; later ones leave them as they are
    const int rs_mode = gen ? 0 : (t == t0 ? 2 : (overlap ? 0 : 1));
    rc = day_begin_launch(people, layers, params, work, t, t > t0 ? 1 : 0, 1, rs_mode, st, &r);
    if (rc) return rc;
    r.infect_blocks = infect_blocks;  // the previous day's flow counts
    r.infect_t = t - 1;
    if (imports) {
      for (int j = 0; j < imports->n; ++j) {
        if (imports->days[j] != t) continue;
        const int32_t v = imports->variants[j];
        rc = choose_impl(people, params, work, t, STREAM_IMPORT | (uint32_t)v, imports->n_imports[j], 1,
                         CVK_LAYER_IMPORT, v, st);
        if (rc) return rc;
      }
    }
    if (diag) {  // interventions that test and trace, then update_states_post (sim.py:591-596)
      rc = diag_day(people, layers, params, work, imports, t, st);
      if (rc) return rc;
    }
    rc = edge_pass_impl(people, layers, params, work, t, nullptr, st, &r, gen);  // block 0 reduces day_begin's rows
    if (rc) return rc;
    // tomorrow's resample (once today's edges have been read) with today's infections
    rc = infect_impl(people, layers, params, work, t, 1, 1, st, gen || overlap, &infect_blocks,
                     overlap && t + 1 < t1 ? t + 1 : -1);
    if (rc) return rc;
  }
  // close the last day: stocks and NAb update, and (generated layers) the last day's contacts, weights 1
  DayReduce r;
  rc = day_begin_launch(people, layers, params, work, t1 - 1, 0, 0, gen ? 2 : 0, st, &r);
  if (rc) return rc;
  r.infect_blocks = infect_blocks;
  r.infect_t = t1 - 1;
  return day_reduce_launch(work, r, st);
}

static int check_run_args(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                          int32_t t0, int32_t t1, const cvk_imports* imports) {
  int rc = validate(people, params, work, t0);
  if (rc) return rc;
  if (t1 < t0 || t1 > work->npts) return fail(CVK_E_INVALID, "bad day range");
  if (params->n_layers > 0 && !layers) return fail(CVK_E_INVALID, "layers is null");
  return check_diag_schedule(people, layers, params, work, imports);
}

extern "C" int cvk_graph_create(const cvk_people* people, cvk_layer* layers, const cvk_params* params,
                                cvk_work* work, int32_t t0, int32_t t1, int64_t pop_infected,
                                const cvk_imports* imports, void* stream, void** graph_exec_out);
extern "C" int cvk_graph_launch(void* graph_exec, void* stream);
extern "C" int cvk_graph_destroy(void* graph_exec);

extern "C" int cvk_run_days(const cvk_people* people, cvk_layer* layers, const cvk_params* params, cvk_work* work,
                            int32_t t0, int32_t t1, int64_t pop_infected, const cvk_imports* imports,
                            int32_t use_graph, void* stream) {
  int rc = check_run_args(people, layers, params, work, t0, t1, imports);
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
  int rc = check_run_args(people, layers, params, work, t0, t1, imports);
  if (rc) return rc;
  if (t1 <= t0) return fail(CVK_E_INVALID, "cvk_graph_create: empty day range");
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) return fail(CVK_E_INVALID, "cvk_graph_create: graph capture needs a non-default stream");
  int64_t budget = 0;
  rc = edge_dynamic_lds(&budget);  // not a stream operation: done before capture
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
