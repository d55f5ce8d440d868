// Batched day loop: S independent sims share one launch per phase per day.
//
// Storage: agents of all sims concatenated (sim s owns [a_off[s], a_off[s+1])); each layer's edge
// lists concatenated over sims with endpoints given as indices within the sim; one infectious and
// one susceptible bitmap slice per sim (16-byte aligned). Work maps:
//   agent pass   workgroup b handles AGENT_BLOCK * APT agents of sim block_sim[b] (blocks never
//                straddle sims), so a block reads one sim's parameters with scalar loads
//   edge pass    tiles of one (sim, layer) segment; a persistent grid pulls tiles in order from an
//                atomic counter and reloads its LDS bitmap when the next tile belongs to another sim
//   infect pass  the day's targets of all sims in one list; each lane finds its sim by binary search
//                over the agent offsets (LDS) and uses that sim's parameters
//   choose       segmented 8-pass radix select (per-sim histograms and thresholds)
// Draws and keys use indices within the sim, so sim s of a batch reproduces cvk_run_days for s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "cvk_host.h"
#include "cvk_kernels.h"

namespace cvk {

constexpr int RS_CHUNK = 2048;  // edges per resample work item

struct BatchSim {
  int64_t a_off;     // storage index of local agent 0
  int64_t n;         // agents
  int64_t bm_off;    // first 32-bit word of the sim's bitmap slices
  uint32_t b_off;    // first agent-pass block of the sim
  uint32_t n_blocks; // agent-pass blocks of the sim
};

struct BatchCommon {
  cvk_people P;
  int64_t n_total;
  const BatchSim* sims;
  const cvk_params* params;
  int32_t n_sims, n_layers, npts, pad;
  uint32_t* results;     // [S][npts][CVK_R_COLS]
  double* results_f;     // [S][npts][CVK_RF_COLS]
  uint32_t* partials;    // [CVK_PARTIAL_COLS][part_stride]
  double* partials_f;    // [CVK_RF_COLS][part_stride]
  int64_t part_stride;
  uint64_t* rec;
  uint16_t* bound;
  uint32_t* bitmap;
  uint64_t* winner;
  uint32_t* touched;     // bitmap over storage indices: targets holding a key
  uint32_t* ctl;         // global counters (CVK_CTL_* indices)
};

// ---------------------------------------------------------------------------
// agent pass + resample
// ---------------------------------------------------------------------------
struct BatchDayArgs {
  int32_t t, prologue, transitions, resample;
  uint32_t agent_blocks;
  int32_t n_rs_chunks;
  const int32_t* block_sim;   // [agent_blocks]
  const int32_t* rs_chunk_seg;     // [n_rs_chunks] segment (s * L + l) of each resample chunk
  const int64_t* rs_chunk_first;   // [n_rs_chunks] first edge (within the segment) of the chunk
  const SegView* segs;        // [S * L]
};

// The resample work items [k0, n) with stride `step` (2048-edge chunks of dynamic segments) for day t.
__device__ inline void batch_resample_chunks(const BatchCommon& C, const SegView* segs, const int32_t* chunk_seg,
                                             const int64_t* chunk_first, int32_t k0, int32_t n, int32_t step,
                                             int32_t t, bool weights) {
  for (int32_t k = k0; k < n; k += step) {
    const int32_t seg = chunk_seg[k];
    const int32_t s = seg / C.n_layers;
    const SegView& L = segs[seg];
    const int64_t e0 = chunk_first[k];
    const int64_t e1 = std::min<int64_t>(e0 + RS_CHUNK, L.n);
    const uint64_t seed = C.params[s].seed;
    const uint32_t na = (uint32_t)C.sims[s].n;
    const uint32_t thr = resample_thr(na);
    for (int64_t q = e0 / 2 + threadIdx.x; q < (e1 + 1) / 2; q += AGENT_BLOCK)  // chunks start at even edges
      resample_pair(seed, t, (uint32_t)L.l, na, thr, q, L.n, (int32_t*)L.p1, (int32_t*)L.p2,
                    weights ? (float*)L.beta : nullptr);
  }
}

template <int APT>
__global__ __launch_bounds__(AGENT_BLOCK) void batch_day_begin_kernel(BatchCommon C, BatchDayArgs a) {
  if (blockIdx.x >= a.agent_blocks) {  // resample: one 2048-edge chunk of one dynamic segment per iteration
    batch_resample_chunks(C, a.segs, a.rs_chunk_seg, a.rs_chunk_first, (int32_t)(blockIdx.x - a.agent_blocks),
                          a.n_rs_chunks, (int32_t)(gridDim.x - a.agent_blocks), a.t, a.resample == 2);
    return;
  }
  const int32_t s = a.block_sim[blockIdx.x];
  const BatchSim& B = C.sims[s];
  AgentSpan S;
  S.g0 = B.a_off;
  S.li0 = (int64_t)(blockIdx.x - B.b_off) * (AGENT_BLOCK * APT);
  S.n_local = B.n;
  S.n_stride = C.n_total;
  S.bitmap = (uint64_t*)(C.bitmap + B.bm_off);
  S.coarse = nullptr;  // the batched edge pass stages sims' whole bitmaps
  S.rec = C.rec + B.a_off;
  S.bound = C.bound + B.a_off;
  S.aux = nullptr;
  S.partials = C.partials;
  S.partials_f = C.partials_f;
  S.part_stride = C.part_stride;
  S.part_col = blockIdx.x;
  agent_block<APT, false>(C.P, C.params[s], S, a.t, a.prologue, a.transitions);
}

struct BatchReduce {
  int32_t active, t, prologue, transitions;
};

// Sums sim s's partial rows and writes its day rows (block-wide).
__device__ void batch_reduce_sim(const BatchCommon& C, const BatchReduce& r, int32_t s) {
  __shared__ uint32_t s_tot[N_SUM_COLS];
  __shared__ double s_totf[N_AGENT_FCOLS];
  const BatchSim& B = C.sims[s];
  sum_partials<N_SUM_COLS, N_AGENT_FCOLS>(C.partials, C.partials_f, C.part_stride, B.b_off, B.n_blocks, s_tot,
                                            s_totf);
  write_agent_rows(C.results + (int64_t)s * C.npts * CVK_R_COLS, C.results_f + (int64_t)s * C.npts * CVK_RF_COLS,
                   s_tot, s_totf, r.t, r.prologue, r.transitions);
}

__global__ __launch_bounds__(1024) void batch_reduce_kernel(BatchCommon C, BatchReduce r) {
  for (int32_t s = blockIdx.x; s < C.n_sims; s += gridDim.x) batch_reduce_sim(C, r, s);
}

// ---------------------------------------------------------------------------
// edge pass
// ---------------------------------------------------------------------------
struct BatchEdgeArgs {
  const SimView* views;        // [S]
  const SegView* segs;         // [S * L]
  const int64_t* seg_tile_start;  // [S * L + 1] first wave tile of each segment
  const int32_t* tile_seg;     // [n_tiles] segment of each wave tile
  int64_t n_tiles;
  int32_t lds_words;           // bitmap words the LDS stage holds (larger sims read L2)
  int32_t t;
  uint32_t gen_mask;           // bit l: layer l's endpoints of day t come from the resample stream (GEN)
  uint32_t src_mask;           // bit l: layer l has source lists (every sim); walked sims skip its tiles
  // source work items (SRC_ITEM_AGENTS agents of one listed segment each) of the walk phase
  const SrcView* srcs;         // [S * L]
  const int32_t* item_seg;     // [n_items] segment of each item
  const int64_t* seg_item_start;  // [S * L + 1] first item of each segment
  int64_t n_items;             // 0: no walk phase
  const uint32_t* walk;        // [S] nonzero: sim s walks its listed layers today
  BatchReduce reduce;
};

// Day t's rates of every sim (block s) into the views the edge pass reads: beta_v of the sim, beta_layer of
// its segments (change_beta schedules; the kernels read no other copy of them).
__global__ __launch_bounds__(64) void batch_rates_kernel(SimView* views, SegView* segs, const float* bv_days,
                                                         const float* bl_days, int32_t n_days, int32_t n_layers,
                                                         int32_t t) {
  const int32_t s = blockIdx.x;
  const int64_t row = (int64_t)s * n_days + (t < n_days ? t : n_days - 1);
  if (threadIdx.x < MAXV) views[s].beta_v[threadIdx.x] = bv_days[row * CVK_MAX_VARIANTS + threadIdx.x];
  if ((int)threadIdx.x < n_layers)
    segs[(int64_t)s * n_layers + threadIdx.x].beta_layer = bl_days[row * CVK_MAX_LAYERS + threadIdx.x];
}

// The day's walk decision of every sim (block s): walk its listed layers unless automatic and the agent
// pass set more than max_permille infectious bits per thousand of its agents (the partial column DC_BITS
// of the sim's agent-pass blocks), as the single-sim edge pass decides.
__global__ __launch_bounds__(256) void batch_walk_kernel(BatchCommon C, int32_t automatic, int32_t max_permille,
                                                         uint32_t* walk) {
  __shared__ uint32_t s_w[4];
  const int32_t s = blockIdx.x;
  const BatchSim bs = C.sims[s];
  uint32_t cnt = 0;
  for (uint32_t b = threadIdx.x; b < bs.n_blocks; b += 256)
    cnt += C.partials[(int64_t)DC_BITS * C.part_stride + bs.b_off + b];
  cnt = wave_sum(cnt);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t inf = (uint64_t)s_w[0] + s_w[1] + s_w[2] + s_w[3];
    walk[s] = !automatic || inf * 1000u <= (uint64_t)max_permille * (uint64_t)bs.n ? 1u : 0u;
  }
}

// Workgroup b owns the wave tiles [T*b/G, T*(b+1)/G) of the day, ordered by (sim, layer). It walks
// them one sim at a time: the sim's infectious bitmap is staged in LDS (when it fits), then the
// block's waves take the sim's tiles from an LDS counter, as in the single-sim pass. A wave's exact
// queue is drained when its tiles change segment and when the sim is done. GEN: the dynamic layers'
// endpoints are computed from the day's resample stream instead of read (a.gen_mask).
template <int BLOCK, int K, bool GEN>
__global__ __launch_bounds__(BLOCK) void batch_edge_kernel(BatchCommon C, BatchEdgeArgs a) {
  extern __shared__ uint32_t s_mem[];
  __shared__ uint32_t s_next;
  __shared__ float s_beta[MAXV];
  const WaveQ Q = wave_queues(s_mem + a.lds_words + (threadIdx.x >> 6) * EDGE_WAVE_WORDS, s_beta);
  if (a.reduce.active)  // the agent pass's counts, one sim per block
    for (int32_t s = blockIdx.x; s < C.n_sims; s += gridDim.x) batch_reduce_sim(C, a.reduce, s);
  if (a.n_items) {  // walk phase: source items over the grid's waves, items of non-walking sims skipped
    // the waves walk different sims, so each keeps its current sim's beta_v in its own LDS row
    __shared__ float s_wbeta[BLOCK / 64][MAXV];
    WaveQ QW = Q;
    QW.beta = s_wbeta[threadIdx.x >> 6];
    uint64_t cnt[2] = {0, 0};
    uint32_t n_w = 0;
    int32_t w_seg = -1, w_sim = -1;
    SimView SW;
    SegView LW;
    const int64_t stride = (int64_t)gridDim.x * (BLOCK / 64);
    for (int64_t it = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); it < a.n_items; it += stride) {
      const int32_t seg = __builtin_amdgcn_readfirstlane(a.item_seg[it]);
      const int32_t s = seg / C.n_layers;
      if (!__builtin_amdgcn_readfirstlane(a.walk[s])) continue;
      if (seg != w_seg) {
        if (n_w) flush_exact(SW, LW, a.t, QW, n_w);  // with the queued sim's beta_v still in place
        n_w = 0;
        w_seg = seg;
        SW = uniform_copy(a.views[s]);
        LW = uniform_copy(a.segs[seg]);
        if (s != w_sim) {
          w_sim = s;
          wave_sync_lds();
          if (lane_id() < (uint32_t)MAXV) s_wbeta[threadIdx.x >> 6][lane_id()] = gmem(a.views)[s].beta_v[lane_id()];
          wave_sync_lds();
        }
      }
      const SrcView R = uniform_copy(a.srcs[seg]);
      source_item_pass<BITS_L2>(SW, LW, R, a.t, nullptr, QW, (it - a.seg_item_start[seg]) * (int64_t)SRC_ITEM_AGENTS,
                                n_w, cnt);
    }
    if (n_w) flush_exact(SW, LW, a.t, QW, n_w);
  }
  const int64_t first = a.n_tiles * blockIdx.x / gridDim.x, last = a.n_tiles * (blockIdx.x + 1) / gridDim.x;
  int64_t tile0 = first;
  while (tile0 < last) {  // block-uniform: one sim of the range per iteration
    const int32_t s = a.tile_seg[tile0] / C.n_layers;
    const int64_t sim_end = a.seg_tile_start[(int64_t)(s + 1) * C.n_layers];
    const int64_t sub_last = sim_end < last ? sim_end : last;
    const SimView S = uniform_copy(a.views[s]);
    const int32_t nw = (int32_t)(((int64_t)S.n_agents + 127) / 128 * 4);
    const bool in_lds = nw <= a.lds_words;
    __syncthreads();  // every wave is done with the previous sim's bitmap and counter
    if (in_lds) {
      const uint4* src = (const uint4*)S.bitmap;
      uint4* dst = (uint4*)s_mem;
      for (int j = threadIdx.x; j < (nw >> 2); j += BLOCK) dst[j] = src[j];
    }
    if (threadIdx.x == 0) s_next = 0;
    if (threadIdx.x < MAXV) s_beta[threadIdx.x] = gmem(a.views)[s].beta_v[threadIdx.x];
    __syncthreads();
    uint32_t n_x = 0;
    int32_t q_seg = -1;
    SegView L = uniform_copy(a.segs[a.tile_seg[tile0]]);  // segment of the wave's tiles and queued tests
    if (GEN) L.gen = (int32_t)((a.gen_mask >> L.l) & 1u);
    for (;;) {
      uint32_t k = 0;
      if (lane_id() == 0) k = atomicAdd(&s_next, 1u);
      k = __builtin_amdgcn_readfirstlane(__shfl(k, 0, 64));
      const int64_t tile = tile0 + k;
      if (tile >= sub_last) break;
      const int32_t seg = a.tile_seg[tile];
      if (((a.src_mask >> (seg % C.n_layers)) & 1u) && a.walk[s]) continue;  // walked above
      WaveTile<K> T;
      T.seg = seg;
      T.e0 = (tile - a.seg_tile_start[seg]) * (int64_t)(64 * 4 * K);
      if (seg != q_seg) {
        if (n_x) flush_exact(S, L, a.t, Q, n_x);
        n_x = 0;
        q_seg = seg;
        L = uniform_copy(a.segs[seg]);
        if (GEN) L.gen = (int32_t)((a.gen_mask >> L.l) & 1u);
      }
      load_wave_edges<K, GEN>(S, L, a.t, T);
      if (in_lds) wave_tile_pass<BITS_LDS, K>(S, L, a.t, s_mem, Q, T, n_x);
      else wave_tile_pass<BITS_L2, K>(S, L, a.t, nullptr, Q, T, n_x);
    }
    if (n_x) flush_exact(S, L, a.t, Q, n_x);
    tile0 = sub_last;
  }
}

// ---------------------------------------------------------------------------
// infect pass
// ---------------------------------------------------------------------------
struct BatchInfectArgs {
  const int64_t* a_off;   // [S + 1]
  const SegView* segs;    // [S * L]
  int32_t* log;
  int64_t log_capacity;
  int32_t t, count_flows, allow_beds, n_variants;
  uint32_t infect_blocks;      // workgroups [0, infect_blocks) infect; the rest resample day rs_t
  uint32_t gen_mask;           // bit l: layer l's sources come from the resample stream of day t
  int32_t rs_t;
  int32_t n_rs_chunks;
  const int32_t* rs_chunk_seg;
  const int64_t* rs_chunk_first;
};

// Occupancy hint 3 waves/SIMD, as for the single-sim infect pass: resample workgroups share the kernel.
__global__ __launch_bounds__(INFECT_BLOCK, 3) void batch_infect_kernel(BatchCommon C, BatchInfectArgs ia) {
  if (blockIdx.x >= ia.infect_blocks) {  // the next day's resample
    batch_resample_chunks(C, ia.segs, ia.rs_chunk_seg, ia.rs_chunk_first, (int32_t)(blockIdx.x - ia.infect_blocks),
                          ia.n_rs_chunks, (int32_t)(gridDim.x - ia.infect_blocks), ia.rs_t, false);
    return;
  }
  __shared__ int32_t s_off[CVK_BATCH_MAX_SIMS + 1];  // agent offsets (total agents < 2^31)
  for (int j = threadIdx.x; j <= C.n_sims; j += INFECT_BLOCK) s_off[j] = (int32_t)ia.a_off[j];
  __syncthreads();
  const int32_t t = ia.t;
  __shared__ int32_t s_list[INFECT_BLOCK / 64][INFECT_WAVE_LIST];
  int32_t* list = s_list[threadIdx.x >> 6];
  const int64_t n_words = (C.n_total + 31) / 32;
  const int64_t stride = (int64_t)ia.infect_blocks * INFECT_BLOCK;
  const int64_t n_sweeps = (n_words + stride - 1) / stride;
  for (int64_t sw = 0; sw < n_sweeps; ++sw) {
    const int64_t w = sw * stride + (int64_t)blockIdx.x * INFECT_BLOCK + threadIdx.x;
    const uint32_t total = compact_touched(C.touched, w, n_words, C.n_total, list);
    for (uint32_t base = 0; base < total; base += 64) {  // wave-uniform: one target per lane per round
      InfectOut o = {KEY_NONE, false, false, false, false, 0u, 0u};
      int32_t source = -1, s = 0;
      int64_t il = 0;
      const uint32_t jj = base + lane_id();
      const int32_t gi = jj < total ? list[jj] : -1;
      if (gi >= 0) {
        const int64_t g = gi;
        int32_t lo = 0, hi = C.n_sims - 1;  // the sim with s_off[s] <= g < s_off[s + 1]
        while (lo < hi) {
          const int32_t mid = (lo + hi + 1) >> 1;
          if (s_off[mid] <= g) lo = mid;
          else hi = mid - 1;
        }
        s = lo;
        il = g - s_off[s];
        const cvk_params& prm = C.params[s];
        const uint32_t* row = C.results + ((int64_t)s * C.npts + t) * CVK_R_COLS;
        const bool hosp_max = ia.allow_beds && prm.n_beds_hosp >= 0 && (int64_t)row[CVK_R_MID_SEVERE] > prm.n_beds_hosp;
        const bool icu_max = ia.allow_beds && prm.n_beds_icu >= 0 && (int64_t)row[CVK_R_MID_CRITICAL] > prm.n_beds_icu;
        o = infect_target(C.P, prm, C.winner, g, (uint32_t)il, C.n_total, t, hosp_max, icu_max);
        if (o.inf && (int32_t)o.l < C.n_layers) {
          const SegView& L = ia.segs[s * C.n_layers + (int32_t)o.l];
          source = key_source(o.key, L.p1, L.p2, (ia.gen_mask >> o.l) & 1u, prm.seed, t,
                              (uint32_t)(s_off[s + 1] - s_off[s]));
        }
      }
      // infection log, people.py:509-511 (indices within the sim, sim in bits 12+ of the last word)
      const uint64_t m = __ballot(o.inf);
      if (m) {
        unsigned long long base = 0;
        if (lane_id() == 0)
          base = atomicAdd((unsigned long long*)&C.ctl[CVK_CTL_LOG_COUNT], (unsigned long long)__popcll(m));
        base = __shfl(base, 0, 64);
        if (o.inf) {
          const unsigned long long pos = base + lanes_below(m);
          if ((int64_t)pos < ia.log_capacity) {
            int4 rec;
            rec.x = source;
            rec.y = (int32_t)il;
            rec.z = t;
            rec.w = (int32_t)(o.l | (o.v << 8) | ((uint32_t)s << 12));
            ((int4*)ia.log)[pos] = rec;
          } else {
            C.ctl[CVK_CTL_LOG_OVERFLOW] = 1;
          }
        }
      }
      // flows: lanes peeled by sim (targets of one sim are adjacent), one atomic per column per sim
      bool pending = ia.count_flows && o.inf;
      while (true) {
        const uint64_t pm = __ballot(pending);
        if (!pm) break;
        const int lead = __ffsll((unsigned long long)pm) - 1;
        const int32_t sl = __shfl(s, lead, 64);
        const bool mine = pending && s == sl;
        uint32_t cnt[N_INFECT_COLS];
        cnt[0] = (uint32_t)__popcll(__ballot(mine));
        cnt[1] = (uint32_t)__popcll(__ballot(mine && o.reinf));
        for (int vv = 0; vv < MAXV; ++vv) {
          if (vv < ia.n_variants) {
            cnt[2 + vv] = (uint32_t)__popcll(__ballot(mine && o.v == (uint32_t)vv));
            cnt[2 + MAXV + vv] = (uint32_t)__popcll(__ballot(mine && o.symp && o.v == (uint32_t)vv));
            cnt[2 + 2 * MAXV + vv] = (uint32_t)__popcll(__ballot(mine && o.sev && o.v == (uint32_t)vv));
          } else {
            cnt[2 + vv] = cnt[2 + MAXV + vv] = cnt[2 + 2 * MAXV + vv] = 0;
          }
        }
        if ((int)lane_id() == lead) {
          uint32_t* row = C.results + ((int64_t)sl * C.npts + t) * CVK_R_COLS;
          for (int k = 0; k < N_INFECT_COLS; ++k)
            if (cnt[k]) atomicAdd(&row[infect_col(k)], cnt[k]);
        }
        pending = pending && !mine;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the list is rewritten by the next sweep
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ---------------------------------------------------------------------------
// segmented choose (seeds, imports): per-sim k smallest selection keys
// ---------------------------------------------------------------------------
struct ChooseState {
  uint32_t* hist;      // [S][256]
  uint64_t* prefix;    // [S]
  uint32_t* remain;    // [S]
  uint32_t* all;       // [S] everyone eligible is chosen
  uint32_t* shortfall; // [S] fewer eligible than requested
  const int64_t* k;    // [S] this call's counts
};

__global__ void batch_choose_init_kernel(ChooseState cs, int32_t n_sims) {
  const int32_t s = blockIdx.x;
  if (s >= n_sims) return;
  cs.hist[(int64_t)s * 256 + threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    cs.prefix[s] = 0;
    cs.remain[s] = (uint32_t)cs.k[s];
    cs.all[s] = 0;
  }
}

template <int APT>
__global__ __launch_bounds__(AGENT_BLOCK) void batch_choose_hist_kernel(BatchCommon C, ChooseState cs,
                                                                        const int32_t* block_sim, int32_t t,
                                                                        uint32_t stream_id, int32_t sus_only, int pass) {
  __shared__ uint32_t s_hist[256];
  const int32_t s = block_sim[blockIdx.x];
  if (cs.k[s] == 0 || cs.all[s]) return;  // block-uniform
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const BatchSim& B = C.sims[s];
  const uint64_t seed = C.params[s].seed;
  const uint64_t prefix = cs.prefix[s];
  const int shift = 56 - 8 * pass;
  const int64_t li0 = (int64_t)(blockIdx.x - B.b_off) * (AGENT_BLOCK * APT);
  for (int r = 0; r < APT; ++r) {
    const int64_t li = li0 + (int64_t)r * AGENT_BLOCK + threadIdx.x;
    if (li >= B.n) continue;
    if (sus_only && !(C.P.flags[B.a_off + li] & CVK_SUSCEPTIBLE)) continue;
    const uint64_t key = selection_key(seed, t, stream_id, (uint32_t)li);
    if (pass > 0 && (key >> (shift + 8)) != (prefix >> (shift + 8))) continue;
    atomicAdd(&s_hist[(key >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  if (s_hist[threadIdx.x]) atomicAdd(&cs.hist[(int64_t)s * 256 + threadIdx.x], s_hist[threadIdx.x]);
}

// One wave per sim: picks the byte of the k-th smallest key among the candidates of this pass.
__global__ __launch_bounds__(64) void batch_choose_select_kernel(ChooseState cs, int32_t n_sims, int pass) {
  const int32_t s = blockIdx.x;
  if (s >= n_sims || cs.k[s] == 0 || cs.all[s]) return;
  uint32_t* h = cs.hist + (int64_t)s * 256;
  const uint32_t lane = lane_id();
  uint32_t b[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = h[lane * 4 + j];
  const uint32_t lane_total = (b[0] + b[1]) + (b[2] + b[3]);
  uint32_t incl = lane_total;  // inclusive scan over lanes
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if ((int)lane >= off) incl += y;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  uint32_t remain = cs.remain[s];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; ++j) h[lane * 4 + j] = 0;
  if (pass == 0 && remain >= total) {
    if (lane == 0) {
      cs.all[s] = 1;
      if (remain > total) cs.shortfall[s] = 1;
    }
    return;
  }
  // the first bin whose cumulative count reaches `remain`
  uint32_t before = incl - lane_total;
  int digit = -1;
  uint32_t cum_before_digit = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (digit < 0 && before + b[j] >= remain) {
      digit = (int)(lane * 4 + j);
      cum_before_digit = before;
    }
    before += b[j];
  }
  const uint64_t found = __ballot(digit >= 0);
  const int first = __ffsll((unsigned long long)found) - 1;
  if ((int)lane == first) {
    const int shift = 56 - 8 * pass;
    cs.prefix[s] |= (uint64_t)digit << shift;
    cs.remain[s] = remain - cum_before_digit;
  }
}

template <int APT>
__global__ __launch_bounds__(AGENT_BLOCK) void batch_choose_mark_kernel(BatchCommon C, ChooseState cs,
                                                                        const int32_t* block_sim, int32_t t,
                                                                        uint32_t stream_id, int32_t sus_only,
                                                                        uint32_t layer_code, uint32_t variant) {
  const int32_t s = block_sim[blockIdx.x];
  if (cs.k[s] == 0) return;
  const BatchSim& B = C.sims[s];
  const uint64_t seed = C.params[s].seed;
  const uint64_t thr = cs.all[s] ? ~0ull : cs.prefix[s];
  const int64_t li0 = (int64_t)(blockIdx.x - B.b_off) * (AGENT_BLOCK * APT);
  for (int r = 0; r < APT; ++r) {
    const int64_t li = li0 + (int64_t)r * AGENT_BLOCK + threadIdx.x;
    if (li >= B.n) continue;
    const int64_t g = B.a_off + li;
    if (sus_only && !(C.P.flags[g] & CVK_SUSCEPTIBLE)) continue;
    const uint64_t key = selection_key(seed, t, stream_id, (uint32_t)li);
    if (key > thr) continue;
    const uint64_t wk = make_key(0, variant, layer_code, 0, (uint64_t)li);
    const uint64_t old = atomicMin((unsigned long long*)&C.winner[g], (unsigned long long)wk);
    if (old == KEY_NONE) atomicOr(&C.touched[g >> 5], 1u << (g & 31));
  }
}

// ---------------------------------------------------------------------------
// fills (graph-safe replacements of memset)
// ---------------------------------------------------------------------------
__global__ void fill_u64_kernel(uint64_t* p, int64_t n, uint64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ===========================================================================
// host side
// ===========================================================================
template <class T>
static int dev_alloc(T** p, int64_t count) {
  *p = nullptr;
  if (count <= 0) return CVK_OK;
  if (hipMalloc((void**)p, (size_t)count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return fail(CVK_E_HIP, "cvk_batch: device allocation failed");
  }
  return CVK_OK;
}

template <class T>
static int upload(T* dst, const T* src, int64_t count) {
  if (count <= 0) return CVK_OK;
  if (hipMemcpy(dst, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_batch: upload failed");
  return CVK_OK;
}

struct Batch {
  int device = -1;
  int n_cu = 1, max_lds = 65536;
  int32_t S = 0, L = 0, V = 0, npts = 0;
  int64_t N = 0;
  cvk_people P;
  std::vector<int64_t> a_off;          // [S + 1]
  std::vector<cvk_params> params;
  std::vector<int32_t> dynamic;
  std::vector<SegView> segs;           // [S * L]
  std::vector<SimView> views;          // [S]
  std::vector<BatchSim> sims;          // [S]
  cvk_batch_config cfg{1024, 4, 1, 2};
  int64_t n_tiles = 0, n_rs_chunks = 0;
  int64_t dyn_edges = 0;  // edges of all sims' dynamic layers
  uint32_t agent_blocks = 0;
  int64_t part_stride = 0;
  int64_t bitmap_words = 0;
  int64_t log_capacity = 0;
  int64_t max_sim_words = 0;
  // device
  BatchSim* d_sims = nullptr;
  cvk_params* d_params = nullptr;
  SegView* d_segs = nullptr;
  SimView* d_views = nullptr;
  int64_t* d_a_off = nullptr;
  int64_t* d_seg_tile_start = nullptr;
  int32_t* d_tile_seg = nullptr;
  int32_t* d_block_sim = nullptr;
  int32_t* d_rs_chunk_seg = nullptr;
  int64_t* d_rs_chunk_first = nullptr;
  uint64_t* d_rec = nullptr;
  uint16_t* d_bound = nullptr;
  uint32_t* d_bitmap = nullptr;
  uint64_t* d_winner = nullptr;
  uint32_t* d_touched = nullptr;
  uint32_t* d_ctl = nullptr;
  uint32_t* d_results = nullptr;
  double* d_results_f = nullptr;
  uint32_t* d_partials = nullptr;
  double* d_partials_f = nullptr;
  int32_t* d_log = nullptr;
  uint32_t* d_hist = nullptr;
  uint64_t* d_prefix = nullptr;
  uint32_t* d_remain = nullptr;
  uint32_t* d_all = nullptr;
  uint32_t* d_short = nullptr;
  int64_t* d_k = nullptr;        // [k_slots][S] counts for choose calls
  // source lists of the static layers (walked on quiet days)
  uint32_t src_mask = 0;
  int64_t n_items = 0;
  SrcView* d_srcs = nullptr;           // [S * L]
  int32_t* d_item_seg = nullptr;       // [n_items]
  int64_t* d_seg_item_start = nullptr;  // [S * L + 1]
  uint32_t* d_walk = nullptr;          // [S]
  // per-day rates (change_beta): [S][n_beta_days][CVK_MAX_VARIANTS] and [S][n_beta_days][CVK_MAX_LAYERS]
  int32_t n_beta_days = 0;
  float* d_bv_days = nullptr;
  float* d_bl_days = nullptr;
  int64_t k_slots = 0;
  // edge-pass timing (cvk_batch_time_edges): an event pair per day around the edge launch, eager runs only
  bool time_edges = false;
  std::vector<hipEvent_t> ev_begin, ev_end;
  std::vector<uint8_t> ev_used;

  ~Batch() {
    void* ptrs[] = {d_sims, d_params, d_segs, d_views, d_a_off, d_seg_tile_start, d_tile_seg, d_block_sim,
                    d_rs_chunk_seg, d_rs_chunk_first, d_rec, d_bound, d_bitmap, d_winner, d_touched, d_ctl,
                    d_results, d_results_f, d_partials, d_partials_f, d_log, d_hist, d_prefix, d_remain,
                    d_all, d_short, d_k, d_srcs, d_item_seg, d_seg_item_start, d_walk, d_bv_days, d_bl_days};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    for (hipEvent_t e : ev_begin) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_end) (void)hipEventDestroy(e);
  }

  BatchCommon common() const {
    BatchCommon c;
    memset(&c, 0, sizeof(c));
    c.P = P;
    c.n_total = N;
    c.sims = d_sims;
    c.params = d_params;
    c.n_sims = S;
    c.n_layers = L;
    c.npts = npts;
    c.results = d_results;
    c.results_f = d_results_f;
    c.partials = d_partials;
    c.partials_f = d_partials_f;
    c.part_stride = part_stride;
    c.rec = d_rec;
    c.bound = d_bound;
    c.bitmap = d_bitmap;
    c.winner = d_winner;
    c.touched = d_touched;
    c.ctl = d_ctl;
    return c;
  }
};

static int64_t words_for(int64_t n) { return ((n + 127) / 128) * 4; }

static bool config_ok(const cvk_batch_config& c) {
  return (c.edge_block == 512 || c.edge_block == 1024) && (c.edge_chunks == 2 || c.edge_chunks == 4) &&
         (c.edge_blocks_per_cu == 1 || c.edge_blocks_per_cu == 2) &&
         (c.agents_per_thread == 1 || c.agents_per_thread == 2 || c.agents_per_thread == 4);
}

// Tables that depend on the launch shape: agent blocks per sim, edge tiles per segment.
static int apply_config(Batch& b, const cvk_batch_config& c) {
  if (!config_ok(c)) return fail(CVK_E_INVALID, "cvk_batch_set_config: unsupported launch shape");
  const int64_t span = (int64_t)AGENT_BLOCK * c.agents_per_thread;
  std::vector<int32_t> block_sim;
  uint32_t blocks = 0;
  for (int32_t s = 0; s < b.S; ++s) {
    const uint32_t nb = (uint32_t)((b.sims[s].n + span - 1) / span);
    b.sims[s].b_off = blocks;
    b.sims[s].n_blocks = nb;
    for (uint32_t k = 0; k < nb; ++k) block_sim.push_back(s);
    blocks += nb;
  }
  if ((int64_t)blocks > b.part_stride) return fail(CVK_E_CAPACITY, "cvk_batch_set_config: partial rows too small");
  const int64_t tile = 64 * 4 * (int64_t)c.edge_chunks;  // wave tiles
  std::vector<int64_t> seg_tile_start((size_t)b.S * b.L + 1);
  std::vector<int32_t> tile_seg;
  int64_t tiles = 0;
  for (int32_t seg = 0; seg < b.S * b.L; ++seg) {
    seg_tile_start[seg] = tiles;
    const int64_t nt = (b.segs[seg].n + tile - 1) / tile;
    for (int64_t k = 0; k < nt; ++k) tile_seg.push_back(seg);
    tiles += nt;
  }
  seg_tile_start[(size_t)b.S * b.L] = tiles;
  if (b.d_block_sim) (void)hipFree(b.d_block_sim);
  if (b.d_seg_tile_start) (void)hipFree(b.d_seg_tile_start);
  if (b.d_tile_seg) (void)hipFree(b.d_tile_seg);
  b.d_block_sim = nullptr;
  b.d_seg_tile_start = nullptr;
  b.d_tile_seg = nullptr;
  int rc = dev_alloc(&b.d_block_sim, (int64_t)block_sim.size());
  if (!rc) rc = upload(b.d_block_sim, block_sim.data(), (int64_t)block_sim.size());
  if (!rc) rc = dev_alloc(&b.d_seg_tile_start, (int64_t)seg_tile_start.size());
  if (!rc) rc = upload(b.d_seg_tile_start, seg_tile_start.data(), (int64_t)seg_tile_start.size());
  if (!rc) rc = dev_alloc(&b.d_tile_seg, (int64_t)tile_seg.size());
  if (!rc) rc = upload(b.d_tile_seg, tile_seg.data(), (int64_t)tile_seg.size());
  if (rc) return rc;
  b.agent_blocks = blocks;
  b.n_tiles = tiles;
  b.cfg = c;
  if (hipMemcpy(b.d_sims, b.sims.data(), b.sims.size() * sizeof(BatchSim), hipMemcpyHostToDevice) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_batch_set_config: upload failed");
  return CVK_OK;
}

template <int BLOCK, int CHUNKS>
static int edge_lds_budget(int max_lds, int64_t* budget) {
  static thread_local int configured = -1;
  static thread_local int64_t value = 0;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (configured != dev) {
    const void* fns[2] = {(const void*)batch_edge_kernel<BLOCK, CHUNKS, false>,
                          (const void*)batch_edge_kernel<BLOCK, CHUNKS, true>};
    int64_t b = max_lds;
    for (const void* fn : fns) {  // both instantiations get the same dynamic LDS limit
      hipFuncAttributes attr;
      if (hipFuncGetAttributes(&attr, fn) != hipSuccess) return fail(CVK_E_HIP, "cannot query the batch edge pass");
      b = std::min<int64_t>(b, (int64_t)max_lds - (int64_t)attr.sharedSizeBytes);
    }
    for (const void* fn : fns)
      if (b <= 0 || hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b) != hipSuccess)
        return fail(CVK_E_HIP, "cannot raise the LDS limit of the batch edge pass");
    value = b;
    configured = dev;
  }
  *budget = value;
  return CVK_OK;
}

static int edge_budget_for(const Batch& b, int64_t* budget) {
  const cvk_batch_config& c = b.cfg;
  if (c.edge_block == 1024 && c.edge_chunks == 4) return edge_lds_budget<1024, 4>(b.max_lds, budget);
  if (c.edge_block == 1024) return edge_lds_budget<1024, 2>(b.max_lds, budget);
  if (c.edge_chunks == 4) return edge_lds_budget<512, 4>(b.max_lds, budget);
  return edge_lds_budget<512, 2>(b.max_lds, budget);
}

static int launch_day_begin(Batch& b, int32_t t, int prologue, int transitions, int resample, hipStream_t st) {
  BatchDayArgs a;
  memset(&a, 0, sizeof(a));
  a.t = t;
  a.prologue = prologue;
  a.transitions = transitions;
  a.resample = resample;
  a.agent_blocks = b.agent_blocks;
  a.n_rs_chunks = resample ? (int32_t)b.n_rs_chunks : 0;
  a.block_sim = b.d_block_sim;
  a.rs_chunk_seg = b.d_rs_chunk_seg;
  a.rs_chunk_first = b.d_rs_chunk_first;
  a.segs = b.d_segs;
  const uint32_t rs_blocks = a.n_rs_chunks ? (uint32_t)std::min<int64_t>(a.n_rs_chunks, 8 * b.n_cu) : 0u;
  const dim3 grid(b.agent_blocks + rs_blocks);
  const BatchCommon C = b.common();
  switch (b.cfg.agents_per_thread) {
    case 1: hipLaunchKernelGGL(batch_day_begin_kernel<1>, grid, dim3(AGENT_BLOCK), 0, st, C, a); break;
    case 4: hipLaunchKernelGGL(batch_day_begin_kernel<4>, grid, dim3(AGENT_BLOCK), 0, st, C, a); break;
    default: hipLaunchKernelGGL(batch_day_begin_kernel<2>, grid, dim3(AGENT_BLOCK), 0, st, C, a); break;
  }
  return check_launch("cvk_batch day_begin");
}

// walk: the day's walk decisions are in b.d_walk (batch_walk_kernel ran after the agent pass).
static int launch_edge(Batch& b, int32_t t, const BatchReduce& red, hipStream_t st, bool gen, bool walk) {
  int64_t budget = 0;
  int rc = edge_budget_for(b, &budget);
  if (rc) return rc;
  const cvk_batch_config& c = b.cfg;
  const int64_t queue_bytes = (int64_t)(c.edge_block / 64) * EDGE_WAVE_WORDS * 4;
  // two workgroups per CU split the CU's LDS
  const int64_t per_block = std::min<int64_t>(budget, (int64_t)b.max_lds / c.edge_blocks_per_cu - 1024);
  int64_t lds_words = ((per_block - queue_bytes) / 4) & ~3ll;
  lds_words = std::max<int64_t>(0, std::min<int64_t>(lds_words, b.max_sim_words));
  BatchEdgeArgs a;
  memset(&a, 0, sizeof(a));
  a.views = b.d_views;
  a.segs = b.d_segs;
  a.seg_tile_start = b.d_seg_tile_start;
  a.tile_seg = b.d_tile_seg;
  a.n_tiles = b.n_tiles;
  a.lds_words = (int32_t)lds_words;
  a.t = t;
  a.reduce = red;
  if (walk) {
    a.src_mask = b.src_mask;
    a.srcs = b.d_srcs;
    a.item_seg = b.d_item_seg;
    a.seg_item_start = b.d_seg_item_start;
    a.n_items = b.n_items;
    a.walk = b.d_walk;
  }
  if (gen)
    for (int l = 0; l < b.L; ++l)
      if (b.dynamic[l]) a.gen_mask |= 1u << l;
  const size_t shmem = (size_t)(lds_words * 4 + queue_bytes);
  const dim3 grid((unsigned)std::max(1, b.n_cu * c.edge_blocks_per_cu));
  const BatchCommon C = b.common();
  const bool g = a.gen_mask != 0;
  if (c.edge_block == 1024 && c.edge_chunks == 4) {
    if (g) hipLaunchKernelGGL((batch_edge_kernel<1024, 4, true>), grid, dim3(1024), shmem, st, C, a);
    else hipLaunchKernelGGL((batch_edge_kernel<1024, 4, false>), grid, dim3(1024), shmem, st, C, a);
  } else if (c.edge_block == 1024) {
    if (g) hipLaunchKernelGGL((batch_edge_kernel<1024, 2, true>), grid, dim3(1024), shmem, st, C, a);
    else hipLaunchKernelGGL((batch_edge_kernel<1024, 2, false>), grid, dim3(1024), shmem, st, C, a);
  } else if (c.edge_chunks == 4) {
    if (g) hipLaunchKernelGGL((batch_edge_kernel<512, 4, true>), grid, dim3(512), shmem, st, C, a);
    else hipLaunchKernelGGL((batch_edge_kernel<512, 4, false>), grid, dim3(512), shmem, st, C, a);
  } else {
    if (g) hipLaunchKernelGGL((batch_edge_kernel<512, 2, true>), grid, dim3(512), shmem, st, C, a);
    else hipLaunchKernelGGL((batch_edge_kernel<512, 2, false>), grid, dim3(512), shmem, st, C, a);
  }
  return check_launch("cvk_batch edge_pass");
}

// resample_t >= 0: the launch also resamples the dynamic layers for day resample_t (extra workgroups),
// and the sources of dynamic-layer infections come from the resample stream; gen: they come from it
// because the day's dynamic contacts were generated, not written.
static int launch_infect(Batch& b, int32_t t, int count_flows, hipStream_t st, int32_t resample_t = -1,
                         bool gen = false) {
  BatchInfectArgs ia;
  memset(&ia, 0, sizeof(ia));
  ia.a_off = b.d_a_off;
  ia.segs = b.d_segs;
  ia.log = b.d_log;
  ia.log_capacity = b.log_capacity;
  ia.t = t;
  ia.count_flows = count_flows;
  ia.allow_beds = count_flows;
  ia.n_variants = b.V;
  {  // as the single-sim infect pass: one sweep's workgroups, between one and three per CU
    const int64_t n_cu = std::max(1, b.n_cu);
    const int64_t want = (b.N / 32 + INFECT_BLOCK) / INFECT_BLOCK;
    ia.infect_blocks = (uint32_t)std::min<int64_t>(std::max<int64_t>(want, n_cu), 3 * n_cu);
  }
  uint32_t rs_blocks = 0;
  if (resample_t >= 0 && b.n_rs_chunks > 0) {
    ia.rs_t = resample_t;
    ia.n_rs_chunks = (int32_t)b.n_rs_chunks;
    ia.rs_chunk_seg = b.d_rs_chunk_seg;
    ia.rs_chunk_first = b.d_rs_chunk_first;
    for (int l = 0; l < b.L; ++l)
      if (b.dynamic[l]) ia.gen_mask |= 1u << l;
    rs_blocks = (uint32_t)std::min<int64_t>(b.n_rs_chunks, 8 * b.n_cu);
  }
  if (gen)
    for (int l = 0; l < b.L; ++l)
      if (b.dynamic[l]) ia.gen_mask |= 1u << l;
  hipLaunchKernelGGL(batch_infect_kernel, dim3(ia.infect_blocks + rs_blocks), dim3(INFECT_BLOCK), 0, st, b.common(),
                     ia);
  return check_launch("cvk_batch infect_pass");
}

// k counts for choose call `slot` are at d_k + slot * S (uploaded before any capture).
static int launch_choose(Batch& b, int64_t slot, int32_t t, uint32_t stream_id, int32_t sus_only, int32_t layer_code,
                         int32_t variant, hipStream_t st) {
  ChooseState cs;
  cs.hist = b.d_hist;
  cs.prefix = b.d_prefix;
  cs.remain = b.d_remain;
  cs.all = b.d_all;
  cs.shortfall = b.d_short;
  cs.k = b.d_k + slot * b.S;
  const BatchCommon C = b.common();
  hipLaunchKernelGGL(batch_choose_init_kernel, dim3(b.S), dim3(256), 0, st, cs, b.S);
  const dim3 grid(b.agent_blocks);
  for (int pass = 0; pass < 8; ++pass) {
    switch (b.cfg.agents_per_thread) {
      case 1: hipLaunchKernelGGL(batch_choose_hist_kernel<1>, grid, dim3(AGENT_BLOCK), 0, st, C, cs, b.d_block_sim, t, stream_id, sus_only, pass); break;
      case 4: hipLaunchKernelGGL(batch_choose_hist_kernel<4>, grid, dim3(AGENT_BLOCK), 0, st, C, cs, b.d_block_sim, t, stream_id, sus_only, pass); break;
      default: hipLaunchKernelGGL(batch_choose_hist_kernel<2>, grid, dim3(AGENT_BLOCK), 0, st, C, cs, b.d_block_sim, t, stream_id, sus_only, pass); break;
    }
    hipLaunchKernelGGL(batch_choose_select_kernel, dim3(b.S), dim3(64), 0, st, cs, b.S, pass);
  }
  switch (b.cfg.agents_per_thread) {
    case 1: hipLaunchKernelGGL(batch_choose_mark_kernel<1>, grid, dim3(AGENT_BLOCK), 0, st, C, cs, b.d_block_sim, t, stream_id, sus_only, (uint32_t)layer_code, (uint32_t)variant); break;
    case 4: hipLaunchKernelGGL(batch_choose_mark_kernel<4>, grid, dim3(AGENT_BLOCK), 0, st, C, cs, b.d_block_sim, t, stream_id, sus_only, (uint32_t)layer_code, (uint32_t)variant); break;
    default: hipLaunchKernelGGL(batch_choose_mark_kernel<2>, grid, dim3(AGENT_BLOCK), 0, st, C, cs, b.d_block_sim, t, stream_id, sus_only, (uint32_t)layer_code, (uint32_t)variant); break;
  }
  return check_launch("cvk_batch choose");
}

// Uploads the choose counts: slot 0 = seeds, slot 1 + j = import event j. Not a stream operation.
static int upload_counts(Batch& b, const int64_t* pop_infected, const cvk_batch_imports* imp) {
  const int64_t n_ev = imp ? imp->n_events : 0;
  const int64_t slots = 1 + n_ev;
  std::vector<int64_t> k((size_t)(slots * b.S), 0);
  for (int32_t s = 0; s < b.S; ++s) {
    const int64_t v = pop_infected ? pop_infected[s] : 0;
    if (v < 0 || v > b.sims[s].n) return fail(CVK_E_INVALID, "cvk_batch: pop_infected outside [0, n]");
    k[s] = v;
  }
  for (int64_t j = 0; j < n_ev; ++j) {
    if (imp->variants[j] < 0 || imp->variants[j] >= b.V) return fail(CVK_E_INVALID, "cvk_batch: bad import variant");
    for (int32_t s = 0; s < b.S; ++s) {
      const int64_t v = imp->counts[j * b.S + s];
      if (v < 0 || v > b.sims[s].n) return fail(CVK_E_INVALID, "cvk_batch: import count outside [0, n]");
      k[(size_t)((1 + j) * b.S + s)] = v;
    }
  }
  if (slots > b.k_slots) {
    if (b.d_k) (void)hipFree(b.d_k);
    b.d_k = nullptr;
    int rc = dev_alloc(&b.d_k, slots * b.S);
    if (rc) return rc;
    b.k_slots = slots;
  }
  return upload(b.d_k, k.data(), slots * b.S);
}

static int run_days_impl(Batch& b, int32_t t0, int32_t t1, bool seeds, const cvk_batch_imports* imp, hipStream_t st,
                         bool captured = false) {
  int rc;
  const bool timed = b.time_edges && !captured;
  // as in the single-sim loop: the dynamic layers' endpoints are generated inside the edge and infect
  // passes and written once, with the last day's contacts, when the run closes (gen); otherwise day
  // t+1's resample runs as extra workgroups of day t's infect pass (overlap)
  const bool gen = b.n_rs_chunks > 0 && generate_dynamic_for(b.dyn_edges, false);
  const bool overlap = !gen && overlap_resample_enabled() && b.n_rs_chunks > 0 && t1 - t0 > 1;
  for (int32_t t = t0; t < t1; ++t) {
    if (t == 0 && seeds) {
      rc = launch_choose(b, 0, 0, STREAM_CHOOSE, 0, CVK_LAYER_SEED, 0, st);
      if (rc) return rc;
      rc = launch_infect(b, 0, 0, st);
      if (rc) return rc;
    }
    rc = launch_day_begin(b, t, t > t0 ? 1 : 0, 1, gen ? 0 : (t == t0 ? 2 : (overlap ? 0 : 1)), st);  // 2: weights too
    if (rc) return rc;
    if (imp) {
      for (int32_t j = 0; j < imp->n_events; ++j) {
        if (imp->days[j] != t) continue;
        const int32_t v = imp->variants[j];
        rc = launch_choose(b, 1 + j, t, STREAM_IMPORT | (uint32_t)v, 1, CVK_LAYER_IMPORT, v, st);
        if (rc) return rc;
      }
    }
    BatchReduce red{1, t, t > t0 ? 1 : 0, 1};
    if (b.n_beta_days > 0) {  // the day's rates (change_beta)
      hipLaunchKernelGGL(batch_rates_kernel, dim3(b.S), dim3(64), 0, st, b.d_views, b.d_segs,
                         (const float*)b.d_bv_days, (const float*)b.d_bl_days, b.n_beta_days, b.L, t);
      rc = check_launch("cvk_batch rates");
      if (rc) return rc;
    }
    const int mode = source_lists_mode();
    const bool walk = b.n_items > 0 && mode != 0;
    if (walk) {  // which sims walk their listed layers today
      hipLaunchKernelGGL(batch_walk_kernel, dim3(b.S), dim3(256), 0, st, b.common(), mode == 2 ? 1 : 0,
                         source_permille(), b.d_walk);
      rc = check_launch("cvk_batch walk decision");
      if (rc) return rc;
    }
    if (timed && hipEventRecord(b.ev_begin[t], st) != hipSuccess) return fail(CVK_E_HIP, "cvk_batch: event record failed");
    rc = launch_edge(b, t, red, st, gen, walk);
    if (rc) return rc;
    if (timed) {
      if (hipEventRecord(b.ev_end[t], st) != hipSuccess) return fail(CVK_E_HIP, "cvk_batch: event record failed");
      b.ev_used[t] = 1;
    }
    rc = launch_infect(b, t, 1, st, overlap && t + 1 < t1 ? t + 1 : -1, gen);
    if (rc) return rc;
  }
  rc = launch_day_begin(b, t1 - 1, 0, 0, gen ? 2 : 0, st);  // generated: the last day's contacts, weights 1
  if (rc) return rc;
  BatchReduce red{1, t1 - 1, 0, 0};
  hipLaunchKernelGGL(batch_reduce_kernel, dim3(std::min(b.S, 8 * b.n_cu)), dim3(1024), 0, st, b.common(), red);
  return check_launch("cvk_batch reduce");
}

static bool seeds_any(const Batch& b, const int64_t* pop_infected) {
  if (!pop_infected) return false;
  for (int32_t s = 0; s < b.S; ++s)
    if (pop_infected[s] > 0) return true;
  return false;
}

}  // namespace cvk

using namespace cvk;

extern "C" int cvk_batch_create(const cvk_batch_desc* d, int32_t npts, int64_t log_capacity, void** handle_out) {
  if (!handle_out) return fail(CVK_E_INVALID, "cvk_batch_create: null handle");
  *handle_out = nullptr;
  if (!d || d->n_sims < 1 || d->n_sims > CVK_BATCH_MAX_SIMS) return fail(CVK_E_INVALID, "cvk_batch_create: n_sims outside [1, 4096]");
  if (d->n_layers < 0 || d->n_layers > CVK_MAX_LAYERS) return fail(CVK_E_INVALID, "cvk_batch_create: n_layers outside [0, 8]");
  if (d->n_variants < 1 || d->n_variants > CVK_MAX_VARIANTS) return fail(CVK_E_INVALID, "cvk_batch_create: n_variants outside [1, 8]");
  if (npts < 1 || log_capacity < 0) return fail(CVK_E_INVALID, "cvk_batch_create: bad npts or log capacity");
  if (!d->agent_offsets_host || !d->params_host || (d->n_layers > 0 && (!d->edge_starts_host || !d->edge_counts_host || !d->dynamic_host)))
    return fail(CVK_E_INVALID, "cvk_batch_create: null host table");
  const cvk_people& P = d->people;
  if (P.n != d->n_agents_total || P.n <= 0 || P.n >= (1ll << 31) - 1)
    return fail(CVK_E_INVALID, "cvk_batch_create: people.n must equal n_agents_total (< 2^31 - 1)");
  Batch* b = new Batch();
  std::unique_ptr<Batch> guard(b);
  if (hipGetDevice(&b->device) != hipSuccess) return fail(CVK_E_HIP, "cvk_batch_create: no HIP device");
  (void)hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device);
  (void)hipDeviceGetAttribute(&b->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device);
  b->S = d->n_sims;
  b->L = d->n_layers;
  b->V = d->n_variants;
  b->N = d->n_agents_total;
  b->npts = npts;
  b->P = P;
  b->log_capacity = log_capacity;
  b->a_off.assign(d->agent_offsets_host, d->agent_offsets_host + b->S + 1);
  if (b->a_off[0] != 0 || b->a_off[b->S] != b->N) return fail(CVK_E_INVALID, "cvk_batch_create: agent offsets must span [0, n_agents_total]");
  b->params.assign(d->params_host, d->params_host + b->S);
  b->dynamic.assign(CVK_MAX_LAYERS, 0);
  for (int l = 0; l < b->L; ++l) b->dynamic[l] = d->dynamic_host[l];
  int64_t words = 0;
  for (int32_t s = 0; s < b->S; ++s) {
    const int64_t n = b->a_off[s + 1] - b->a_off[s];
    const cvk_params& p = b->params[s];
    if (n <= 0) return fail(CVK_E_INVALID, "cvk_batch_create: every sim needs at least one agent");
    if (p.n_agents != n) return fail(CVK_E_INVALID, "cvk_batch_create: params.n_agents does not match the offsets");
    if (p.n_layers != b->L || p.n_variants != b->V || (p.use_waning != 0) != (d->use_waning != 0))
      return fail(CVK_E_INVALID, "cvk_batch_create: sims must share n_layers, n_variants and use_waning");
    if (p.use_waning && (!p.nab_kin || p.nab_kin_len <= 0)) return fail(CVK_E_INVALID, "cvk_batch_create: use_waning requires nab_kin");
    BatchSim bs;
    memset(&bs, 0, sizeof(bs));
    bs.a_off = b->a_off[s];
    bs.n = n;
    bs.bm_off = words;
    words += words_for(n);
    b->max_sim_words = std::max(b->max_sim_words, words_for(n));
    b->sims.push_back(bs);
  }
  b->bitmap_words = words;
  // segments: sim-major, layers in order
  for (int l = 0; l < b->L; ++l) {
    int64_t need = 0;
    for (int32_t s = 0; s < b->S; ++s) {
      const int64_t e0 = d->edge_starts_host[(size_t)l * b->S + s];
      const int64_t ne = d->edge_counts_host[(size_t)l * b->S + s];
      if (e0 < 0 || ne < 0) return fail(CVK_E_INVALID, "cvk_batch_create: negative edge start or count");
      // the 16-byte endpoint loads need every sim's run to start on a 4-edge boundary
      if (e0 & 3) return fail(CVK_E_ALIGN, "cvk_batch_create: every sim's edge start must be a multiple of 4");
      need += ne;
    }
    if (need > 0 && (!d->p1[l] || !d->p2[l] || !d->beta[l])) return fail(CVK_E_INVALID, "cvk_batch_create: null layer array");
    if (((uintptr_t)d->p1[l] | (uintptr_t)d->p2[l]) & 15) return fail(CVK_E_ALIGN, "cvk_batch_create: p1/p2 must be 16-byte aligned");
    if ((uintptr_t)d->beta[l] & 7) return fail(CVK_E_ALIGN, "cvk_batch_create: beta must be 8-byte aligned");
  }
  std::vector<int32_t> rs_seg;
  std::vector<int64_t> rs_first;
  b->segs.resize((size_t)b->S * b->L);
  for (int32_t s = 0; s < b->S; ++s) {
    for (int l = 0; l < b->L; ++l) {
      const int64_t e0 = d->edge_starts_host[(size_t)l * b->S + s];
      SegView& G = b->segs[(size_t)s * b->L + l];
      memset(&G, 0, sizeof(G));
      G.p1 = d->p1[l] + e0;
      G.p2 = d->p2[l] + e0;
      G.beta = d->beta[l] + e0;
      G.unif = nullptr;
      G.n = d->edge_counts_host[(size_t)l * b->S + s];
      G.beta_layer = b->params[s].beta_layer[l];
      G.iso_factor = b->params[s].iso_factor[l];
      G.quar_factor = b->params[s].quar_factor[l];
      G.beta_ones = (d->beta_ones[l] || b->dynamic[l]) ? 1 : 0;
      G.l = l;
      if (b->dynamic[l]) b->dyn_edges += G.n;
      if (b->dynamic[l])
        for (int64_t e = 0; e < G.n; e += RS_CHUNK) {
          rs_seg.push_back(s * b->L + l);
          rs_first.push_back(e);
        }
    }
  }
  b->n_rs_chunks = (int64_t)rs_seg.size();
  // source lists: a static layer is listed when every sim with edges in it has lists
  std::vector<SrcView> srcs((size_t)b->S * b->L);
  memset(srcs.data(), 0, srcs.size() * sizeof(SrcView));
  if (d->sources) {
    for (int l = 0; l < b->L; ++l) {
      if (b->dynamic[l]) continue;
      bool all = true, any = false;
      for (int32_t s = 0; s < b->S && all; ++s) {
        const cvk_sources& R = d->sources[(size_t)s * b->L + l];
        const int64_t ne = b->segs[(size_t)s * b->L + l].n;
        if (ne == 0) continue;
        if (!R.indptr || !R.target || !R.edge_dir || R.n_slots < 0 || R.n_slots > 2 * ne) all = false;
        any = true;
      }
      if (!all || !any) continue;
      b->src_mask |= 1u << l;
      for (int32_t s = 0; s < b->S; ++s) {
        const cvk_sources& R = d->sources[(size_t)s * b->L + l];
        if (b->segs[(size_t)s * b->L + l].n == 0) continue;
        srcs[(size_t)s * b->L + l] = SrcView{R.indptr, R.target, R.edge_dir};
      }
    }
  }
  std::vector<int32_t> item_seg;
  std::vector<int64_t> seg_item_start((size_t)b->S * b->L + 1, 0);
  for (int32_t seg = 0; seg < b->S * b->L; ++seg) {
    seg_item_start[seg] = (int64_t)item_seg.size();
    if (!((b->src_mask >> (seg % b->L)) & 1u) || !srcs[seg].indptr) continue;
    const int64_t n = b->a_off[seg / b->L + 1] - b->a_off[seg / b->L];
    for (int64_t k = 0; k < (n + SRC_ITEM_AGENTS - 1) / SRC_ITEM_AGENTS; ++k) item_seg.push_back(seg);
  }
  seg_item_start[(size_t)b->S * b->L] = (int64_t)item_seg.size();
  b->n_items = (int64_t)item_seg.size();
  // device buffers
  int64_t max_blocks = 0;
  for (int32_t s = 0; s < b->S; ++s) max_blocks += (b->sims[s].n + AGENT_BLOCK - 1) / AGENT_BLOCK;
  b->part_stride = max_blocks;
  int rc = CVK_OK;
  rc = rc ? rc : dev_alloc(&b->d_sims, b->S);
  rc = rc ? rc : dev_alloc(&b->d_params, b->S);
  rc = rc ? rc : upload(b->d_params, b->params.data(), b->S);
  rc = rc ? rc : dev_alloc(&b->d_segs, std::max<int64_t>(1, (int64_t)b->segs.size()));
  rc = rc ? rc : upload(b->d_segs, b->segs.data(), (int64_t)b->segs.size());
  rc = rc ? rc : dev_alloc(&b->d_a_off, b->S + 1);
  rc = rc ? rc : upload(b->d_a_off, b->a_off.data(), b->S + 1);
  rc = rc ? rc : dev_alloc(&b->d_rs_chunk_seg, std::max<int64_t>(1, b->n_rs_chunks));
  rc = rc ? rc : upload(b->d_rs_chunk_seg, rs_seg.data(), b->n_rs_chunks);
  rc = rc ? rc : dev_alloc(&b->d_rs_chunk_first, std::max<int64_t>(1, b->n_rs_chunks));
  rc = rc ? rc : upload(b->d_rs_chunk_first, rs_first.data(), b->n_rs_chunks);
  rc = rc ? rc : dev_alloc(&b->d_rec, b->N);
  rc = rc ? rc : dev_alloc(&b->d_bound, b->N);
  rc = rc ? rc : dev_alloc(&b->d_bitmap, words);
  rc = rc ? rc : dev_alloc(&b->d_winner, b->N);
  rc = rc ? rc : dev_alloc(&b->d_touched, (b->N + 31) / 32);
  rc = rc ? rc : dev_alloc(&b->d_ctl, CVK_CTL_WORDS);
  rc = rc ? rc : dev_alloc(&b->d_results, (int64_t)b->S * npts * CVK_R_COLS);
  rc = rc ? rc : dev_alloc(&b->d_results_f, (int64_t)b->S * npts * CVK_RF_COLS);
  rc = rc ? rc : dev_alloc(&b->d_partials, (int64_t)CVK_PARTIAL_COLS * max_blocks);
  rc = rc ? rc : dev_alloc(&b->d_partials_f, (int64_t)CVK_RF_COLS * max_blocks);
  rc = rc ? rc : dev_alloc(&b->d_log, std::max<int64_t>(1, log_capacity) * 4);
  rc = rc ? rc : dev_alloc(&b->d_hist, (int64_t)b->S * 256);
  rc = rc ? rc : dev_alloc(&b->d_prefix, b->S);
  rc = rc ? rc : dev_alloc(&b->d_remain, b->S);
  rc = rc ? rc : dev_alloc(&b->d_all, b->S);
  rc = rc ? rc : dev_alloc(&b->d_short, b->S);
  rc = rc ? rc : dev_alloc(&b->d_srcs, std::max<int64_t>(1, (int64_t)srcs.size()));
  rc = rc ? rc : upload(b->d_srcs, srcs.data(), (int64_t)srcs.size());
  rc = rc ? rc : dev_alloc(&b->d_item_seg, std::max<int64_t>(1, b->n_items));
  rc = rc ? rc : upload(b->d_item_seg, item_seg.data(), b->n_items);
  rc = rc ? rc : dev_alloc(&b->d_seg_item_start, (int64_t)seg_item_start.size());
  rc = rc ? rc : upload(b->d_seg_item_start, seg_item_start.data(), (int64_t)seg_item_start.size());
  rc = rc ? rc : dev_alloc(&b->d_walk, b->S);
  if (!rc && d->n_beta_days > 0) {
    if (!d->beta_v_days || !d->beta_layer_days) return fail(CVK_E_INVALID, "cvk_batch_create: null rate tables");
    b->n_beta_days = d->n_beta_days;
    const int64_t rows = (int64_t)b->S * d->n_beta_days;
    rc = dev_alloc(&b->d_bv_days, rows * CVK_MAX_VARIANTS);
    rc = rc ? rc : upload(b->d_bv_days, d->beta_v_days, rows * CVK_MAX_VARIANTS);
    rc = rc ? rc : dev_alloc(&b->d_bl_days, rows * CVK_MAX_LAYERS);
    rc = rc ? rc : upload(b->d_bl_days, d->beta_layer_days, rows * CVK_MAX_LAYERS);
  }
  if (rc) return rc;
  // per-sim views for the edge pass (pointers offset to each sim's slices)
  b->views.resize(b->S);
  for (int32_t s = 0; s < b->S; ++s) {
    SimView& V = b->views[s];
    memset(&V, 0, sizeof(V));
    const cvk_params& p = b->params[s];
    const BatchSim& bs = b->sims[s];
    V.rec = b->d_rec + bs.a_off;
    V.bound = b->d_bound + bs.a_off;
    V.winner = b->d_winner + bs.a_off;
    V.sus_imm = P.sus_imm + bs.a_off;
    V.sus_stride = b->N;
    V.bitmap = b->d_bitmap + bs.bm_off;
    V.coarse = nullptr;
    V.touched = b->d_touched;
    V.a_off = bs.a_off;
    V.seed = p.seed;
    V.n_agents = (uint32_t)bs.n;
    V.n_variants = p.n_variants;
    V.use_sus_imm = p.use_waning;
    V.asymp_factor = p.asymp_factor;
    V.vl_hi = p.vl_hi;
    V.vl_lo = p.vl_lo;
    for (int v = 0; v < MAXV; ++v) V.beta_v[v] = p.beta_v[v];
  }
  rc = dev_alloc(&b->d_views, b->S);
  if (!rc) rc = upload(b->d_views, b->views.data(), b->S);
  if (rc) return rc;
  rc = apply_config(*b, b->cfg);
  if (rc) return rc;
  *handle_out = guard.release();
  return CVK_OK;
}

extern "C" int cvk_batch_destroy(void* handle) {
  if (!handle) return fail(CVK_E_INVALID, "cvk_batch_destroy: null handle");
  delete (Batch*)handle;
  return CVK_OK;
}

extern "C" int cvk_batch_set_config(void* handle, const cvk_batch_config* config) {
  if (!handle || !config) return fail(CVK_E_INVALID, "cvk_batch_set_config: null argument");
  return apply_config(*(Batch*)handle, *config);
}

extern "C" int cvk_batch_get_config(void* handle, cvk_batch_config* config_out) {
  if (!handle || !config_out) return fail(CVK_E_INVALID, "cvk_batch_get_config: null argument");
  *config_out = ((Batch*)handle)->cfg;
  return CVK_OK;
}

extern "C" int cvk_batch_reset(void* handle, void* stream) {
  if (!handle) return fail(CVK_E_INVALID, "cvk_batch_reset: null handle");
  Batch& b = *(Batch*)handle;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)std::max(1, 4 * b.n_cu)), block(256);
  hipLaunchKernelGGL(fill_u64_kernel, grid, block, 0, st, b.d_winner, b.N, KEY_NONE);
  hipLaunchKernelGGL(fill_u32_kernel, grid, block, 0, st, b.d_touched, (b.N + 31) / 32, 0u);
  hipLaunchKernelGGL(fill_u32_kernel, grid, block, 0, st, b.d_ctl, (int64_t)CVK_CTL_WORDS, 0u);
  hipLaunchKernelGGL(fill_u32_kernel, grid, block, 0, st, b.d_results, (int64_t)b.S * b.npts * CVK_R_COLS, 0u);
  hipLaunchKernelGGL(fill_u64_kernel, grid, block, 0, st, (uint64_t*)b.d_results_f, (int64_t)b.S * b.npts * CVK_RF_COLS, 0ull);
  hipLaunchKernelGGL(fill_u32_kernel, grid, block, 0, st, b.d_short, (int64_t)b.S, 0u);
  return check_launch("cvk_batch_reset");
}

static int batch_run_args(void* handle, int32_t t0, int32_t t1) {
  if (!handle) return fail(CVK_E_INVALID, "cvk_batch: null handle");
  Batch& b = *(Batch*)handle;
  if (t0 < 0 || t1 < t0 || t1 > b.npts) return fail(CVK_E_INVALID, "cvk_batch: bad day range");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != b.device) return fail(CVK_E_INVALID, "cvk_batch: wrong current device");
  return CVK_OK;
}

extern "C" int cvk_batch_graph_create(void* handle, int32_t t0, int32_t t1, const int64_t* pop_infected_host,
                                      const cvk_batch_imports* imports, void* stream, void** graph_exec_out) {
  if (!graph_exec_out) return fail(CVK_E_INVALID, "cvk_batch_graph_create: null output handle");
  *graph_exec_out = nullptr;
  int rc = batch_run_args(handle, t0, t1);
  if (rc) return rc;
  if (t1 <= t0) return fail(CVK_E_INVALID, "cvk_batch_graph_create: empty day range");
  Batch& b = *(Batch*)handle;
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) return fail(CVK_E_INVALID, "cvk_batch_graph_create: graph capture needs a non-default stream");
  rc = upload_counts(b, pop_infected_host, imports);
  if (rc) return rc;
  int64_t budget = 0;
  rc = edge_budget_for(b, &budget);  // function attributes are set before capture
  if (rc) return rc;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_batch_graph_create: hipStreamBeginCapture failed");
  rc = run_days_impl(b, t0, t1, seeds_any(b, pop_infected_host), imports, st, true);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess || !graph) return fail(CVK_E_HIP, "cvk_batch_graph_create: hipStreamEndCapture failed");
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(CVK_E_HIP, "cvk_batch_graph_create: hipGraphInstantiate failed");
  *graph_exec_out = (void*)exec;
  return CVK_OK;
}

extern "C" int cvk_batch_run_days(void* handle, int32_t t0, int32_t t1, const int64_t* pop_infected_host,
                                  const cvk_batch_imports* imports, int32_t use_graph, void* stream) {
  int rc = batch_run_args(handle, t0, t1);
  if (rc) return rc;
  if (t1 == t0) return CVK_OK;
  Batch& b = *(Batch*)handle;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) {
    rc = upload_counts(b, pop_infected_host, imports);
    if (rc) return rc;
    return run_days_impl(b, t0, t1, seeds_any(b, pop_infected_host), imports, st);
  }
  void* exec = nullptr;
  rc = cvk_batch_graph_create(handle, t0, t1, pop_infected_host, imports, stream, &exec);
  if (rc) return rc;
  if (hipGraphLaunch((hipGraphExec_t)exec, st) != hipSuccess) rc = fail(CVK_E_HIP, "cvk_batch_run_days: graph launch failed");
  (void)hipStreamSynchronize(st);
  (void)hipGraphExecDestroy((hipGraphExec_t)exec);
  return rc;
}

extern "C" int cvk_batch_copy_results(void* handle, uint32_t* results_host, double* results_f_host) {
  if (!handle) return fail(CVK_E_INVALID, "cvk_batch_copy_results: null handle");
  Batch& b = *(Batch*)handle;
  if (hipDeviceSynchronize() != hipSuccess) return fail(CVK_E_HIP, "cvk_batch_copy_results: device error");
  if (results_host && hipMemcpy(results_host, b.d_results, (size_t)b.S * b.npts * CVK_R_COLS * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_batch_copy_results: copy failed");
  if (results_f_host && hipMemcpy(results_f_host, b.d_results_f, (size_t)b.S * b.npts * CVK_RF_COLS * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_batch_copy_results: copy failed");
  return CVK_OK;
}

extern "C" int cvk_batch_results(void* handle, const uint32_t** results_out, const double** results_f_out) {
  if (!handle) return fail(CVK_E_INVALID, "cvk_batch_results: null handle");
  Batch& b = *(Batch*)handle;
  if (results_out) *results_out = b.d_results;
  if (results_f_out) *results_f_out = b.d_results_f;
  return CVK_OK;
}

extern "C" int cvk_batch_copy_log(void* handle, int32_t* log_host, int64_t capacity, int64_t* count_out) {
  if (!handle || !count_out) return fail(CVK_E_INVALID, "cvk_batch_copy_log: null argument");
  Batch& b = *(Batch*)handle;
  uint32_t ctl[CVK_CTL_WORDS];
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(ctl, b.d_ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_batch_copy_log: copy failed");
  const int64_t written = (int64_t)(((uint64_t)ctl[CVK_CTL_LOG_COUNT + 1] << 32) | ctl[CVK_CTL_LOG_COUNT]);
  const int64_t stored = std::min(written, b.log_capacity);
  *count_out = stored;
  if (log_host && capacity > 0 && stored > 0) {
    const int64_t m = std::min(stored, capacity);
    if (hipMemcpy(log_host, b.d_log, (size_t)m * 16, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(CVK_E_HIP, "cvk_batch_copy_log: copy failed");
  }
  if (written > b.log_capacity) return fail(CVK_E_CAPACITY, "cvk_batch_copy_log: the log overflowed its capacity");
  return CVK_OK;
}

extern "C" int cvk_batch_time_edges(void* handle, int32_t enable) {
  if (!handle) return fail(CVK_E_INVALID, "cvk_batch_time_edges: null handle");
  Batch& b = *(Batch*)handle;
  if (enable && b.ev_begin.empty()) {
    b.ev_begin.assign(b.npts, nullptr);
    b.ev_end.assign(b.npts, nullptr);
    for (int32_t t = 0; t < b.npts; ++t)
      if (hipEventCreate(&b.ev_begin[t]) != hipSuccess || hipEventCreate(&b.ev_end[t]) != hipSuccess)
        return fail(CVK_E_HIP, "cvk_batch_time_edges: hipEventCreate failed");
  }
  b.ev_used.assign(b.npts, 0);
  b.time_edges = enable != 0;
  return CVK_OK;
}

extern "C" int cvk_batch_edge_times(void* handle, float* ms_out, int32_t n) {
  if (!handle || !ms_out) return fail(CVK_E_INVALID, "cvk_batch_edge_times: null argument");
  Batch& b = *(Batch*)handle;
  if (hipDeviceSynchronize() != hipSuccess) return fail(CVK_E_HIP, "cvk_batch_edge_times: device error");
  const int32_t m = std::min<int32_t>(n, (int32_t)b.ev_used.size());
  for (int32_t t = 0; t < m; ++t) {
    ms_out[t] = -1.0f;
    if (b.ev_used[t] && hipEventElapsedTime(&ms_out[t], b.ev_begin[t], b.ev_end[t]) != hipSuccess)
      return fail(CVK_E_HIP, "cvk_batch_edge_times: hipEventElapsedTime failed");
  }
  return m;
}
