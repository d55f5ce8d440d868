// Testing, contact tracing, quarantine and isolation on the device (gfx950, wave64).
//
//   test_kernel          test_prob.apply (interventions.py:921-962) + People.test (people.py:589-617):
//                        one pass over the agents; the test, true-positive and loss draws of agent i
//                        are Philox words keyed by (i, t, TEST | slot). Options: subtarget (a bitmap
//                        and values), ILI (the day's uniform choice of agents, by a radix select of
//                        keys, cvk_select.h) and swab delay (swab_hist_kernel + swab_prob_kernel: the
//                        symptomatic agents by days since onset, then the probability per day count)
//   trace_cases_kernel   contact_tracing.select_cases (interventions.py:1069-1086): bitmap of the
//                        agents diagnosed today (or, presumptive, exposed and tested today); under a
//                        capacity only the cases a radix select of keys chose
//   trace_edges_kernel   Layer.find_contacts (base.py:1808, utils.py:131-147) for every traced layer at once: each
//                        edge with a case at one end marks the other end's layer bit (atomicOr;
//                        only the few edges at a case write)
//   trace_notify_kernel  binomial_filter per (contact, layer) and notify_contacts
//                        (interventions.py:1125-1145): known_contact, date_known_contact and the
//                        quarantine scheduled to start on day int(t + trace_time)
//   states_post_kernel   update_states_post (people.py:189-196): check_diagnosed, check_quar,
//                        check_enter_iso, then the edge records of agents whose quarantine or
//                        isolation changed (the transmission step of the same day reads them)
//   diag_reduce_kernel   the day's per-block counts (tests, diagnoses, quarantined, isolated) into the
//                        day's result row: every agent pass writes its block's counts to its own
//                        entries, so no result word is the target of one atomic per block
//
// Scheduled quarantines (People._pending_quarantine, people.py:116, 620-641) are a ring of per-agent
// end days: row d % ring holds the largest end day among the quarantines scheduled to start on day d.
// check_quar walks the day's list in order; the first entry of an agent starts or extends its
// quarantine and the later ones only extend it, and whether it may start does not change within the
// walk, so the walk's result is that of the single entry with the largest end day (and the agent is
// counted once), which is what the ring keeps.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cvk_host.h"
#include "cvk_kernels.h"
#include "cvk_select.h"

namespace cvk {
namespace {

constexpr int DIAG_BLOCK = 256;
constexpr int DIAG_APT = 2;                 // agents per thread and pass (contiguous 256-agent chunks)
constexpr int DIAG_SPAN = DIAG_BLOCK * DIAG_APT;
constexpr int TRACE_EDGE_BLOCK = 256;
constexpr int TRACE_EDGE_GRID = 2048;       // grid-stride blocks of the edge scan
constexpr int COARSE_MAX_WORDS = 2048;      // 8 KiB of LDS (a coarse bit per 4 agents at 262k agents, 64 at 4M)
constexpr int N_PART = 4;                   // per-block count columns: tests, diagnoses, quarantined, isolated

__host__ __device__ inline int coarse_shift(int64_t n) {
  int s = 0;
  while (((n + ((int64_t)1 << s) - 1) >> s) > (int64_t)COARSE_MAX_WORDS * 32) ++s;
  return s;
}
__host__ __device__ inline int64_t coarse_words(int64_t n) {
  const int s = coarse_shift(n);
  return (((n + ((int64_t)1 << s) - 1) >> s) + 31) / 32;
}

// Adds this block's wave counts c[0..NC) (wave-uniform, from ballots) to its entries of the
// per-block count columns cols[k] (one writer per entry; kernels of a day run in stream order).
template <int NC>
__device__ inline void block_part(const uint32_t (&c)[NC], uint32_t* part, const int (&cols)[NC]) {
  __shared__ uint32_t s[DIAG_BLOCK / 64][NC];
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < NC; ++k) s[w][k] = c[k];
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < DIAG_BLOCK / 64; ++j) tot += s[j][threadIdx.x];
    if (tot) part[(int64_t)cols[threadIdx.x] * CVK_DIAG_MAX_BLOCKS + blockIdx.x] += tot;
  }
}

// Agent passes stride over spans of DIAG_SPAN agents; agent index of chunk r of the span at base.
__device__ inline int64_t span_stride() { return (int64_t)gridDim.x * DIAG_SPAN; }

// get_quar_inds (interventions.py:682-710) for one agent.
__device__ inline bool quar_test(const cvk_diag_op& op, uint32_t f, const cvk_diag& D, int64_t i, int32_t t) {
  switch (op.quar_policy) {
    case CVK_QP_START: return D.date_quarantined[i] == (float)(t - 1);
    case CVK_QP_END: return D.date_end_quarantine[i] == (float)(t + 1);
    case CVK_QP_BOTH: return D.date_quarantined[i] == (float)(t - 1) || D.date_end_quarantine[i] == (float)(t + 1);
    case CVK_QP_DAILY: return (f & CVK_QUARANTINED) != 0;
    case CVK_QP_DAYS: {
      const float dq = D.date_quarantined[i];
      bool q = false;
      for (int k = 0; k < op.n_quar_days; ++k) q = q || dq == (float)(t - 1 - op.quar_days[k]);
      return q;
    }
    default: return false;
  }
}

// The key of agent i in a uniform choice of agents (ILI, tracing capacity): cvk_select.h's item key.
__device__ inline uint64_t agent_key(uint64_t seed, int32_t t, uint32_t stream_id, int64_t i) {
  return item_key(seed, t, stream_id, i);
}

// ili_thr: the day's k-th smallest ILI key (agents at or below it have influenza-like symptoms), or
// null; swab_prob: the day's probability by days since onset (swab delay), or null.
__global__ __launch_bounds__(DIAG_BLOCK) void test_kernel(cvk_people P, cvk_diag D, cvk_diag_op op, uint64_t seed,
                                                          int32_t t, const uint64_t* ili_thr, const double* swab_prob) {
  const float tf = (float)t;
  const uint32_t stream = STREAM_TEST | (uint32_t)op.slot;
  const uint32_t ili_stream = STREAM_ILI | (uint32_t)op.slot;
  const uint64_t thr = ili_thr ? *ili_thr : 0ull;
  uint32_t c[1] = {0};
  for (int64_t base = (int64_t)blockIdx.x * DIAG_SPAN; base < P.n; base += span_stride()) {
#pragma unroll
    for (int r = 0; r < DIAG_APT; ++r) {
      const int64_t i = base + r * DIAG_BLOCK + threadIdx.x;
      bool tested = false;
      if (i < P.n) {
        const uint32_t f = P.flags[i];
        if (!(f & CVK_DIAGNOSED)) {  // test_probs[diag_inds] = 0
          // test_probs set piece by piece in the reference's order: symptomatic, ILI, asymptomatic,
          // the two quarantine groups (ILI agents are in neither), then the subtarget
          double p;
          if (op.sub_mask && ((op.sub_mask[i >> 5] >> (i & 31)) & 1u)) {
            p = op.sub_vals ? op.sub_vals[i] : op.sub_val;
          } else if (f & CVK_SYMPTOMATIC) {
            if (quar_test(op, f, D, i, t))
              p = op.symp_quar_prob;
            else if (swab_prob) {
              const int32_t st = (int32_t)(tf - P.date_symptomatic[i]);  // 0 <= st < t < n_swab_bins
              p = st >= 0 && st < op.n_swab_bins ? swab_prob[st] : 0.0;
            } else
              p = op.symp_prob;
          } else if (ili_thr && agent_key(seed, t, ili_stream, i) <= thr) {
            p = op.symp_prob;
          } else {
            p = quar_test(op, f, D, i, t) ? op.asymp_quar_prob : op.asymp_prob;
          }
          if (p > 0.0) {  // binomial_arr: u < p (never for p <= 0 or NaN)
            const u32x4 w = philox_seeded(seed, (uint32_t)i, (uint32_t)t, stream, 0);
            if (u01(w.x, w.y) < p) {
              tested = true;
              P.flags[i] = f | CVK_TESTED;
              D.date_tested[i] = tf;
              // true positive among the infectious, not yet holding a diagnosis date, not lost
              if ((f & CVK_INFECTIOUS) && u01(w.z, w.w) < op.sensitivity && isnan(P.date_diagnosed[i])) {
                const u32x4 w2 = philox_seeded(seed, (uint32_t)i, (uint32_t)t, stream, 1);
                if (u01(w2.x, w2.y) < op.keep_prob) {
                  P.date_diagnosed[i] = (float)((double)t + op.test_delay);
                  D.date_pos_test[i] = tf;
                }
              }
            }
          }
        }
      }
      c[0] += (uint32_t)__popcll(__ballot(tested));
    }
  }
  const int cols[1] = {0};
  block_part<1>(c, D.part, cols);
}

__global__ __launch_bounds__(DIAG_BLOCK) void fill_u32_kernel(uint32_t* p, int32_t n) {
  for (int32_t k = blockIdx.x * DIAG_BLOCK + threadIdx.x; k < n; k += gridDim.x * DIAG_BLOCK) p[k] = 0u;
}

// The symptomatic agents by days since onset s = t - date_symptomatic (hist[s]) and all of them
// (hist[bins]), as np.bincount and len in test_prob.apply (interventions.py:925-927).
__global__ __launch_bounds__(DIAG_BLOCK) void swab_hist_kernel(cvk_people P, int32_t t, uint32_t* hist, int32_t bins) {
  __shared__ uint32_t s_total;
  if (threadIdx.x == 0) s_total = 0;
  __syncthreads();
  const float tf = (float)t;
  uint32_t mine = 0;
  for (int64_t i = (int64_t)blockIdx.x * DIAG_BLOCK + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * DIAG_BLOCK) {
    if (!(P.flags[i] & CVK_SYMPTOMATIC)) continue;
    const int32_t s = (int32_t)(tf - P.date_symptomatic[i]);
    if (s >= 0 && s < bins) atomicAdd(hist + s, 1u);
    ++mine;
  }
  mine = wave_sum(mine);
  if (lane_id() == 0 && mine) atomicAdd(&s_total, mine);
  __syncthreads();
  if (threadIdx.x == 0 && s_total) atomicAdd(hist + bins, s_total);
}

// Per day count s: base[s] * (1 / (count[s] / n_symp)), in the reference's operation order
// (interventions.py:926-933: inv_count = bincount / len, count = 1 / inv_count, pdf * prob * count).
__global__ __launch_bounds__(DIAG_BLOCK) void swab_prob_kernel(const uint32_t* hist, const double* base, int32_t bins,
                                                               double* prob) {
  const double total = (double)hist[bins];
  for (int32_t s = threadIdx.x; s < bins; s += DIAG_BLOCK) {
    const uint32_t h = hist[s];
    prob[s] = h ? base[s] * (1.0 / ((double)h / total)) : 0.0;
  }
}

// Bitmap of today's cases, and the coarse map (bit k: some case among agents [k << shift, ...)).
__device__ inline bool trace_case(const cvk_people& P, const cvk_diag& D, int32_t presumptive, float tf, int64_t i) {
  return presumptive ? ((P.flags[i] & CVK_EXPOSED) && D.date_tested[i] == tf) : P.date_diagnosed[i] == tf;
}

// Eligible: today's cases; key: the tracing capacity's choice key.
struct CaseKey {
  cvk_people P;
  cvk_diag D;
  uint64_t seed;
  int32_t presumptive, t;
  uint32_t stream_id;
  __device__ bool operator()(int64_t i, uint64_t& key) const {
    if (!trace_case(P, D, presumptive, (float)t, i)) return false;
    key = agent_key(seed, t, stream_id, i);
    return true;
  }
};

// cap_mode: 0 every case, 1 the cases whose capacity key is at most *cap_thr, 2 none (capacity 0).
__global__ __launch_bounds__(DIAG_BLOCK) void trace_cases_kernel(cvk_people P, cvk_diag D, int32_t presumptive,
                                                                 int32_t t, int32_t shift, int32_t cap_mode,
                                                                 const uint64_t* cap_thr, uint64_t seed,
                                                                 uint32_t cap_stream) {
  const float tf = (float)t;
  const int64_t n_words = (P.n + 63) / 64;
  const uint64_t thr = cap_mode == 1 ? *cap_thr : 0ull;
  for (int64_t base = (int64_t)blockIdx.x * DIAG_SPAN; base < P.n; base += span_stride()) {
#pragma unroll
    for (int r = 0; r < DIAG_APT; ++r) {
      const int64_t i = base + r * DIAG_BLOCK + threadIdx.x;
      bool is_case = false;
      if (i < P.n && cap_mode != 2) {
        is_case = trace_case(P, D, presumptive, tf, i);
        if (is_case && cap_mode == 1) is_case = agent_key(seed, t, cap_stream, i) <= thr;
      }
      const uint64_t b = __ballot(is_case);
      const int64_t word = (base + r * DIAG_BLOCK + (threadIdx.x & ~63u)) >> 6;
      if (lane_id() == 0 && word < n_words) {
        D.cases[word] = b;
        if (b) {
          if (shift >= 6) {
            const int64_t k = (word << 6) >> shift;
            atomicOr(D.coarse + (k >> 5), 1u << (k & 31));
          } else {
            const int g = 1 << shift;
            uint64_t cb = 0;
            for (int j = 0; j < 64 / g; ++j)
              if ((b >> (j * g)) & ((g == 64 ? 0ull : (1ull << g)) - 1ull)) cb |= 1ull << j;
            const int64_t k0 = (word << 6) >> shift;  // a multiple of 64 / g
            if (g == 1) {
              atomicOr(D.coarse + (k0 >> 5), (uint32_t)cb);
              // the upper half exists only where the map has a word for it (the last bitmap word of
              // an n with n % 64 in 1..32 has none)
              if ((cb >> 32) != 0) atomicOr(D.coarse + (k0 >> 5) + 1, (uint32_t)(cb >> 32));
            } else {
              atomicOr(D.coarse + (k0 >> 5), (uint32_t)cb << (k0 & 31));
            }
          }
        }
      }
    }
  }
}

struct TraceLayers {
  const int32_t* p1[CVK_MAX_LAYERS];
  const int32_t* p2[CVK_MAX_LAYERS];
  int64_t n[CVK_MAX_LAYERS];
  uint32_t layer[CVK_MAX_LAYERS];  // layer index (the contact bit)
  int32_t nl;
};

// Marks the partners of today's cases, every traced layer in turn (grid-stride, 4 edges per lane).
// The coarse map sits in LDS; only endpoints whose coarse bit is set read the case bitmap.
__global__ __launch_bounds__(TRACE_EDGE_BLOCK) void trace_edges_kernel(TraceLayers L, const uint64_t* cases,
                                                                       const uint32_t* coarse, int32_t shift,
                                                                       int32_t n_cwords, uint32_t* contacts) {
  extern __shared__ uint32_t s_coarse[];
  int any = 0;
  for (int k = threadIdx.x; k < n_cwords; k += TRACE_EDGE_BLOCK) {
    const uint32_t v = coarse[k];
    s_coarse[k] = v;
    any |= v != 0u;
  }
  if (!__syncthreads_or(any)) return;  // nobody to trace today
  auto is_case = [&](int32_t a) -> bool {
    const uint32_t k = (uint32_t)a >> shift;
    if (!((s_coarse[k >> 5] >> (k & 31u)) & 1u)) return false;
    return (cases[(uint32_t)a >> 6] >> ((uint32_t)a & 63u)) & 1ull;
  };
  const int64_t stride = (int64_t)gridDim.x * TRACE_EDGE_BLOCK;
#pragma unroll 1
  for (int l = 0; l < CVK_MAX_LAYERS; ++l) {
    if (l >= L.nl) break;
    const int32_t* p1 = L.p1[l];
    const int32_t* p2 = L.p2[l];
    const int64_t n = L.n[l];
    const uint32_t bit = 1u << L.layer[l];
    // two quads per lane and pass: all four loads are in flight before the first lookup
    for (int64_t q = (int64_t)blockIdx.x * TRACE_EDGE_BLOCK + threadIdx.x; q * 4 < n; q += 2 * stride) {
      const int64_t e0 = q * 4, e1 = (q + stride) * 4;
      const v4i a0 = load4(p1, e0, n), b0 = load4(p2, e0, n);
      const v4i a1 = load4(p1, e1, n), b1 = load4(p2, e1, n);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (e0 + k < n) {
          if (is_case(a0[k])) atomicOr(contacts + b0[k], bit);
          if (is_case(b0[k])) atomicOr(contacts + a0[k], bit);
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (e1 + k < n) {
          if (is_case(a1[k])) atomicOr(contacts + b1[k], bit);
          if (is_case(b1[k])) atomicOr(contacts + a1[k], bit);
        }
      }
    }
  }
}

// notify_contacts for agent i (a contact of today's cases in the layers of its contact mask).
__device__ inline void notify_agent(const cvk_people& P, const cvk_diag& D, const cvk_diag_op& op, uint64_t seed,
                                    int32_t t, int64_t i) {
  const uint32_t m = D.contacts[i];
  if (!m) return;
  D.contacts[i] = 0;
  const uint32_t f = P.flags[i];
  if (f & CVK_DEAD) return;  // np.setdiff1d(contact_inds, is_dead)
  float dkc = D.date_known_contact[i];
  bool any = false;
  for (uint32_t mm = m; mm; mm &= mm - 1) {
    const int l = __ffs(mm) - 1;
    const u32x4 w = philox_seeded(seed, (uint32_t)i, (uint32_t)t,
                                  STREAM_TRACE | ((uint32_t)op.slot << 8) | (uint32_t)l, 0);
    if (!(u01(w.x, w.y) < op.trace_prob[l])) continue;  // binomial_filter
    any = true;
    const double s = (double)t + op.trace_time[l];
    dkc = fminf(dkc, (float)s);  // np.fmin: a NaN date gives way
    const int32_t start = (int32_t)s;
    const int32_t end = start + op.trace_period[l];
    int32_t* pend = D.pending + (int64_t)(start % D.ring) * P.n + i;
    if (end > *pend) *pend = end;
  }
  if (any) {
    P.flags[i] = f | CVK_KNOWN_CONTACT;
    D.date_known_contact[i] = dkc;
  }
}

// The coarse map has been read by the scan: clear it for the next tracing op.
__device__ inline void clear_coarse(const cvk_diag& D, int32_t n_cwords) {
  for (int64_t k = (int64_t)blockIdx.x * DIAG_BLOCK + threadIdx.x; k < n_cwords; k += (int64_t)gridDim.x * DIAG_BLOCK)
    D.coarse[k] = 0u;
}

__global__ __launch_bounds__(DIAG_BLOCK) void trace_notify_kernel(cvk_people P, cvk_diag D, cvk_diag_op op,
                                                                  uint64_t seed, int32_t t, int32_t n_cwords) {
  clear_coarse(D, n_cwords);
  for (int64_t base = (int64_t)blockIdx.x * DIAG_SPAN; base < P.n; base += span_stride()) {
#pragma unroll
    for (int r = 0; r < DIAG_APT; ++r) {
      const int64_t i = base + r * DIAG_BLOCK + threadIdx.x;
      if (i < P.n) notify_agent(P, D, op, seed, t, i);
    }
  }
}

// NOTIFY: the day's last op is a tracing op whose notify pass is done here, agent by agent, before
// the agent's update_states_post (each agent's steps touch only its own state, so the order between
// agents does not matter; the quarantine it schedules for today is read back by the same thread).
template <bool NOTIFY>
__global__ __launch_bounds__(DIAG_BLOCK) void states_post_kernel(cvk_people P, cvk_params prm, cvk_work W,
                                                                 cvk_diag D, int32_t t, cvk_diag_op op,
                                                                 int32_t n_cwords) {
  if (NOTIFY) clear_coarse(D, n_cwords);
  const float tf = (float)t;
  float quar_max = 0.0f, iso_max = 0.0f;
  for (int l = 0; l < prm.n_layers; ++l) {
    quar_max = fmaxf(quar_max, prm.quar_factor[l]);
    iso_max = fmaxf(iso_max, prm.iso_factor[l]);
  }
  int32_t* pend_row = D.pending + (int64_t)(t % D.ring) * P.n;
  uint32_t c[3] = {0, 0, 0};
  for (int64_t base = (int64_t)blockIdx.x * DIAG_SPAN; base < P.n; base += span_stride()) {
#pragma unroll
    for (int r = 0; r < DIAG_APT; ++r) {
      const int64_t i = base + r * DIAG_BLOCK + threadIdx.x;
      bool new_diag = false, new_quar = false, new_iso = false;
      if (i < P.n) {
        if (NOTIFY) notify_agent(P, D, op, prm.seed, t, i);
        const uint32_t f0 = P.flags[i];
        uint32_t f = f0;
        const float dd = P.date_diagnosed[i];
        // check_diagnosed (people.py:315-332): positive tests counted, then diagnoses that are due
        if (!(f & CVK_DIAGNOSED)) {
          const float dpt = D.date_pos_test[i];
          if (!isnan(dpt) && tf >= dpt) {
            D.date_pos_test[i] = __int_as_float(0x7FC00000);
            new_diag = true;
          }
          if (!isnan(dd) && tf >= dd) f |= CVK_DIAGNOSED;
        }
        // check_quar (people.py:335-358)
        const int32_t e = pend_row[i];
        float deq = 0.0f;
        bool deq_loaded = false;
        if (e != CVK_PENDING_NONE) {
          pend_row[i] = CVK_PENDING_NONE;
          if (f & CVK_QUARANTINED) {
            deq = D.date_end_quarantine[i];
            deq_loaded = true;
            if ((float)e > deq) {  // max(date_end_quarantine, end_day)
              deq = (float)e;
              D.date_end_quarantine[i] = deq;
            }
          } else if (!(f & (CVK_DEAD | CVK_RECOVERED | CVK_DIAGNOSED | CVK_ISOLATED))) {
            f |= CVK_QUARANTINED;
            D.date_quarantined[i] = tf;
            deq = (float)e;
            deq_loaded = true;
            D.date_end_quarantine[i] = deq;
            new_quar = true;
          }
        }
        if (f & CVK_QUARANTINED) {
          if (!deq_loaded) deq = D.date_end_quarantine[i];
          if (dd == tf) {  // diagnosed today: quarantine ends, isolation begins
            deq = tf;
            D.date_end_quarantine[i] = tf;
          }
          if (!isnan(deq) && tf >= deq) f &= ~CVK_QUARANTINED;
        }
        // check_enter_iso (people.py:361-365)
        if (dd == tf) {
          f |= CVK_ISOLATED;
          P.date_end_isolation[i] = P.date_recovered[i];
          new_iso = true;
        }
        if (f != f0) {
          P.flags[i] = f;
          if ((f ^ f0) & (CVK_QUARANTINED | CVK_ISOLATED)) {  // factors of today's transmission changed
            const float rt = P.rel_trans[i], rs = P.rel_sus[i];
            W.rec[i] = edge_record(f, rt, rs);
            W.bound[i] = edge_bound(f, rt, rs, prm, quar_max, iso_max);
          }
        }
      }
      c[0] += (uint32_t)__popcll(__ballot(new_diag));
      c[1] += (uint32_t)__popcll(__ballot(new_quar));
      c[2] += (uint32_t)__popcll(__ballot(new_iso));
    }
  }
  const int cols[3] = {1, 2, 3};
  block_part<3>(c, D.part, cols);
}

// The day's per-block counts into row t (one block), then cleared.
constexpr int DIAG_REDUCE_BLOCK = 1024;
__global__ __launch_bounds__(DIAG_REDUCE_BLOCK) void diag_reduce_kernel(uint32_t* part, int32_t n_blocks,
                                                                        uint32_t* row) {
  __shared__ uint32_t s[DIAG_REDUCE_BLOCK / 64][N_PART];
  uint32_t acc[N_PART];
#pragma unroll
  for (int k = 0; k < N_PART; ++k) {
    acc[k] = 0;
    for (int b = threadIdx.x; b < n_blocks; b += DIAG_REDUCE_BLOCK) {
      uint32_t* p = part + (int64_t)k * CVK_DIAG_MAX_BLOCKS + b;
      acc[k] += *p;
      *p = 0;
    }
    acc[k] = wave_sum(acc[k]);
  }
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < N_PART; ++k) s[threadIdx.x >> 6][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < N_PART) {
    uint32_t tot = 0;
    for (int j = 0; j < DIAG_REDUCE_BLOCK / 64; ++j) tot += s[j][threadIdx.x];
    const int cols[N_PART] = {CVK_R_NEW_TESTS, CVK_R_NEW_DIAGNOSES, CVK_R_NEW_QUARANTINED, CVK_R_NEW_ISOLATED};
    row[cols[threadIdx.x]] = tot;
  }
}

int agent_grid(int64_t n) {
  const int64_t g = (n + DIAG_SPAN - 1) / DIAG_SPAN;
  return (int)(g < CVK_DIAG_MAX_BLOCKS ? g : CVK_DIAG_MAX_BLOCKS);
}

bool diag_ok(const cvk_diag* D) {
  return D && D->date_tested && D->date_pos_test && D->date_known_contact && D->date_quarantined &&
         D->date_end_quarantine && D->pending && D->cases && D->coarse && D->contacts && D->part && D->ring >= 1;
}

int check_common(const cvk_people* P, const cvk_params* prm, const cvk_work* W, const cvk_diag* D, int32_t t) {
  if (!P || !prm || !W) return fail(CVK_E_INVALID, "null people, params or work");
  if (!P->flags || !P->date_diagnosed || !P->date_end_isolation || !P->date_recovered || !P->rel_trans ||
      !P->rel_sus)
    return fail(CVK_E_INVALID, "people arrays missing");
  if (P->n <= 0 || P->n > INT32_MAX) return fail(CVK_E_INVALID, "agent count out of range");
  if (!W->results || !W->rec || !W->bound) return fail(CVK_E_INVALID, "workspace arrays missing");
  if (t < 0 || t >= W->npts) return fail(CVK_E_INVALID, "day out of range");
  if (!diag_ok(D)) return fail(CVK_E_INVALID, "cvk_diag arrays missing or ring < 1");
  if ((uintptr_t)D->cases & 7) return fail(CVK_E_ALIGN, "cvk_diag.cases must be 8-byte aligned");
  return CVK_OK;
}

int check_op(const cvk_diag_op* op, const cvk_diag* D, const cvk_params* prm, const cvk_layer* layers,
             int64_t n_agents, int32_t npts) {
  if (!op) return fail(CVK_E_INVALID, "null op");
  if (op->slot < 0 || op->slot > 255) return fail(CVK_E_INVALID, "op slot must be 0..255");
  if (op->kind == CVK_OP_TEST) {
    if (op->quar_policy < CVK_QP_NONE || op->quar_policy > CVK_QP_DAYS)
      return fail(CVK_E_INVALID, "unknown quarantine testing policy");
    if (op->n_quar_days < 0 || op->n_quar_days > CVK_MAX_QUAR_DAYS)
      return fail(CVK_E_INVALID, "n_quar_days out of range");
    if (op->ili_counts) {
      if (op->n_ili_days < 0) return fail(CVK_E_INVALID, "n_ili_days < 0");
      for (int r = 0; r < op->n_ili_days; ++r)
        if (op->ili_counts[r] < 0 || op->ili_counts[r] > n_agents)
          return fail(CVK_E_INVALID, "ili_counts must lie in [0, n_agents]");
      if (!D->select) return fail(CVK_E_INVALID, "ILI draws need cvk_diag.select");
    }
    if (op->n_swab_bins < 0) return fail(CVK_E_INVALID, "n_swab_bins < 0");
    if (op->n_swab_bins > 0) {
      if (!op->swab_base) return fail(CVK_E_INVALID, "swab delay without swab_base");
      if (op->n_swab_bins < npts) return fail(CVK_E_INVALID, "n_swab_bins must be at least the sim's days");
      if (op->n_swab_bins > D->swab_bins || !D->swab_hist || !D->swab_prob)
        return fail(CVK_E_CAPACITY, "cvk_diag.swab_bins too small for the op's swab delay");
    }
    if (op->sub_vals && !op->sub_mask) return fail(CVK_E_INVALID, "sub_vals without sub_mask");
    return CVK_OK;
  }
  if (op->kind != CVK_OP_TRACE) return fail(CVK_E_INVALID, "unknown op kind");
  if (op->capacity > 0 && !D->select) return fail(CVK_E_INVALID, "a tracing capacity needs cvk_diag.select");
  for (int l = 0; l < prm->n_layers; ++l) {
    if (!(op->trace_prob[l] > 0.0)) continue;
    if (!(op->trace_time[l] >= 0.0) || op->trace_time[l] >= (double)INT32_MAX / 2)
      return fail(CVK_E_INVALID, "trace_time must be >= 0");
    if ((int64_t)op->trace_time[l] + 1 >= D->ring) return fail(CVK_E_INVALID, "cvk_diag.ring too small for trace_time");
    if (!layers) return fail(CVK_E_INVALID, "null layers");
    if (layers[l].n_edges > 0 && (!layers[l].p1 || !layers[l].p2)) return fail(CVK_E_INVALID, "layer arrays missing");
    if (((uintptr_t)layers[l].p1 | (uintptr_t)layers[l].p2) & 15)
      return fail(CVK_E_ALIGN, "traced layers: p1/p2 must be 16-byte aligned");
  }
  return CVK_OK;
}

// defer_notify: leave the tracing op's notify pass to the day's update_states_post launch.
int apply_impl(const cvk_people* P, const cvk_layer* layers, const cvk_params* prm, cvk_work* W, const cvk_diag* D,
               const cvk_diag_op* op, int32_t t, hipStream_t st, bool defer_notify = false) {
  if (t < op->start_day || (op->end_day >= 0 && t > op->end_day)) return CVK_OK;
  const int grid = agent_grid(P->n);
  if (op->kind == CVK_OP_TEST) {
    const uint64_t* ili_thr = nullptr;
    const int64_t r = (int64_t)t - op->start_day;
    if (op->ili_counts && r < op->n_ili_days && op->ili_counts[r] > 0) {
      select_kth(AllItemsKey{prm->seed, t, STREAM_ILI | (uint32_t)op->slot}, P->n, op->ili_counts[r], D->select,
                 (uint32_t*)(D->select + 2), st);
      ili_thr = D->select;
    }
    const double* swab_prob = nullptr;
    if (op->n_swab_bins > 0) {
      const int32_t bins = op->n_swab_bins;
      fill_u32_kernel<<<1, DIAG_BLOCK, 0, st>>>(D->swab_hist, bins + 1);
      swab_hist_kernel<<<grid, DIAG_BLOCK, 0, st>>>(*P, t, D->swab_hist, bins);
      swab_prob_kernel<<<1, DIAG_BLOCK, 0, st>>>(D->swab_hist, op->swab_base, bins, D->swab_prob);
      swab_prob = D->swab_prob;
    }
    test_kernel<<<grid, DIAG_BLOCK, 0, st>>>(*P, *D, *op, prm->seed, t, ili_thr, swab_prob);
    return check_launch("test_kernel");
  }
  TraceLayers L{};
  int nl = 0;
  int64_t quads = 0;
  for (int l = 0; l < prm->n_layers; ++l) {
    if (!(op->trace_prob[l] > 0.0) || layers[l].n_edges <= 0) continue;  // trace_prob == 0: layer skipped
    L.p1[nl] = layers[l].p1;
    L.p2[nl] = layers[l].p2;
    L.n[nl] = layers[l].n_edges;
    L.layer[nl] = (uint32_t)l;
    quads = std::max<int64_t>(quads, (layers[l].n_edges + 3) / 4);
    ++nl;
  }
  if (nl == 0) return CVK_OK;
  L.nl = nl;
  const int shift = coarse_shift(P->n);
  const int64_t n_cwords = coarse_words(P->n);
  const uint32_t cap_stream = STREAM_TRACE_CAP | (uint32_t)op->slot;
  int32_t cap_mode = op->capacity < 0 ? 0 : op->capacity == 0 ? 2 : 1;
  if (cap_mode == 1)
    select_kth(CaseKey{*P, *D, prm->seed, op->presumptive, t, cap_stream}, P->n, op->capacity, D->select,
               (uint32_t*)(D->select + 2), st);
  trace_cases_kernel<<<grid, DIAG_BLOCK, 0, st>>>(*P, *D, op->presumptive, t, shift, cap_mode, D->select, prm->seed,
                                                  cap_stream);
  int rc = check_launch("trace_cases_kernel");
  if (rc) return rc;
  const int64_t gx = (quads + TRACE_EDGE_BLOCK - 1) / TRACE_EDGE_BLOCK;
  trace_edges_kernel<<<(unsigned)std::min<int64_t>(gx, TRACE_EDGE_GRID), TRACE_EDGE_BLOCK, (size_t)n_cwords * 4,
                       st>>>(L, D->cases, D->coarse, shift, (int32_t)n_cwords, D->contacts);
  rc = check_launch("trace_edges_kernel");
  if (rc || defer_notify) return rc;
  trace_notify_kernel<<<grid, DIAG_BLOCK, 0, st>>>(*P, *D, *op, prm->seed, t, (int32_t)n_cwords);
  return check_launch("trace_notify_kernel");
}

// Whether op runs its contact scan on day t (active, with a traced layer that has edges).
bool traces_on(const cvk_diag_op& op, const cvk_layer* layers, const cvk_params* prm, int32_t t) {
  if (op.kind != CVK_OP_TRACE || t < op.start_day || (op.end_day >= 0 && t > op.end_day)) return false;
  for (int l = 0; l < prm->n_layers; ++l)
    if (op.trace_prob[l] > 0.0 && layers[l].n_edges > 0) return true;
  return false;
}

// notify: a tracing op whose notify pass this launch does first (its scan has run), or null.
int post_impl(const cvk_people* P, const cvk_params* prm, cvk_work* W, const cvk_diag* D, int32_t t, hipStream_t st,
              const cvk_diag_op* notify = nullptr) {
  const int grid = agent_grid(P->n);
  if (notify)
    states_post_kernel<true><<<grid, DIAG_BLOCK, 0, st>>>(*P, *prm, *W, *D, t, *notify, (int32_t)coarse_words(P->n));
  else
    states_post_kernel<false><<<grid, DIAG_BLOCK, 0, st>>>(*P, *prm, *W, *D, t, cvk_diag_op{}, 0);
  const int rc = check_launch("states_post_kernel");
  if (rc) return rc;
  diag_reduce_kernel<<<1, DIAG_REDUCE_BLOCK, 0, st>>>(D->part, grid, W->results + (int64_t)t * CVK_R_COLS);
  return check_launch("diag_reduce_kernel");
}

}  // namespace

// The day loop's testing and tracing step (after the imports, before transmission).
int diag_day(const cvk_people* P, const cvk_layer* layers, const cvk_params* prm, cvk_work* W,
             const cvk_imports* imports, int32_t t, hipStream_t st) {
  if (!imports || !imports->diag) return CVK_OK;
  // the day's last op, when it traces, has its notify pass fused into update_states_post
  const int n = imports->n_diag_ops;
  const cvk_diag_op* last = n > 0 ? imports->diag_ops + (n - 1) : nullptr;
  const bool fuse = last && traces_on(*last, layers, prm, t);
  for (int j = 0; j < n; ++j) {
    const int rc = apply_impl(P, layers, prm, W, imports->diag, imports->diag_ops + j, t, st, fuse && j == n - 1);
    if (rc) return rc;
  }
  return post_impl(P, prm, W, imports->diag, t, st, fuse ? last : nullptr);
}

int check_diag_schedule(const cvk_people* P, const cvk_layer* layers, const cvk_params* prm, const cvk_work* W,
                        const cvk_imports* imports) {
  if (!imports || !imports->diag) return CVK_OK;
  int rc = check_common(P, prm, W, imports->diag, 0);
  if (rc) return rc;
  if (imports->n_diag_ops < 0 || (imports->n_diag_ops > 0 && !imports->diag_ops))
    return fail(CVK_E_INVALID, "bad diag_ops");
  for (int j = 0; j < imports->n_diag_ops; ++j) {
    rc = check_op(imports->diag_ops + j, imports->diag, prm, layers, P->n, W->npts);
    if (rc) return rc;
  }
  return CVK_OK;
}

}  // namespace cvk

using namespace cvk;

extern "C" int64_t cvk_diag_coarse_words(int64_t n) { return n > 0 ? coarse_words(n) : 0; }

extern "C" int cvk_diag_apply(const cvk_people* people, const cvk_layer* layers, const cvk_params* params,
                              cvk_work* work, const cvk_diag* diag, const cvk_diag_op* op, int32_t t, void* stream) {
  int rc = check_common(people, params, work, diag, t);
  if (rc) return rc;
  rc = check_op(op, diag, params, layers, people->n, work->npts);
  if (rc) return rc;
  return apply_impl(people, layers, params, work, diag, op, t, (hipStream_t)stream);
}

extern "C" int cvk_states_post(const cvk_people* people, const cvk_params* params, cvk_work* work,
                               const cvk_diag* diag, int32_t t, void* stream) {
  const int rc = check_common(people, params, work, diag, t);
  if (rc) return rc;
  return post_impl(people, params, work, diag, t, (hipStream_t)stream);
}
