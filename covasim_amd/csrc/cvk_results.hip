// Results on the device: finalisation of the day loop's per-day count tables (covasim/sim.py:
// 764-850 finalize + compute_states, :872-895 compute_doubling, :898-1001 compute_r_eff 'daily')
// and the reduction over many sims of covasim/run.py:220-303 (MultiSim.reduce).
//
// The finalisation reads the [sim][day][column] count table the day loop accumulates and writes
// double rows [sim][row][day]: one workgroup per sim, a lane per flow / stock / by-variant series
// for the running sums, a lane per day for the derived states, then one lane each for the doubling
// time and r_eff (short serial recurrences). r_eff's mean infectious duration is a reduction over
// the sim's agents, done as fixed-order per-chunk partial sums so the result is deterministic.
// The reduction over sims sorts each (row, day) column of up to 4096 sims in LDS (bitonic) and
// interpolates quantiles as numpy's 'linear' method does, or sums with numpy's pairwise order.
#include <algorithm>
#include <math.h>
#include <string.h>
#include <vector>

#include "cvk_device.h"
#include "cvk_host.h"

namespace cvk {
namespace {

constexpr int FIN_BLOCK = 256;
constexpr int64_t REFF_CHUNK = 16384;
constexpr int PW_BLOCK = 128;  // numpy's pairwise-summation block

__device__ inline int flow_col(int f) {
  switch (f) {
    case 0: return CVK_R_NEW_INFECTIONS;
    case 1: return CVK_R_NEW_REINFECTIONS;
    case 2: return CVK_R_NEW_INFECTIOUS;
    case 3: return CVK_R_NEW_SYMPTOMATIC;
    case 4: return CVK_R_NEW_SEVERE;
    case 5: return CVK_R_NEW_CRITICAL;
    case 6: return CVK_R_NEW_RECOVERIES;
    case 7: return CVK_R_NEW_DEATHS;
    case 8: return CVK_R_NEW_TESTS;
    case 9: return CVK_R_NEW_DIAGNOSES;
    case 10: return CVK_R_NEW_KNOWN_DEATHS;
    case 11: return CVK_R_NEW_QUARANTINED;
    case 12: return CVK_R_NEW_ISOLATED;
    default: return -1;  // doses, vaccinated: not produced
  }
}

__device__ inline int stock_col(int k) { return CVK_R_N_SUSCEPTIBLE + k; }  // same order, n_susceptible..n_vaccinated

constexpr int FLOW_RECOVERIES = 6, FLOW_DEATHS = 7, FLOW_INFECTIONS = 0, FLOW_TESTS = 8, FLOW_DIAGNOSES = 9;
constexpr int STOCK_EXPOSED = 1, STOCK_INFECTIOUS = 2, STOCK_RECOVERED = 6, STOCK_VACCINATED = 12;

// numpy's pairwise summation of get(lo), .., get(lo + n - 1) (the order of np.add.reduce over a
// contiguous run): below 8 a running sum from -0.0, up to 128 eight interleaved accumulators
// combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) and then the remainder, above that the two
// halves (split at a multiple of 8) summed separately.
template <class Get>
__device__ double pairwise_sum(const Get& get, int64_t lo, int64_t n) {
  if (n < 8) {
    double res = -0.0;
    for (int64_t i = 0; i < n; ++i) res += get(lo + i);
    return res;
  }
  if (n <= PW_BLOCK) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = get(lo + j);
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += get(lo + i + j);
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += get(lo + i);
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_sum(get, lo, n2) + pairwise_sum(get, lo + n2, n - n2);
}

struct ReffPart {
  double outcome_sum, inf_sum;
  int64_t count, pad;
};

// Chunk c of the agents (sims split into REFF_CHUNK-agent chunks, chunk_first[s] = first chunk of
// sim s): sums over agents with a defined recovery date and over agents with a defined death date
// (an agent with both counts twice, as the reference's concatenation does).
__global__ __launch_bounds__(FIN_BLOCK) void reff_partials_kernel(int32_t S, const int64_t* chunk_first,
                                                                  const int64_t* agent_off, const float* d_inf,
                                                                  const float* d_rec, const float* d_dead,
                                                                  ReffPart* parts) {
  __shared__ double s_o[FIN_BLOCK], s_i[FIN_BLOCK];
  __shared__ int64_t s_n[FIN_BLOCK];
  const int64_t c = blockIdx.x;
  int lo = 0, hi = S;  // chunk_first[lo] <= c < chunk_first[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (chunk_first[mid] <= c) lo = mid;
    else hi = mid;
  }
  const int s = lo;
  const int64_t a0 = agent_off[s] + (c - chunk_first[s]) * REFF_CHUNK;
  const int64_t a1 = std::min<int64_t>(a0 + REFF_CHUNK, agent_off[s + 1]);
  double so = 0.0, si = 0.0;
  int64_t n = 0;
  for (int64_t i = a0 + threadIdx.x; i < a1; i += FIN_BLOCK) {
    const float r = d_rec[i], d = d_dead[i], f = d_inf[i];
    if (r == r) {
      so += (double)r;
      si += (double)f;
      ++n;
    }
    if (d == d) {
      so += (double)d;
      si += (double)f;
      ++n;
    }
  }
  s_o[threadIdx.x] = so;
  s_i[threadIdx.x] = si;
  s_n[threadIdx.x] = n;
  __syncthreads();
  for (int w = FIN_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_o[threadIdx.x] += s_o[threadIdx.x + w];
      s_i[threadIdx.x] += s_i[threadIdx.x + w];
      s_n[threadIdx.x] += s_n[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[c] = ReffPart{s_o[0], s_i[0], s_n[0], 0};
}

struct FinArgs {
  int32_t S, npts, nv, use_waning, smoothing, window;
  double max_doubling;
  const uint32_t* results;
  const double* results_f;
  const double* imports;
  const int64_t* agent_off;
  const int64_t* chunk_first;
  const double* pop_infected;
  const int32_t* initial_period;
  const ReffPart* parts;
  double* out;
  double* out_v;
  double* scratch;  // [S][2][npts]
};

__global__ __launch_bounds__(FIN_BLOCK) void finalize_kernel(FinArgs a) {
  const int s = blockIdx.x;
  const int T = a.npts, nv = a.nv;
  const uint32_t* R = a.results + (int64_t)s * T * CVK_R_COLS;
  const double* RF = a.results_f + (int64_t)s * T * CVK_RF_COLS;
  double* O = a.out + (int64_t)s * CVK_F_ROWS * T;
  double* OV = a.out_v + (int64_t)s * CVK_FV_ROWS * nv * T;
  auto row = [&](int r) { return O + (int64_t)r * T; };
  auto vrow = [&](int r, int v) { return OV + ((int64_t)r * nv + v) * T; };
  const double pop = (double)(a.agent_off[s + 1] - a.agent_off[s]);
  const double pi = a.pop_infected[s];

  // running sums and copies: flows, stocks, by-variant series
  const int n_tasks = CVK_F_N_FLOWS + CVK_F_N_STOCKS + 6 * nv;
  for (int task = threadIdx.x; task < n_tasks; task += FIN_BLOCK) {
    if (task < CVK_F_N_FLOWS) {
      const int f = task, col = flow_col(f);
      double* nw = row(CVK_F_NEW + f);
      double* cm = row(CVK_F_CUM + f);
      double acc = 0.0;
      for (int t = 0; t < T; ++t) {
        const double v = col >= 0 ? (double)R[(int64_t)t * CVK_R_COLS + col] : 0.0;
        nw[t] = v;
        acc += v;
        cm[t] = f == FLOW_INFECTIONS ? acc + pi : acc;
      }
    } else if (task < CVK_F_N_FLOWS + CVK_F_N_STOCKS) {
      const int k = task - CVK_F_N_FLOWS;
      double* st = row(CVK_F_STOCK + k);
      for (int t = 0; t < T; ++t) st[t] = (double)R[(int64_t)t * CVK_R_COLS + stock_col(k)];
    } else {
      const int j = task - CVK_F_N_FLOWS - CVK_F_N_STOCKS, k = j / nv, v = j % nv;
      const int col = CVK_R_BYVAR + k * CVK_MAX_VARIANTS + v;
      double acc = 0.0;
      for (int t = 0; t < T; ++t) {
        const double x = (double)R[(int64_t)t * CVK_R_COLS + col];
        if (k < 4) {
          vrow(CVK_FV_NEW + k, v)[t] = x;
          acc += x;
          // sim.py:785-786 adds pop_infected to the whole (n_variants, npts) cumulative array
          vrow(CVK_FV_CUM + k, v)[t] = k == 0 ? acc + pi : acc;
        } else {
          vrow(k == 4 ? CVK_FV_N_EXPOSED : CVK_FV_N_INFECTIOUS, v)[t] = x;
        }
      }
    }
  }
  __syncthreads();

  // derived states, one lane per day (compute_states)
  const double count_recov = a.use_waning ? 0.0 : 1.0;
  for (int t = threadIdx.x; t < T; t += FIN_BLOCK) {
    const double cd = row(CVK_F_CUM + FLOW_DEATHS)[t], crec = row(CVK_F_CUM + FLOW_RECOVERIES)[t];
    const double ne = row(CVK_F_STOCK + STOCK_EXPOSED)[t], ni = row(CVK_F_STOCK + STOCK_INFECTIOUS)[t];
    const double nr = row(CVK_F_STOCK + STOCK_RECOVERED)[t];
    const double n_alive = pop - cd;
    const double n_sus = n_alive - ne - count_recov * crec;
    row(CVK_F_N_ALIVE)[t] = n_alive;
    row(CVK_F_N_NAIVE)[t] = pop - cd - nr - ne;
    row(CVK_F_STOCK + 0)[t] = n_sus;
    row(CVK_F_N_PREINFECTIOUS)[t] = ne - ni;
    row(CVK_F_N_REMOVED)[t] = count_recov * crec + cd;
    row(CVK_F_PREVALENCE)[t] = ne / n_alive;
    row(CVK_F_INCIDENCE)[t] = row(CVK_F_NEW + FLOW_INFECTIONS)[t] / n_sus;
    row(CVK_F_FRAC_VACCINATED)[t] = row(CVK_F_STOCK + STOCK_VACCINATED)[t] / n_alive;
    row(CVK_F_N_IMPORTS)[t] = a.imports ? a.imports[(int64_t)s * T + t] : 0.0;
    row(CVK_F_POP_NABS)[t] = RF[(int64_t)t * CVK_RF_COLS + CVK_RF_NAB_SUM] /
                             (double)R[(int64_t)t * CVK_R_COLS + CVK_R_N_ALIVE_NAB];
    row(CVK_F_POP_PROTECTION)[t] = RF[(int64_t)t * CVK_RF_COLS + CVK_RF_SUS_IMM_SUM] / ((double)nv * pop);
    row(CVK_F_POP_SYMP_PROTECTION)[t] = RF[(int64_t)t * CVK_RF_COLS + CVK_RF_SYMP_IMM_SUM] / ((double)nv * pop);
    // compute_yield (sim.py:853-868): positives per test, and relative to the undiagnosed alive
    const double nt = row(CVK_F_NEW + FLOW_TESTS)[t];
    const double ty = nt != 0.0 ? row(CVK_F_NEW + FLOW_DIAGNOSES)[t] / nt : 0.0;
    row(CVK_F_TEST_YIELD)[t] = ty;
    row(CVK_F_REL_TEST_YIELD)[t] = ni != 0.0 ? ty / (ni / (n_alive - row(CVK_F_CUM + FLOW_DIAGNOSES)[t])) : 0.0;
    const double inv_sus = 1.0 / n_sus, inv_alive = 1.0 / n_alive;
    for (int v = 0; v < nv; ++v) {
      const double x = vrow(CVK_FV_NEW + 0, v)[t];
      vrow(CVK_FV_INCIDENCE, v)[t] = x * inv_sus;
      vrow(CVK_FV_PREVALENCE, v)[t] = x * inv_alive;
    }
  }
  __syncthreads();

  if (threadIdx.x == 0) {  // compute_doubling
    const double* cum = row(CVK_F_CUM + FLOW_INFECTIONS);
    double* dt = row(CVK_F_DOUBLING_TIME);
    const int w = a.window;
    for (int t = 0; t < T; ++t) {
      double v = NAN;
      if (t >= w) {
        const double now = cum[t], prev = cum[t - w];
        if (prev > 0 && now > prev) {
          const double d = (double)w * 0.6931471805599453 / log(now / prev);
          v = d < a.max_doubling ? d : a.max_doubling;
        }
      }
      dt[t] = v;
    }
  }
  if (threadIdx.x == 64) {  // compute_r_eff, method 'daily'
    double so = 0.0, si = 0.0;
    int64_t n = 0;
    for (int64_t c = a.chunk_first[s]; c < a.chunk_first[s + 1]; ++c) {
      so += a.parts[c].outcome_sum;
      si += a.parts[c].inf_sum;
      n += a.parts[c].count;
    }
    const double mean_inf = n > 0 ? so / (double)n - si / (double)n : 0.0;
    double* raw = a.scratch + (int64_t)s * 2 * T;
    double* tmp = raw + T;
    const double* nw = row(CVK_F_NEW + FLOW_INFECTIONS);
    const double* imp = row(CVK_F_N_IMPORTS);
    const double* ninf = row(CVK_F_STOCK + STOCK_INFECTIOUS);
    for (int t = 0; t < T; ++t) raw[t] = mean_inf * (ninf[t] > 0 ? (nw[t] - imp[t]) / ninf[t] : 0.0);
    double* out = row(CVK_F_R_EFF);
    if (T >= 3) {
      const int ip = a.initial_period[s];
      auto at = [raw](int64_t i) { return raw[i]; };
      for (int i = 0; i < ip; ++i) raw[i] = pairwise_sum(at, i, ip - i) / (double)(ip - i);
      // repeated convolution with (0.25, 0.5, 0.25), zero beyond the ends (np.convolve 'same')
      for (int t = 0; t < T; ++t) out[t] = raw[t];
      for (int r = 0; r < a.smoothing; ++r) {
        for (int t = 0; t < T; ++t) {
          double acc;
          if (t == 0) acc = out[0] * 0.5 + out[1] * 0.25;
          else if (t == T - 1) acc = out[T - 2] * 0.25 + out[T - 1] * 0.5;
          else acc = out[t - 1] * 0.25 + out[t] * 0.5 + out[t + 1] * 0.25;
          tmp[t] = acc;
        }
        for (int t = 0; t < T; ++t) out[t] = tmp[t];
      }
      const int sm = std::min(a.smoothing, T);
      for (int t = 0; t < sm; ++t) {
        out[t] = raw[t];
        out[T - 1 - t] = raw[T - 1 - t];
      }
    } else {
      for (int t = 0; t < T; ++t) out[t] = raw[t];
    }
  }
}

// numpy 'linear' quantile of a sorted run v[0..n): virtual index (n - 1) q; neighbours floor and
// floor + 1, or the last element twice at or past the end (index -1, so the weight is vi + 1);
// numpy's two-sided lerp even then, so infinities give what numpy gives.
__device__ inline double linear_quantile(const double* v, int n, double q) {
  const double vi = (double)(n - 1) * q;
  int ia, ib;
  double prev;
  if (vi >= (double)(n - 1)) {
    ia = ib = n - 1;
    prev = -1.0;
  } else if (vi < 0.0) {
    ia = ib = 0;
    prev = 0.0;
  } else {
    prev = floor(vi);
    ia = (int)prev;
    ib = ia + 1;
  }
  const double a = v[ia], b = v[ib], g = vi - prev;
  const double diff = b - a;
  return g >= 0.5 ? b - diff * (1.0 - g) : a + diff * g;
}

__global__ __launch_bounds__(FIN_BLOCK) void reduce_quantile_kernel(int32_t S, int32_t P, int64_t n_cols,
                                                                    const double* values, double q_lo, double q_hi,
                                                                    double* center, double* low, double* high) {
  extern __shared__ double s_v[];
  __shared__ int s_nan;
  const int64_t c = blockIdx.x;
  if (threadIdx.x == 0) s_nan = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += FIN_BLOCK) {
    const double x = i < S ? values[(int64_t)i * n_cols + c] : INFINITY;
    if (x != x) s_nan = 1;
    s_v[i] = x;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += FIN_BLOCK) {
        const int l = i ^ j;
        if (l > i) {
          const double x = s_v[i], y = s_v[l];
          const bool up = (i & k) == 0;
          if (up ? (x > y) : (x < y)) {
            s_v[i] = y;
            s_v[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    if (s_nan) {
      center[c] = NAN;
      low[c] = NAN;
      high[c] = NAN;
    } else {
      center[c] = linear_quantile(s_v, S, 0.5);
      low[c] = linear_quantile(s_v, S, q_lo);
      high[c] = linear_quantile(s_v, S, q_hi);
    }
  }
}

// np.mean and np.std (ddof 0) along the sims of each column, bounds * std either side: the mean is
// the pairwise sum / S, the variance the pairwise sum of (x - mean)^2 (each rounded as numpy's
// subtract and multiply do) / S.
__global__ __launch_bounds__(FIN_BLOCK) void reduce_mean_kernel(int32_t S, int64_t n_cols, const double* values,
                                                                double bounds, double* center, double* low,
                                                                double* high) {
  const int64_t c = (int64_t)blockIdx.x * FIN_BLOCK + threadIdx.x;
  if (c >= n_cols) return;
  auto x = [=](int64_t i) { return values[i * n_cols + c]; };
  const double mean = pairwise_sum(x, 0, S) / (double)S;
  auto dev2 = [=](int64_t i) {
    const double d = values[i * n_cols + c] - mean;
    return d * d;
  };
  const double sd = sqrt(pairwise_sum(dev2, 0, S) / (double)S);
  center[c] = mean;
  low[c] = mean - bounds * sd;
  high[c] = mean + bounds * sd;
}

inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

struct FinWork {
  int64_t* agent_off;
  int64_t* chunk_first;
  double* pop_infected;
  int32_t* initial_period;
  ReffPart* parts;
  double* scratch;
  int64_t n_chunks;
};

int64_t fin_layout(const cvk_finalize_desc* d, FinWork* w, char* base) {
  const int S = d->n_sims;
  int64_t chunks = 0;
  for (int s = 0; s < S; ++s) {
    const int64_t n = d->agent_offsets_host[s + 1] - d->agent_offsets_host[s];
    chunks += std::max<int64_t>(1, (n + REFF_CHUNK - 1) / REFF_CHUNK);
  }
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  char* ao = take(8 * (S + 1));
  char* cf = take(8 * (S + 1));
  char* pi = take(8 * S);
  char* ip = take(4 * S);
  char* parts = take((int64_t)sizeof(ReffPart) * chunks);
  char* scratch = take(8 * 2 * (int64_t)S * d->npts);
  if (w) {
    w->agent_off = (int64_t*)ao;
    w->chunk_first = (int64_t*)cf;
    w->pop_infected = (double*)pi;
    w->initial_period = (int32_t*)ip;
    w->parts = (ReffPart*)parts;
    w->scratch = (double*)scratch;
    w->n_chunks = chunks;
  }
  return off;
}

int check_desc(const cvk_finalize_desc* d) {
  if (!d || d->n_sims < 1 || d->n_sims > CVK_BATCH_MAX_SIMS || d->npts < 1 || d->n_variants < 1 ||
      d->n_variants > CVK_MAX_VARIANTS || d->r_eff_smoothing < 0 || d->doubling_window < 1 ||
      !d->pop_infected_host || !d->initial_period_host || !d->agent_offsets_host)
    return fail(CVK_E_INVALID, "cvk_finalize: invalid descriptor");
  for (int s = 0; s < d->n_sims; ++s)
    if (d->agent_offsets_host[s + 1] <= d->agent_offsets_host[s] || d->initial_period_host[s] < 0 ||
        d->initial_period_host[s] > d->npts)
      return fail(CVK_E_INVALID, "cvk_finalize: agent offsets must increase and initial periods lie in [0, npts]");
  return CVK_OK;
}

}  // namespace
}  // namespace cvk

using namespace cvk;

extern "C" int64_t cvk_finalize_work_bytes(const cvk_finalize_desc* desc) {
  if (check_desc(desc)) return 0;
  return fin_layout(desc, nullptr, nullptr);
}

extern "C" int cvk_finalize(const cvk_finalize_desc* desc, const uint32_t* results, const double* results_f,
                            const double* imports, const float* date_infectious, const float* date_recovered,
                            const float* date_dead, double* out, double* out_variant, void* work, int64_t work_bytes,
                            void* stream) {
  int rc = check_desc(desc);
  if (rc) return rc;
  if (!results || !results_f || !date_infectious || !date_recovered || !date_dead || !out || !out_variant)
    return fail(CVK_E_INVALID, "cvk_finalize: null table, date array or output");
  if (!work || ((uintptr_t)work & 255)) return fail(CVK_E_ALIGN, "cvk_finalize: workspace must be 256-byte aligned");
  if (work_bytes < fin_layout(desc, nullptr, nullptr)) return fail(CVK_E_CAPACITY, "cvk_finalize: workspace too small");
  FinWork w;
  fin_layout(desc, &w, (char*)work);
  const int S = desc->n_sims;
  std::vector<int64_t> cf(S + 1, 0);
  std::vector<double> pi(S);
  for (int s = 0; s < S; ++s) {
    const int64_t n = desc->agent_offsets_host[s + 1] - desc->agent_offsets_host[s];
    cf[s + 1] = cf[s] + std::max<int64_t>(1, (n + REFF_CHUNK - 1) / REFF_CHUNK);
    pi[s] = desc->pop_infected_host[s];
  }
  hipStream_t st = (hipStream_t)stream;
  if (hipMemcpyAsync(w.agent_off, desc->agent_offsets_host, 8 * (S + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(w.chunk_first, cf.data(), 8 * (S + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(w.pop_infected, pi.data(), 8 * S, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(w.initial_period, desc->initial_period_host, 4 * S, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(CVK_E_HIP, "cvk_finalize: upload failed");
  hipLaunchKernelGGL(reff_partials_kernel, dim3((uint32_t)w.n_chunks), dim3(FIN_BLOCK), 0, st, S,
                     (const int64_t*)w.chunk_first, (const int64_t*)w.agent_off, date_infectious, date_recovered,
                     date_dead, w.parts);
  FinArgs a;
  memset(&a, 0, sizeof(a));
  a.S = S;
  a.npts = desc->npts;
  a.nv = desc->n_variants;
  a.use_waning = desc->use_waning;
  a.smoothing = desc->r_eff_smoothing;
  a.window = desc->doubling_window;
  a.max_doubling = desc->max_doubling_time;
  a.results = results;
  a.results_f = results_f;
  a.imports = imports;
  a.agent_off = w.agent_off;
  a.chunk_first = w.chunk_first;
  a.pop_infected = w.pop_infected;
  a.initial_period = w.initial_period;
  a.parts = w.parts;
  a.out = out;
  a.out_v = out_variant;
  a.scratch = w.scratch;
  hipLaunchKernelGGL(finalize_kernel, dim3(S), dim3(FIN_BLOCK), 0, st, a);
  if ((rc = check_launch("cvk_finalize"))) return rc;
  if (hipStreamSynchronize(st) != hipSuccess) return fail(CVK_E_HIP, "cvk_finalize: device error");
  return CVK_OK;
}

extern "C" int cvk_reduce(int32_t n_sims, int64_t n_cols, const double* values, int32_t use_mean, double q_low,
                          double q_high, double bounds, double* center, double* low, double* high, void* stream) {
  if (n_sims < 1 || n_sims > CVK_BATCH_MAX_SIMS || n_cols < 0 || n_cols > (1ll << 31) ||
      (n_cols > 0 && (!values || !center || !low || !high)))
    return fail(CVK_E_INVALID, "cvk_reduce: invalid sizes or null pointer");
  if (!use_mean && !(q_low >= 0.0 && q_low <= 1.0 && q_high >= 0.0 && q_high <= 1.0))
    return fail(CVK_E_INVALID, "cvk_reduce: quantiles must lie in [0, 1]");
  if (n_cols == 0) return CVK_OK;
  hipStream_t st = (hipStream_t)stream;
  if (use_mean) {
    hipLaunchKernelGGL(reduce_mean_kernel, dim3((uint32_t)((n_cols + FIN_BLOCK - 1) / FIN_BLOCK)), dim3(FIN_BLOCK), 0,
                       st, n_sims, n_cols, values, bounds, center, low, high);
    return check_launch("cvk_reduce");
  }
  int P = 1;
  while (P < n_sims) P <<= 1;
  hipLaunchKernelGGL(reduce_quantile_kernel, dim3((uint32_t)n_cols), dim3(FIN_BLOCK), (size_t)P * sizeof(double), st,
                     n_sims, P, n_cols, values, q_low, q_high, center, low, high);
  return check_launch("cvk_reduce");
}
