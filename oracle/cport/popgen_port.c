/*
 * Population generator, CPU restatement (TEST INFRASTRUCTURE: the checker of the device
 * generator covasim_amd/csrc/cvk_popgen.hip and the CPU baseline of population building).
 *
 * Restates covasim/population.py:143-226 (make_randpop ages and sexes), :241-284
 * (make_random_contacts), :287-329 (make_microstructured_contacts), :332-364
 * (make_hybrid_contacts) and people.py:139-160 + utils.py:409-426 (prognoses, rel_trans) with
 * counter-based Philox draws, written as plain sequential loops: households are drawn one after
 * another until they cover the population, the age order is a sequential stable counting sort,
 * layers are walked member by member. The draw rules are the ones stated with cvk_popgen in the
 * library's C header; this file shares no code with the library.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void port_philox(uint32_t c[4], uint64_t seed); /* covasim_port.c */

#define POP_AGE 0x70000000u
#define POP_AGE_NOISE 0x70000100u
#define POP_PERSON 0x70000200u
#define POP_HH_SIZE 0x70000300u
#define POP_HH_PERM 0x70000400u
#define POP_COUNT 0x70000500u
#define POP_EDGE 0x70000600u
#define POP_EDGE_RETRY 0x70000700u
#define KMAX 65536

typedef struct {
  uint64_t seed;
  int32_t n_age_bins, n_hh_sizes, hh_size_min, correlate, n_prog_bins, pad;
  double age_cdf[32], age_lo[32], age_width[32], hh_cdf[64];
  double age_noise;
  double prog_cutoff[16], trans_or[16];
  float symp_prob[16], severe_prob[16], crit_prob[16], death_prob[16], rel_sus[16];
  double nb_p0, nb_q, nb_r, nb_step;
} pop_tables;

typedef struct {
  int32_t kind; /* 0 Poisson, 1 negative binomial */
  int32_t pad;
  double p0, lam, r, q;
} pop_counts;

static void draw4(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) {
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
  port_philox(out, seed);
}

static double unit(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

/* first table entry whose cumulative probability reaches u (numpy searchsorted 'left'), capped */
static int pick(const double* cdf, int n, double u) {
  for (int j = 0; j < n; ++j)
    if (!(cdf[j] < u)) return j;
  return n - 1;
}

static int64_t inverse_count(const pop_counts* c, double u) {
  int64_t k = 0;
  double p = c->p0, F = p;
  while (u > F && k < KMAX) {
    if (c->kind == 0) {
      k += 1;
      p = p * c->lam / (double)k;
    } else {
      const double ratio = ((double)k + c->r) / ((double)k + 1.0);
      p = p * ratio * c->q;
      k += 1;
    }
    F = F + p;
  }
  return k;
}

/* numpy: np.round(k / 2.0), ties to even */
static int64_t half_even(int64_t k) {
  const double h = (double)k / 2.0;
  const double f = floor(h);
  if (h - f != 0.5) return (int64_t)h;
  return ((int64_t)f % 2 == 0) ? (int64_t)f : (int64_t)f + 1;
}

void pop_ages(const pop_tables* g, int64_t n, float* age, uint8_t* key) {
  for (int64_t i = 0; i < n; ++i) {
    uint32_t w[4], z[4];
    draw4(g->seed, (uint32_t)i, 0u, POP_AGE, 0u, w);
    const int b = pick(g->age_cdf, g->n_age_bins, unit(w[0], w[1]));
    const double a = g->age_lo[b] + g->age_width[b] * unit(w[2], w[3]);
    age[i] = (float)a;
    if (key) {
      draw4(g->seed, (uint32_t)i, 0u, POP_AGE_NOISE, 0u, z);
      const double s = ((double)z[0] + (double)z[1] + (double)z[2] + (double)z[3]) / 4294967296.0;
      double k = floor(a + (s - 2.0) * g->age_noise) + 64.0;
      if (k < 0.0) k = 0.0;
      if (k > 255.0) k = 255.0;
      key[i] = (uint8_t)k;
    }
  }
}

static uint32_t feistel(uint64_t seed, uint32_t x, uint32_t b) {
  const uint32_t mask = (1u << b) - 1u;
  uint32_t L = x >> b, R = x & mask;
  for (uint32_t r = 0; r < 4; ++r) {
    uint32_t w[4];
    draw4(seed, R, r, POP_HH_PERM, 0u, w);
    const uint32_t t = L ^ (w[0] & mask);
    L = R;
    R = t;
  }
  return (L << b) | R;
}

/* Households of n agents. Returns the edge count; writes edges only when e_cap covers them.
 * *n_households = H. age_out gets the ages in final agent order. */
int64_t pop_households(const pop_tables* g, int64_t n, const float* age_in, const uint8_t* key, float* age_out,
                       int32_t* p1, int32_t* p2, int64_t e_cap, int64_t* n_households) {
  int64_t cap = 1024, H = 0, covered = 0;
  int64_t* size = (int64_t*)malloc(sizeof(int64_t) * cap);
  while (covered < n) {
    if (H == cap) {
      cap *= 2;
      size = (int64_t*)realloc(size, sizeof(int64_t) * cap);
    }
    uint32_t w[4];
    draw4(g->seed, (uint32_t)H, (uint32_t)((uint64_t)H >> 32), POP_HH_SIZE, 0u, w);
    size[H] = g->hh_size_min + pick(g->hh_cdf, g->n_hh_sizes, unit(w[0], w[1]));
    covered += size[H];
    ++H;
  }
  if (H > 0) size[H - 1] -= covered - n;
  *n_households = H;

  /* agents in household order: stable counting sort by key, or identity */
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  if (g->correlate) {
    int64_t count[257];
    memset(count, 0, sizeof(count));
    for (int64_t i = 0; i < n; ++i) count[key[i] + 1]++;
    for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
    for (int64_t i = 0; i < n; ++i) order[count[key[i]]++] = (int32_t)i;
  } else {
    for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
  }

  /* position of each household, then the first agent of each position */
  int64_t* where = (int64_t*)malloc(sizeof(int64_t) * (H > 0 ? H : 1));
  int64_t* at = (int64_t*)malloc(sizeof(int64_t) * (H > 0 ? H : 1)); /* household at position j */
  uint32_t nb = 0;
  for (uint64_t x = H > 1 ? (uint64_t)(H - 1) : 0; x; x >>= 1) ++nb;
  uint32_t b = (nb + 1) / 2;
  if (b < 1) b = 1;
  for (int64_t h = 0; h < H; ++h) {
    int64_t j = h;
    if (g->correlate) {
      uint32_t y = (uint32_t)h;
      do {
        y = feistel(g->seed, y, b);
      } while ((int64_t)y >= H);
      j = y;
    }
    where[h] = j;
    at[j] = h;
  }
  int64_t* old_start = (int64_t*)malloc(sizeof(int64_t) * (H > 0 ? H : 1));
  int64_t run = 0;
  for (int64_t h = 0; h < H; ++h) {
    old_start[h] = run;
    run += size[h];
  }
  int64_t agent = 0, e = 0;
  for (int64_t j = 0; j < H; ++j) {
    const int64_t h = at[j], m = size[h];
    for (int64_t k = 0; k < m; ++k) age_out[agent + k] = age_in[order[old_start[h] + k]];
    for (int64_t a = 0; a < m; ++a)
      for (int64_t c = a + 1; c < m; ++c) {
        if (e < e_cap) {
          p1[e] = (int32_t)(agent + a);
          p2[e] = (int32_t)(agent + c);
        }
        ++e;
      }
    agent += m;
  }
  free(size);
  free(order);
  free(where);
  free(at);
  free(old_start);
  return e;
}

int64_t pop_members(int64_t n, const float* age, double lo, double hi, int32_t* mapping) {
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double a = (double)age[i];
    if (a >= lo && a < hi) mapping[k++] = (int32_t)i;
  }
  return k;
}

/* Edge counts per member (offsets[n_members + 1]); returns the layer's edge count. */
int64_t pop_contacts_plan(const pop_tables* g, int32_t slot, int64_t n_members, const pop_counts* c,
                          int64_t* offsets) {
  int64_t total = 0;
  for (int64_t m = 0; m < n_members; ++m) {
    uint32_t w[4];
    draw4(g->seed, (uint32_t)m, (uint32_t)((uint64_t)m >> 32), POP_COUNT | (uint32_t)slot, 0u, w);
    offsets[m] = total;
    total += half_even(inverse_count(c, unit(w[0], w[1])));
  }
  offsets[n_members] = total;
  return total;
}

static uint32_t lemire(uint64_t seed, uint32_t slot, int64_t e, uint32_t word, uint32_t n) {
  const uint32_t thr = (uint32_t)((0x100000000ull - n) % n);
  uint64_t m = (uint64_t)word * n;
  for (uint32_t k = 0; k < 255 && (uint32_t)m < thr; ++k) {
    uint32_t r[4];
    draw4(seed, (uint32_t)e, (uint32_t)((uint64_t)e >> 32), POP_EDGE_RETRY | slot, k, r);
    m = (uint64_t)r[0] * n;
  }
  return (uint32_t)(m >> 32);
}

void pop_contacts_fill(const pop_tables* g, int32_t slot, int64_t n_members, const int32_t* mapping,
                       const int64_t* offsets, int32_t* p1, int32_t* p2) {
  for (int64_t m = 0; m < n_members; ++m)
    for (int64_t e = offsets[m]; e < offsets[m + 1]; ++e) p1[e] = mapping ? mapping[m] : (int32_t)m;
  const int64_t n_edges = offsets[n_members];
  for (int64_t e = 0; e < n_edges; ++e) {
    uint32_t w[4];
    const uint64_t q = (uint64_t)e >> 2;
    draw4(g->seed, (uint32_t)q, (uint32_t)(q >> 32), POP_EDGE | (uint32_t)slot, 0u, w);
    const uint32_t r = lemire(g->seed, (uint32_t)slot, e, w[e & 3], (uint32_t)n_members);
    p2[e] = mapping ? mapping[r] : (int32_t)r;
  }
}

void pop_people(const pop_tables* g, int64_t n, const float* age, uint8_t* sex, float* symp, float* severe,
                float* crit, float* death, float* rel_sus, float* rel_trans) {
  pop_counts nb;
  nb.kind = 1;
  nb.pad = 0;
  nb.p0 = g->nb_p0;
  nb.lam = 0.0;
  nb.r = g->nb_r;
  nb.q = g->nb_q;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t w[4];
    draw4(g->seed, (uint32_t)i, 0u, POP_PERSON, 0u, w);
    int bin = -1;
    for (int k = 0; k < g->n_prog_bins; ++k)
      if ((double)age[i] >= g->prog_cutoff[k]) ++bin;
    if (bin < 0) bin = 0;
    const int64_t draws = inverse_count(&nb, unit(w[1], w[2]));
    sex[i] = (uint8_t)(w[0] & 1u);
    symp[i] = g->symp_prob[bin];
    severe[i] = g->severe_prob[bin];
    crit[i] = g->crit_prob[bin];
    death[i] = g->death_prob[bin];
    rel_sus[i] = g->rel_sus[bin];
    rel_trans[i] = (float)(g->trans_or[bin] * ((double)draws * g->nb_step));
  }
}
