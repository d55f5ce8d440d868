// Device-side helpers shared by the covasim_amd kernels: counter-based Philox4x32-10,
// uniform/normal transforms, duration sampling and the transmission-key encoding.
//
// All float arithmetic here is written in the exact operation order of the
// reference (covasim/utils.py, covasim/people.py); the library is compiled with
// -ffp-contract=off so no multiply-add is fused behind our back.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/covasim_amd.h"

namespace cvk {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): published multipliers and
// Weyl key increments; 10 rounds.
// ---------------------------------------------------------------------------
constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;
constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;

// Stream ids carried in counter word c2 (identical table in oracle/philox.py).
constexpr uint32_t STREAM_TX = 0x10000000u;        // | variant << 8 | layer
constexpr uint32_t STREAM_INFECT = 0x20000000u;
constexpr uint32_t STREAM_CHOOSE = 0x30000000u;
constexpr uint32_t STREAM_IMPORT = 0x40000000u;    // | variant
constexpr uint32_t STREAM_RESAMPLE = 0x50000000u;  // | layer
constexpr uint32_t STREAM_RESAMPLE_RETRY = 0x60000000u;  // | layer
constexpr uint32_t RESAMPLE_MAX_RETRY = 255;
// Population generation (cvk_popgen.hip); the low byte of COUNT / EDGE / EDGE_RETRY is the layer slot.
constexpr uint32_t STREAM_POP_AGE = 0x70000000u;
constexpr uint32_t STREAM_POP_AGE_NOISE = 0x70000100u;
constexpr uint32_t STREAM_POP_PERSON = 0x70000200u;
constexpr uint32_t STREAM_POP_HH_SIZE = 0x70000300u;
constexpr uint32_t STREAM_POP_HH_PERM = 0x70000400u;
constexpr uint32_t STREAM_POP_COUNT = 0x70000500u;
constexpr uint32_t STREAM_POP_EDGE = 0x70000600u;
constexpr uint32_t STREAM_POP_EDGE_RETRY = 0x70000700u;
// Edge moves of clip_edges (cvk_clip.hip); the low byte is the layer slot.
constexpr uint32_t STREAM_CLIP = 0x80000000u;
constexpr uint32_t STREAM_PARTIAL = 0xB0000000u;  // | layer: edge choice of a partial resample (frac < 1)
// Testing (| op slot) and contact tracing (| op slot << 8 | layer), cvk_diag.hip.
constexpr uint32_t STREAM_TEST = 0x90000000u;
constexpr uint32_t STREAM_TRACE = 0xA0000000u;
constexpr uint32_t STREAM_ILI = 0xC0000000u;        // | op slot: the agents with influenza-like symptoms
constexpr uint32_t STREAM_TRACE_CAP = 0xD0000000u;  // | op slot: the cases traced under a capacity

// Per-infection draw slots (counter word c3 of STREAM_INFECT).
constexpr uint32_t SLOT_EXP2INF = 0;
constexpr uint32_t SLOT_BERN_SYMP_SEV = 1;
constexpr uint32_t SLOT_BERN_CRIT_DEATH = 2;
constexpr uint32_t SLOT_DUR_A = 3;
constexpr uint32_t SLOT_DUR_B = 4;
constexpr uint32_t SLOT_DUR_C = 5;
constexpr uint32_t SLOT_DUR_D = 6;
constexpr uint32_t SLOT_NAB = 7;

struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
  }
  return u32x4{c0, c1, c2, c3};
}

__host__ __device__ inline u32x4 philox_seeded(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3) {
  return philox4x32_10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 53-bit double in [0,1): ((a>>5)*2^26 + (b>>6)) / 2^53.
__host__ __device__ inline double u01(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}


__device__ inline double box_muller(double u0, double u1) {
  return sqrt(-2.0 * log1p(-u0)) * cos((2.0 * 3.141592653589793) * u1);
}

// sample() of utils.py:156-236 given a standard normal z and uniform u0.
__device__ inline double sample_dur(const cvk_dur& d, double z, double u0) {
  switch (d.dist) {
    case CVK_DIST_UNIFORM:
      return d.par1 + (d.par2 - d.par1) * u0;
    case CVK_DIST_NORMAL:
      return d.par1 + d.par2 * z;
    case CVK_DIST_NORMAL_POS:
      return fabs(d.par1 + d.par2 * z);
    case CVK_DIST_NORMAL_INT:
      return rint(fabs(d.par1 + d.par2 * z));
    case CVK_DIST_LOGNORMAL:
      return d.par1 > 0 ? exp(d.mu + d.sigma * z) : 0.0;
    case CVK_DIST_LOGNORMAL_INT:
      return d.par1 > 0 ? rint(exp(d.mu + d.sigma * z)) : 0.0;
    default:
      return 0.0;
  }
}

// Duration draw of one infection-outcome slot.
__device__ inline double draw_dur(const cvk_dur& d, uint64_t seed, uint32_t agent, int32_t t, uint32_t slot) {
  const u32x4 w = philox_seeded(seed, agent, (uint32_t)t, STREAM_INFECT, slot);
  const double u0 = u01(w.x, w.y);
  const double u1 = u01(w.z, w.w);
  return sample_dur(d, box_muller(u0, u1), u0);
}

__device__ inline void draw_pair(uint64_t seed, uint32_t agent, int32_t t, uint32_t slot, double& u0,
                                 double& u1) {
  const u32x4 w = philox_seeded(seed, agent, (uint32_t)t, STREAM_INFECT, slot);
  u0 = u01(w.x, w.y);
  u1 = u01(w.z, w.w);
}

// ---------------------------------------------------------------------------
// Infection keys: the smallest key per target wins (People.infect keeps the
// first occurrence, people.py:465; layers and variants are visited in order,
// sim.py:622-649). Layout:
//   bit 63      class: 0 = seed/import (registered before transmission), 1 = transmission
//   bits 62..59 variant
//   bits 58..55 layer (CVK_LAYER_SEED / CVK_LAYER_IMPORT for class 0)
//   bit 54      direction (0: p1 -> p2, 1: p2 -> p1)
//   bits 53..0  edge index (class 1) or agent index (class 0)
// ---------------------------------------------------------------------------
constexpr uint64_t KEY_NONE = ~0ull;

__host__ __device__ inline uint64_t make_key(uint32_t cls, uint32_t v, uint32_t l, uint32_t dir, uint64_t e) {
  return ((uint64_t)cls << 63) | ((uint64_t)v << 59) | ((uint64_t)l << 55) | ((uint64_t)dir << 54) |
         (e & ((1ull << 54) - 1));
}
__host__ __device__ inline uint32_t key_class(uint64_t k) { return (uint32_t)(k >> 63); }
__host__ __device__ inline uint32_t key_variant(uint64_t k) { return (uint32_t)(k >> 59) & 0xF; }
__host__ __device__ inline uint32_t key_layer(uint64_t k) { return (uint32_t)(k >> 55) & 0xF; }
__host__ __device__ inline uint32_t key_dir(uint64_t k) { return (uint32_t)(k >> 54) & 1; }
__host__ __device__ inline uint64_t key_index(uint64_t k) { return k & ((1ull << 54) - 1); }

// vstate accessors
__host__ __device__ inline uint32_t exp_var(uint32_t vs) { return vs & 0xF; }
__host__ __device__ inline uint32_t rec_var(uint32_t vs) { return (vs >> 4) & 0xF; }
__host__ __device__ inline uint32_t inf_var(uint32_t f) { return (f >> CVK_INFVAR_SHIFT) & 0xF; }

}  // namespace cvk
