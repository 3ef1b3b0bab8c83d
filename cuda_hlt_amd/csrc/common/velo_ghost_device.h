// Device functions of the ghost-track filter shared by the feature,
// inference and training kernels and by consolidation (the inline filter):
// the ten track features and the perceptron's forward pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "velo_device.h"
#include "velo_ghost.h"

namespace velo {

// Offsets of the parameter blocks (W1[32][10], b1, W2[32][32], b2, w3, b3)
constexpr uint32_t kGhostW1 = 0;
constexpr uint32_t kGhostB1 = kGhostW1 + VELO_GHOST_HIDDEN * VELO_GHOST_FEATURES;
constexpr uint32_t kGhostW2 = kGhostB1 + VELO_GHOST_HIDDEN;
constexpr uint32_t kGhostB2 = kGhostW2 + VELO_GHOST_HIDDEN * VELO_GHOST_HIDDEN;
constexpr uint32_t kGhostW3 = kGhostB2 + VELO_GHOST_HIDDEN;
constexpr uint32_t kGhostB3 = kGhostW3 + VELO_GHOST_HIDDEN;
static_assert(kGhostB3 + 1 == VELO_GHOST_PARAMS, "parameter count");

constexpr float kGhostPi = 3.14159265358979f;

// Module of a VP LHCbID: sensor = bits 18-25, four sensors per module.
__host__ __device__ inline uint32_t lhcbid_module(uint32_t id) { return ((id >> 18) & 0xFFu) >> 2; }

// The ten features of one track.  hits(h, &x, &y, &z, &id) gives hit h in
// track order (the first three are the seed triplet); the state is the
// beamline state of the straight-line fit (x, y at z = zb) and its chi2/ndof.
template<typename Hits>
__device__ inline void track_features(
  const Hits& hits, uint32_t nh, float xb, float yb, float tx, float ty, float zb, float chi2, float* f)
{
  uint32_t first = 0xFFFFFFFFu, last = 0, pair_mask = 0;
  float sum_r = 0.0f, max_r = 0.0f;
  float phi0 = 0.0f, phi1 = 0.0f, phi2 = 0.0f;
  const auto add_hit = [&](uint32_t h) {
    float x, y, z;
    uint32_t id;
    hits(h, &x, &y, &z, &id);
    const uint32_t m = lhcbid_module(id);
    first = m < first ? m : first;
    last = m > last ? m : last;
    pair_mask |= 1u << ((m >> 1) & 31u);
    const float dz = z - zb;
    const float dx = x - (xb + tx * dz);
    const float dy = y - (yb + ty * dz);
    const float r = sqrtf(dx * dx + dy * dy);
    sum_r += r;
    max_r = r > max_r ? r : max_r;
  };
  if (nh <= VELO_MAX_TRACK_SIZE) {
    // unrolled: a caller's per-hit arrays keep constant indices (registers)
#pragma clang loop unroll(full)
    for (uint32_t h = 0; h < VELO_MAX_TRACK_SIZE; ++h) {
      if (h >= nh) break;
      add_hit(h);
    }
  }
  else {
    for (uint32_t h = 0; h < nh; ++h) add_hit(h);
  }
  // phi of the seed hits, each read again (no per-hit index into locals)
  const auto hit_phi = [&](uint32_t h) {
    float x, y, z;
    uint32_t id;
    hits(h, &x, &y, &z, &id);
    return atan2_poly(y, x);
  };
  if (nh > 0) phi0 = hit_phi(0);
  if (nh > 1) phi1 = hit_phi(1);
  if (nh > 2) phi2 = hit_phi(2);
  if (nh == 0) first = 0;
  // largest pairwise |dphi| (wrapped to [0, pi]) among the seed hits
  const auto wrapped = [](float a, float b) {
    const float d = fabsf(a - b);
    return d > kGhostPi ? 2.0f * kGhostPi - d : d;
  };
  float spread = 0.0f;
  if (nh >= 2) spread = wrapped(phi0, phi1);
  if (nh >= 3) {
    const float d02 = wrapped(phi0, phi2), d12 = wrapped(phi1, phi2);
    spread = d02 > spread ? d02 : spread;
    spread = d12 > spread ? d12 : spread;
  }
  const uint32_t span = nh ? (last >> 1) - (first >> 1) + 1u : 0u;
  f[0] = (float) nh;
  f[1] = (float) first;
  f[2] = (float) last;
  f[3] = tx;
  f[4] = ty;
  f[5] = chi2;
  f[6] = nh ? sum_r / (float) nh : 0.0f;
  f[7] = max_r;
  f[8] = (float) (span - (uint32_t) __builtin_popcount(pair_mask));
  f[9] = spread;
}

// Logit of p(ghost) for normalised inputs x; w holds the parameters as FP32
// (widened from the model's FP16), accumulation in FP32 in index order, one
// fused multiply-add per weight (the order and roundings of ghost_logit_model).
__device__ __forceinline__ float ghost_logit(const float* w, const float* x, float* h1, float* h2)
{
#pragma unroll
  for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) {
    float s = w[kGhostB1 + j];
#pragma unroll
    for (uint32_t i = 0; i < VELO_GHOST_FEATURES; ++i) s = fmaf(w[kGhostW1 + j * VELO_GHOST_FEATURES + i], x[i], s);
    h1[j] = s > 0.0f ? s : 0.0f;
  }
#pragma unroll
  for (uint32_t k = 0; k < VELO_GHOST_HIDDEN; ++k) {
    float s = w[kGhostB2 + k];
#pragma unroll
    for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) s = fmaf(w[kGhostW2 + k * VELO_GHOST_HIDDEN + j], h1[j], s);
    h2[k] = s > 0.0f ? s : 0.0f;
  }
  float z = w[kGhostB3];
#pragma unroll
  for (uint32_t k = 0; k < VELO_GHOST_HIDDEN; ++k) z = fmaf(w[kGhostW3 + k], h2[k], z);
  return z;
}

// The model's FP16 parameters as 32-bit pairs in the constant address space:
// the pointer is the same for every lane, so the compiler fetches the pairs
// with scalar loads into SGPRs, and each multiply-add becomes one
// v_fma_mix_f32 (the FP16 weight widened exactly, FP32 activation and
// accumulator, one rounding) with no per-lane copy of the weights.  Valid
// for kernels that do not write the model.
typedef const __attribute__((address_space(4))) uint32_t* ghost_pairs_t;
typedef const __attribute__((address_space(4))) float* ghost_floats_t;

// The address passes through an empty asm statement, so a loop over tracks
// cannot hoist the weights out of its body into hundreds of registers: each
// evaluation streams them through SGPRs again (from the scalar cache).
__device__ __forceinline__ ghost_pairs_t ghost_pairs(const velo_ghost_model* m)
{
  uint64_t addr = (uint64_t) (uintptr_t) m->weights;
  asm volatile("" : "+s"(addr));
  return (ghost_pairs_t) (uintptr_t) addr;
}

// Parameter p of the pair table (FP16 widened to FP32).
__device__ __forceinline__ float ghost_param(ghost_pairs_t w, uint32_t p)
{
  const uint32_t pair = w[p >> 1];
  const uint16_t bits = (p & 1u) ? (uint16_t) (pair >> 16) : (uint16_t) (pair & 0xFFFFu);
  return __half2float(__ushort_as_half(bits));
}

// Normalised inputs (f - mean) * scale of one track's features, in place.
__device__ __forceinline__ void ghost_normalise(const velo_ghost_model* m, float* f)
{
  const ghost_floats_t mean = (ghost_floats_t) (const void*) m->mean;
  const ghost_floats_t scale = (ghost_floats_t) (const void*) m->scale;
#pragma unroll
  for (uint32_t k = 0; k < VELO_GHOST_FEATURES; ++k) f[k] = (f[k] - mean[k]) * scale[k];
}

// The logit of ghost_logit (same operations in the same order, so the same
// result) evaluated from the model in device memory.  Each second-layer
// unit's ReLU output is added to the logit as soon as it is formed, so no
// per-lane array is indexed at run time.
__device__ __forceinline__ float ghost_logit_model(const velo_ghost_model* m, const float* x)
{
  static_assert(kGhostW1 % 2 == 0 && kGhostB1 % 2 == 0 && kGhostW2 % 2 == 0 && VELO_GHOST_FEATURES % 2 == 0 &&
                  VELO_GHOST_HIDDEN % 2 == 0,
                "parameter blocks start on FP16 pairs");
  const ghost_pairs_t w = ghost_pairs(m);
  float h1[VELO_GHOST_HIDDEN];
#pragma unroll
  for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) {
    float s = ghost_param(w, kGhostB1 + j);
#pragma unroll
    for (uint32_t i = 0; i < VELO_GHOST_FEATURES; ++i) s = fmaf(ghost_param(w, kGhostW1 + j * VELO_GHOST_FEATURES + i), x[i], s);
    h1[j] = s > 0.0f ? s : 0.0f;
  }
  float z = ghost_param(w, kGhostB3);
#pragma unroll 2
  for (uint32_t k = 0; k < VELO_GHOST_HIDDEN; ++k) {
    float s = ghost_param(w, kGhostB2 + k);
#pragma unroll
    for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) s = fmaf(ghost_param(w, kGhostW2 + k * VELO_GHOST_HIDDEN + j), h1[j], s);
    const float h2 = s > 0.0f ? s : 0.0f;
    z = fmaf(ghost_param(w, kGhostW3 + k), h2, z);
  }
  return z;
}

__device__ inline float ghost_sigmoid(float z)
{
  if (z >= 0.0f) return 1.0f / (1.0f + expf(-z));
  const float e = expf(z);
  return e / (1.0f + e);
}

// log(1 + exp(z)) without overflow
__device__ inline float ghost_softplus(float z) { return z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z)); }

} // namespace velo
