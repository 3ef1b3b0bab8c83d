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
  for (uint32_t h = 0; h < nh; ++h) {
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
    if (h == 0) phi0 = atan2_poly(y, x);
    else if (h == 1) phi1 = atan2_poly(y, x);
    else if (h == 2) phi2 = atan2_poly(y, x);
  }
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
// (widened from the model's FP16), accumulation in FP32 in index order.
__device__ __forceinline__ float ghost_logit(const float* w, const float* x, float* h1, float* h2)
{
#pragma unroll
  for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) {
    float s = w[kGhostB1 + j];
#pragma unroll
    for (uint32_t i = 0; i < VELO_GHOST_FEATURES; ++i) s += w[kGhostW1 + j * VELO_GHOST_FEATURES + i] * x[i];
    h1[j] = s > 0.0f ? s : 0.0f;
  }
#pragma unroll
  for (uint32_t k = 0; k < VELO_GHOST_HIDDEN; ++k) {
    float s = w[kGhostB2 + k];
#pragma unroll
    for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) s += w[kGhostW2 + k * VELO_GHOST_HIDDEN + j] * h1[j];
    h2[k] = s > 0.0f ? s : 0.0f;
  }
  float z = w[kGhostB3];
#pragma unroll
  for (uint32_t k = 0; k < VELO_GHOST_HIDDEN; ++k) z += w[kGhostW3 + k] * h2[k];
  return z;
}

// The same logit without keeping the second hidden layer: each unit's
// ReLU output is added to the logit as soon as it is formed (identical
// operations in identical order, so identical results), which keeps the
// inference paths free of per-lane arrays indexed at run time.
__device__ __forceinline__ float ghost_logit_infer(const float* w, const float* x)
{
  float h1[VELO_GHOST_HIDDEN];
#pragma unroll
  for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) {
    float s = w[kGhostB1 + j];
#pragma unroll
    for (uint32_t i = 0; i < VELO_GHOST_FEATURES; ++i) s += w[kGhostW1 + j * VELO_GHOST_FEATURES + i] * x[i];
    h1[j] = s > 0.0f ? s : 0.0f;
  }
  float z = w[kGhostB3];
  // one second-layer unit at a time: its 32 weights in flight, not all 1024
#pragma unroll 1
  for (uint32_t k = 0; k < VELO_GHOST_HIDDEN; ++k) {
    float s = w[kGhostB2 + k];
#pragma unroll
    for (uint32_t j = 0; j < VELO_GHOST_HIDDEN; ++j) s += w[kGhostW2 + k * VELO_GHOST_HIDDEN + j] * h1[j];
    const float h2 = s > 0.0f ? s : 0.0f;
    z += w[kGhostW3 + k] * h2;
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
