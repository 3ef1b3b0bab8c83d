// Float functions of the beamline PV finder, for device and host code.
// Built with -ffp-contract=off, so every operation below rounds once as
// written and the host and device values are identical.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace velo {

// e^x: Cody-Waite reduction x = k ln2 + r with ln2 split so that k * ln2_hi
// is exact for |k| < 2^15, a degree-7 Taylor polynomial of e^r (|r| <= ~ln2/2,
// truncation below 6e-9 relative) in Horner form, then scaling by 2^k.
__host__ __device__ inline float pv_exp(float x)
{
  if (!(x == x)) return x;               // NaN
  if (x > 88.72284f) return INFINITY;    // above ln(FLT_MAX)
  if (x < -104.0f) return 0.0f;          // below half the smallest subnormal
  constexpr float kLog2e = 1.44269504f;
  constexpr float kLn2Hi = 0.693145751953125f;     // 0x3f317200
  constexpr float kLn2Lo = 1.42860677e-06f;
  const float k = rintf(x * kLog2e);
  const float r = (x - k * kLn2Hi) - k * kLn2Lo;
  float p = 1.0f / 5040.0f;
  p = 1.0f / 720.0f + r * p;
  p = 1.0f / 120.0f + r * p;
  p = 1.0f / 24.0f + r * p;
  p = 1.0f / 6.0f + r * p;
  p = 0.5f + r * p;
  p = 1.0f + r * p;
  p = 1.0f + r * p;
  return ldexpf(p, (int) k);
}

// Integral of the approximate unit Gaussian up to x (pv_beamline_histo.cu:3-11):
// a polynomial in xi = x / sqrt(7) on [-sqrt(7), sqrt(7)].
__host__ __device__ inline float pv_gauss_integral(float x)
{
  const float a = sqrtf(7.0f); // sqrt(2 * order + 3), order 2
  const float xi = x / a;
  const float eta = 1.0f - xi * xi;
  return 0.5f + xi * (0.5f + eta * (0.25f + eta * 0.1875f));
}

} // namespace velo
