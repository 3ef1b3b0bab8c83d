// velo_kalman_fit (VeloKalmanFilter.cu:6-73, VeloKalmanFilter.cuh:22-91)
// for gfx950: one workgroup per event (or persistent), one lane per track,
// the filter run over the track's consolidated hits in registers.
#include "velo_device.h"

namespace velo {

constexpr float kParamWInverted = 0.000252083f; // VeloDefinitions.cuh:74

// One step of one projection: prediction to zhit, gain, update of (x, tx)
// and their covariance; returns the chi2 contribution.
__device__ inline float kalman_step(
  float z, float zhit, float xhit, float whit, float& x, float& tx, float& cov_xx, float& cov_xtx, float& cov_txtx)
{
  const float dz = zhit - z;
  const float predx = x + dz * tx;
  const float dz_t_covtxtx = dz * cov_txtx;
  const float predcov_xtx = cov_xtx + dz_t_covtxtx;
  const float dx_t_covxtx = dz * cov_xtx;
  const float predcov_xx = cov_xx + 2.0f * dx_t_covxtx + dz * dz_t_covtxtx;
  const float predcov_txtx = cov_txtx;
  const float r_gain = 1.0f / ((1.0f / whit) + predcov_xx);
  const float kx = predcov_xx * r_gain;
  const float ktx = predcov_xtx * r_gain;
  const float r = xhit - predx;
  x = predx + kx * r;
  tx = tx + ktx * r;
  cov_xx = (1.0f - kx) * predcov_xx;
  cov_xtx = (1.0f - kx) * predcov_xtx;
  cov_txtx = predcov_txtx - ktx * predcov_xtx;
  return r * r * r_gain;
}

__global__ void __launch_bounds__(1024) velo_kalman_fit_kernel(velo_kalman_fit_args a)
{
  const uint32_t N = a.n_events;
  for_each_event(a.tuning, N, [&](const uint32_t e) {
    const uint32_t T = (uint32_t) a.dev_atomics_velo[2 * N];
    const uint32_t Ht = a.dev_velo_track_hit_number[T];
    const uint32_t n = (uint32_t) a.dev_atomics_velo[e];
    const uint32_t toff = (uint32_t) a.dev_atomics_velo[N + e];
    const float* hx = reinterpret_cast<const float*>(a.dev_velo_track_hits);
    const float* hy = hx + Ht;
    const float* hz = hy + Ht;
    const float* S = reinterpret_cast<const float*>(a.dev_velo_states);
    float* O = reinterpret_cast<float*>(a.dev_velo_kalman_beamline_states);
    const uint8_t* bwd_in = reinterpret_cast<const uint8_t*>(a.dev_velo_states + 12 * sizeof(float) * (size_t) T);
    uint8_t* bwd_out = reinterpret_cast<uint8_t*>(a.dev_velo_kalman_beamline_states + 12 * sizeof(float) * (size_t) T);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      const size_t t = toff + i;
      const uint32_t h0 = a.dev_velo_track_hit_number[t];
      const int nh = (int) (a.dev_velo_track_hit_number[t + 1] - h0);
      const float* x = hx + h0;
      const float* y = hy + h0;
      const float* z = hz + h0;
      const float beam_tx = S[2 * (size_t) T + t], beam_ty = S[3 * (size_t) T + t];
      const float beam_z = S[11 * (size_t) T + t];
      const bool backward = beam_z > z[0];
      const int direction = backward ? 1 : -1; // upstream fit
      const float slope2 = beam_tx * beam_tx + beam_ty * beam_ty;
      const float noise2 = (float) (1e-8 + 7e-6 * (double) slope2);
      int first = 0, last = nh - 1, step = 1;
      if ((z[last] - z[first]) * (float) direction < 0.0f) {
        first = nh - 1;
        last = 0;
        step = -1;
      }
      float sx = x[first], sy = y[first], sz = z[first], stx = beam_tx, sty = beam_ty;
      float c00 = kParamWInverted, c11 = kParamWInverted, c20 = 0.0f, c31 = 0.0f, c22 = 1.0f, c33 = 1.0f;
      float chi2 = 0.0f;
      for (int h = first + step; h != last + step; h += step) {
        c22 += noise2;
        c33 += noise2;
        const float zh = z[h];
        chi2 += kalman_step(sz, zh, x[h], kParamW, sx, stx, c00, c20, c22);
        chi2 += kalman_step(sz, zh, y[h], kParamW, sy, sty, c11, c31, c33);
        sz = zh;
      }
      c22 += noise2;
      c33 += noise2;
      O[0 * (size_t) T + t] = sx;
      O[1 * (size_t) T + t] = sy;
      O[2 * (size_t) T + t] = stx;
      O[3 * (size_t) T + t] = sty;
      O[4 * (size_t) T + t] = c00;
      O[5 * (size_t) T + t] = c20;
      O[6 * (size_t) T + t] = c22;
      O[7 * (size_t) T + t] = c11;
      O[8 * (size_t) T + t] = c31;
      O[9 * (size_t) T + t] = c33;
      O[10 * (size_t) T + t] = chi2;
      O[11 * (size_t) T + t] = sz;
      bwd_out[t] = bwd_in[t];
    }
  });
}

} // namespace velo

using namespace velo;

extern "C" int velo_kalman_fit_launch(const velo_kalman_fit_args* a, void* stream)
{
  if (!a || !a->dev_atomics_velo || !a->dev_velo_track_hit_number || !a->dev_velo_track_hits ||
      !a->dev_velo_states || !a->dev_velo_kalman_beamline_states)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  const uint32_t block = a->tuning.block_dim ? a->tuning.block_dim : 256;
  if (block % kWave || block > 1024) return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(velo_kalman_fit_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(block), 0,
                     (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}
