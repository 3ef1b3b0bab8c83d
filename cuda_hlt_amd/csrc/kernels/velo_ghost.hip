// Learned ghost-track filter for gfx950: track features, inference, and the
// training step (gradient of the weighted cross-entropy + Adam).
//
// The network is tiny (1441 parameters, 1376 multiply-adds per track), so
// every kernel gives one lane to one track; nothing here is worth a
// matrix-core tile.  Inference reads the FP16 weights through scalar loads
// (one mixed-precision FMA per weight); training keeps them in LDS widened
// to FP32, the same values in the same order.
// Training reductions are deterministic: each 64-sample tile writes its own
// partial gradient row, and one pass sums the rows in tile order.
#include <hip/hip_fp16.h>

#include "velo_device.h"
#include "velo_ghost.h"
#include "velo_ghost_device.h"

namespace velo {

constexpr uint32_t kGhostTile = 64; // samples per workgroup of the gradient kernel
constexpr uint32_t kGradRow = VELO_GHOST_PARAMS + 1;

__device__ inline void load_weights(float* w, const velo_ghost_model* m)
{
  for (uint32_t i = threadIdx.x; i < VELO_GHOST_PARAMS; i += blockDim.x)
    w[i] = __half2float(__ushort_as_half(m->weights[i]));
  __syncthreads();
}

// ---- features of consolidated tracks ----------------------------------------
struct ConsolidatedHits {
  const float *x, *y, *z;
  const uint32_t* id;
  uint32_t base;
  __device__ void operator()(uint32_t h, float* hx, float* hy, float* hz, uint32_t* hid) const
  {
    *hx = x[base + h];
    *hy = y[base + h];
    *hz = z[base + h];
    *hid = id[base + h];
  }
};

__global__ void __launch_bounds__(1024) track_features_kernel(velo_track_features_args a)
{
  const uint32_t N = a.n_events;
  for_each_event(a.tuning, N, [&](const uint32_t e) {
    const uint32_t T = (uint32_t) a.dev_atomics_velo[2 * N];
    const uint32_t Ht = a.dev_velo_track_hit_number[T];
    const uint32_t n = (uint32_t) a.dev_atomics_velo[e];
    const uint32_t toff = (uint32_t) a.dev_atomics_velo[N + e];
    const float* cx = reinterpret_cast<const float*>(a.dev_velo_track_hits);
    const float* S = reinterpret_cast<const float*>(a.dev_velo_states);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      const uint32_t t = toff + i;
      const uint32_t ho = a.dev_velo_track_hit_number[t];
      const uint32_t nh = a.dev_velo_track_hit_number[t + 1] - ho;
      ConsolidatedHits hits {cx, cx + Ht, cx + 2 * (size_t) Ht,
                             reinterpret_cast<const uint32_t*>(cx + 3 * (size_t) Ht), ho};
      float f[VELO_GHOST_FEATURES];
      track_features(hits, nh, S[0 * (size_t) T + t], S[1 * (size_t) T + t], S[2 * (size_t) T + t],
                     S[3 * (size_t) T + t], S[11 * (size_t) T + t], S[10 * (size_t) T + t], f);
      for (uint32_t k = 0; k < VELO_GHOST_FEATURES; ++k) a.dev_features[(size_t) t * VELO_GHOST_FEATURES + k] = f[k];
    }
  });
}

// ---- inference -----------------------------------------------------------------
// One lane per track; the weights come through scalar loads (ghost_logit_model).
__global__ void __launch_bounds__(256) ghost_infer_kernel(velo_ghost_infer_args a)
{
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  float x[VELO_GHOST_FEATURES];
  for (uint32_t k = 0; k < VELO_GHOST_FEATURES; ++k) x[k] = a.dev_features[(size_t) s * VELO_GHOST_FEATURES + k];
  ghost_normalise(a.dev_model, x);
  a.dev_score[s] = ghost_sigmoid(ghost_logit_model(a.dev_model, x));
}

// ---- training: gradient of the weighted cross-entropy ----------------------------
__global__ void __launch_bounds__(kGhostTile) ghost_grad_partials_kernel(velo_ghost_grad_args a)
{
  constexpr uint32_t F = VELO_GHOST_FEATURES, H = VELO_GHOST_HIDDEN;
  __shared__ float w[VELO_GHOST_PARAMS];
  __shared__ float s_x[kGhostTile][F + 1];
  __shared__ float s_h1[kGhostTile][H + 1];
  __shared__ float s_h2[kGhostTile][H + 1];
  __shared__ float s_d1[kGhostTile][H + 1];
  __shared__ float s_d2[kGhostTile][H + 1];
  __shared__ float s_dz[kGhostTile];
  __shared__ float s_loss[kGhostTile];
  load_weights(w, a.dev_model);

  const uint32_t l = threadIdx.x;
  const uint32_t s = blockIdx.x * kGhostTile + l;
  const bool valid = s < a.batch_size;
  float x[F];
  float y = 0.0f;
  if (valid) {
    const uint32_t idx = a.dev_batch[s];
    for (uint32_t k = 0; k < F; ++k)
      x[k] = (a.dev_features[(size_t) idx * F + k] - a.dev_model->mean[k]) * a.dev_model->scale[k];
    y = a.dev_labels[idx];
  }
  else {
    for (uint32_t k = 0; k < F; ++k) x[k] = 0.0f;
  }
  float h1[H], h2[H];
  const float z = ghost_logit(w, x, h1, h2);
  const float p = ghost_sigmoid(z);
  const float ws = y > 0.5f ? a.pos_weight : 1.0f;
  const float inv_b = 1.0f / (float) a.batch_size;
  const float dz = valid ? ws * (p - y) * inv_b : 0.0f;
  const float loss = valid ? ws * (y * ghost_softplus(-z) + (1.0f - y) * ghost_softplus(z)) * inv_b : 0.0f;
  float d2[H];
  for (uint32_t k = 0; k < H; ++k) d2[k] = h2[k] > 0.0f ? dz * w[kGhostW3 + k] : 0.0f;
  for (uint32_t j = 0; j < H; ++j) {
    float g = 0.0f;
    for (uint32_t k = 0; k < H; ++k) g += w[kGhostW2 + k * H + j] * d2[k];
    s_d1[l][j] = h1[j] > 0.0f ? g : 0.0f;
    s_h1[l][j] = h1[j];
    s_h2[l][j] = h2[j];
    s_d2[l][j] = d2[j];
  }
  for (uint32_t k = 0; k < F; ++k) s_x[l][k] = x[k];
  s_dz[l] = dz;
  s_loss[l] = loss;
  __syncthreads();

  // parameter p of this tile: sum over the tile's samples in sample order
  float* out = a.dev_partials + (size_t) blockIdx.x * kGradRow;
  for (uint32_t q = l; q < kGradRow; q += kGhostTile) {
    float g = 0.0f;
    if (q < kGhostB1) {
      const uint32_t j = q / F, i = q % F;
      for (uint32_t r = 0; r < kGhostTile; ++r) g += s_d1[r][j] * s_x[r][i];
    }
    else if (q < kGhostW2) {
      for (uint32_t r = 0; r < kGhostTile; ++r) g += s_d1[r][q - kGhostB1];
    }
    else if (q < kGhostB2) {
      const uint32_t k = (q - kGhostW2) / H, j = (q - kGhostW2) % H;
      for (uint32_t r = 0; r < kGhostTile; ++r) g += s_d2[r][k] * s_h1[r][j];
    }
    else if (q < kGhostW3) {
      for (uint32_t r = 0; r < kGhostTile; ++r) g += s_d2[r][q - kGhostB2];
    }
    else if (q < kGhostB3) {
      for (uint32_t r = 0; r < kGhostTile; ++r) g += s_dz[r] * s_h2[r][q - kGhostW3];
    }
    else if (q == kGhostB3) {
      for (uint32_t r = 0; r < kGhostTile; ++r) g += s_dz[r];
    }
    else {
      for (uint32_t r = 0; r < kGhostTile; ++r) g += s_loss[r];
    }
    out[q] = g;
  }
}

__global__ void __launch_bounds__(256) ghost_grad_sum_kernel(const float* partials, uint32_t tiles, float* grad)
{
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kGradRow) return;
  float g = 0.0f;
  for (uint32_t t = 0; t < tiles; ++t) g += partials[(size_t) t * kGradRow + q];
  grad[q] = g;
}

__global__ void __launch_bounds__(256) ghost_adam_kernel(velo_ghost_adam_args a, float bc1, float bc2)
{
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= VELO_GHOST_PARAMS) return;
  const float g = a.dev_grad[q];
  const float m = a.beta1 * a.dev_m[q] + (1.0f - a.beta1) * g;
  const float v = a.beta2 * a.dev_v[q] + (1.0f - a.beta2) * g * g;
  a.dev_m[q] = m;
  a.dev_v[q] = v;
  const float mhat = m / bc1;
  const float vhat = v / bc2;
  const float master = a.dev_master[q] - a.lr * mhat / (sqrtf(vhat) + a.eps);
  a.dev_master[q] = master;
  a.dev_model->weights[q] = __half_as_ushort(__float2half_rn(master));
}

static uint32_t features_block(const velo_tuning& t)
{
  const uint32_t b = t.block_dim ? t.block_dim : 256;
  return (b % kWave || b > 1024) ? 0 : b;
}

} // namespace velo

using namespace velo;

extern "C" {

int velo_track_features_launch(const velo_track_features_args* a, void* stream)
{
  if (!a || !a->dev_atomics_velo || !a->dev_velo_track_hit_number || !a->dev_velo_track_hits ||
      !a->dev_velo_states || !a->dev_features)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  const uint32_t block = features_block(a->tuning);
  if (!block) return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(track_features_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(block), 0,
                     (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

int velo_ghost_infer_launch(const velo_ghost_infer_args* a, void* stream)
{
  if (!a || !a->dev_features || !a->dev_model || !a->dev_score) return (int) hipErrorInvalidValue;
  if (a->n == 0) return 0;
  hipLaunchKernelGGL(ghost_infer_kernel, dim3((a->n + 255) / 256), dim3(256), 0, (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

size_t velo_ghost_partials_size(uint32_t batch_size)
{
  const size_t tiles = (batch_size + kGhostTile - 1) / kGhostTile;
  return (tiles ? tiles : 1) * kGradRow;
}

int velo_ghost_grad_launch(const velo_ghost_grad_args* a, void* stream)
{
  if (!a || !a->dev_features || !a->dev_labels || !a->dev_batch || !a->dev_model || !a->dev_partials ||
      !a->dev_grad || a->batch_size == 0)
    return (int) hipErrorInvalidValue;
  const uint32_t tiles = (a->batch_size + kGhostTile - 1) / kGhostTile;
  hipStream_t s = (hipStream_t) stream;
  hipLaunchKernelGGL(ghost_grad_partials_kernel, dim3(tiles), dim3(kGhostTile), 0, s, *a);
  hipLaunchKernelGGL(ghost_grad_sum_kernel, dim3((kGradRow + 255) / 256), dim3(256), 0, s, a->dev_partials, tiles,
                     a->dev_grad);
  return (int) hipGetLastError();
}

int velo_ghost_adam_launch(const velo_ghost_adam_args* a, void* stream)
{
  if (!a || !a->dev_grad || !a->dev_master || !a->dev_m || !a->dev_v || !a->dev_model || a->step == 0)
    return (int) hipErrorInvalidValue;
  const double bc1 = 1.0 - std::pow((double) a->beta1, (double) a->step);
  const double bc2 = 1.0 - std::pow((double) a->beta2, (double) a->step);
  hipLaunchKernelGGL(ghost_adam_kernel, dim3((VELO_GHOST_PARAMS + 255) / 256), dim3(256), 0, (hipStream_t) stream, *a,
                     (float) bc1, (float) bc2);
  return (int) hipGetLastError();
}

} // extern "C"
