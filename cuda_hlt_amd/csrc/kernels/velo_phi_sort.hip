// calculate_phi_and_sort (CalculatePhiAndSort.cu:6-52) and fill_candidates
// (FillCandidates.cu:8-97) for gfx950.
//
// One workgroup per event, wave w owns modules w, w+W, ...  Phi sort: when
// the event fits (tuning.lds_tile hits, default 1024) the workgroup first
// copies its x, y, z, id slots to LDS in one coalesced pass; the wave then
// computes the module's phis into its LDS slice, ranks every hit against
// all others (stable: ties broken by slot, CalculatePhi.cu:47-53) and
// writes x, y, z, id from LDS to their sorted positions.  Larger events
// gather by the permutation from global memory into registers instead, and
// modules larger than the register budget (512 hits; the reference caps a
// module at 500) fall back to the reference's slot-by-slot permutation
// through global memory.
#include "velo_device.h"

namespace velo {

constexpr uint32_t kSortCap = 512;
constexpr uint32_t kSortChunks = kSortCap / kWave;
constexpr uint32_t kSortBlock = 256;
constexpr uint32_t kHoleBits = 0x7FC00000u; // quiet NaN
constexpr uint32_t kDefaultSortStage = 1024;  // hits of an event staged in LDS
constexpr size_t kMaxSortLds = 65536;

__device__ inline float module_phi(uint32_t m, float x, float y) { return (m & 1u) ? phi_odd(x, y) : phi_even(x, y); }

// phi as an unsigned key that orders like the float; -0 and +0 share the
// key of +0 (they compare equal), so the key of 0 does not keep the sign.
__device__ inline uint32_t phi_key(float p)
{
  const uint32_t b = __float_as_uint(p + 0.0f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline uint32_t key_phi_bits(uint32_t k) { return (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k; }

// Sorted position of each of a module's n hits, by phi with ties in slot
// order (CalculatePhi.cu:47-53): perm[pos] = i.  keys[i] is phi_key of hit
// i.  Counting the strictly smaller keys gives every hit its position when
// no two phis are equal; equal phis share a count, which leaves a position
// unwritten, and the module is then ranked again with the slot tie-break.
__device__ inline uint32_t keys_below(uint4 q, uint32_t k)
{
  return (uint32_t) (q.x < k) + (uint32_t) (q.y < k) + (uint32_t) (q.z < k) + (uint32_t) (q.w < k);
}

__device__ inline void rank_module(const uint32_t* keys, uint16_t* perm, uint32_t n_any, uint32_t lane)
{
  const uint32_t n = __builtin_amdgcn_readfirstlane(n_any); // the same on every lane
  const uint4* k4 = reinterpret_cast<const uint4*>(keys);
  for (uint32_t p = lane; p < n; p += kWave) perm[p] = 0xFFFFu;
  wave_sync();
#pragma unroll 1
  for (uint32_t i0 = 0; i0 < n; i0 += kWave) {
    const uint32_t i = i0 + lane;
    const uint32_t ki = i < n ? keys[i] : 0u;
    uint32_t r = 0, j = 0;
    // four 16-byte reads in flight per step
    for (; j + 16 <= n; j += 16) {
      const uint4 q0 = k4[j / 4], q1 = k4[j / 4 + 1], q2 = k4[j / 4 + 2], q3 = k4[j / 4 + 3];
      r += keys_below(q0, ki) + keys_below(q1, ki) + keys_below(q2, ki) + keys_below(q3, ki);
    }
    for (; j + 4 <= n; j += 4) r += keys_below(k4[j / 4], ki);
    for (; j < n; ++j) r += (uint32_t) (keys[j] < ki);
    if (i < n) perm[r] = (uint16_t) i;
  }
  wave_sync();
  bool clash = false;
  for (uint32_t p = lane; p < n; p += kWave) clash |= perm[p] == 0xFFFFu;
  if (ballot(clash) == 0) return;
  for (uint32_t i = lane; i < n; i += kWave) {
    const uint32_t ki = keys[i];
    uint32_t r = 0;
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t kj = keys[j];
      r += (kj < ki) || (kj == ki && j < i);
    }
    perm[r] = (uint16_t) i;
  }
  wave_sync();
}

// At most 96 VGPRs: five waves per SIMD, as many as the LDS of a
// 256-thread workgroup allows.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5))) calculate_phi_and_sort_kernel(velo_calculate_phi_and_sort_args a)
{
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t waves = blockDim.x / kWave;
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint32_t total = a.dev_module_cluster_start[VELO_N_MODULES * a.n_events];
    uint32_t* X = a.dev_velo_cluster_container;
    uint32_t* Y = X + total;
    uint32_t* Z = Y + total;
    uint32_t* ID = Z + total;
    uint32_t* PHI = ID + total;
    uint32_t* TMP = PHI + total;
    uint32_t* keys = lds + wave * kSortCap;
    uint16_t* perm = reinterpret_cast<uint16_t*>(lds + waves * kSortCap) + wave * kSortCap;
    // the event's x, y, z, id (slots 0-3), when they fit
    uint32_t* ev = lds + waves * kSortCap + waves * kSortCap / 2;
    const uint32_t cap = a.tuning.lds_tile;
    const uint32_t off = a.dev_module_cluster_start[e * VELO_N_MODULES];
    const uint32_t ev_hits = a.dev_module_cluster_start[(e + 1) * VELO_N_MODULES] - off;
    const bool staged = ev_hits <= cap;
    if (staged) {
      for (uint32_t i = threadIdx.x; i < ev_hits; i += blockDim.x) {
        ev[i] = X[off + i];
        ev[cap + i] = Y[off + i];
        ev[2 * cap + i] = Z[off + i];
        ev[3 * cap + i] = ID[off + i];
      }
      __syncthreads();
    }

    for (uint32_t m = wave; m < VELO_N_MODULES; m += waves) {
      const uint32_t s = a.dev_module_cluster_start[e * VELO_N_MODULES + m];
      const uint32_t n = a.dev_module_cluster_num[e * VELO_N_MODULES + m];
      // Slots estimated but not filled ("holes") read as NaN in every sorted
      // slot, so no search can ever select them (the reference leaves them
      // uninitialised).
      const uint32_t hole_end = a.dev_module_cluster_start[e * VELO_N_MODULES + m + 1];
      for (uint32_t p = s + n + lane; p < hole_end; p += kWave) {
        X[p] = kHoleBits;
        Y[p] = kHoleBits;
        Z[p] = kHoleBits;
        PHI[p] = kHoleBits;
        TMP[p] = kHoleBits;
      }
      if (n == 0) continue;
      if (staged && n <= kSortCap) {
        const uint32_t r = s - off;
        for (uint32_t i = lane; i < n; i += kWave)
          keys[i] = phi_key(module_phi(m, __uint_as_float(ev[r + i]), __uint_as_float(ev[cap + r + i])));
        wave_sync();
        rank_module(keys, perm, n, lane);
        for (uint32_t p = lane; p < n; p += kWave) {
          const uint32_t i = perm[p];
          uint32_t bits = key_phi_bits(keys[i]);
          if (bits == 0u) bits = __float_as_uint(module_phi(m, __uint_as_float(ev[r + i]), __uint_as_float(ev[cap + r + i])));
          a.dev_hit_permutation[s + p] = s + i;
          PHI[s + p] = bits;
        }
        // SortByPhi.cu:21-33 order: X->5, Y->0, Z->1, ID->2; every read is from LDS
        for (uint32_t p = lane; p < n; p += kWave) {
          const uint32_t q = r + perm[p];
          TMP[s + p] = ev[q];
          X[s + p] = ev[cap + q];
          Y[s + p] = ev[2 * cap + q];
          Z[s + p] = ev[3 * cap + q];
        }
      }
      else if (n <= kSortCap) {
        // each lane keeps its hits' x, y, z, id and phi in registers (one
        // coalesced read) and scatters them to their sorted slots
        uint32_t vx[kSortChunks], vy[kSortChunks], vz[kSortChunks], vid[kSortChunks];
        float vp[kSortChunks];
  #pragma unroll
        for (uint32_t c = 0; c < kSortChunks; ++c) {
          const uint32_t i = c * kWave + lane;
          if (i < n) {
            vx[c] = X[s + i];
            vy[c] = Y[s + i];
            vz[c] = Z[s + i];
            vid[c] = ID[s + i];
            vp[c] = module_phi(m, __uint_as_float(vx[c]), __uint_as_float(vy[c]));
            keys[i] = phi_key(vp[c]);
          }
        }
        wave_sync();
        rank_module(keys, perm, n, lane);
        // the keys' slots take the inverse permutation: keys[i] = sorted position of hit i
        for (uint32_t p = lane; p < n; p += kWave) keys[perm[p]] = p;
        wave_sync();
  #pragma unroll
        for (uint32_t c = 0; c < kSortChunks; ++c) {
          const uint32_t i = c * kWave + lane;
          if (i < n) {
            const uint32_t r = s + keys[i];
            a.dev_hit_permutation[r] = s + i;
            PHI[r] = __float_as_uint(vp[c]);
            TMP[r] = vx[c];
            X[r] = vy[c];
            Y[r] = vz[c];
            Z[r] = vid[c];
          }
        }
      }
      else {
        // phis staged in slot 5 (it receives X only at the end)
        for (uint32_t i = lane; i < n; i += kWave)
          TMP[s + i] = __float_as_uint(module_phi(m, __uint_as_float(X[s + i]), __uint_as_float(Y[s + i])));
        wave_sync();
        for (uint32_t i = lane; i < n; i += kWave) {
          const float p = __uint_as_float(TMP[s + i]);
          uint32_t pos = 0;
          for (uint32_t j = 0; j < n; ++j) {
            const float q = __uint_as_float(TMP[s + j]);
            pos += (p > q) || (p == q && i > j);
          }
          a.dev_hit_permutation[s + pos] = s + i;
          PHI[s + pos] = __float_as_uint(p);
        }
        wave_sync();
        // SortByPhi.cu:21-33 order: X->5, Y->0, Z->1, ID->2
        for (uint32_t p = lane; p < n; p += kWave) TMP[s + p] = X[a.dev_hit_permutation[s + p]];
        wave_sync();
        for (uint32_t p = lane; p < n; p += kWave) X[s + p] = Y[a.dev_hit_permutation[s + p]];
        wave_sync();
        for (uint32_t p = lane; p < n; p += kWave) Y[s + p] = Z[a.dev_hit_permutation[s + p]];
        wave_sync();
        for (uint32_t p = lane; p < n; p += kWave) Z[s + p] = ID[a.dev_hit_permutation[s + p]];
      }
      wave_sync();
    }
  });
}

// fill_candidates: the event's phi slot is copied to LDS when it fits
// (tuning.lds_tile floats); the waves take 64-hit passes over modules 2-49.
constexpr uint32_t kDefaultPhiTile = 4096;

// The hits of modules 2-49 are taken as one range, 64 per pass, whatever
// module they belong to (a module holds 6 to 58 hits on average, so passes
// per module would leave lanes idle); each lane finds its module from the
// event's module table.  mstart: 53 event-relative module starts, mnum: 52
// hit counts.
template<typename A>
__device__ __forceinline__ void fill_event(
  const A& phis, const uint32_t* mstart, const uint32_t* mnum, int16_t* h0c, int16_t* h2c)
{
  const uint32_t waves = blockDim.x / kWave;
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  const uint32_t lo = mstart[2], hi = mstart[50];
  uint32_t mb = 2; // module of the pass's first hit (passes of a wave move forward)
  for (uint32_t base = lo + wave * kWave; base < hi; base += waves * kWave) {
    while (mb < 49 && mstart[mb + 1] <= base) ++mb;
    const uint32_t h1 = base + lane;
    if (h1 >= hi) continue;
    uint32_t m = mb;
    while (m < 49 && mstart[m + 1] <= h1) ++m;
    if (h1 - mstart[m] >= mnum[m]) continue; // an unfilled slot
    const uint32_t m0s = mstart[m + 2], m0n = mnum[m + 2];
    const uint32_t m2s = mstart[m - 2], m2n = mnum[m - 2];
    const float phi = phis(h1);
    int f0 = -1, l0 = -1, f2 = -1, l2 = -1;
    if (m0n > 0) {
      const auto w0 = [&](int i) { return phis(m0s + i); };
      f0 = bs_first_candidate(w0, m0n, phi, kPhiExtrapolation);
      if (f0 != -1) {
        const auto w0f = [&](int i) { return phis(m0s + f0 + i); };
        l0 = bs_second_candidate(w0f, m0n - (uint32_t) f0, phi, kPhiExtrapolation);
        f0 += (int) m0s;
        l0 = l0 == 0 ? f0 + 1 : f0 + l0;
      }
    }
    if (m2n > 0) {
      const auto w2 = [&](int i) { return phis(m2s + i); };
      f2 = bs_first_candidate(w2, m2n, phi, kPhiExtrapolation);
      if (f2 != -1) {
        const auto w2f = [&](int i) { return phis(m2s + f2 + i); };
        l2 = bs_second_candidate(w2f, m2n - (uint32_t) f2, phi, kPhiExtrapolation);
        f2 += (int) m2s;
        l2 = l2 == 0 ? f2 + 1 : f2 + l2;
      }
    }
    h0c[2 * h1] = (int16_t) f0;
    h0c[2 * h1 + 1] = (int16_t) l0;
    h2c[2 * h1] = (int16_t) f2;
    h2c[2 * h1 + 1] = (int16_t) l2;
  }
}

__global__ void __launch_bounds__(1024) fill_candidates_kernel(velo_fill_candidates_args a)
{
  extern __shared__ float s_phi[];
  __shared__ uint32_t s_mstart[VELO_N_MODULES + 1];
  __shared__ uint32_t s_mnum[VELO_N_MODULES];
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t total = a.dev_module_cluster_start[VELO_N_MODULES * a.n_events];
    const uint32_t* starts = a.dev_module_cluster_start + e * VELO_N_MODULES;
    const uint32_t* nums = a.dev_module_cluster_num + e * VELO_N_MODULES;
    const uint32_t off = starts[0];
    const uint32_t ev_hits = starts[VELO_N_MODULES] - off;
    const float* phis = reinterpret_cast<const float*>(a.dev_velo_cluster_container + 4 * (size_t) total + off);
    int16_t* h0c = a.dev_h0_candidates + 2 * (size_t) off;
    int16_t* h2c = a.dev_h2_candidates + 2 * (size_t) off;
    if (threadIdx.x <= VELO_N_MODULES) s_mstart[threadIdx.x] = starts[threadIdx.x] - off;
    if (threadIdx.x < VELO_N_MODULES) s_mnum[threadIdx.x] = nums[threadIdx.x];
    if (ev_hits <= a.tuning.lds_tile) {
      for (uint32_t i = threadIdx.x; i < ev_hits; i += blockDim.x) s_phi[i] = phis[i];
      __syncthreads();
      const auto acc = [&](int i) { return s_phi[i]; };
      fill_event(acc, s_mstart, s_mnum, h0c, h2c);
    }
    else {
      __syncthreads();
      fill_event(ArrayAt {phis}, s_mstart, s_mnum, h0c, h2c);
    }
  });
}

} // namespace velo

using namespace velo;

extern "C" {

int velo_calculate_phi_and_sort_launch(const velo_calculate_phi_and_sort_args* a, void* stream)
{
  if (!a || !a->dev_module_cluster_start || !a->dev_module_cluster_num || !a->dev_velo_cluster_container ||
      !a->dev_hit_permutation)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  const uint32_t block = a->tuning.block_dim ? a->tuning.block_dim : kSortBlock;
  if (block % kWave || block > 1024) return (int) hipErrorInvalidValue;
  velo_calculate_phi_and_sort_args args = *a;
  const size_t sort_lds = (size_t) (block / kWave) * kSortCap * (sizeof(float) + sizeof(uint16_t));
  // event staging: as requested, within 64 KB of LDS per workgroup
  const size_t room = sort_lds < kMaxSortLds ? (kMaxSortLds - sort_lds) / (4 * sizeof(uint32_t)) : 0;
  const size_t want = args.tuning.lds_tile ? args.tuning.lds_tile : kDefaultSortStage;
  args.tuning.lds_tile = (uint32_t) (want < room ? want : room);
  const size_t shmem = sort_lds + (size_t) args.tuning.lds_tile * 4 * sizeof(uint32_t);
  hipLaunchKernelGGL(calculate_phi_and_sort_kernel, dim3(event_grid(args.tuning, args.n_events)), dim3(block), shmem,
                     (hipStream_t) stream, args);
  return (int) hipGetLastError();
}

int velo_fill_candidates_launch(const velo_fill_candidates_args* a, void* stream)
{
  if (!a || !a->dev_velo_cluster_container || !a->dev_module_cluster_start || !a->dev_module_cluster_num ||
      !a->dev_h0_candidates || !a->dev_h2_candidates)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  const uint32_t block = a->tuning.block_dim ? a->tuning.block_dim : 256;
  if (block % kWave || block > 1024) return (int) hipErrorInvalidValue;
  velo_fill_candidates_args args = *a;
  if (!args.tuning.lds_tile) args.tuning.lds_tile = kDefaultPhiTile;
  hipLaunchKernelGGL(fill_candidates_kernel, dim3(event_grid(args.tuning, args.n_events)), dim3(block), args.tuning.lds_tile * sizeof(float),
                     (hipStream_t) stream, args);
  return (int) hipGetLastError();
}

} // extern "C"
