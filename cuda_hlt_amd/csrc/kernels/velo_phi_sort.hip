// calculate_phi_and_sort (CalculatePhiAndSort.cu:6-52) and fill_candidates
// (FillCandidates.cu:8-97) for gfx950.
//
// One workgroup per event.  Phi sort takes the event's cluster slots as one
// flat list (a few per thread, whatever module they belong to): each
// thread keeps its hits' x, y, z, id in registers and their phi keys in
// LDS, then, after one workgroup barrier, ranks each hit among its module's
// keys (stable: ties broken by slot, CalculatePhi.cu:47-53) and writes it
// to its sorted slot.  fill_candidates takes the hits of modules 2-49 as
// one range, 64 per pass.
#include "velo_device.h"

namespace velo {

constexpr uint32_t kSortBlock = 256;
constexpr uint32_t kHoleBits = 0x7FC00000u; // quiet NaN

__device__ inline float module_phi(uint32_t m, float x, float y) { return (m & 1u) ? phi_odd(x, y) : phi_even(x, y); }

// phi as an unsigned key that orders like the float; -0 and +0 share the
// key of +0 (they compare equal), so the key of 0 does not keep the sign.
__device__ inline uint32_t phi_key(float p)
{
  const uint32_t b = __float_as_uint(p + 0.0f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline uint32_t key_phi_bits(uint32_t k) { return (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k; }

// Flat sort: the slots of a group of whole modules are taken as one list,
// kSortRounds per thread.  Pass 1 reads each filled slot's x, y, z, id into
// registers and its phi key into LDS (and fills the holes); after a
// workgroup barrier, pass 2 ranks each hit among its module's keys (ties in
// slot order) and writes it to its sorted slot.  An event whose slots
// exceed kSortRounds x workgroup size is taken in several module groups; a
// single module larger than that is sorted through global memory.
constexpr uint32_t kSortRounds = 2;

// Largest m in [0, 51] with mst[m] <= i (the module holding slot i).
__device__ inline uint32_t module_of_slot(const uint32_t* mst, uint32_t i)
{
  uint32_t lo = 0, hi = VELO_N_MODULES - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (mst[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Module m of more than kSortRounds x workgroup size slots, by wave 0,
// through global memory (phis staged in slot 5, which receives X last).
__device__ inline void sort_module_global(
  uint32_t m, uint32_t s, uint32_t n, uint32_t* X, uint32_t* Y, uint32_t* Z, uint32_t* ID, uint32_t* PHI, uint32_t* TMP,
  uint32_t* perm)
{
  const uint32_t lane = lane_id();
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
    perm[s + pos] = s + i;
    PHI[s + pos] = __float_as_uint(p);
  }
  wave_sync();
  // SortByPhi.cu:21-33 order: X->5, Y->0, Z->1, ID->2
  for (uint32_t p = lane; p < n; p += kWave) TMP[s + p] = X[perm[s + p]];
  wave_sync();
  for (uint32_t p = lane; p < n; p += kWave) X[s + p] = Y[perm[s + p]];
  wave_sync();
  for (uint32_t p = lane; p < n; p += kWave) Y[s + p] = Z[perm[s + p]];
  wave_sync();
  for (uint32_t p = lane; p < n; p += kWave) Z[s + p] = ID[perm[s + p]];
}

__global__ void __launch_bounds__(1024) calculate_phi_and_sort_kernel(velo_calculate_phi_and_sort_args a)
{
  extern __shared__ __attribute__((aligned(16))) uint32_t keys[]; // kSortRounds x workgroup size
  __shared__ uint32_t s_mst[VELO_N_MODULES + 1];
  __shared__ uint32_t s_mnum[VELO_N_MODULES];
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t total = a.dev_module_cluster_start[VELO_N_MODULES * a.n_events];
    uint32_t* X = a.dev_velo_cluster_container;
    uint32_t* Y = X + total;
    uint32_t* Z = Y + total;
    uint32_t* ID = Z + total;
    uint32_t* PHI = ID + total;
    uint32_t* TMP = PHI + total;
    const uint32_t off = a.dev_module_cluster_start[e * VELO_N_MODULES];
    if (threadIdx.x <= VELO_N_MODULES) s_mst[threadIdx.x] = a.dev_module_cluster_start[e * VELO_N_MODULES + threadIdx.x] - off;
    if (threadIdx.x < VELO_N_MODULES) s_mnum[threadIdx.x] = a.dev_module_cluster_num[e * VELO_N_MODULES + threadIdx.x];
    __syncthreads();
    const uint32_t cap = kSortRounds * blockDim.x;
    const uint32_t lane = lane_id();
    for (uint32_t m0 = 0; m0 < VELO_N_MODULES;) {
      // the group: modules [m0, m1) whose slots fit cap
      const uint32_t lim = s_mst[m0] + cap;
      const uint32_t m1 =
        (uint32_t) __popcll(ballot(lane <= VELO_N_MODULES && s_mst[lane <= VELO_N_MODULES ? lane : 0] <= lim)) - 1u;
      if (m1 == m0) {
        // one module of more than cap slots (its holes are filled too)
        if (threadIdx.x < kWave) {
          const uint32_t s = off + s_mst[m0], n = s_mnum[m0];
          for (uint32_t p = s + n + lane; p < off + s_mst[m0 + 1]; p += kWave) {
            X[p] = kHoleBits;
            Y[p] = kHoleBits;
            Z[p] = kHoleBits;
            PHI[p] = kHoleBits;
            TMP[p] = kHoleBits;
          }
          sort_module_global(m0, s, n, X, Y, Z, ID, PHI, TMP, a.dev_hit_permutation);
        }
        __syncthreads();
        ++m0;
        continue;
      }
      const uint32_t base = s_mst[m0], end = s_mst[m1];
      uint32_t vx[kSortRounds], vy[kSortRounds], vz[kSortRounds], vid[kSortRounds], vm[kSortRounds];
#pragma unroll
      for (uint32_t q = 0; q < kSortRounds; ++q) {
        const uint32_t i = base + q * blockDim.x + threadIdx.x;
        vm[q] = 0xFFFFFFFFu;
        if (i < end) {
          const uint32_t m = module_of_slot(s_mst, i);
          if (i - s_mst[m] < s_mnum[m]) {
            vx[q] = X[off + i];
            vy[q] = Y[off + i];
            vz[q] = Z[off + i];
            vid[q] = ID[off + i];
            keys[i - base] = phi_key(module_phi(m, __uint_as_float(vx[q]), __uint_as_float(vy[q])));
            vm[q] = m;
          }
          else {
            // Slots estimated but not filled ("holes") read as NaN in every
            // sorted slot, so no search can ever select them (the reference
            // leaves them uninitialised).
            X[off + i] = kHoleBits;
            Y[off + i] = kHoleBits;
            Z[off + i] = kHoleBits;
            PHI[off + i] = kHoleBits;
            TMP[off + i] = kHoleBits;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (uint32_t q = 0; q < kSortRounds; ++q) {
        if (vm[q] != 0xFFFFFFFFu) {
          const uint32_t i = base + q * blockDim.x + threadIdx.x;
          const uint32_t m = vm[q];
          const uint32_t s = s_mst[m] - base, n = s_mnum[m];
          const uint32_t k = keys[i - base];
          // rank in the module, ties in slot order (CalculatePhi.cu:47-53)
          uint32_t below = 0, equal = 0;
          for (uint32_t j = s; j < s + n; ++j) {
            const uint32_t kj = keys[j];
            below += kj < k ? 1u : 0u;
            equal += kj == k ? 1u : 0u;
          }
          if (equal > 1) {
            for (uint32_t j = s; j < i - base; ++j) below += keys[j] == k ? 1u : 0u;
          }
          const uint32_t r = off + base + s + below;
          uint32_t bits = key_phi_bits(k);
          if (bits == 0u) bits = __float_as_uint(module_phi(m, __uint_as_float(vx[q]), __uint_as_float(vy[q])));
          a.dev_hit_permutation[r] = off + i;
          PHI[r] = bits;
          // SortByPhi.cu:21-33 order: X->5, Y->0, Z->1, ID->2
          TMP[r] = vx[q];
          X[r] = vy[q];
          Y[r] = vz[q];
          Z[r] = vid[q];
        }
      }
      __syncthreads();
      m0 = m1;
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
  const velo_calculate_phi_and_sort_args args = *a;
  const size_t shmem = (size_t) kSortRounds * block * sizeof(uint32_t);
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
