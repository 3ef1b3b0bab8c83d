// Device-side definitions shared by the gfx950 kernels: constants, bit
// layouts, the phi function, the binary searches and wave64 primitives.
//
// Float expressions are written one operation per statement where the order
// matters; the kernels are compiled with -ffp-contract=off so that each
// operation rounds once, as the CPU oracle does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "velo_hlt.h"

namespace velo {

// VeloDefinitions.cuh:45-77
constexpr float kPhiExtrapolation = 0.062f;
constexpr float kForwardPhiTolerance = 0.052f;
constexpr float kMaxSlope = 0.4f;
constexpr float kTolerance = 0.6f;
constexpr float kMaxScatterSeeding = 0.004f;
constexpr float kMaxScatterForwarding = 0.004f;
constexpr uint32_t kMaxSkippedModules = 1u;
constexpr float kParamW = 3966.94f;
constexpr float kMaxChi2 = 20.0f;
constexpr float kPixelSize = 0.055f;
constexpr float kFltMax = 3.402823466e+38f;

// ClusteringDefinitions.cuh:14-21
constexpr uint32_t kMaskBottom = 0xFFFEFFFFu;
constexpr uint32_t kMaskTop = 0xFFFF7FFFu;
constexpr uint32_t kMaskTopLeft = 0x7FFF7FFFu;
constexpr uint32_t kMaskBottomRight = 0xFFFEFFFEu;
constexpr uint32_t kMaskLtrTopRight = 0x7FFF0000u;
constexpr uint32_t kMaskRtlBottomLeft = 0x0000FFFEu;
constexpr int kMaxClusteringIterations = 12;

constexpr uint32_t kWave = 64;

// ---------------------------------------------------------------------------
// phi: atan2 from a minimax polynomial, atan(t) = t * P(t^2) on [0, 1].
// The same coefficients and operation order are used by the CPU oracle.
// ---------------------------------------------------------------------------
__host__ __device__ inline float atan2_poly(float y, float x)
{
  constexpr float c0 = 1.0f, c1 = -0.3333328366279602f, c2 = 0.1999826729297638f,
                  c3 = -0.14261946082115173f, c4 = 0.10942086577415466f, c5 = -0.0837746188044548f,
                  c6 = 0.05755440890789032f, c7 = -0.03080674819648266f, c8 = 0.010728039778769016f,
                  c9 = -0.001754164695739746f;
  constexpr float kPi = 3.14159274101257324f;
  constexpr float kPi2 = 1.57079637050628662f;
  const float ax = fabsf(x);
  const float ay = fabsf(y);
  const float mx = ax > ay ? ax : ay;
  const float mn = ax > ay ? ay : ax;
  float a;
  if (mx == 0.0f) {
    a = 0.0f;
  }
  else {
    const float t = mn / mx;
    const float s = t * t;
    float p = c9;
    p = p * s; p = p + c8;
    p = p * s; p = p + c7;
    p = p * s; p = p + c6;
    p = p * s; p = p + c5;
    p = p * s; p = p + c4;
    p = p * s; p = p + c3;
    p = p * s; p = p + c2;
    p = p * s; p = p + c1;
    p = p * s; p = p + c0;
    a = t * p;
  }
  if (ay > ax) a = kPi2 - a;
  if (signbit(x)) a = kPi - a;
  if (signbit(y)) a = -a;
  return a;
}

// VeloTools.cu:3-10: odd modules in (-pi, pi], even modules in [0, 2pi).
__host__ __device__ inline float phi_odd(float x, float y) { return atan2_poly(y, x); }
__host__ __device__ inline float phi_even(float x, float y)
{
  const float phi = atan2_poly(y, x);
  const float shift = phi < 0.0f ? 6.28318548202514648f : 0.0f;
  return phi + shift;
}

// VeloTools.cuh:6-26 (array_size 0 still inspects array[0]).  `a(i)`
// returns element i (a pointer or an LDS/global accessor).
template<typename A>
__device__ inline int bs_first_candidate(const A& a, uint32_t size, float value, float margin)
{
  bool found = false;
  int l = 0;
  int r = (int) (size - 1u);
  while (l < r) {
    const int m = (l + r) / 2;
    const float el = a(m);
    const float d = value - el;
    found |= fabsf(d) < margin;
    const float lo = value - margin;
    if (lo > el) l = m + 1;
    else r = m;
  }
  const float d = value - a(l);
  found |= fabsf(d) < margin;
  return found ? l : -1;
}

// VeloTools.cuh:28-45
template<typename A>
__device__ inline int bs_second_candidate(const A& a, uint32_t size, float value, float margin)
{
  int l = 0;
  int r = (int) (size - 1u);
  while (l < r) {
    const int m = (l + r) / 2;
    const float hi = value + margin;
    if (hi > a(m)) l = m + 1;
    else r = m;
  }
  const float d = value - a(l);
  return fabsf(d) < margin ? l + 1 : l;
}

// Element accessor over a plain array, offset by `base`.
struct ArrayAt {
  const float* p;
  __device__ float operator()(int i) const { return p[i]; }
};

// ---------------------------------------------------------------------------
// raw format (ClusteringDefinitions.cu:3-21)
// ---------------------------------------------------------------------------
struct RawEvent {
  uint32_t n_banks;
  const uint32_t* offsets;
  const char* payload;
  __device__ explicit RawEvent(const char* ev)
  {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(ev);
    n_banks = w[0];
    offsets = w + 1;
    payload = ev + 4 * (n_banks + 2);
  }
};

struct RawBank {
  uint32_t sensor_index;
  uint32_t sp_count;
  const uint32_t* sp_word;
  __device__ RawBank(const RawEvent& e, uint32_t b)
  {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(e.payload + e.offsets[b]);
    sensor_index = w[0];
    sp_count = w[1];
    sp_word = w + 2;
  }
};

__host__ __device__ inline uint32_t lhcb_id(uint32_t sensor, uint32_t cx, uint32_t cy)
{
  return 0x80000000u + ((sensor << 18) | ((cx / 256u) << 16) | ((cx % 256u) << 8) | cy);
}

// EstimateInputSize.cu:71-79
__device__ inline uint32_t isolated_cluster_count(uint32_t sp)
{
  const bool p0 = (sp & 0x88u) && !(sp & 0x44u) && (sp & 0x33u);
  const bool p1 = (sp & 0xCCu) && !(sp & 0x22u) && (sp & 0x11u);
  return 1u + (uint32_t) (p0 | p1);
}

// MaskedVeloClustering.cu:4-46
__device__ inline uint32_t current_mask(uint32_t p)
{
  return ((p & kMaskTop) << 1) | ((p & kMaskBottom) >> 1) | ((p & kMaskBottomRight) << 15) |
         ((p & kMaskTopLeft) >> 15) | (p >> 16) | (p >> 17) | (p << 16) | (p << 17);
}
__device__ inline uint32_t mask_ltr(uint32_t p) { return ((p & kMaskLtrTopRight) >> 15) | (p >> 16) | (p >> 17); }
__device__ inline uint32_t mask_rtl(uint32_t p) { return ((p & kMaskRtlBottomLeft) << 15) | (p << 16) | (p << 17); }
__device__ inline uint32_t cluster_current_mask(uint32_t p)
{
  return ((p & kMaskTop) << 1) | ((p & kMaskBottomRight) << 15) | (p << 16) | (p << 17);
}
__device__ inline uint32_t cand_mask_rc(uint32_t p, uint32_t rp)
{
  return ((p & kMaskBottom) >> 1) & (((p & kMaskTopLeft) >> 15) | (rp << 17)) & ((p >> 16) | (rp << 16)) &
         ((p >> 17) | ((rp & kMaskRtlBottomLeft) << 15));
}

// ---------------------------------------------------------------------------
// wave64 primitives
// ---------------------------------------------------------------------------
__device__ inline uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of m in lanes below the current one
__device__ inline uint32_t count_below(uint64_t m)
{
  return __builtin_amdgcn_mbcnt_hi((uint32_t) (m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t) m, 0u));
}

// Exclusive prefix sum across the wave of small non-negative values (< 4).
__device__ inline uint32_t wave_prefix_small(uint32_t v, uint32_t* total)
{
  const uint64_t b0 = ballot(v & 1u);
  const uint64_t b1 = ballot(v & 2u);
  *total = (uint32_t) __popcll(b0) + 2u * (uint32_t) __popcll(b1);
  return count_below(b0) + 2u * count_below(b1);
}

__device__ inline uint32_t wave_inclusive_scan(uint32_t v);

// Sum over the wave (the same value on every lane): the last lane of the
// DPP scan.
__device__ inline uint32_t wave_sum(uint32_t v) { return __builtin_amdgcn_readlane(wave_inclusive_scan(v), 63); }

// Inclusive prefix sum across the wave by DPP (no LDS): shifts by 1, 2, 4
// and 8 lanes within each row of 16, then lane 15 of rows 0 and 2 added to
// rows 1 and 3, then lane 31 added to rows 2 and 3.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v)
{
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xf, 0xf, false); // row_shr:1
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xf, 0xf, false); // row_shr:2
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xf, 0xf, false); // row_shr:4
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xf, 0xf, false); // row_shr:8
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x142, 0xa, 0xf, false); // row_bcast:15
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x143, 0xc, 0xf, false); // row_bcast:31
  return v;
}

// Inclusive maximum across the wave (the same DPP pattern).
__device__ inline uint32_t wave_inclusive_max(uint32_t v)
{
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xf, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x142, 0xa, 0xf, false));
  v = max(v, (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x143, 0xc, 0xf, false));
  return v;
}

// OR over the wave (the same value on every lane)
__device__ inline uint32_t wave_or(uint32_t v)
{
  v |= (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xf, 0xf, false);
  v |= (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xf, 0xf, false);
  v |= (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xf, 0xf, false);
  v |= (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xf, 0xf, false);
  v |= (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x142, 0xa, 0xf, false);
  v |= (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x143, 0xc, 0xf, false);
  return __builtin_amdgcn_readlane(v, 63);
}

// Runs k = 0..nk-1 of a flat list (SPs of a bank, hits of a track): run k
// holds positions [start[k], start[k+1]), start[nk] = the list's length,
// and the starts increase strictly (no empty run).  Lane j of a wave takes
// position f0 + j; flat_run gives its run k and the run's range [s, e)
// from the run starts that fall inside the 64 positions (one LDS read, a
// wave OR of their bits, popcounts; no search).
struct FlatRun {
  uint32_t k, s, e;
};

// kcur (the same on every lane): the run holding position f0; *knext: the
// run holding position f0 + 64.
__device__ inline FlatRun flat_run(const uint32_t* start, uint32_t nk, uint32_t f0, uint32_t kcur, uint32_t* knext)
{
  const uint32_t lane = lane_id();
  const uint32_t kk = kcur + lane;
  // starts increase strictly, so lanes 1.. hold starts above f0 and the
  // ones inside the chunk come first
  const uint32_t sl = kk <= nk ? start[kk] : 0xFFFFFFFFu;
  const uint32_t p = sl - f0;
  const bool inside = lane > 0 && p < kWave;
  const uint32_t lo = inside && p < 32u ? 1u << p : 0u;
  const uint32_t hi = inside && p >= 32u ? 1u << (p - 32u) : 0u;
  const uint64_t bits = ((uint64_t) wave_or(hi) << 32) | (uint64_t) wave_or(lo);
  const uint64_t upto = lane == kWave - 1 ? ~0ull : (2ull << lane) - 1ull;
  const uint64_t below = bits & upto, above = bits & ~upto;
  const uint32_t inner = (uint32_t) __popcll(bits);
  uint32_t next; // first run start past the chunk
  if (inner + 1 < kWave) next = __builtin_amdgcn_readlane(sl, inner + 1);
  else next = kcur + kWave <= nk ? start[kcur + kWave] : 0xFFFFFFFFu;
  FlatRun r;
  r.k = kcur + (uint32_t) __popcll(below);
  r.s = below ? f0 + 63u - (uint32_t) __clzll(below) : __builtin_amdgcn_readlane(sl, 0);
  r.e = above ? f0 + (uint32_t) __ffsll((unsigned long long) above) - 1u : next;
  *knext = kcur + inner + (next == f0 + kWave ? 1u : 0u);
  return r;
}

// The run holding position f (f below the list's length).
__device__ inline uint32_t flat_run_first(const uint32_t* start, uint32_t nk, uint32_t f)
{
  uint32_t c = 0;
  for (uint32_t k0 = 0; k0 < nk; k0 += kWave) {
    const uint32_t k = k0 + lane_id();
    c += (uint32_t) __popcll(ballot(k < nk && start[k] <= f));
  }
  return c - 1u;
}

// Orders this wave's LDS and global accesses (a wave-local barrier).
__device__ inline void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t event_index(const uint32_t* event_list, uint32_t e)
{
  return event_list ? event_list[e] : e;
}

__device__ inline void set_status(uint32_t* status, uint32_t e, uint32_t bits)
{
  if (status && bits) atomicOr(status + e, bits);
}

// Runs body(e) for the events of this workgroup.  One workgroup per event
// (t.grid_dim == 0): e = blockIdx.x.  Persistent (t.grid_dim > 0): the
// workgroup takes positions from t.dev_work_counter[0] until they run out,
// e = t.dev_work_order[k]; the last workgroup to leave zeroes both counter
// words (word 1 counts finished workgroups) so the next launch starts
// from zero.  body may use __syncthreads and LDS: iterations are separated
// by a workgroup barrier.
template<typename F>
__device__ inline void for_each_event(const velo_tuning& t, uint32_t n_events, F&& body)
{
  if (t.grid_dim == 0) {
    body(blockIdx.x);
    return;
  }
  __shared__ uint32_t s_position;
  for (;;) {
    if (threadIdx.x == 0) s_position = atomicAdd(t.dev_work_counter, 1u);
    __syncthreads();
    const uint32_t k = s_position;
    __syncthreads();
    if (k >= n_events) break;
    body(t.dev_work_order ? t.dev_work_order[k] : k);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(t.dev_work_counter + 1, 1u) == gridDim.x - 1) {
      atomicExch(t.dev_work_counter, 0u);
      atomicExch(t.dev_work_counter + 1, 0u);
    }
  }
}

// Grid of a per-event launch: n_events workgroups, or the persistent
// grid_dim (never more workgroups than events).
inline uint32_t event_grid(const velo_tuning& t, uint32_t n_events)
{
  return t.grid_dim ? (t.grid_dim < n_events ? t.grid_dim : n_events) : n_events;
}

// Host-side check of the persistent fields.
inline bool schedule_ok(const velo_tuning& t)
{
  return t.reserved == 0 && (t.grid_dim == 0 || t.dev_work_counter != nullptr);
}

} // namespace velo
