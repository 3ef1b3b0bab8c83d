// Prefix sums between the VELO steps (PrefixSumHandler.cuh:15-62,
// PrefixSumSingleBlock.cu:130-171) and the track consolidation
// (ConsolidateVelo.cu:6-170) for gfx950.
//
// Exclusive scan of n u32 in three launches: per-tile sums (tile = 256
// threads x 8 elements), one workgroup scanning the tile sums, then each
// tile scanned in LDS and offset.  HBM traffic: 4 B read + 4 B read + 4 B
// written per element.
#include <hip/hip_fp16.h>

#include <mutex>

#include "velo_device.h"
#include "velo_ghost_device.h"

namespace velo {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread over a 256-thread workgroup.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* wave_sums)
{
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t waves = blockDim.x / kWave;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) wave_sums[wave] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < waves; ++w) {
    const uint32_t s = wave_sums[w];
    if (w < wave) before += s;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// Element count of a scan: the capacity, or the count the device holds.
__device__ inline uint32_t scan_count(const uint32_t* count, uint32_t capacity)
{
  if (!count) return capacity;
  const uint32_t c = *count;
  return c < capacity ? c : capacity;
}

__global__ void __launch_bounds__(kScanThreads)
  scan_reduce_kernel(const uint32_t* a, uint32_t* aux, uint32_t capacity, const uint32_t* count)
{
  __shared__ uint32_t wave_sums[kScanThreads / kWave];
  const uint32_t n = scan_count(count, capacity);
  const uint32_t base = blockIdx.x * kScanTile;
  uint32_t s = 0;
  for (uint32_t k = 0; k < kScanItems; ++k) {
    const uint32_t i = base + k * kScanThreads + threadIdx.x;
    if (i < n) s += a[i];
  }
  s = wave_sum(s);
  if (lane_id() == 0) wave_sums[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < kScanThreads / kWave; ++w) t += wave_sums[w];
    aux[blockIdx.x] = t;
  }
}

// Single workgroup: exclusive scan of m elements in place, total to *total_out
// (when non-null).  Processes chunks of 1024 with a running carry.
__global__ void __launch_bounds__(1024)
  scan_single_block_kernel(uint32_t* total_out, const uint32_t* in, uint32_t* out, uint32_t m)
{
  __shared__ uint32_t wave_sums[1024 / kWave];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < m; base += blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < m ? in[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, &tot, wave_sums);
    if (i < m) out[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Scans each tile and adds its offset; the tile holding the last element
// (tile 0 when the scan is empty) also writes the total to a[n].
__global__ void __launch_bounds__(kScanThreads)
  scan_tiles_kernel(uint32_t* a, const uint32_t* aux, uint32_t capacity, const uint32_t* count)
{
  __shared__ uint32_t tile[kScanTile];
  __shared__ uint32_t wave_sums[kScanThreads / kWave];
  const uint32_t n = scan_count(count, capacity);
  const uint32_t last_tile = n ? (n - 1) / kScanTile : 0u;
  if (blockIdx.x > last_tile) return;
  const uint32_t base = blockIdx.x * kScanTile;
  for (uint32_t k = 0; k < kScanItems; ++k) {
    const uint32_t j = k * kScanThreads + threadIdx.x;
    tile[j] = base + j < n ? a[base + j] : 0u;
  }
  __syncthreads();
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (uint32_t k = 0; k < kScanItems; ++k) {
    v[k] = tile[threadIdx.x * kScanItems + k];
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, &tot, wave_sums) + aux[blockIdx.x];
  for (uint32_t k = 0; k < kScanItems; ++k) {
    tile[threadIdx.x * kScanItems + k] = run;
    run += v[k];
  }
  __syncthreads();
  for (uint32_t k = 0; k < kScanItems; ++k) {
    const uint32_t j = k * kScanThreads + threadIdx.x;
    if (base + j < n) a[base + j] = tile[j];
  }
  if (blockIdx.x == last_tile && threadIdx.x == 0) a[n] = aux[blockIdx.x] + tot;
}

__global__ void store_zero_kernel(uint32_t* p) { *p = 0u; }

// K11: one workgroup per event (PrefixSumSingleBlock.cu:153-171).
__global__ void copy_track_hit_number_kernel(velo_copy_velo_track_hit_number_args a)
{
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const int32_t acc = a.dev_atomics_storage[a.n_events + e];
    const int32_t n = a.dev_atomics_storage[e];
    const uint8_t* tracks = static_cast<const uint8_t*>(a.dev_tracks) + (size_t) e * VELO_MAX_TRACKS * VELO_TRACK_HITS_BYTES;
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x)
      a.dev_velo_track_hit_number[acc + i] = *reinterpret_cast<const uint16_t*>(tracks + (size_t) i * VELO_TRACK_HITS_BYTES);
  });
}

// K15: one workgroup per event, one lane per track.  The event's sorted x,
// y, z, id are copied to LDS when they fit (tuning.lds_tile hits) so the
// per-track gathers hit LDS.
constexpr uint32_t kDefaultConsolidateHits = 3072;
constexpr uint32_t kMaxConsolidateHits = 3584;

struct SlotView {
  const float *x, *y, *z;
  const uint32_t* id;
};

// Hits of one track (TrackHits indices) read through a slot view.
template<typename V>
struct TrackHitsOf {
  const V& hv;
  const uint16_t* t;
  __device__ void operator()(uint32_t h, float* x, float* y, float* z, uint32_t* id) const
  {
    const uint32_t hi = t[1 + h];
    *x = hv.x[hi];
    *y = hv.y[hi];
    *z = hv.z[hi];
    *id = hv.id[hi];
  }
};

// Filter: score tracks with the ghost model at a.dev_ghost_model
// write_hits: each lane stores its track's hits (else the flat pass below
// stores them).
template<bool Filter, typename V>
__device__ __forceinline__ void consolidate_event(
  const V& hv, const velo_consolidate_velo_tracks_args& a, uint32_t e, uint32_t n, uint32_t toff,
  uint32_t total_tracks, uint32_t total_hits, bool write_hits)
{
  float* cx = reinterpret_cast<float*>(a.dev_velo_track_hits);
  float* cy = cx + total_hits;
  float* cz = cy + total_hits;
  uint32_t* cid = reinterpret_cast<uint32_t*>(cz + total_hits);
  float* S = reinterpret_cast<float*>(a.dev_velo_states);
  uint8_t* backward = reinterpret_cast<uint8_t*>(a.dev_velo_states + 12 * sizeof(float) * (size_t) total_tracks);
  const uint8_t* tracks =
    static_cast<const uint8_t*>(a.dev_tracks) + (size_t) e * VELO_MAX_TRACKS * VELO_TRACK_HITS_BYTES;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint16_t* t = reinterpret_cast<const uint16_t*>(tracks + (size_t) i * VELO_TRACK_HITS_BYTES);
    // the whole row (count and 26 hit slots) is read at once, so the hit
    // loops below wait for one memory round trip instead of one per hit
    uint32_t hh[VELO_MAX_TRACK_SIZE];
#pragma clang loop unroll(full)
    for (uint32_t h = 0; h < VELO_MAX_TRACK_SIZE; ++h) hh[h] = t[1 + h];
    const uint32_t nh = t[0] < VELO_MAX_TRACK_SIZE ? t[0] : VELO_MAX_TRACK_SIZE;
    const uint32_t ho = a.dev_velo_track_hit_number[toff + i];
    float s0 = 0.0f, sx = 0.0f, sz = 0.0f, sxz = 0.0f, sz2 = 0.0f;
    float u0 = 0.0f, uy = 0.0f, uz = 0.0f, uyz = 0.0f, uz2 = 0.0f;
#pragma clang loop unroll(full)
    for (uint32_t h = 0; h < VELO_MAX_TRACK_SIZE; ++h) {
      if (h >= nh) break;
      const uint32_t hi = hh[h];
      const float x = hv.x[hi], y = hv.y[hi], z = hv.z[hi];
      if (write_hits) {
        cx[ho + h] = x;
        cy[ho + h] = y;
        cz[ho + h] = z;
        cid[ho + h] = hv.id[hi];
      }
      const float wx = kParamW;
      const float wx_t_x = wx * x;
      const float wx_t_z = wx * z;
      s0 += wx;
      sx += wx_t_x;
      sz += wx_t_z;
      sxz += wx_t_x * z;
      sz2 += wx_t_z * z;
      const float wy = kParamW;
      const float wy_t_y = wy * y;
      const float wy_t_z = wy * z;
      u0 += wy;
      uy += wy_t_y;
      uz += wy_t_z;
      uyz += wy_t_y * z;
      uz2 += wy_t_z * z;
    }
    const float dens = 1.0f / (sz2 * s0 - sz * sz);
    const float tx = (sxz * s0 - sx * sz) * dens;
    float x0 = (sx * sz2 - sxz * sz) * dens;
    const float denu = 1.0f / (uz2 * u0 - uz * uz);
    const float ty = (uyz * u0 - uy * uz) * denu;
    float y0 = (uy * uz2 - uyz * uz) * denu;
    const float zb = -(x0 * tx + y0 * ty) / (tx * tx + ty * ty);
    const float z_first = hv.z[hh[0]];
    const bool bwd = zb > z_first;
    const float m00 = s0;
    const float m11 = u0;
    const float m20 = sz - zb * s0;
    const float m31 = uz - zb * u0;
    const float m22 = sz2 - 2.0f * zb * sz + zb * zb * s0;
    const float m33 = uz2 - 2.0f * zb * uz + zb * zb * u0;
    const float den20 = 1.0f / (m22 * m00 - m20 * m20);
    const float den31 = 1.0f / (m33 * m11 - m31 * m31);
    float ch = 0.0f;
    int ndof = -4;
#pragma clang loop unroll(full)
    for (uint32_t h = 0; h < VELO_MAX_TRACK_SIZE; ++h) {
      if (h >= nh) break;
      const uint32_t hi = hh[h];
      const float z = hv.z[hi];
      const float x = x0 + tx * z;
      const float y = y0 + ty * z;
      const float dx = x - hv.x[hi];
      const float dy = y - hv.y[hi];
      ch += dx * dx * kParamW + dy * dy * kParamW;
      ndof += 2;
    }
    const float chi2 = ch / (float) ndof;
    x0 = x0 + tx * zb;
    y0 = y0 + ty * zb;
    const size_t ti = toff + i;
    const size_t T = total_tracks;
    S[0 * T + ti] = x0;
    S[1 * T + ti] = y0;
    S[2 * T + ti] = tx;
    S[3 * T + ti] = ty;
    S[4 * T + ti] = m22 * den20;
    S[5 * T + ti] = -m20 * den20;
    S[6 * T + ti] = m00 * den20;
    S[7 * T + ti] = m33 * den31;
    S[8 * T + ti] = -m31 * den31;
    S[9 * T + ti] = m11 * den31;
    S[10 * T + ti] = chi2;
    S[11 * T + ti] = zb;
    backward[ti] = bwd ? 1 : 0;
    uint8_t accept = 1;
    if constexpr (Filter) {
      const velo_ghost_model* model = static_cast<const velo_ghost_model*>(a.dev_ghost_model);
      float f[VELO_GHOST_FEATURES];
      track_features(TrackHitsOf<V> {hv, t}, nh, x0, y0, tx, ty, zb, chi2, f);
      ghost_normalise(model, f);
      const float p = ghost_sigmoid(ghost_logit_model(model, f));
      accept = p < model->threshold ? 1 : 0;
      if (a.dev_ghost_score) a.dev_ghost_score[ti] = p;
    }
    if (a.dev_accepted_velo_tracks) a.dev_accepted_velo_tracks[ti] = accept;
  }
}

// The consolidated hits of an event's tracks, one output slot per lane:
// consecutive lanes store consecutive slots.  ends[0..n]: the tracks' first
// slots relative to the event's first (ends[n] = its hit count); a lane
// finds its track from the track starts inside its 64 slots (flat_run).
template<typename V>
__device__ inline void store_track_hits(
  const V& hv, const uint8_t* tracks, const uint32_t* ends, uint32_t n, uint32_t ho0, float* cx, float* cy, float* cz,
  uint32_t* cid)
{
  const uint32_t waves = blockDim.x / kWave, wave = threadIdx.x / kWave, lane = lane_id();
  const uint32_t count = ends[n];
  const uint32_t chunks = (count + kWave - 1) / kWave;
  const uint32_t per_wave = (chunks + waves - 1) / waves;
  const uint32_t c0 = wave * per_wave < chunks ? wave * per_wave : chunks;
  const uint32_t c1 = c0 + per_wave < chunks ? c0 + per_wave : chunks;
  if (c0 >= c1) return;
  uint32_t kcur = flat_run_first(ends, n, c0 * kWave);
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t f = c * kWave + lane;
    const FlatRun r = flat_run(ends, n, c * kWave, kcur, &kcur);
    if (f < count) {
      const uint16_t* t = reinterpret_cast<const uint16_t*>(tracks + (size_t) r.k * VELO_TRACK_HITS_BYTES);
      const uint32_t hi = t[1 + (f - r.s)];
      cx[ho0 + f] = hv.x[hi];
      cy[ho0 + f] = hv.y[hi];
      cz[ho0 + f] = hv.z[hi];
      cid[ho0 + f] = hv.id[hi];
    }
  }
}

// Events with up to this many tracks store their hits in the flat pass.
constexpr uint32_t kFlatStoreTracks = 512;

template<bool Filter>
__device__ __forceinline__ void consolidate_body(const velo_consolidate_velo_tracks_args& a)
{
  extern __shared__ uint32_t s_hits[];
  __shared__ uint32_t s_ends[kFlatStoreTracks + 1];
  const uint32_t N = a.n_events;
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t total_tracks = (uint32_t) a.dev_atomics_velo[2 * N];
    const uint32_t total_hits = a.dev_velo_track_hit_number[total_tracks];
    const uint32_t n = (uint32_t) a.dev_atomics_velo[e];
    const uint32_t toff = (uint32_t) a.dev_atomics_velo[N + e];
    const uint32_t total = a.dev_module_cluster_start[VELO_N_MODULES * N];
    const uint32_t off = a.dev_module_cluster_start[e * VELO_N_MODULES];
    const uint32_t ev_hits = a.dev_module_cluster_start[(e + 1) * VELO_N_MODULES] - off;
    SlotView g;
    g.y = reinterpret_cast<const float*>(a.dev_velo_cluster_container + off);
    g.z = reinterpret_cast<const float*>(a.dev_velo_cluster_container + total + off);
    g.id = a.dev_velo_cluster_container + 2 * (size_t) total + off;
    g.x = reinterpret_cast<const float*>(a.dev_velo_cluster_container + 5 * (size_t) total + off);
    const uint32_t cap = a.tuning.lds_tile;
    if (n == 0) return;
    // the tracks' first output slots (flat store pass), when every track
    // has a hit and there are not too many tracks.  The filtering instance
    // stores each track's hits in its own loop: its lanes spend most of the
    // time on the network, and a pass after it would only lengthen the event.
    const uint32_t ho0 = a.dev_velo_track_hit_number[toff];
    bool empty_track = false;
    if (!Filter && n <= kFlatStoreTracks) {
      for (uint32_t i = threadIdx.x; i <= n; i += blockDim.x) {
        const uint32_t hn = a.dev_velo_track_hit_number[toff + i];
        s_ends[i] = hn - ho0;
        if (i < n) empty_track |= a.dev_velo_track_hit_number[toff + i + 1] <= hn;
      }
    }
    const bool staged = ev_hits <= cap;
    if (staged) {
      for (uint32_t i = threadIdx.x; i < ev_hits; i += blockDim.x) {
        s_hits[i] = __float_as_uint(g.x[i]);
        s_hits[cap + i] = __float_as_uint(g.y[i]);
        s_hits[2 * cap + i] = __float_as_uint(g.z[i]);
        s_hits[3 * cap + i] = g.id[i];
      }
    }
    // the workgroup barrier after the staging, for every instance
    const bool any_empty = __syncthreads_or(empty_track) != 0;
    const bool flat = !Filter && n <= kFlatStoreTracks && !any_empty;
    float* cx = reinterpret_cast<float*>(a.dev_velo_track_hits);
    float* cy = cx + total_hits;
    float* cz = cy + total_hits;
    uint32_t* cid = reinterpret_cast<uint32_t*>(cz + total_hits);
    const uint8_t* tracks =
      static_cast<const uint8_t*>(a.dev_tracks) + (size_t) e * VELO_MAX_TRACKS * VELO_TRACK_HITS_BYTES;
    if (staged) {
      SlotView l;
      l.x = reinterpret_cast<const float*>(s_hits);
      l.y = reinterpret_cast<const float*>(s_hits + cap);
      l.z = reinterpret_cast<const float*>(s_hits + 2 * cap);
      l.id = s_hits + 3 * cap;
      consolidate_event<Filter>(l, a, e, n, toff, total_tracks, total_hits, !flat);
      if (flat) store_track_hits(l, tracks, s_ends, n, ho0, cx, cy, cz, cid);
    }
    else {
      consolidate_event<Filter>(g, a, e, n, toff, total_tracks, total_hits, !flat);
      if (flat) store_track_hits(g, tracks, s_ends, n, ho0, cx, cy, cz, cid);
    }
  });
}

// At most 80 VGPRs: six waves per SIMD, as many as the LDS allows for
// 512-thread workgroups.  The filtering instance needs no LDS of its own
// (the network's weights come through scalar loads) and more registers for
// the network: at most 128 VGPRs.
template<bool Filter>
__global__ void consolidate_kernel(velo_consolidate_velo_tracks_args a);

template<>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6)))
consolidate_kernel<false>(velo_consolidate_velo_tracks_args a)
{
  consolidate_body<false>(a);
}

template<>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4)))
consolidate_kernel<true>(velo_consolidate_velo_tracks_args a)
{
  consolidate_body<true>(a);
}

} // namespace velo

using namespace velo;

// Scratch of velo_prefix_sum_noalloc_launch: the per-tile sums of its scans
// (256 KiB on every GPU the library runs on, held until the process ends).
constexpr uint32_t kNoAllocMaxTiles = 65536;
static_assert((uint64_t) kNoAllocMaxTiles * kScanTile == VELO_PREFIX_SUM_NOALLOC_MAX_ELEMENTS,
              "allocation-free scan bound");
static __device__ __align__(256) uint32_t scratch[kNoAllocMaxTiles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device: host code never passes the array's name itself to a kernel.
hipError_t scratch_on_current_device(uint32_t** out)
{
  int dev = 0;
  hipError_t rc = hipGetDevice(&dev);
  if (rc != hipSuccess) return rc;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess) return rc;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}
} // namespace

extern "C" {

int velo_prefix_sum_noalloc_launch(uint32_t* dev_array, uint32_t array_size, const uint32_t* dev_count, void* stream)
{
  if (!dev_array || array_size > VELO_PREFIX_SUM_NOALLOC_MAX_ELEMENTS) return (int) hipErrorInvalidValue;
  uint32_t* aux = nullptr;
  const hipError_t rc = scratch_on_current_device(&aux);
  if (rc != hipSuccess) return (int) rc;
  velo_prefix_sum_args a {};
  a.dev_array = dev_array;
  a.dev_auxiliary_array = aux;
  a.array_size = array_size;
  a.tuning.dev_count = dev_count;
  return velo_prefix_sum_launch(&a, stream);
}

uint32_t velo_prefix_sum_aux_size(uint32_t n)
{
  const uint32_t s = (n + kScanTile - 1) / kScanTile;
  return s > 0 ? s : 1u;
}

int velo_prefix_sum_launch(const velo_prefix_sum_args* a, void* stream)
{
  if (!a || !a->dev_array || !a->dev_auxiliary_array) return (int) hipErrorInvalidValue;
  hipStream_t s = (hipStream_t) stream;
  const uint32_t n = a->array_size;
  if (n == 0) {
    hipLaunchKernelGGL(store_zero_kernel, dim3(1), dim3(1), 0, s, a->dev_array);
    return (int) hipGetLastError();
  }
  const uint32_t tiles = velo_prefix_sum_aux_size(n);
  const uint32_t* count = a->tuning.dev_count;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(tiles), dim3(kScanThreads), 0, s, a->dev_array, a->dev_auxiliary_array,
                     n, count);
  hipLaunchKernelGGL(scan_single_block_kernel, dim3(1), dim3(1024), 0, s, (uint32_t*) nullptr,
                     a->dev_auxiliary_array, a->dev_auxiliary_array, tiles);
  hipLaunchKernelGGL(scan_tiles_kernel, dim3(tiles), dim3(kScanThreads), 0, s, a->dev_array, a->dev_auxiliary_array,
                     n, count);
  return (int) hipGetLastError();
}

int velo_copy_and_prefix_sum_single_block_launch(const velo_copy_and_prefix_sum_single_block_args* a, void* stream)
{
  if (!a || !a->dev_total_sum || (a->array_size && (!a->dev_input_array || !a->dev_output_array)))
    return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(scan_single_block_kernel, dim3(1), dim3(1024), 0, (hipStream_t) stream, a->dev_total_sum,
                     a->dev_input_array, a->dev_output_array, a->array_size);
  return (int) hipGetLastError();
}

int velo_copy_velo_track_hit_number_launch(const velo_copy_velo_track_hit_number_args* a, void* stream)
{
  if (!a || !a->dev_tracks || !a->dev_atomics_storage || !a->dev_velo_track_hit_number)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  const uint32_t block = a->tuning.block_dim ? a->tuning.block_dim : 256;
  hipLaunchKernelGGL(copy_track_hit_number_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(block), 0, (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

int velo_consolidate_velo_tracks_launch(const velo_consolidate_velo_tracks_args* a, void* stream)
{
  if (!a || !a->dev_atomics_velo || !a->dev_tracks || !a->dev_velo_track_hit_number ||
      !a->dev_velo_cluster_container || !a->dev_module_cluster_start || !a->dev_velo_track_hits ||
      !a->dev_velo_states)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  const uint32_t block = a->tuning.block_dim ? a->tuning.block_dim : 256;
  if (block % kWave || block > 1024) return (int) hipErrorInvalidValue;
  velo_consolidate_velo_tracks_args args = *a;
  if (!args.tuning.lds_tile) args.tuning.lds_tile = kDefaultConsolidateHits;
  // the staged hits stay within 56 KB of LDS per workgroup
  if (args.tuning.lds_tile > kMaxConsolidateHits) return (int) hipErrorInvalidValue;
  const dim3 grid(event_grid(args.tuning, args.n_events));
  const size_t shmem = 4 * sizeof(uint32_t) * args.tuning.lds_tile;
  if (args.dev_ghost_model)
    hipLaunchKernelGGL(consolidate_kernel<true>, grid, dim3(block), shmem, (hipStream_t) stream, args);
  else
    hipLaunchKernelGGL(consolidate_kernel<false>, grid, dim3(block), shmem, (hipStream_t) stream, args);
  return (int) hipGetLastError();
}

} // extern "C"
