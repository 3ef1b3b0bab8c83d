// global_event_cut (GlobalEventCut.cu:3-55) for gfx950.
//
// Two launches.  gec_flags_kernel gives each event one wave: lanes stride
// over the SciFi bank offsets (bank sizes) and the UT banks (hit counts),
// the wave sums them, and lane 0 stores the keep flag in the event's
// dev_event_list word.  gec_compact_kernel, one workgroup, then turns the
// flags into the list of kept events in increasing order, a tile of 1024
// flags at a time (ballot + per-wave counts in LDS), in place: a kept
// event's position is never above its own index, and every flag of a tile
// is read before any position of it is written.
#include "velo_device.h"

namespace velo {

constexpr uint32_t kCompactBlock = 1024;

// Bank count of the raw event at byte offset `off` of a buffer of `size`
// bytes, or -1 when its header or offset table reaches past the end.
__device__ inline int64_t checked_banks(const char* raw, size_t size, uint32_t off)
{
  if ((size_t) off + 8 > size || (off & 3u)) return -1;
  const uint32_t n = *reinterpret_cast<const uint32_t*>(raw + off);
  if ((size_t) off + 4 * ((size_t) n + 2) > size) return -1;
  return n;
}

__global__ void __launch_bounds__(1024) gec_flags_kernel(velo_global_event_cut_args a)
{
  const uint32_t lane = threadIdx.x % kWave;
  const uint32_t e = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (e >= a.n_events) return;

  // SciFi: bank bytes without the source word (GlobalEventCut.cu:16-31)
  bool ok = true;
  uint32_t scifi_bytes = 0;
  {
    const uint32_t off = a.dev_scifi_raw_input_offsets[e];
    const int64_t nb = checked_banks(a.dev_scifi_raw_input, a.scifi_events_size, off);
    if (nb < 0) {
      ok = false;
    }
    else {
      const uint32_t* bank_off = reinterpret_cast<const uint32_t*>(a.dev_scifi_raw_input + off) + 1;
      const size_t payload = (size_t) off + 4 * ((size_t) nb + 2);
      uint32_t bad = 0;
      for (uint32_t i = lane; i < (uint32_t) nb; i += kWave) {
        const uint32_t lo = bank_off[i], hi = bank_off[i + 1];
        scifi_bytes += hi - lo - 4u;
        bad |= (hi < lo || payload + hi > a.scifi_events_size) ? 1u : 0u;
      }
      scifi_bytes = wave_sum(scifi_bytes);
      ok = wave_sum(bad) == 0;
    }
  }
  const uint32_t n_scifi = scifi_bytes / 2u - 2u;

  // UT: hits per bank, low 16 bits of the word after the source id
  // (GlobalEventCut.cu:37-46, UTRaw.cuh:10-18)
  uint32_t n_ut = 0;
  if (ok) {
    const uint32_t off = a.dev_ut_raw_input_offsets[e];
    const int64_t nb = checked_banks(a.dev_ut_raw_input, a.ut_events_size, off);
    if (nb < 0) {
      ok = false;
    }
    else {
      const uint32_t* bank_off = reinterpret_cast<const uint32_t*>(a.dev_ut_raw_input + off) + 1;
      const size_t payload = (size_t) off + 4 * ((size_t) nb + 2);
      uint32_t hits = 0, bad = 0;
      for (uint32_t i = lane; i < (uint32_t) nb; i += kWave) {
        const size_t at = payload + bank_off[i];
        if ((at & 3u) || at + 8 > a.ut_events_size) {
          bad = 1;
          continue;
        }
        hits += *reinterpret_cast<const uint32_t*>(a.dev_ut_raw_input + at + 4) & 0xFFFFu;
      }
      // 32-bit sums as in the reference; they cannot wrap below 65537 banks
      n_ut = wave_sum(hits);
      ok = wave_sum(bad) == 0;
    }
  }
  const uint32_t max = VELO_GEC_MAX_SCIFI_UT_CLUSTERS;
  const bool keep = ok && n_scifi < max && n_ut < max && n_ut + n_scifi < max;
  if (lane == 0) a.dev_event_list[e] = keep ? 1u : 0u;
}

__global__ void __launch_bounds__(kCompactBlock) gec_compact_kernel(uint32_t* list, uint32_t* count, uint32_t n)
{
  __shared__ uint32_t wave_count[kCompactBlock / kWave];
  const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  uint32_t base = 0;
  for (uint32_t tile = 0; tile < n; tile += kCompactBlock) {
    const uint32_t i = tile + threadIdx.x;
    const bool keep = i < n && list[i] != 0;
    const uint64_t ballot = __ballot(keep);
    if (lane == 0) wave_count[wave] = (uint32_t) __popcll(ballot);
    __syncthreads(); // every flag of the tile is read
    uint32_t before = 0, total = 0;
    for (uint32_t w = 0; w < kCompactBlock / kWave; ++w) {
      before += w < wave ? wave_count[w] : 0u;
      total += wave_count[w];
    }
    if (keep) list[base + before + (uint32_t) __popcll(ballot & __lanemask_lt())] = i;
    base += total;
    __syncthreads(); // wave_count is rewritten by the next tile
  }
  if (threadIdx.x == 0) *count = base;
}

} // namespace velo

using namespace velo;

extern "C" int velo_global_event_cut_launch(const velo_global_event_cut_args* a, void* stream)
{
  if (!a || !a->dev_number_of_selected_events || (!a->dev_event_list && a->n_events)) return (int) hipErrorInvalidValue;
  if (a->n_events && (!a->dev_ut_raw_input || !a->dev_ut_raw_input_offsets || !a->dev_scifi_raw_input ||
                      !a->dev_scifi_raw_input_offsets))
    return (int) hipErrorInvalidValue;
  const uint32_t block = a->tuning.block_dim ? a->tuning.block_dim : 256;
  if (block % kWave || block > 1024) return (int) hipErrorInvalidValue;
  if (a->n_events) {
    const uint32_t per_block = block / kWave;
    hipLaunchKernelGGL(gec_flags_kernel, dim3((a->n_events + per_block - 1) / per_block), dim3(block), 0,
                       (hipStream_t) stream, *a);
  }
  hipLaunchKernelGGL(gec_compact_kernel, dim3(1), dim3(kCompactBlock), 0, (hipStream_t) stream, a->dev_event_list,
                     a->dev_number_of_selected_events, a->n_events);
  return (int) hipGetLastError();
}
