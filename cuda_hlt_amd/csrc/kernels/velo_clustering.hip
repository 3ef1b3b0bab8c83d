// Raw SuperPixel decoding: init_event_list, estimate_input_size and masked
// clustering for gfx950, plus the device constants.
//
// Work layout: one workgroup per event.  The workgroup first copies the raw
// event into LDS (coalesced) and wave 0 builds a flat table of the event's
// banks: the SPs of all banks numbered in bank order, which is module
// order.  estimate_input_size gives each wave a contiguous run of 64-SP
// chunks, resolves each SP's neighbourhood in its bank in LDS, sums the
// estimates per module and lists the candidates in bank, SP, nibble order.
// Masked clustering gives each wave the whole modules nearest its share of
// the SPs, takes up to 64 consecutive SPs per pass whatever module they
// belong to, reads their candidates from the list (or scans the banks
// itself when it has none) and grows one candidate per lane.  Output slots
// come from ballot-based wave prefix sums restarted at each module, so
// clusters are written in bank, SP, cluster order with no atomics and no
// cross-wave synchronisation.
#include <cstdio>
#include <cstring>
#include <vector>

#include "velo_device.h"

namespace velo {

constexpr uint32_t kDefaultBlock = 256;

// Constants.cu:36
__constant__ uint8_t c_candidate_ks[9] = {0, 0, 1, 4, 4, 5, 5, 5, 5};

// Candidate byte of a non-isolated SP (EstimateInputSize.cu:118-183).  Every
// SP of the bank enters the neighbour test, isolated or not (:126, Q1).
__device__ inline uint32_t neighbour_pixels(uint32_t w, const uint32_t* __restrict__ words, uint32_t n)
{
  const uint32_t sp_addr = (w & 0x007FFF00u) >> 8;
  const uint32_t sp = w & 0xFFu;
  uint32_t pixels = (sp & 0x0Fu) | ((sp & 0xF0u) << 2);
  const uint32_t sp_row = sp_addr & 0x3Fu;
  const uint32_t sp_col = sp_addr >> 6;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t o = words[k];
    const uint32_t o_addr = (o & 0x007FFF00u) >> 8;
    const uint32_t o_row = o_addr & 0x3Fu;
    const uint32_t o_col = o_addr >> 6;
    const uint32_t osp = o & 0xFFu;
    const uint32_t dcol = o_col - sp_col; // 0 or 1 for the columns of interest
    const uint32_t drow = o_row - sp_row; // 1, 0 or -1
    if (dcol == 0u && drow == 1u) pixels |= ((osp & 0x01u) | ((osp & 0x10u) << 2)) << 4;
    if (dcol == 1u && drow == 1u) pixels |= (osp & 0x01u) << 16;
    if (dcol == 1u && drow == 0u) pixels |= (osp & 0x0Fu) << 12;
    if (dcol == 1u && drow == 0xFFFFFFFFu) pixels |= (osp & 0x08u) << 8;
    if (dcol == 0u && drow == 0xFFFFFFFFu) pixels |= (osp & 0x80u) >> 2;
  }
  return pixels;
}

__device__ inline uint32_t candidates_from_pixels(uint32_t pixels)
{
  const uint32_t inside = pixels & 0x3CFu;
  const uint32_t mask = (inside << 1) | (inside << 5) | (inside << 6) | (inside << 7);
  const uint32_t working = mask & ~pixels;
  const uint32_t cand = (working >> 1) & (working >> 5) & (working >> 6) & (working >> 7) & pixels;
  return (cand & 0x03u) | ((cand & 0xC0u) >> 4) | ((cand & 0x0Cu) << 2) | ((cand & 0x0300u) >> 2);
}

// The 3x4 SP window around (sp_row, sp_col) built from non-isolated SPs
// (MaskedVeloClustering.cu:221-272).
__device__ inline void load_window(
  int32_t sp_row, int32_t sp_col, const uint32_t* __restrict__ words, uint32_t n, uint32_t pa[3])
{
  pa[0] = pa[1] = pa[2] = 0;
  const int32_t row_lo = sp_row - 2;
  const int32_t col_lo = sp_col - 1;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t o = words[k];
    if (o & 0x80000000u) continue;
    const uint32_t o_addr = (o & 0x007FFF00u) >> 8;
    const int32_t rr = (int32_t) (o_addr & 0x3Fu) - row_lo;
    const int32_t rc = (int32_t) (o_addr >> 6) - col_lo;
    if ((uint32_t) rr < 4u && (uint32_t) rc < 3u) {
      const uint32_t osp = o & 0xFFu;
      const uint32_t bits = ((osp & 0x0Fu) << (4 * rr)) | ((osp & 0xF0u) << (12 + 4 * rr));
      if (rc == 0) pa[0] |= bits;
      else if (rc == 1) pa[1] |= bits;
      else pa[2] |= bits;
    }
  }
}

// ---------------------------------------------------------------------------
// Neighbour lookups in a bank whose SP addresses are strictly increasing
// (checked at run time; other banks use the linear scans above).  They
// return exactly what the linear scans return for such a bank: with unique
// addresses every neighbour position matches at most one SP.
// ---------------------------------------------------------------------------
__device__ inline uint32_t sp_address(uint32_t w) { return (w & 0x007FFF00u) >> 8; }

template<typename P>
__device__ inline uint32_t lower_bound_address(P words, uint32_t n, uint32_t key)
{
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sp_address(words[mid]) < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// neighbour_pixels for the SP at index i of a sorted bank
template<typename P>
__device__ inline uint32_t neighbour_pixels_sorted(uint32_t w, uint32_t i, P words, uint32_t n)
{
  const uint32_t addr = sp_address(w);
  const uint32_t sp = w & 0xFFu;
  uint32_t pixels = (sp & 0x0Fu) | ((sp & 0xF0u) << 2);
  const uint32_t row = addr & 0x3Fu;
  if (i > 0 && row > 0) { // bottom: (col, row - 1)
    const uint32_t o = words[i - 1];
    if (sp_address(o) == addr - 1u) pixels |= ((o & 0xFFu) & 0x80u) >> 2;
  }
  if (i + 1 < n && row < 63) { // top: (col, row + 1)
    const uint32_t o = words[i + 1];
    if (sp_address(o) == addr + 1u) {
      const uint32_t osp = o & 0xFFu;
      pixels |= ((osp & 0x01u) | ((osp & 0x10u) << 2)) << 4;
    }
  }
  // right column: right-bottom (+63), right (+64), top-right (+65); the
  // first SP at or above +63 is a few SPs after i (a binary search after
  // four steps)
  uint32_t j0 = i + 1;
  for (int step = 0; j0 < n && sp_address(words[j0]) < addr + 63u; ++step) {
    if (step == 4) {
      j0 += lower_bound_address(words + j0, n - j0, addr + 63u);
      break;
    }
    ++j0;
  }
  for (uint32_t j = j0; j < n; ++j) {
    const uint32_t o = words[j];
    const uint32_t oa = sp_address(o);
    if (oa > addr + 65u) break;
    const uint32_t osp = o & 0xFFu;
    if (oa == addr + 63u && row > 0) pixels |= (osp & 0x08u) << 8;
    if (oa == addr + 64u) pixels |= (osp & 0x0Fu) << 12;
    if (oa == addr + 65u && row < 63) pixels |= (osp & 0x01u) << 16;
  }
  return pixels;
}

// load_window for a sorted bank: one address range from (col - 1, row - 2)
// to (col + 1, row + 1), rows outside the window skipped.  hint: index of
// the SP at (sp_row, sp_col); the range starts at most a few SPs before it,
// so it is found by stepping back (a binary search after four steps).
template<typename P>
__device__ inline void load_window_sorted(
  int32_t sp_row, int32_t sp_col, P words, uint32_t n, uint32_t hint, uint32_t pa[3])
{
  uint32_t p0 = 0, p1 = 0, p2 = 0;
  const int32_t row_lo = sp_row - 2;
  const int32_t r0 = row_lo < 0 ? 0 : row_lo;
  const int32_t r1 = sp_row + 1 > 63 ? 63 : sp_row + 1;
  const int32_t c0 = sp_col - 1 < 0 ? 0 : sp_col - 1;
  const int32_t c1 = sp_col + 1 > 511 ? 511 : sp_col + 1;
  const uint32_t key_lo = ((uint32_t) c0 << 6) | (uint32_t) r0;
  const uint32_t key_hi = ((uint32_t) c1 << 6) | (uint32_t) r1;
  uint32_t j0 = hint < n ? hint : n;
  for (int step = 0; j0 > 0 && sp_address(words[j0 - 1]) >= key_lo; ++step) {
    if (step == 4) {
      j0 = lower_bound_address(words, j0, key_lo);
      break;
    }
    --j0;
  }
  for (uint32_t j = j0; j < n; ++j) {
    const uint32_t o = words[j];
    const uint32_t oa = sp_address(o);
    if (oa > key_hi) break;
    const int32_t rr = (int32_t) (oa & 0x3Fu) - row_lo;
    if ((o & 0x80000000u) || (uint32_t) rr >= 4u) continue;
    const int32_t rc = (int32_t) (oa >> 6) - (sp_col - 1); // 0..2 inside the key range
    const uint32_t osp = o & 0xFFu;
    const uint32_t bits = ((osp & 0x0Fu) << (4 * rr)) | ((osp & 0xF0u) << (12 + 4 * rr));
    p0 |= rc == 0 ? bits : 0u;
    p1 |= rc == 1 ? bits : 0u;
    p2 |= rc == 2 ? bits : 0u;
  }
  pa[0] = p0;
  pa[1] = p1;
  pa[2] = p2;
}

struct Hit {
  float x, y, z;
  uint32_t id;
};

// A cluster before the local -> global transform: centroid pixel, its
// fractional offsets and the sensor; key = cx | cy << 10 | sensor << 18.
struct HitSeed {
  uint32_t key;
  float fx, fy;
};

// Local -> global, MaskedVeloClustering.cu:122-138 and :431-449.
__device__ inline Hit make_hit(const char* geometry, uint32_t sensor, uint32_t cx, uint32_t cy, float fx, float fy)
{
  const float* local_x = reinterpret_cast<const float*>(geometry);
  const float* x_pitch = local_x + 768;
  const float* ltg = x_pitch + 768 + 16 * sensor;
  const float t = fx * x_pitch[cx];
  const float lx = local_x[cx] + t;
  double dly = (double) cy + 0.5;
  dly = dly + (double) fy;
  dly = dly * (double) kPixelSize;
  const float ly = (float) dly;
  Hit h;
  float a = ltg[0] * lx;
  float b = ltg[1] * ly;
  h.x = a + b;
  h.x = h.x + ltg[9];
  a = ltg[3] * lx;
  b = ltg[4] * ly;
  h.y = a + b;
  h.y = h.y + ltg[10];
  a = ltg[6] * lx;
  b = ltg[7] * ly;
  h.z = a + b;
  h.z = h.z + ltg[11];
  h.id = lhcb_id(sensor, cx, cy);
  return h;
}

// Cluster grown from candidate pixel k (MaskedVeloClustering.cu:274-461).
__device__ inline bool cluster_candidate(
  uint32_t pa0, uint32_t pa1, uint32_t pa2, int32_t sp_row, int32_t sp_col, uint32_t k, uint32_t sensor, HitSeed* out)
{
  const int32_t row_lower_limit = (sp_row - 2) * 4;
  const int32_t col_lower_limit = (sp_col - 1) * 2;
  const uint32_t row = (uint32_t) (sp_row * 4) + (k % 4u);
  const uint32_t col = (uint32_t) (sp_col * 2) + (k >= 4u ? 1u : 0u);
  const uint32_t start = (1u << (row - (uint32_t) row_lower_limit)) << (16u * (col % 2u));
  // The reference grows a frontier with the masks of current_mask /
  // mask_ltr / mask_rtl, which together are the 8-neighbourhood of a pixel
  // in the 6-column x 16-row window.  After step s its cluster is every
  // window pixel within s king moves of the candidate along window pixels;
  // it stops after the first step that adds nothing, or after 12 steps.
  // The same ball is grown here by dilating the whole cluster (rows, then
  // pixel columns) and keeping the window pixels.
  uint32_t c0 = 0, c1 = start, c2 = 0;
  for (int it = 0; it < kMaxClusteringIterations; ++it) {
    const uint32_t v0 = c0 | ((c0 << 1) & 0xFFFEFFFFu) | ((c0 >> 1) & 0xFFFF7FFFu);
    const uint32_t v1 = c1 | ((c1 << 1) & 0xFFFEFFFFu) | ((c1 >> 1) & 0xFFFF7FFFu);
    const uint32_t v2 = c2 | ((c2 << 1) & 0xFFFEFFFFu) | ((c2 >> 1) & 0xFFFF7FFFu);
    // (v1:v0) >> 16 moves pixel column 1 into column 0 and column 2 into 1
    const uint32_t a01 = (uint32_t) ((((uint64_t) v1 << 32) | v0) >> 16);
    const uint32_t a12 = (uint32_t) ((((uint64_t) v2 << 32) | v1) >> 16);
    const uint32_t d0 = (v0 | (v0 << 16) | a01) & pa0;
    const uint32_t d1 = ((v1 | a01 | a12) & pa1) | start;
    const uint32_t d2 = (v2 | a12 | (v2 >> 16)) & pa2;
    const bool grew = ((d0 ^ c0) | (d1 ^ c1) | (d2 ^ c2)) != 0u;
    c0 = d0;
    c1 = d1;
    c2 = d2;
    if (!grew) break;
  }
  if (c2) return false;
  const int n = __popc(c0) + __popc(c1);
  if (n >= 3) {
    const uint32_t x0 = cluster_current_mask(c0) & ~c0;
    const uint32_t x1 = (cluster_current_mask(c1) | mask_ltr(c0)) & ~c1;
    const uint32_t x2 = mask_ltr(c1);
    const uint32_t k0 = cand_mask_rc(x0, x1) & c0;
    const uint32_t k1 = (cand_mask_rc(x1, x2) & c1) ^ start;
    if (k0 | k1) {
      const int32_t neg = (int32_t) ~(0u - start);
      const bool under = start < 4096u;
      uint32_t smaller = under ? (0xFFFu & (uint32_t) neg) : (0xFFFu & (uint32_t) (neg >> 16));
      smaller |= smaller << 16;
      const uint32_t cur = smaller | (under ? 0xFFFF0000u : 0u);
      if (((k0 & smaller) | (k1 & cur)) != 0u) return false;
    }
  }
  // centroid: sums of absolute pixel columns and rows
  const int x = __popc(c0 & 0xFFFFu) * col_lower_limit + __popc(c0 >> 16) * (col_lower_limit + 1) +
                __popc(c1 & 0xFFFFu) * (col_lower_limit + 2) + __popc(c1 >> 16) * (col_lower_limit + 3);
  // sum over pixels of (row_lower_limit + r): the r sum taken bit by bit of r
  const int y = row_lower_limit * n + (__popc(c0 & 0xAAAAAAAAu) + __popc(c1 & 0xAAAAAAAAu)) +
                2 * (__popc(c0 & 0xCCCCCCCCu) + __popc(c1 & 0xCCCCCCCCu)) +
                4 * (__popc(c0 & 0xF0F0F0F0u) + __popc(c1 & 0xF0F0F0F0u)) +
                8 * (__popc(c0 & 0xFF00FF00u) + __popc(c1 & 0xFF00FF00u));
  const float fxq = (float) x / (float) n;
  const float fyq = (float) y / (float) n;
  // x, y >= 0 (no pixel lies left of or below the sensor), n <= 64 and the
  // quotients stay below 1024: the rounded float quotient is within 1e-4 of
  // x / n and exact when n divides x, so its floor is the integer quotient
  const uint32_t cx = (uint32_t) fxq;
  const uint32_t cy = (uint32_t) fyq;
  out->fx = fxq - (float) cx;
  out->fy = fyq - (float) cy;
  out->key = cx | (cy << 10) | (sensor << 18);
  return true;
}

// ---------------------------------------------------------------------------
// K0
// ---------------------------------------------------------------------------
__global__ void init_event_list_kernel(uint32_t* event_list, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) event_list[i] = i;
}

// Largest-first processing order of the selected events: one workgroup
// counts the events per size bucket (64 bytes of raw input each), scans the
// buckets from the largest down and places every event at its bucket's next
// position.  Events of one bucket come in no particular order; only the
// scheduling of persistent launches depends on the order, never an output.
constexpr uint32_t kOrderBuckets = 1024;
constexpr uint32_t kOrderBlock = 1024;

__device__ inline uint32_t order_bucket(const uint32_t* offsets, const uint32_t* event_list, uint32_t i)
{
  const uint32_t ev = event_list ? event_list[i] : i;
  const uint32_t size = offsets[ev + 1] - offsets[ev];
  return min(size >> 6, kOrderBuckets - 1);
}

__global__ void __launch_bounds__(kOrderBlock)
  event_order_kernel(const uint32_t* offsets, const uint32_t* event_list, uint32_t* order, uint32_t n)
{
  __shared__ uint32_t next[kOrderBuckets];
  __shared__ uint32_t wave_total[kOrderBlock / kWave];
  const uint32_t t = threadIdx.x;
  next[t] = 0;
  __syncthreads();
  for (uint32_t i = t; i < n; i += kOrderBlock) atomicAdd(&next[order_bucket(offsets, event_list, i)], 1u);
  __syncthreads();
  // exclusive scan from the largest bucket down: thread t holds bucket 1023 - t
  const uint32_t count = next[kOrderBuckets - 1 - t];
  const uint32_t incl = wave_inclusive_scan(count);
  if (t % kWave == kWave - 1) wave_total[t / kWave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < t / kWave; ++w) before += wave_total[w];
  next[kOrderBuckets - 1 - t] = before + incl - count;
  __syncthreads();
  for (uint32_t i = t; i < n; i += kOrderBlock) order[atomicAdd(&next[order_bucket(offsets, event_list, i)], 1u)] = i;
}

// ---------------------------------------------------------------------------
// Event staging and module iteration (K1 and K5)
// ---------------------------------------------------------------------------
// The raw event is copied word by word into LDS by the whole workgroup when
// it fits (tuning.lds_tile words); otherwise the same code reads it from
// global memory.
constexpr uint32_t kDefaultEventWords = 5632;   // estimate: 22 KB of LDS per workgroup (six per CU)
constexpr uint32_t kDefaultClusterWords = 5632; // clustering: 22 KB (event and candidate list; six per CU)
constexpr uint32_t kClusterEntries = kWave; // candidate entries per wave

// Bank b of an event of nwords words.  A bank whose offset table, header
// or SP words would reach past the event is flagged and skipped, like a
// mislabelled one (the reference reads on past the event).  So is a bank
// of kMaxBankSps SPs or more (a 16 MB bank; VELO banks hold a few hundred
// SPs), so that the SP counts of an event's 208 banks sum below 2^30.
constexpr uint32_t kMaxBankSps = 1u << 22;

template<typename P>
__device__ inline void bank_of(
  P ev, uint32_t n_banks, uint32_t nwords, uint32_t b, uint32_t* base, uint32_t* n, uint32_t* status)
{
  *base = 0;
  *n = 0;
  if (b >= n_banks || b >= VELO_N_SENSORS) return;
  // an event holds fewer than 2^30 words, so none of these sums wraps
  if (nwords < 2u || n_banks > nwords - 2u) {
    *status |= VELO_STATUS_BAD_BANK;
    return;
  }
  const uint32_t h = n_banks + 2 + ev[1 + b] / 4u;
  if (h > nwords - 2u) {
    *status |= VELO_STATUS_BAD_BANK;
    return;
  }
  const uint32_t sensor = ev[h];
  const uint32_t cnt = ev[h + 1];
  if (sensor != b || cnt > nwords - 2u - h || cnt >= kMaxBankSps) {
    *status |= VELO_STATUS_BAD_BANK;
    return;
  }
  if (cnt > 256u) *status |= VELO_STATUS_SP_INDEX_TRUNCATED;
  *base = h + 2;
  *n = cnt;
}

// Copies n words to LDS with the whole workgroup, 4 loads in flight per lane.
__device__ inline void stage_words(uint32_t* dst, const uint32_t* src, uint32_t n)
{
  const uint32_t step = 4 * blockDim.x;
  for (uint32_t base = threadIdx.x; base < n; base += step) {
    uint32_t v[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t i = base + k * blockDim.x;
      v[k] = i < n ? src[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t i = base + k * blockDim.x;
      if (i < n) dst[i] = v[k];
    }
  }
}

// ---------------------------------------------------------------------------
// K1 over one flat SP list per event
// ---------------------------------------------------------------------------
// The SPs of all banks are numbered in bank order (flat index f).  The
// banks that hold SPs are numbered k = 0..nnz-1; bank k holds flat SPs
// [start[k], start[k+1]) and SP f of it is event word f + delta[k].  Wave w
// takes a contiguous run of 64-SP chunks, so every wave does the same work
// whatever the module sizes; within a chunk each lane finds its bank from
// the bank starts that fall inside the chunk (one LDS read and a wave OR,
// no search).  Flat order is bank order, which is module order, so the
// candidate slots are one running count per wave.
struct FlatBanks {
  static constexpr bool kByBank = false;
  uint32_t start[VELO_N_SENSORS + 1]; // by k; start[nnz] = number of SPs
  uint32_t delta[VELO_N_SENSORS];     // by k
  uint32_t unsorted[(VELO_N_SENSORS + 31) / 32]; // bit b: bank b's addresses do not strictly increase
  uint32_t est[VELO_N_MODULES];       // estimate: estimates; clustering: cluster counts
  uint32_t cand[VELO_N_MODULES];
  uint32_t wave_cands[1024 / kWave];
  uint32_t nnz, total, status;
  uint8_t bank[VELO_N_SENSORS];       // by k
};
constexpr uint32_t kFlatZeroWords = (VELO_N_SENSORS + 31) / 32 + 2 * VELO_N_MODULES;

// Clustering's table adds each module's first flat SP, its cluster slots
// and a bank -> k map (estimate_input_size keeps the smaller table: less LDS
// per workgroup, more events per CU).
struct ClusterBanks : FlatBanks {
  static constexpr bool kByBank = true;
  uint32_t mflat[VELO_N_MODULES + 1]; // flat index of each module's first SP; [52] = number of SPs
  uint32_t mst[VELO_N_MODULES + 1];   // the event's module cluster starts and end
  uint8_t kmap[VELO_N_SENSORS];       // bank -> k (0xFF: no SPs)
};

// The flat SP index of a lane: flat_run over the starts of the banks with
// SPs (velo_device.h).
using FlatSp = FlatRun;

__device__ inline FlatSp flat_sp(const FlatBanks& fb, uint32_t nnz, uint32_t f0, uint32_t kcur, uint32_t* knext)
{
  return flat_run(fb.start, nnz, f0, kcur, knext);
}

// The bank holding SP f (f below the number of SPs).
__device__ inline uint32_t flat_bank_of(const FlatBanks& fb, uint32_t nnz, uint32_t f)
{
  return flat_run_first(fb.start, nnz, f);
}

// Bank table of the event in flat form, by wave 0.
template<typename P, typename B>
__device__ inline void build_flat_banks(P ev, uint32_t nwords, B& fb)
{
  const uint32_t lane = lane_id();
  const uint32_t n_banks = nwords ? ev[0] : 0u;
  uint32_t nnz = 0, total = 0, st = 0;
#pragma unroll
  for (uint32_t b0 = 0; b0 < VELO_N_SENSORS; b0 += kWave) {
    const uint32_t b = b0 + lane;
    uint32_t base = 0, n = 0;
    if (b < VELO_N_SENSORS) bank_of(ev, n_banks, nwords, b, &base, &n, &st);
    // a bank holds fewer than 2^22 SPs, so no sum here wraps
    const uint64_t nz = ballot(n != 0);
    const uint32_t inc = wave_inclusive_scan(n);
    const uint32_t s = total + inc - n;
    const uint32_t k = nnz + count_below(nz);
    if constexpr (B::kByBank) {
      if (b < VELO_N_SENSORS) {
        fb.kmap[b] = n ? (uint8_t) k : (uint8_t) 0xFFu;
        if (b % 4u == 0) fb.mflat[b / 4u] = s;
      }
    }
    if (n) {
      fb.start[k] = s;
      fb.delta[k] = base - s;
      fb.bank[k] = (uint8_t) b;
    }
    nnz += (uint32_t) __popcll(nz);
    total += __builtin_amdgcn_readlane(inc, 63);
  }
  st = wave_or(st);
  if (lane == 0) {
    fb.start[nnz] = total;
    if constexpr (B::kByBank) fb.mflat[VELO_N_MODULES] = total;
    fb.nnz = nnz;
    fb.total = total;
    fb.status = st;
  }
}

template<typename P>
__device__ __forceinline__ void estimate_event_flat(
  P ev, FlatBanks& fb, const velo_estimate_input_size_args& a, uint32_t e, uint32_t* stash)
{
  // stash: the event's LDS copy (null when the event is read from global
  // memory).  The estimate pass keeps each non-isolated SP's candidate byte
  // in bits 23-30 of its word there, bits no reader of the copy looks at,
  // so the candidate pass need not compute it again.
  const uint32_t waves = blockDim.x / kWave;
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  const uint32_t nnz = fb.nnz, total = fb.total;
  const uint32_t chunks = (total + kWave - 1) / kWave;
  const uint32_t per_wave = (chunks + waves - 1) / waves;
  const uint32_t c0 = wave * per_wave < chunks ? wave * per_wave : chunks;
  const uint32_t c1 = c0 + per_wave < chunks ? c0 + per_wave : chunks;
  const uint32_t kfirst = c0 < c1 ? flat_bank_of(fb, nnz, c0 * kWave) : 0u;

  // pass 1: estimates and candidate counts, summed per module.  Banks are
  // taken as sorted and checked on the way; an event with an unsorted bank
  // (the flags are then known) runs the pass again.
  uint32_t wave_cands = 0;
  for (uint32_t attempt = 0;; ++attempt) {
    wave_cands = 0;
    uint32_t kcur = kfirst;
    for (uint32_t c = c0; c < c1; ++c) {
      const uint32_t f = c * kWave + lane;
      const FlatSp sp = flat_sp(fb, nnz, c * kWave, kcur, &kcur);
      uint32_t est = 0, nc = 0, mod = 0xFFFFFFFFu;
      if (f < total) {
        const uint32_t d = fb.delta[sp.k];
        const uint32_t b = fb.bank[sp.k];
        mod = b / 4u;
        const uint32_t w = ev[f + d];
        if (attempt == 0 && f + 1 < sp.e && sp_address(w) >= sp_address(ev[f + d + 1]))
          atomicOr(&fb.unsorted[b / 32], 1u << (b % 32));
        if (w & 0x80000000u) {
          est = isolated_cluster_count(w & 0xFFu);
        }
        else {
          const uint32_t base = sp.s + d, n = sp.e - sp.s;
          const bool sorted = attempt == 0 || !((fb.unsorted[b / 32] >> (b % 32)) & 1u);
          const uint32_t cb = candidates_from_pixels(sorted ? neighbour_pixels_sorted(w, f - sp.s, ev + base, n)
                                                            : neighbour_pixels(w, ev + base, n));
          if (stash) stash[f + d] = (w & 0x807FFFFFu) | (cb << 23);
          nc = ((cb & 0x0Fu) ? 1u : 0u) + ((cb & 0xF0u) ? 1u : 0u);
          est = nc;
        }
      }
      // modules are runs of lanes: the last lane of each run adds the scan
      // up to it to its module and takes it off the next run's module
      const uint32_t ie = wave_inclusive_scan(est);
      const uint32_t ic = wave_inclusive_scan(nc);
      const uint32_t nmod = __shfl_down(mod, 1, kWave);
      const bool last = lane == kWave - 1 || nmod != mod;
      if (last && mod != 0xFFFFFFFFu) {
        atomicAdd(&fb.est[mod], ie);
        atomicAdd(&fb.cand[mod], ic);
      }
      if (last && lane != kWave - 1 && mod != 0xFFFFFFFFu && nmod != 0xFFFFFFFFu) {
        atomicSub(&fb.est[nmod], ie);
        atomicSub(&fb.cand[nmod], ic);
      }
      wave_cands += __builtin_amdgcn_readlane(ic, 63);
    }
    if (lane == 0) fb.wave_cands[wave] = wave_cands;
    __syncthreads();
    bool redo = false;
    for (uint32_t i = 0; i < (VELO_N_SENSORS + 31) / 32; ++i) redo |= fb.unsorted[i] != 0;
    if (attempt > 0 || !redo) break;
    __syncthreads();
    for (uint32_t m = threadIdx.x; m < VELO_N_MODULES; m += blockDim.x) {
      fb.est[m] = 0;
      fb.cand[m] = 0;
    }
    __syncthreads();
  }

  for (uint32_t m = threadIdx.x; m < VELO_N_MODULES; m += blockDim.x) {
    a.dev_estimated_input_size[e * VELO_N_MODULES + m] = fb.est[m];
    a.dev_module_cluster_num[e * VELO_N_MODULES + m] = 0;
  }
  const uint32_t before = wave_sum(lane < wave ? fb.wave_cands[lane] : 0u);
  if (threadIdx.x == 0) {
    uint32_t total_cands = 0;
    for (uint32_t w = 0; w < waves; ++w) total_cands += fb.wave_cands[w];
    uint32_t s = fb.status | VELO_STATUS_BANK_ORDER_CHECKED;
    for (uint32_t i = 0; i < (VELO_N_SENSORS + 31) / 32; ++i)
      if (fb.unsorted[i]) s |= VELO_STATUS_UNSORTED_BANK;
    if (total_cands > VELO_MAX_CANDIDATES_EVENT) s |= VELO_STATUS_CANDIDATES_OVERFLOW;
    a.dev_event_candidate_num[e] = total_cands < VELO_MAX_CANDIDATES_EVENT ? total_cands : VELO_MAX_CANDIDATES_EVENT;
    // first writer of the event's status word: store (no memset needed)
    if (a.dev_velo_event_status) a.dev_velo_event_status[e] = s;
  }
  if (!wave_cands) return;

  // pass 2: candidates in bank, SP, low-then-high-nibble order
  uint32_t* cands = a.dev_cluster_candidates + (size_t) e * VELO_MAX_CANDIDATES_EVENT;
  uint32_t pos = before;
  uint32_t kcur = kfirst;
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t f = c * kWave + lane;
    const FlatSp sp = flat_sp(fb, nnz, c * kWave, kcur, &kcur);
    uint32_t cb = 0, b = 0;
    if (f < total) {
      const uint32_t d = fb.delta[sp.k];
      b = fb.bank[sp.k];
      const uint32_t w = ev[f + d];
      if (!(w & 0x80000000u)) {
        if (stash) {
          cb = (w >> 23) & 0xFFu;
        }
        else {
          const uint32_t base = sp.s + d, n = sp.e - sp.s;
          const bool sorted = !((fb.unsorted[b / 32] >> (b % 32)) & 1u);
          cb = candidates_from_pixels(sorted ? neighbour_pixels_sorted(w, f - sp.s, ev + base, n)
                                             : neighbour_pixels(w, ev + base, n));
        }
      }
    }
    const uint32_t cnt = ((cb & 0x0Fu) ? 1u : 0u) + ((cb & 0xF0u) ? 1u : 0u);
    uint32_t tot;
    const uint32_t pre = wave_prefix_small(cnt, &tot);
    uint32_t slot = pos + pre;
    const uint32_t i = f - sp.s;
    if (cb & 0x0Fu) {
      if (slot < VELO_MAX_CANDIDATES_EVENT) cands[slot] = (i << 11) | (b << 3) | c_candidate_ks[cb & 0x0Fu];
      ++slot;
    }
    if ((cb & 0xF0u) && slot < VELO_MAX_CANDIDATES_EVENT)
      cands[slot] = (i << 11) | (b << 3) | (c_candidate_ks[cb >> 4] + 2u);
    pos += tot;
  }
}

__global__ void __launch_bounds__(1024) estimate_input_size_kernel(velo_estimate_input_size_args a)
{
  extern __shared__ uint32_t lds[];
  __shared__ FlatBanks fb;
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t ei = event_index(a.dev_event_list, e);
    const uint32_t off = a.dev_raw_input_offsets[ei];
    const uint32_t nwords = (a.dev_raw_input_offsets[ei + 1] - off) / 4u;
    const uint32_t* gev = reinterpret_cast<const uint32_t*>(a.dev_raw_input + off);
    uint32_t* zero = fb.unsorted;
    for (uint32_t i = threadIdx.x; i < kFlatZeroWords; i += blockDim.x) zero[i] = 0;
    if (nwords <= a.tuning.lds_tile) {
      stage_words(lds, gev, nwords);
      __syncthreads();
      if (threadIdx.x < kWave) build_flat_banks(static_cast<const uint32_t*>(lds), nwords, fb);
      __syncthreads();
      estimate_event_flat(static_cast<const uint32_t*>(lds), fb, a, e, lds);
    }
    else {
      if (threadIdx.x < kWave) build_flat_banks(gev, nwords, fb);
      __syncthreads();
      estimate_event_flat(gev, fb, a, e, nullptr);
    }
  });
}

// ---------------------------------------------------------------------------
// K5: masked clustering
// ---------------------------------------------------------------------------
// The event's SPs are taken in flat bank order (as in estimate_input_size).
// Wave w takes the modules whose SPs fall nearest to its share of the list
// (module boundaries closest to w/W of the SPs), so every wave has about the
// same number of SPs and each module is one wave's.  Per wave pass over up
// to 64 consecutive SPs:
//  1. each lane classifies its SP: isolated (pattern table) or not;
//  2. the candidates of the non-isolated SPs (low nibble first) form the
//     wave's entries, one per lane, so one growth pass serves up to 64
//     candidates whatever SP and module they came from.  They are read from
//     the event's candidate list (written by estimate_input_size in bank,
//     SP, nibble order) when it is given and staged, else derived from the
//     raw banks;
//  3. a wave prefix over the per-SP cluster counts, restarted at each module
//     and carried across passes, gives each cluster its slot in its module,
//     in bank, SP, cluster order.
// The pass takes the SPs whose candidates fit in one wave of entries (at
// least 32 SPs).

// Value of the lane below (lane 0: its own) and of the lane above (lane 63:
// its own), by DPP wave shifts.
__device__ inline uint32_t lane_below(uint32_t v)
{
  return (uint32_t) __builtin_amdgcn_update_dpp((int) v, (int) v, 0x138, 0xf, 0xf, false); // wave_shr:1
}
__device__ inline uint32_t lane_above(uint32_t v)
{
  return (uint32_t) __builtin_amdgcn_update_dpp((int) v, (int) v, 0x130, 0xf, 0xf, false); // wave_shl:1
}

// First module of wave w's share: the module boundary nearest to w/W of the
// event's SPs (boundaries are flat starts of banks 4m).
__device__ inline uint32_t wave_first_module(const ClusterBanks& fb, uint32_t waves, uint32_t w)
{
  if (w == 0) return 0;
  if (w >= waves) return VELO_N_MODULES;
  const uint32_t lane = lane_id();
  const uint32_t t = (uint32_t) (((uint64_t) fb.total * w) / waves);
  const uint32_t ms = lane <= VELO_N_MODULES ? fb.mflat[lane] : 0xFFFFFFFFu;
  const uint32_t lo = (uint32_t) __popcll(ballot(ms <= t)) - 1u; // mflat[0] = 0
  if (lo >= VELO_N_MODULES) return VELO_N_MODULES;
  const uint32_t slo = __builtin_amdgcn_readlane(ms, lo), shi = __builtin_amdgcn_readlane(ms, lo + 1);
  return t - slo <= shi - t ? lo : lo + 1;
}

template<typename P>
__device__ __forceinline__ void cluster_event(
  P ev, ClusterBanks& fb, const velo_masked_clustering_args& a, uint32_t e, uint32_t* entries,
  const uint32_t* cl, const uint32_t* cl_mod, bool check_order)
{
  const uint32_t waves = blockDim.x / kWave;
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  const uint32_t total = a.dev_module_cluster_start[VELO_N_MODULES * a.n_events];
  float* xs = reinterpret_cast<float*>(a.dev_velo_cluster_container);
  float* ys = xs + total;
  float* zs = ys + total;
  uint32_t* ids = a.dev_velo_cluster_container + 3 * (size_t) total;
  const uint32_t nnz = fb.nnz;
  const uint32_t m0 = wave_first_module(fb, waves, wave), m1 = wave_first_module(fb, waves, wave + 1);
  const uint32_t fbeg = fb.mflat[m0], fend = fb.mflat[m1];
  const uint32_t kfirst = fbeg < fend ? flat_bank_of(fb, nnz, fbeg) : 0u;

  if (check_order) {
    uint32_t kcur = kfirst;
    for (uint32_t jb = fbeg; jb < fend; jb += kWave) {
      const uint32_t f = jb + lane;
      const FlatSp sp = flat_sp(fb, nnz, jb, kcur, &kcur);
      if (f < fend && f + 1 < sp.e) {
        const uint32_t d = fb.delta[sp.k];
        const uint32_t b = fb.bank[sp.k];
        if (sp_address(ev[f + d]) >= sp_address(ev[f + d + 1])) atomicOr(&fb.unsorted[b / 32], 1u << (b % 32));
      }
    }
    __syncthreads();
  }

  uint32_t cptr = cl ? cl_mod[m0] : 0u; // next list entry of the wave
  const uint32_t cend = cl ? cl_mod[m1] : 0u;
  uint32_t cmod = 0xFFFFFFFFu, carry = 0; // the module continuing from the last pass, its clusters so far
  uint32_t kcur = kfirst;
  for (uint32_t jb = fbeg; jb < fend;) {
    const uint32_t f = jb + lane;
    uint32_t knext;
    const FlatSp sp = flat_sp(fb, nnz, jb, kcur, &knext);
    uint32_t i = 0, b = 0, w = 0, c = 0, idx = 0, mod = 0xFFFFFFFFu;
    bool live = f < fend;
    if (live) {
      const uint32_t d = fb.delta[sp.k];
      b = fb.bank[sp.k];
      mod = b / 4u;
      i = f - sp.s;
      w = ev[f + d];
      if (w & 0x80000000u) idx = a.dev_velo_sp_patterns[w & 0xFFu];
      else if (!cl) {
        const uint32_t base = sp.s + d, n = sp.e - sp.s;
        const bool sorted = !((fb.unsorted[b / 32] >> (b % 32)) & 1u);
        c = candidates_from_pixels(sorted ? neighbour_pixels_sorted(w, i, ev + base, n)
                                          : neighbour_pixels(w, ev + base, n));
      }
    }
    // per SP: first entry and number of entries; per entry: its SP's lane,
    // SP index, bank, pixel
    uint32_t first = 0, n_own = 0, taken, n_ent;
    uint32_t src = 0, si = 0, sb = 0, kp = 0;
    if (!cl) {
      uint32_t has_lo = (c & 0x0Fu) ? 1u : 0u;
      uint32_t has_hi = (c & 0xF0u) ? 1u : 0u;
      first = wave_prefix_small(has_lo + has_hi, &n_ent);
      const uint32_t end = first + has_lo + has_hi;
      if (end > kWave) {
        live = false;
        has_lo = has_hi = 0;
      }
      n_own = has_lo + has_hi;
      taken = (uint32_t) __popcll(ballot(live));
      n_ent = n_ent > kWave ? __shfl(end, taken - 1u, 64) : n_ent;
      if (has_lo) entries[first] = lane;
      if (has_hi) entries[first + has_lo] = lane | 64u;
      wave_sync();
      const uint32_t dd = lane < n_ent ? entries[lane] : 0u;
      src = dd & 63u;
      const uint32_t sc = __shfl(c, src, 64);
      si = __shfl(i, src, 64);
      sb = __shfl(b, src, 64);
      kp = (dd & 64u) ? c_candidate_ks[sc >> 4] + 2u : c_candidate_ks[sc & 0x0Fu];
    }
    else {
      // entries of SPs [jb, jb + 64) from the list; when 64 entries do not
      // reach past the pass, its SPs end before the SP of entry 64
      const bool ok = cptr + lane < cend;
      const uint32_t ent = ok ? cl[cptr + lane] : 0u;
      si = ent >> 11;
      sb = (ent >> 3) & 0xFFu;
      kp = ent & 7u;
      const uint32_t ks = ok && sb < VELO_N_SENSORS ? fb.kmap[sb] : 0xFFu;
      const uint32_t epos = ks != 0xFFu ? fb.start[ks] + si : 0xFFFFFFFFu;
      uint32_t limit = jb + kWave;
      if (cptr + kWave < cend) {
        const uint32_t e64 = cl[cptr + kWave];
        const uint32_t b64 = (e64 >> 3) & 0xFFu;
        const uint32_t k64 = b64 < VELO_N_SENSORS ? fb.kmap[b64] : 0xFFu;
        const uint32_t p64 = k64 != 0xFFu ? fb.start[k64] + (e64 >> 11) : 0xFFFFFFFFu;
        limit = p64 < limit ? p64 : limit;
      }
      limit = limit < fend ? limit : fend;
      limit = limit > jb ? limit : jb + 1; // a list that does not match the banks still ends
      taken = limit - jb;
      live = lane < taken;
      const bool mine = epos >= jb && epos < limit;
      n_ent = (uint32_t) __popcll(ballot(mine));
      src = mine ? epos - jb : 0u;
      // the SP of each entry records its first entry and count
      entries[lane] = 0u;
      wave_sync();
      const uint32_t prev = lane_below(src);
      const bool lead = mine && (lane == 0 || prev != src);
      const uint32_t next = lane_above(src);
      const bool pair = lead && lane + 1 < n_ent && next == src;
      if (lead) entries[src] = (lane + 1) | (pair ? 0x100u : 0u);
      wave_sync();
      if (live) {
        const uint32_t info = entries[lane];
        first = (info & 0xFFu) - 1u;
        n_own = info ? 1u + (info >> 8) : 0u;
      }
      cptr += n_ent;
    }
    if (!live) mod = 0xFFFFFFFFu;
    const bool iso = live && (w & 0x80000000u);

    // growth: lane t takes entry t
    HitSeed r {};
    uint64_t acc = 0;
    if (n_ent) {
      bool good = false;
      if (lane < n_ent) {
        // a bank with a candidate has SPs (0xFF only for a list that does
        // not match the banks: an empty window)
        const uint32_t kb = sb < VELO_N_SENSORS ? fb.kmap[sb] : 0xFFu;
        const uint32_t base = kb != 0xFFu ? fb.start[kb] + fb.delta[kb] : 0u;
        const uint32_t n = kb != 0xFFu ? fb.start[kb + 1] - fb.start[kb] : 0u;
        // the reference decodes the candidate's SP index as uint8_t (:179)
        const uint32_t cw = ev[base + (si & 0xFFu)];
        const uint32_t caddr = (cw & 0x007FFF00u) >> 8;
        const int32_t sp_row = (int32_t) (caddr & 0x3Fu), sp_col = (int32_t) (caddr >> 6);
        uint32_t pa[3];
        if (!((fb.unsorted[sb / 32] >> (sb % 32)) & 1u)) load_window_sorted(sp_row, sp_col, ev + base, n, si & 0xFFu, pa);
        else load_window(sp_row, sp_col, ev + base, n, pa);
        good = cluster_candidate(pa[0], pa[1], pa[2], sp_row, sp_col, kp, sb, &r);
      }
      acc = ballot(good);
    }

    // clusters per SP; their rank in the pass and slot in the module
    const uint32_t own_ok = n_own ? (uint32_t) __popcll((acc >> first) & (n_own > 1 ? 3u : 1u)) : 0u;
    const uint32_t cnt = iso ? ((idx & 8u) ? 2u : 1u) : own_ok;
    uint32_t pass_clusters;
    const uint32_t rel = wave_prefix_small(cnt, &pass_clusters);
    const uint32_t prev_mod = lane_below(mod);
    const bool seg_start = lane == 0 || prev_mod != mod;
    const uint32_t seg_base = wave_inclusive_max(seg_start ? rel : 0u);
    const uint32_t in_mod = rel - seg_base + (mod == cmod ? carry : 0u);
    uint32_t slot0 = 0xFFFFFFFFu, slot_end = 0;
    if (live) {
      slot0 = fb.mst[mod] + in_mod; // the module's estimated slots end at mst[mod + 1]
      slot_end = fb.mst[mod + 1];
    }
    auto store = [&](uint32_t key, float fx, float fy, uint32_t slot) {
      const Hit h = make_hit(a.dev_velo_geometry, key >> 18, key & 0x3FFu, (key >> 10) & 0xFFu, fx, fy);
      xs[slot] = h.x;
      ys[slot] = h.y;
      zs[slot] = h.z;
      ids[slot] = h.id;
    };
    if (iso) {
      // isolated SP: pre-computed pattern (MaskedVeloClustering.cu:105-167)
      const uint32_t sp_addr = (w & 0x007FFF00u) >> 8, sp = w & 0xFFu;
      const uint32_t sp_row = sp_addr & 0x3Fu, sp_col = sp_addr >> 6, sensor = b << 18;
      if (slot0 < slot_end)
        store((sp_col * 2u + ((idx >> 2) & 1u)) | ((sp_row * 4u + (idx & 3u)) << 10) | sensor,
              a.dev_velo_sp_fx[2 * sp], a.dev_velo_sp_fy[2 * sp], slot0);
      if ((idx & 8u) && slot0 + 1 < slot_end)
        store((sp_col * 2u + ((idx >> 6) & 1u)) | ((sp_row * 4u + ((idx >> 4) & 3u)) << 10) | sensor,
              a.dev_velo_sp_fx[2 * sp + 1], a.dev_velo_sp_fy[2 * sp + 1], slot0 + 1);
    }
    if (n_ent) {
      // an SP's second cluster follows its first when that one was kept
      const uint32_t room = live && slot0 < slot_end ? min(slot_end - slot0, 3u) : 0u;
      const uint32_t packed = __shfl(first | (room << 8), src, 64);
      const uint32_t s0 = __shfl(slot0, src, 64);
      const uint32_t f1 = packed & 0xFFu;
      const uint32_t after = lane > f1 ? (uint32_t) (acc >> f1) & 1u : 0u;
      if (((acc >> lane) & 1u) && after < (packed >> 8)) store(r.key, r.fx, r.fy, s0 + after);
    }
    // cluster counts per module: the last lane of each module's run adds
    // the run's clusters
    const uint32_t next_mod = lane_above(mod);
    if (mod != 0xFFFFFFFFu && (lane == kWave - 1 || next_mod != mod)) atomicAdd(&fb.est[mod], rel + cnt - seg_base);
    cmod = __builtin_amdgcn_readlane(mod, taken - 1u);
    carry = __builtin_amdgcn_readlane(in_mod + cnt, taken - 1u);
    kcur = taken < kWave ? __builtin_amdgcn_readlane(sp.k, taken) : knext;
    jb += taken;
    // the entry list is rewritten by the next pass
    wave_sync();
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    const uint32_t n = lane < VELO_N_MODULES ? fb.est[lane] : 0u;
    if (lane < VELO_N_MODULES) a.dev_module_cluster_num[e * VELO_N_MODULES + lane] = n;
    const bool module_overflow = ballot(n > VELO_MAX_NUMHITS_IN_MODULE) != 0;
    const uint32_t event_hits = wave_sum(n);
    uint32_t status = module_overflow ? VELO_STATUS_MODULE_HITS_OVERFLOW : 0u;
    if (event_hits >= VELO_MAX_HITS_PER_EVENT) status |= VELO_STATUS_EVENT_HITS_OVERFLOW;
    if (lane == 0 && status && a.dev_velo_event_status) atomicOr(a.dev_velo_event_status + e, status);
  }
}

// At most 80 VGPRs: six waves per SIMD.  The kernel is bound by how many
// events a CU holds at once (LDS: the event tile and each wave's entry list)
// and by the latency of each pass, so it keeps no per-wave cluster staging
// and every lane writes its own clusters.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6)))
masked_clustering_kernel(velo_masked_clustering_args a)
{
  extern __shared__ uint32_t lds[];
  __shared__ ClusterBanks fb;
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t ei = event_index(a.dev_event_list, e);
    const uint32_t off = a.dev_raw_input_offsets[ei];
    const uint32_t nwords = (a.dev_raw_input_offsets[ei + 1] - off) / 4u;
    const uint32_t* gev = reinterpret_cast<const uint32_t*>(a.dev_raw_input + off);
    // each wave's candidate entry list follows the event tile
    uint32_t* entries = lds + a.tuning.lds_tile + (threadIdx.x / kWave) * kClusterEntries;
    uint32_t* zero = fb.unsorted;
    for (uint32_t i = threadIdx.x; i < kFlatZeroWords; i += blockDim.x) zero[i] = 0;
    for (uint32_t m = threadIdx.x; m <= VELO_N_MODULES; m += blockDim.x)
      fb.mst[m] = a.dev_module_cluster_start[e * VELO_N_MODULES + m];
    // The bank-order check is skipped only where estimate_input_size checked
    // the event's banks (BANK_ORDER_CHECKED) and found them sorted (no
    // UNSORTED_BANK); any other status word keeps the check.
    const bool check_order =
      !a.dev_velo_event_status ||
      (a.dev_velo_event_status[e] & (VELO_STATUS_BANK_ORDER_CHECKED | VELO_STATUS_UNSORTED_BANK)) !=
        VELO_STATUS_BANK_ORDER_CHECKED;
    // the event's candidate list follows the event in LDS when it is given,
    // complete (below the list's capacity) and fits
    const bool list = a.dev_cluster_candidates && a.dev_event_candidate_num;
    const uint32_t ncand = list ? a.dev_event_candidate_num[e] : 0u;
    if (nwords <= a.tuning.lds_tile) {
      stage_words(lds, gev, nwords);
      uint32_t* cl = nullptr;
      uint32_t* cl_mod = nullptr;
      if (list && ncand < VELO_MAX_CANDIDATES_EVENT &&
          (size_t) nwords + ncand + VELO_N_MODULES + 1 <= a.tuning.lds_tile) {
        cl = lds + nwords;
        cl_mod = cl + ncand;
        stage_words(cl, a.dev_cluster_candidates + (size_t) e * VELO_MAX_CANDIDATES_EVENT, ncand);
      }
      __syncthreads();
      if (threadIdx.x < kWave) build_flat_banks(static_cast<const uint32_t*>(lds), nwords, fb);
      // the other waves (or wave 0 alone) find each module's first entry
      const uint32_t helper = blockDim.x > kWave ? kWave : 0u;
      if (cl && threadIdx.x >= helper) {
        for (uint32_t m = threadIdx.x - helper; m <= VELO_N_MODULES; m += blockDim.x - helper) {
          // bank field bits 3-10, module = bank / 4
          uint32_t lo = 0, hi = ncand;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (((cl[mid] >> 5) & 0x3Fu) < m) lo = mid + 1;
            else hi = mid;
          }
          cl_mod[m] = lo;
        }
      }
      __syncthreads();
      cluster_event(static_cast<const uint32_t*>(lds), fb, a, e, entries, cl, cl_mod, check_order);
    }
    else {
      if (threadIdx.x < kWave) build_flat_banks(gev, nwords, fb);
      __syncthreads();
      cluster_event(gev, fb, a, e, entries, nullptr, nullptr, check_order);
    }
  });
}

// ---------------------------------------------------------------------------
// SP pattern tables (ClusteringCommon.cpp:3-101), built once on the host
// ---------------------------------------------------------------------------
static void build_sp_tables(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  std::memset(patterns, 0, 256);
  std::memset(sizes, 0, 256);
  for (int i = 0; i < 512; ++i) fx[i] = fy[i] = 0.f;
  for (uint32_t sp = 0; sp < 256; ++sp) {
    // label the 8 pixels (bit b: col b/4, row b%4) by 8-connected components,
    // components ordered by their lowest bit
    int label[8];
    for (int b = 0; b < 8; ++b) label[b] = -1;
    int n_labels = 0;
    for (int b = 0; b < 8; ++b) {
      if (!(sp & (1u << b)) || label[b] >= 0) continue;
      int stack[8], top = 0;
      stack[top++] = b;
      label[b] = n_labels;
      while (top) {
        const int p = stack[--top];
        const int pr = p % 4, pc = p / 4;
        for (int q = 0; q < 8; ++q) {
          if (!(sp & (1u << q)) || label[q] >= 0) continue;
          const int qr = q % 4, qc = q / 4;
          if (abs(qr - pr) <= 1 && abs(qc - pc) <= 1) {
            label[q] = n_labels;
            stack[top++] = q;
          }
        }
      }
      ++n_labels;
    }
    for (int l = 0; l < n_labels && l < 2; ++l) {
      uint32_t x = 0, y = 0, n = 0;
      for (int b = 0; b < 8; ++b)
        if (label[b] == l) {
          x += (uint32_t) (b / 4);
          y += (uint32_t) (b % 4);
          ++n;
        }
      const uint32_t cx = x / n, cy = y / n;
      sizes[sp] |= (uint8_t) (n << (4 * l));
      patterns[sp] |= (uint8_t) (((cx << 2) | cy) << (4 * l));
      patterns[sp] |= (uint8_t) (l << 3);
      fx[2 * sp + l] = (float) x / (float) n - (float) cx;
      fy[2 * sp + l] = (float) y / (float) n - (float) cy;
    }
  }
}

static uint32_t block_of(const velo_tuning& t)
{
  uint32_t b = t.block_dim ? t.block_dim : kDefaultBlock;
  if (b % kWave || b > 1024) return 0;
  return b;
}

} // namespace velo

using namespace velo;

extern "C" {

void velo_sp_tables(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  build_sp_tables(patterns, sizes, fx, fy);
}

int velo_constants_init(const void* geometry, size_t geometry_bytes, const float* module_zs, velo_constants* out)
{
  // Constants.cu:27-31
  static const float default_zs[VELO_N_MODULES] = {
    -287.5f, -275.f, -262.5f, -250.f, -237.5f, -225.f, -212.5f, -200.f, -137.5f, -125.f, -62.5f, -50.f, -37.5f,
    -25.f,   -12.5f, 0.f,     12.5f,  25.f,    37.5f,  50.f,    62.5f,  75.f,    87.5f,  100.f,  112.5f, 125.f,
    137.5f,  150.f,  162.5f,  175.f,  187.5f,  200.f,  212.5f,  225.f,  237.5f,  250.f,  262.5f, 275.f,  312.5f,
    325.f,   387.5f, 400.f,   487.5f, 500.f,   587.5f, 600.f,   637.5f, 650.f,   687.5f, 700.f,  737.5f, 750.f};
  if (!out || !geometry || geometry_bytes != VELO_GEOMETRY_BYTES) return (int) hipErrorInvalidValue;
  std::memset(out, 0, sizeof(*out));
  uint8_t patterns[256], sizes[256];
  float fx[512], fy[512];
  build_sp_tables(patterns, sizes, fx, fy);
  static const uint8_t ks[9] = {0, 0, 1, 4, 4, 5, 5, 5, 5};
  hipError_t err = hipSuccess;
#define VELO_TRY(x) \
  do {              \
    err = (x);      \
    if (err != hipSuccess) goto fail; \
  } while (0)
  VELO_TRY(hipMalloc((void**) &out->dev_velo_sp_patterns, 256));
  VELO_TRY(hipMalloc((void**) &out->dev_velo_sp_fx, 512 * sizeof(float)));
  VELO_TRY(hipMalloc((void**) &out->dev_velo_sp_fy, 512 * sizeof(float)));
  VELO_TRY(hipMalloc((void**) &out->dev_velo_candidate_ks, 9));
  VELO_TRY(hipMalloc((void**) &out->dev_velo_module_zs, VELO_N_MODULES * sizeof(float)));
  VELO_TRY(hipMalloc((void**) &out->dev_velo_geometry, VELO_GEOMETRY_BYTES));
  VELO_TRY(hipMemcpy(out->dev_velo_sp_patterns, patterns, 256, hipMemcpyHostToDevice));
  VELO_TRY(hipMemcpy(out->dev_velo_sp_fx, fx, sizeof(fx), hipMemcpyHostToDevice));
  VELO_TRY(hipMemcpy(out->dev_velo_sp_fy, fy, sizeof(fy), hipMemcpyHostToDevice));
  VELO_TRY(hipMemcpy(out->dev_velo_candidate_ks, ks, 9, hipMemcpyHostToDevice));
  VELO_TRY(hipMemcpy(out->dev_velo_module_zs, module_zs ? module_zs : default_zs, VELO_N_MODULES * sizeof(float),
                     hipMemcpyHostToDevice));
  VELO_TRY(hipMemcpy(out->dev_velo_geometry, geometry, VELO_GEOMETRY_BYTES, hipMemcpyHostToDevice));
#undef VELO_TRY
  return 0;
fail:
  velo_constants_free(out);
  return (int) err;
}

int velo_constants_free(velo_constants* c)
{
  if (!c) return (int) hipErrorInvalidValue;
  void* ptrs[] = {c->dev_velo_sp_patterns, c->dev_velo_sp_fx, c->dev_velo_sp_fy, c->dev_velo_candidate_ks,
                  c->dev_velo_module_zs, c->dev_velo_geometry};
  for (void* p : ptrs)
    if (p) (void) hipFree(p);
  std::memset(c, 0, sizeof(*c));
  return 0;
}

int velo_init_event_list_launch(const velo_init_event_list_args* a, void* stream)
{
  if (!a || (!a->dev_event_list && a->n_events)) return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  const uint32_t block = 256;
  hipLaunchKernelGGL(init_event_list_kernel, dim3((a->n_events + block - 1) / block), dim3(block), 0,
                     (hipStream_t) stream, a->dev_event_list, a->n_events);
  return (int) hipGetLastError();
}

int velo_event_order_launch(const velo_event_order_args* a, void* stream)
{
  if (!a || (a->n_events && (!a->dev_raw_input_offsets || !a->dev_event_order))) return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  hipLaunchKernelGGL(event_order_kernel, dim3(1), dim3(kOrderBlock), 0, (hipStream_t) stream,
                     a->dev_raw_input_offsets, a->dev_event_list, a->dev_event_order, a->n_events);
  return (int) hipGetLastError();
}

int velo_estimate_input_size_launch(const velo_estimate_input_size_args* in, void* stream)
{
  if (!in || !in->dev_raw_input || !in->dev_raw_input_offsets || !in->dev_estimated_input_size ||
      !in->dev_module_cluster_num || !in->dev_event_candidate_num || !in->dev_cluster_candidates)
    return (int) hipErrorInvalidValue;
  if (in->n_events == 0) return 0;
  if (!schedule_ok(in->tuning)) return (int) hipErrorInvalidValue;
  velo_estimate_input_size_args a = *in;
  const uint32_t block = block_of(a.tuning);
  if (!block) return (int) hipErrorInvalidValue;
  if (!a.tuning.lds_tile) a.tuning.lds_tile = kDefaultEventWords;
  const size_t shmem = (size_t) a.tuning.lds_tile * sizeof(uint32_t);
  hipLaunchKernelGGL(estimate_input_size_kernel, dim3(event_grid(a.tuning, a.n_events)), dim3(block), shmem, (hipStream_t) stream, a);
  return (int) hipGetLastError();
}

int velo_masked_clustering_launch(const velo_masked_clustering_args* in, void* stream)
{
  if (!in || !in->dev_raw_input || !in->dev_raw_input_offsets || !in->dev_module_cluster_start ||
      !in->dev_module_cluster_num || !in->dev_velo_cluster_container || !in->dev_velo_geometry ||
      !in->dev_velo_sp_patterns || !in->dev_velo_sp_fx || !in->dev_velo_sp_fy)
    return (int) hipErrorInvalidValue;
  if (in->n_events == 0) return 0;
  if (!schedule_ok(in->tuning)) return (int) hipErrorInvalidValue;
  velo_masked_clustering_args a = *in;
  const uint32_t block = block_of(a.tuning);
  if (!block) return (int) hipErrorInvalidValue;
  if (!a.tuning.lds_tile) a.tuning.lds_tile = kDefaultClusterWords;
  const size_t shmem =
    ((size_t) a.tuning.lds_tile + (block / kWave) * kClusterEntries) * sizeof(uint32_t);
  hipLaunchKernelGGL(masked_clustering_kernel, dim3(event_grid(a.tuning, a.n_events)), dim3(block), shmem, (hipStream_t) stream, a);
  return (int) hipGetLastError();
}

} // extern "C"
