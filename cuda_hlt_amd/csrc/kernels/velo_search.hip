// search_by_triplet (SearchByTriplet.cu:7-111, ProcessModules.cu:8-163,
// TrackSeeding.cu:7-200, TrackForwarding.cu:12-182) and weak_tracks_adder
// (WeakTracksAdder.cu:8-157) for gfx950.
//
// One wave (64 lanes) per event.  The 24 module-pair steps run in sequence;
// within a step every tracks-to-follow entry (forwarding) and every h1
// candidate (seeding) gets a lane.  Appends (tracks, tracks-to-follow,
// weak tracks) take their slots from a ballot prefix in entry order, so the
// output is deterministic and equal to the CPU oracle's sequential order.
//
// State: hit_used as a bitmask in LDS (events up to 8192 estimated hits;
// larger events use the global byte array directly); the event's module
// table (first hit, count, z) in registers, lane m holding module m; x, y,
// z, phi of the six modules a step searches (f .. f-5 for first module f) in
// an LDS window that is a ring over hit numbers when they fit, else the
// four modules that hold h1, h2 and the forwarding candidates in the window
// and the two h0 modules in global memory, else every read of the step in
// global memory (the step body is instantiated for each mode, so window
// reads need no per-access test and issue together); the
// tracks-to-follow ring lives in dev_tracks_to_follow (2000 entries of 8
// bytes per event):
//   word0 = h[n-2] | h[n-1] << 16
//   word1 = (tracklet ? first hit : track number) | n << 16 | skipped << 28
//           | tracklet << 31
// so forwarding never has to look a track up.
#include "velo_device.h"

namespace velo {

constexpr uint32_t kUsedBits = 8192;
constexpr uint32_t kUsedWords = kUsedBits / 32;
constexpr uint32_t kTtfTrackletFlag = 0x80000000u;

// The small accessors below are forced inline: left to the inliner they
// were called out of line from the kernel, with the set kept in scratch.
struct UsedSet {
  uint32_t* bits;
  uint8_t* bytes;
  bool in_lds;
  __device__ __forceinline__ bool test(int h) const
  {
    return in_lds ? ((bits[(uint32_t) h >> 5] >> ((uint32_t) h & 31u)) & 1u) != 0u : bytes[h] != 0;
  }
  __device__ __forceinline__ void set(uint32_t h) const
  {
    if (in_lds) atomicOr(&bits[h >> 5], 1u << (h & 31u));
    else bytes[h] = 1;
  }
};

struct ModuleData {
  uint32_t start, num;
  float z;
};

// Window modes of a step: its six modules in global memory, all six in the
// LDS window, or the four that hold h1, h2 and the forwarding candidates
// (md[2] .. md[5]) in the window and the h0 modules (md[0], md[1]) in
// global memory, for steps whose six modules do not fit.
constexpr int kWindowNone = 0, kWindowAll = 1, kWindowNoH0 = 2;

// Dynamic LDS of search_by_triplet: the hit window, four arrays of `cap`.
extern __shared__ float s_window[];

// Opaque copy of a register value.  Keeps the optimiser from turning
// "LDS value, else global load" into one flat load through a selected
// pointer (which gfx950 code generation rejects for this LDS symbol);
// emits no instruction.
__device__ inline float launder(float v)
{
  asm volatile("" : "+v"(v));
  return v;
}

// The window is a ring over hit numbers: hit h sits at h mod cap (cap a
// power of two), so consecutive steps, whose windows share four of their
// six modules, copy only the hits they do not share.
struct EventHits {
  const float *gx, *gy, *gz, *gphi; // event-relative global slots
  uint32_t cap;                     // window capacity (hits), a power of two
  uint32_t lo, n;                   // window = hits [lo, lo + n)
  __device__ uint32_t at(uint32_t h, uint32_t r) const { return r < n ? (h & (cap - 1u)) : 0u; }
  __device__ float x(uint32_t h) const
  {
    const uint32_t r = h - lo;
    float v = launder(s_window[at(h, r)]);
    if (r >= n) v = gx[h];
    return v;
  }
  __device__ float y(uint32_t h) const
  {
    const uint32_t r = h - lo;
    float v = launder(s_window[cap + at(h, r)]);
    if (r >= n) v = gy[h];
    return v;
  }
  __device__ float z(uint32_t h) const
  {
    const uint32_t r = h - lo;
    float v = launder(s_window[2 * cap + at(h, r)]);
    if (r >= n) v = gz[h];
    return v;
  }
  // The step's hits through the window (W > 0: it holds them) or global
  // memory (W = 0); one mode for a whole step, so loads need no per-access
  // select and are free to be issued together.
  template<int W>
  __device__ float vx(uint32_t h) const
  {
    if constexpr (W != 0) return s_window[h & (cap - 1u)];
    else return gx[h];
  }
  template<int W>
  __device__ float vy(uint32_t h) const
  {
    if constexpr (W != 0) return s_window[cap + (h & (cap - 1u))];
    else return gy[h];
  }
  template<int W>
  __device__ float vz(uint32_t h) const
  {
    if constexpr (W != 0) return s_window[2 * cap + (h & (cap - 1u))];
    else return gz[h];
  }
  template<int W>
  __device__ float vphi(uint32_t h) const
  {
    if constexpr (W != 0) return s_window[3 * cap + (h & (cap - 1u))];
    else return gphi[h];
  }
  __device__ void copy(uint32_t from, uint32_t to)
  {
    for (uint32_t h = from + lane_id(); h < to; h += kWave) {
      const uint32_t q = h & (cap - 1u);
      s_window[q] = gx[h];
      s_window[cap + q] = gy[h];
      s_window[2 * cap + q] = gz[h];
      s_window[3 * cap + q] = gphi[h];
    }
  }
  // Makes hits [lo_, hi_) the window when they fit, copying those not in the
  // current window (wave-wide).
  __device__ void stage(uint32_t lo_, uint32_t hi_)
  {
    const uint32_t old_lo = lo, old_hi = lo + n;
    const uint32_t want = hi_ - lo_ <= cap ? hi_ - lo_ : 0u;
    if (want) {
      if (n == 0 || hi_ <= old_lo || lo_ >= old_hi) {
        copy(lo_, hi_);
      }
      else {
        if (lo_ < old_lo) copy(lo_, old_lo);
        if (hi_ > old_hi) copy(old_hi, hi_);
      }
    }
    lo = lo_;
    n = want;
    wave_sync();
  }
};

// Plain global view of an event's sorted x, y, z (weak_tracks_adder).
struct GlobalHits {
  const float *gx, *gy, *gz;
  __device__ float x(uint32_t h) const { return gx[h]; }
  __device__ float y(uint32_t h) const { return gy[h]; }
  __device__ float z(uint32_t h) const { return gz[h]; }
};

// The event's module table, lane m holding module m (m < 52): its first hit
// (event-relative), hit count and z; read once per event.
struct ModuleTable {
  uint32_t start, num;
  float z;
};

__device__ inline void load_modules(ModuleData md[6], uint32_t first, const ModuleTable& t)
{
#pragma unroll
  for (uint32_t i = 0; i < 6; ++i) {
    const uint32_t m = first - i;
    md[i].start = __builtin_amdgcn_readlane(t.start, m);
    md[i].num = __builtin_amdgcn_readlane(t.num, m);
    md[i].z = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(t.z), m));
  }
}

// Best (h0, h2) for one h1 (TrackSeeding.cu:266-328); h0 then h2 ascending,
// so equal scatters resolve to the lexicographically smallest pair.
template<int W>
__device__ inline float seed_best(
  uint32_t h1, const EventHits& H, const int16_t* h0c, const int16_t* h2c, const UsedSet& used, uint32_t* b0,
  uint32_t* b2)
{
  constexpr int W0 = W == kWindowAll ? kWindowAll : kWindowNone; // h0 modules: in the window only in mode all
  const float h1x = H.template vx<W>(h1), h1y = H.template vy<W>(h1), h1z = H.template vz<W>(h1);
  const int f0 = h0c[2 * h1], l0 = h0c[2 * h1 + 1];
  const int f2 = h2c[2 * h1], l2 = h2c[2 * h1 + 1];
  float best = kFltMax;
  uint32_t bb0 = 0, bb2 = 0;
  // The reference evaluates every condition for every (h0, h2); the slope
  // cut depends on h0 alone and the scatter only matters once both
  // tolerance cuts pass, so they are tested first.  Same arithmetic, same
  // iteration order, same result.
  for (int h0 = f0; h0 < l0; ++h0) {
    if (used.test(h0)) continue;
    const float h0x = H.template vx<W0>(h0), h0y = H.template vy<W0>(h0), h0z = H.template vz<W0>(h0);
    const float dmax = kMaxSlope * (h0z - h1z);
    if (!(fabsf(h1x - h0x) < dmax && fabsf(h1y - h0y) < dmax)) continue;
    const float dz10 = h1z - h0z;
    for (int h2 = f2; h2 < l2; ++h2) {
      if (used.test(h2)) continue;
      const float h2x = H.template vx<W>(h2), h2y = H.template vy<W>(h2), h2z = H.template vz<W>(h2);
      const float z2_tz = (h2z - h0z) / dz10;
      const float x = h0x + (h1x - h0x) * z2_tz;
      const float y = h0y + (h1y - h0y) * z2_tz;
      const float dx = x - h2x;
      const float dy = y - h2y;
      if (!(fabsf(dx) < kTolerance && fabsf(dy) < kTolerance)) continue;
      const float dz21 = h2z - h1z;
      const float scatter_denom2 = 1.0f / (dz21 * dz21);
      const float scatter_num = (dx * dx) + (dy * dy);
      const float scatter = scatter_num * scatter_denom2;
      if (scatter < kMaxScatterSeeding && scatter < best) {
        best = scatter;
        bb0 = (uint32_t) h0;
        bb2 = (uint32_t) h2;
      }
    }
  }
  *b0 = bb0;
  *b2 = bb2;
  return best;
}

// TrackForwarding.cuh:28-58 for the odd (md[4]) and the even (md[5])
// module at once.  bs_first_candidate keeps its probe sequence (its "found"
// flag depends on the elements it probes).  bs_second_candidate searches
// the module from the first candidate on for the first phi not below
// phi + tolerance (else the last slot): in a module sorted by phi that
// position is the same whether the search starts at the first candidate or
// at the module's first hit, clamped to the first candidate, so it runs on
// the whole module alongside the first search.  The four searches share one
// loop, so their LDS reads are in flight together.
template<int W>
__device__ inline void forward_windows(
  const ModuleData& mo, const ModuleData& me, float tx, float ty, const EventHits& H, float h0x, float h0y,
  float h0z, int* of, int* ol, int* ef, int* el)
{
  const float dzo = mo.z - h0z;
  const float xo = h0x + tx * dzo;
  const float yo = h0y + ty * dzo;
  const float dze = me.z - h0z;
  const float xe = h0x + tx * dze;
  const float ye = h0y + ty * dze;
  const float po = phi_odd(xo, yo);
  const float pe = phi_even(xe, ye);
  const float m = kForwardPhiTolerance;
  bool fo = false, fe = false;
  // first candidate (l, r) and the second search (u, v) on each module
  int lo = 0, ro = (int) (mo.num - 1u), le = 0, re = (int) (me.num - 1u);
  int uo = 0, vo = ro, ue = 0, ve = re;
  const float loo = po - m, loe = pe - m;
  const float hio = po + m, hie = pe + m;
  while (lo < ro || le < re || uo < vo || ue < ve) {
    if (lo < ro) {
      const int k = (lo + ro) / 2;
      const float v = H.template vphi<W>(mo.start + (uint32_t) k);
      const float d = po - v;
      fo |= fabsf(d) < m;
      if (loo > v) lo = k + 1;
      else ro = k;
    }
    if (le < re) {
      const int k = (le + re) / 2;
      const float v = H.template vphi<W>(me.start + (uint32_t) k);
      const float d = pe - v;
      fe |= fabsf(d) < m;
      if (loe > v) le = k + 1;
      else re = k;
    }
    if (uo < vo) {
      const int k = (uo + vo) / 2;
      if (hio > H.template vphi<W>(mo.start + (uint32_t) k)) uo = k + 1;
      else vo = k;
    }
    if (ue < ve) {
      const int k = (ue + ve) / 2;
      if (hie > H.template vphi<W>(me.start + (uint32_t) k)) ue = k + 1;
      else ve = k;
    }
  }
  // the second search from the first candidate on
  const int so = uo > lo ? uo : lo, se = ue > le ? ue : le;
  const float vo_first = H.template vphi<W>(mo.start + (uint32_t) lo);
  const float ve_first = H.template vphi<W>(me.start + (uint32_t) le);
  const float vo_second = H.template vphi<W>(mo.start + (uint32_t) so);
  const float ve_second = H.template vphi<W>(me.start + (uint32_t) se);
  {
    const float d_o = po - vo_first, d_e = pe - ve_first;
    fo |= fabsf(d_o) < m;
    fe |= fabsf(d_e) < m;
  }
  int lasto = -1, laste = -1;
  if (fo) {
    const float d = po - vo_second;
    const int l = fabsf(d) < m ? so - lo + 1 : so - lo;
    const int first = lo + (int) mo.start;
    lasto = l == 0 ? first + 1 : first + l;
    *of = first;
  }
  else {
    *of = -1;
  }
  if (fe) {
    const float d = pe - ve_second;
    const int l = fabsf(d) < m ? se - le + 1 : se - le;
    const int first = le + (int) me.start;
    laste = l == 0 ? first + 1 : first + l;
    *ef = first;
  }
  else {
    *ef = -1;
  }
  *ol = lasto;
  *el = laste;
}

// The first kTtfLds entries of a step stay in LDS (double-buffered: the
// step's entries and those pushed for the next step); the rest go through
// the global ring.
constexpr uint32_t kTtfLds = 128;

struct Ring {
  uint32_t* base; // 2 words per entry
  uint2* cur;     // LDS: entries [0, kTtfLds) of the step being forwarded
  uint2* next;    // LDS: entries [0, kTtfLds) pushed for the next step
  __device__ __forceinline__ uint2 entry(uint32_t begin, uint32_t idx) const { return idx < kTtfLds ? cur[idx] : get(begin + idx); }
  __device__ __forceinline__ void push(uint32_t begin, uint32_t slot, uint32_t w0, uint32_t w1) const
  {
    if (slot < kTtfLds) next[slot] = make_uint2(w0, w1);
    else put(begin + slot, w0, w1);
  }
  __device__ void swap()
  {
    uint2* t = cur;
    cur = next;
    next = t;
  }
  __device__ __forceinline__ uint2 get(uint32_t pos) const
  {
    const uint32_t p = pos % VELO_TTF_MODULO;
    return make_uint2(base[2 * p], base[2 * p + 1]);
  }
  __device__ __forceinline__ void put(uint32_t pos, uint32_t w0, uint32_t w1) const
  {
    const uint32_t p = pos % VELO_TTF_MODULO;
    base[2 * p] = w0;
    base[2 * p + 1] = w1;
  }
};

struct SearchState {
  uint32_t n_tracks;
  uint32_t n_weak;
  uint32_t n_tracklets;
  uint32_t status;
  uint32_t next_begin; // ring position of the first entry of the next step
  uint32_t next_n;     // entries already pushed for the next step
  uint32_t cap;        // entries the next step may receive
};

// Appends the 3-hit weak track (h0, h1, h2) of lanes with `want`, in lane order.
__device__ inline void append_weak(bool want, uint32_t h0, uint32_t h1, uint32_t h2, uint8_t* weak, SearchState& st)
{
  const uint64_t m = ballot(want);
  if (!m) return;
  const uint32_t slot = st.n_weak + count_below(m);
  if (want && slot < VELO_MAX_WEAK_TRACKS) {
    uint16_t* w = reinterpret_cast<uint16_t*>(weak + (size_t) slot * VELO_TRACKLET_BYTES);
    w[0] = (uint16_t) h0;
    w[1] = (uint16_t) h1;
    w[2] = (uint16_t) h2;
  }
  const uint32_t cnt = (uint32_t) __popcll(m);
  if (st.n_weak + cnt > VELO_MAX_WEAK_TRACKS) {
    st.status |= VELO_STATUS_WEAK_OVERFLOW;
    st.n_weak = VELO_MAX_WEAK_TRACKS;
  }
  else {
    st.n_weak += cnt;
  }
}

// Pushes (w0, w1) of lanes with `want` to the next step's entries, in lane order.
__device__ inline void push_ttf(bool want, uint32_t w0, uint32_t w1, const Ring& ring, SearchState& st)
{
  const uint64_t m = ballot(want);
  if (!m) return;
  const uint32_t slot = st.next_n + count_below(m);
  if (want && slot < st.cap) ring.push(st.next_begin, slot, w0, w1);
  const uint32_t cnt = (uint32_t) __popcll(m);
  if (st.next_n + cnt > st.cap) {
    st.status |= VELO_STATUS_TTF_OVERFLOW;
    st.next_n = st.cap;
  }
  else {
    st.next_n += cnt;
  }
}

// LDS of one seeding batch: per staged h1 the first (h1, h0) item and the
// best packed key found so far.
struct SeedScratch {
  uint32_t item_begin[kWave + 1];
  uint64_t best[kWave];
  uint32_t mark[kWave]; // per item pass: (pass tag << 8) | (h1 slot + 1) at each h1's first item
  uint32_t tag;         // item passes so far (tags of the marks; mark[] starts zeroed)
};

// h1 hits staged for a seeding batch with their h0 and h2 windows (first |
// last << 16 as fill_candidates writes them, read once from global memory).
struct SeedStage {
  uint16_t h1[2 * kWave];
  uint32_t c0[2 * kWave];
  uint32_t c2[2 * kWave];
};
__device__ inline uint32_t window_pair(const int16_t* c, uint32_t h)
{
  return *reinterpret_cast<const uint32_t*>(c + 2 * h);
}
__device__ inline int window_first(uint32_t w) { return (int) (int16_t) (w & 0xFFFFu); }
__device__ inline int window_last(uint32_t w) { return (int) (int16_t) (w >> 16); }

// Items of one seeding batch must fit the key's 17-bit item field.
constexpr uint32_t kMaxSeedItems = 1u << 17;

// Best h2 for the item (h1, h0): the inner loop of seed_best for one h0.
// Returns the packed key (scatter bits << 33 | item << 16 | h2), or ~0
// when no h2 qualifies.  Scatters are non-negative (sign bit 0), so their
// bit patterns order like the floats; among equal scatters the lower item
// (smaller h0), then the first h2, wins, as in seed_best's loop order.
template<int W>
__device__ inline uint64_t seed_item(
  uint32_t h1, int h0, uint32_t item, const EventHits& H, int f2, int l2, const UsedSet& used)
{
  constexpr int W0 = W == kWindowAll ? kWindowAll : kWindowNone; // h0 modules: in the window only in mode all
  if (used.test(h0)) return ~0ull;
  const float h1x = H.template vx<W>(h1), h1y = H.template vy<W>(h1), h1z = H.template vz<W>(h1);
  const float h0x = H.template vx<W0>(h0), h0y = H.template vy<W0>(h0), h0z = H.template vz<W0>(h0);
  const float dmax = kMaxSlope * (h0z - h1z);
  if (!(fabsf(h1x - h0x) < dmax && fabsf(h1y - h0y) < dmax)) return ~0ull;
  const float dz10 = h1z - h0z;
  float best = kFltMax;
  uint32_t bb2 = 0;
  for (int h2 = f2; h2 < l2; ++h2) {
    if (used.test(h2)) continue;
    const float h2x = H.template vx<W>(h2), h2y = H.template vy<W>(h2), h2z = H.template vz<W>(h2);
    const float z2_tz = (h2z - h0z) / dz10;
    const float x = h0x + (h1x - h0x) * z2_tz;
    const float y = h0y + (h1y - h0y) * z2_tz;
    const float dx = x - h2x;
    const float dy = y - h2y;
    if (!(fabsf(dx) < kTolerance && fabsf(dy) < kTolerance)) continue;
    const float dz21 = h2z - h1z;
    const float scatter_denom2 = 1.0f / (dz21 * dz21);
    const float scatter_num = (dx * dx) + (dy * dy);
    const float scatter = scatter_num * scatter_denom2;
    if (scatter < kMaxScatterSeeding && scatter < best) {
      best = scatter;
      bb2 = (uint32_t) h2;
    }
  }
  if (!(best < kFltMax)) return ~0ull;
  return ((uint64_t) __float_as_uint(best) << 33) | ((uint64_t) item << 16) | bb2;
}

// Seeds the first k (<= 64) staged h1 hits.  The (h1, h0) pairs of the
// batch are spread over the lanes (one pair per lane, its h2 window in a
// loop); each pair's best is folded into its h1 with an LDS 64-bit minimum.
// The result per h1 equals seed_best's, which runs instead (one h1 per
// lane) for a batch with more pairs than the key can number.
template<int W>
__device__ inline uint32_t seed_batch(
  const SeedStage& stage, uint32_t k, const EventHits& H, const int16_t* h0c, const int16_t* h2c, const UsedSet& used,
  const Ring& ring, SearchState& st, SeedScratch& sc)
{
  const uint32_t lane = lane_id();
  uint32_t h1 = 0, c = 0;
  int f0 = 0;
  if (lane < k) {
    h1 = stage.h1[lane];
    f0 = window_first(stage.c0[lane]);
    const int l0 = window_last(stage.c0[lane]);
    c = l0 > f0 ? (uint32_t) (l0 - f0) : 0u;
  }
  const uint32_t inc = wave_inclusive_scan(c);
  const uint32_t items = __builtin_amdgcn_readlane(inc, kWave - 1);
  if (items >= kMaxSeedItems) {
    bool emit = false;
    uint32_t b0 = 0, b2 = 0;
    if (lane < k) emit = seed_best<W>(h1, H, h0c, h2c, used, &b0, &b2) < kFltMax;
    const uint32_t cnt = (uint32_t) __popcll(ballot(emit));
    st.n_tracklets += cnt;
    push_ttf(emit, h1 | (b2 << 16), b0 | (3u << 16) | kTtfTrackletFlag, ring, st);
    return cnt;
  }
  sc.item_begin[lane] = inc - c;
  if (lane == kWave - 1) sc.item_begin[kWave] = inc;
  sc.best[lane] = ~0ull;
  wave_sync();
  // every staged h1 has at least one item (its h0 window is not empty), so
  // the h1 of item i is the last one whose first item is at or before i:
  // each h1 marks its first item, and a running maximum carries the marks
  uint32_t owner = 0;
  for (uint32_t base = 0; base < items; base += kWave) {
    const uint32_t i = base + lane;
    const uint32_t my_begin = inc - c;
    const uint32_t tag = ++sc.tag;
    if (lane < k && c > 0 && my_begin - base < kWave) sc.mark[my_begin - base] = (tag << 8) | (lane + 1u);
    wave_sync();
    const uint32_t mk = sc.mark[lane];
    const uint32_t here = (mk >> 8) == (tag & 0xFFFFFFu) ? (mk & 0xFFu) : 0u;
    owner = max(wave_inclusive_max(here), __builtin_amdgcn_readlane(owner, kWave - 1));
    wave_sync(); // the marks are rewritten by the next pass
    if (i < items) {
      const uint32_t j = owner - 1u;
      const uint32_t hj = stage.h1[j];
      const int h0 = window_first(stage.c0[j]) + (int) (i - sc.item_begin[j]);
      const uint32_t c2 = stage.c2[j];
      const uint64_t key = seed_item<W>(hj, h0, i, H, window_first(c2), window_last(c2), used);
      if (key != ~0ull) atomicMin(&sc.best[j], key);
    }
  }
  wave_sync();
  bool emit = false;
  uint32_t b0 = 0, b2 = 0;
  if (lane < k) {
    const uint64_t key = sc.best[lane];
    emit = key != ~0ull;
    const uint32_t item = (uint32_t) (key >> 16) & 0xFFFFu;
    b0 = (uint32_t) (f0 + (int) (item - sc.item_begin[lane]));
    b2 = (uint32_t) key & 0xFFFFu;
  }
  wave_sync(); // the scratch is reused by the next batch
  const uint32_t cnt = (uint32_t) __popcll(ballot(emit));
  st.n_tracklets += cnt;
  push_ttf(emit, h1 | (b2 << 16), b0 | (3u << 16) | kTtfTrackletFlag, ring, st);
  return cnt;
}

// TrackSeeding.cu:7-363: h1 candidates of module_data[2] then [3].
// The h0 / h2 windows of the first 64 h1 hits of each seeding module, read
// at the start of a step so that their latency overlaps forwarding.
struct SeedPrefetch {
  uint32_t c0[2], c2[2];
};

__device__ inline SeedPrefetch prefetch_seed_windows(const ModuleData* md, const int16_t* h0c, const int16_t* h2c)
{
  const uint32_t lane = lane_id();
  SeedPrefetch p;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    p.c0[s] = p.c2[s] = 0u;
    if (lane < md[2 + s].num) {
      p.c0[s] = window_pair(h0c, md[2 + s].start + lane);
      p.c2[s] = window_pair(h2c, md[2 + s].start + lane);
    }
  }
  return p;
}

template<int W>
__device__ inline uint32_t seeding(
  const ModuleData* md, SeedStage& stage, const EventHits& H, const int16_t* h0c, const int16_t* h2c,
  const UsedSet& used, const Ring& ring, SearchState& st, SeedScratch& sc, const SeedPrefetch& pf)
{
  const uint32_t lane = lane_id();
  uint32_t staged = 0, tracklets = 0;
#pragma unroll
  for (int side = 2; side <= 3; ++side) {
    const uint32_t side_start = md[side].start, side_num = md[side].num;
    for (uint32_t base = 0; base < side_num; base += kWave) {
      const uint32_t r = base + lane;
      const uint32_t h1 = side_start + r;
      bool cand = false;
      uint32_t c0 = 0, c2 = 0;
      if (r < side_num) {
        if (base == 0) {
          c0 = pf.c0[side - 2];
          c2 = pf.c2[side - 2];
        }
        else {
          c0 = window_pair(h0c, h1);
          c2 = window_pair(h2c, h1);
        }
        cand = !used.test((int) h1) && window_first(c0) != -1 && window_first(c2) != -1;
      }
      const uint64_t m = ballot(cand);
      if (cand) {
        const uint32_t slot = staged + count_below(m);
        stage.h1[slot] = (uint16_t) h1;
        stage.c0[slot] = c0;
        stage.c2[slot] = c2;
      }
      staged += (uint32_t) __popcll(m);
      wave_sync();
      if (staged >= kWave) {
        tracklets += seed_batch<W>(stage, kWave, H, h0c, h2c, used, ring, st, sc);
        wave_sync();
        const uint32_t rest = staged - kWave;
        uint16_t carry = 0;
        uint32_t carry0 = 0, carry2 = 0;
        if (lane < rest) {
          carry = stage.h1[kWave + lane];
          carry0 = stage.c0[kWave + lane];
          carry2 = stage.c2[kWave + lane];
        }
        wave_sync();
        if (lane < rest) {
          stage.h1[lane] = carry;
          stage.c0[lane] = carry0;
          stage.c2[lane] = carry2;
        }
        staged = rest;
        wave_sync();
      }
    }
  }
  if (staged) tracklets += seed_batch<W>(stage, staged, H, h0c, h2c, used, ring, st, sc);
  wave_sync();
  return tracklets;
}

// TrackForwarding.cu:41-182 over the entries [cur_begin, cur_begin + cur_n).
template<int W>
__device__ inline void forwarding(
  const ModuleData* md, uint32_t cur_begin, uint32_t cur_n, const EventHits& H, const UsedSet& used, const Ring& ring,
  uint8_t* tracks, uint8_t* weak, SearchState& st)
{
  const uint32_t lane = lane_id();
  for (uint32_t base = 0; base < cur_n; base += kWave) {
    const uint32_t idx = base + lane;
    const bool active = idx < cur_n;
    uint2 ent = make_uint2(0, 0);
    if (active) ent = ring.entry(cur_begin, idx);
    const bool tracklet = (ent.y & kTtfTrackletFlag) != 0u;
    const uint32_t skipped = (ent.y >> 28) & 7u;
    const uint32_t nh = (ent.y >> 16) & 0x1Fu;
    const uint32_t ha = ent.x & 0xFFFFu; // hits[n-2]
    const uint32_t hb = ent.x >> 16;     // hits[n-1]
    const uint32_t low = ent.y & 0xFFFFu; // first hit or track number

    bool found = false;
    uint32_t best_h2 = 0;
    // the entry's last two hits are in the window for every lane (the
    // common case): plain LDS reads
    const bool all_in_window = ballot(active && (ha - H.lo >= H.n || hb - H.lo >= H.n)) == 0;
    if (active) {
      float h0x, h0y, h0z, h1x, h1y, h1z;
      if (W != kWindowNone && all_in_window) {
        h0x = H.template vx<kWindowAll>(ha);
        h0y = H.template vy<kWindowAll>(ha);
        h0z = H.template vz<kWindowAll>(ha);
        h1x = H.template vx<kWindowAll>(hb);
        h1y = H.template vy<kWindowAll>(hb);
        h1z = H.template vz<kWindowAll>(hb);
      }
      else {
        h0x = H.x(ha), h0y = H.y(ha), h0z = H.z(ha);
        h1x = H.x(hb), h1y = H.y(hb), h1z = H.z(hb);
      }
      const float td = 1.0f / (h1z - h0z);
      const float txn = h1x - h0x;
      const float tyn = h1y - h0y;
      const float tx = txn * td;
      const float ty = tyn * td;
      int of, ol, ef, el;
      forward_windows<W>(md[4], md[5], tx, ty, H, h0x, h0y, h0z, &of, &ol, &ef, &el);
      const int n_odd = ol - of;
      const int n_all = n_odd + (el - ef);
      float best = kFltMax;
      for (int j = 0; j < n_all; ++j) {
        const int h2 = j < n_odd ? of + j : ef + j - n_odd;
        const float h2x = H.template vx<W>(h2), h2y = H.template vy<W>(h2), h2z = H.template vz<W>(h2);
        const float dz = h2z - h0z;
        const float predx = tx * dz;
        const float predy = ty * dz;
        const float xp = h0x + predx;
        const float dx = fabsf(xp - h2x);
        const float yp = h0y + predy;
        const float dy = fabsf(yp - h2y);
        // the scatter (a division) only matters inside both tolerances
        if (!(dx < kTolerance && dy < kTolerance)) continue;
        const float dz21 = h2z - h1z;
        const float scatter_denom2 = 1.0f / (dz21 * dz21);
        const float scatter_num = (dx * dx) + (dy * dy);
        const float scatter = scatter_num * scatter_denom2;
        if (scatter < kMaxScatterForwarding && scatter < best) {
          best = scatter;
          best_h2 = (uint32_t) h2;
        }
      }
      found = best != kFltMax;
    }

    // hits become used when forwarded (TrackForwarding.cu:128-147)
    if (found) {
      used.set(best_h2);
      if (tracklet) {
        used.set(low);
        used.set(ha);
        used.set(hb);
      }
    }

    // new tracks for promoted tracklets, numbered in entry order
    const bool alloc = found && tracklet;
    const uint64_t am = ballot(alloc);
    uint32_t trackno = low;
    bool track_ok = found;
    if (am) {
      const uint32_t slot = st.n_tracks + count_below(am);
      if (alloc) {
        trackno = slot;
        track_ok = slot < VELO_MAX_TRACKS;
      }
      const uint32_t cnt = (uint32_t) __popcll(am);
      if (st.n_tracks + cnt > VELO_MAX_TRACKS) {
        st.status |= VELO_STATUS_TRACKS_OVERFLOW;
        st.n_tracks = VELO_MAX_TRACKS;
      }
      else {
        st.n_tracks += cnt;
      }
    }
    if (found && track_ok) {
      uint16_t* t = reinterpret_cast<uint16_t*>(tracks + (size_t) trackno * VELO_TRACK_HITS_BYTES);
      if (tracklet) {
        t[0] = 4;
        t[1] = (uint16_t) low;
        t[2] = (uint16_t) ha;
        t[3] = (uint16_t) hb;
        t[4] = (uint16_t) best_h2;
      }
      else {
        t[1 + nh] = (uint16_t) best_h2;
        t[0] = (uint16_t) (nh + 1);
      }
    }

    // next step's entries: forwarded tracks and once-skipped tracks, in entry order
    const bool skip = active && !found && skipped < kMaxSkippedModules;
    const bool push = (found && track_ok) || skip;
    uint32_t w0 = 0, w1 = 0;
    if (found) {
      w0 = hb | (best_h2 << 16);
      w1 = trackno | ((nh + 1) << 16);
    }
    else if (skip) {
      w0 = ent.x;
      w1 = (ent.y & 0x8FFFFFFFu) | ((skipped + 1u) << 28);
    }
    push_ttf(push, w0, w1, ring, st);

    // tracklets that cannot be followed any more become weak tracks
    const bool to_weak = active && !found && !skip && nh == 3;
    append_weak(to_weak, low, ha, hb, weak, st);
  }
}

constexpr uint32_t kDefaultWindowHits = 512;

// Window capacity for a requested tile: the largest power of two not above
// it (the window is a ring indexed by hit number mod capacity).
static uint32_t window_hits(uint32_t lds_tile)
{
  const uint32_t t = lds_tile ? lds_tile : kDefaultWindowHits;
  uint32_t p = 1;
  while (p <= t / 2) p *= 2;
  return p;
}

// One step whose modules do not fit the window: every hit from global memory.
__device__ inline uint32_t step_global(
  const ModuleData* md, uint32_t cur_begin, uint32_t cur_n, const EventHits& H, const UsedSet& used,
  const Ring& ring, uint8_t* tracks, uint8_t* weak, SearchState& st, SeedStage& stage, const int16_t* h0c,
  const int16_t* h2c, SeedScratch& sc, const SeedPrefetch& pf)
{
  forwarding<kWindowNone>(md, cur_begin, cur_n, H, used, ring, tracks, weak, st);
  wave_sync();
  return seeding<kWindowNone>(md, stage, H, h0c, h2c, used, ring, st, sc, pf);
}

__global__ void __launch_bounds__(64) search_by_triplet_kernel(velo_search_by_triplet_args a)
{
  __shared__ uint32_t used_bits[kUsedWords];
  __shared__ SeedStage stage;
  __shared__ uint2 ttf_lds[2 * kTtfLds];
  __shared__ SeedScratch seed_scratch;

  const uint32_t N = a.n_events;
  seed_scratch.mark[lane_id()] = 0u;
  if (lane_id() == 0) seed_scratch.tag = 0u;
  wave_sync();
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t lane = lane_id();
    const uint32_t total = a.dev_module_cluster_start[VELO_N_MODULES * N];
    const uint32_t* starts = a.dev_module_cluster_start + e * VELO_N_MODULES;
    const uint32_t* nums = a.dev_module_cluster_num + e * VELO_N_MODULES;
    const uint32_t off = starts[0];
    const uint32_t ev_hits = starts[VELO_N_MODULES] - off;
    const uint32_t cap = a.tuning.lds_tile;
    EventHits H;
    H.gy = reinterpret_cast<const float*>(a.dev_velo_cluster_container + off);
    H.gz = reinterpret_cast<const float*>(a.dev_velo_cluster_container + total + off);
    H.gphi = reinterpret_cast<const float*>(a.dev_velo_cluster_container + 4 * (size_t) total + off);
    H.gx = reinterpret_cast<const float*>(a.dev_velo_cluster_container + 5 * (size_t) total + off);
    H.cap = cap;
    H.lo = 0;
    H.n = 0;
    // window of the six modules a step searches, first .. first-5
    // (contiguous in the container): seeding's h0, h1, h2 and forwarding's
    // phi windows; the tracks' previous hits are read from global memory
    ModuleTable mt {0u, 0u, 0.0f};
    if (lane <= VELO_N_MODULES) mt.start = starts[lane] - off; // lane 52: the event's end
    if (lane < VELO_N_MODULES) {
      mt.num = nums[lane];
      mt.z = a.dev_velo_module_zs[lane];
    }
    // the step's window: all six modules when they fit, else the four from
    // md[5] to md[2] (contiguous: first - 5 .. first - 2), else none
    const auto stage_window = [&](uint32_t first) -> int {
      const uint32_t lo = __builtin_amdgcn_readlane(mt.start, first - 5);
      const uint32_t hi_all = __builtin_amdgcn_readlane(mt.start, first + 1);
      const uint32_t hi_part = __builtin_amdgcn_readlane(mt.start, first - 1);
      const bool all = hi_all - lo <= H.cap;
      H.stage(lo, all ? hi_all : hi_part);
      return all ? kWindowAll : (H.n ? kWindowNoH0 : kWindowNone);
    };
    const int16_t* h0c = a.dev_h0_candidates + 2 * (size_t) off;
    const int16_t* h2c = a.dev_h2_candidates + 2 * (size_t) off;
    uint8_t* tracks = static_cast<uint8_t*>(a.dev_tracks) + (size_t) e * VELO_MAX_TRACKS * VELO_TRACK_HITS_BYTES;
    uint8_t* weak = static_cast<uint8_t*>(a.dev_weak_tracks) + (size_t) e * VELO_MAX_WEAK_TRACKS * VELO_TRACKLET_BYTES;
    Ring ring {a.dev_tracks_to_follow + (size_t) e * VELO_TTF_MODULO * VELO_TTF_ENTRY_WORDS, ttf_lds, ttf_lds + kTtfLds};

    UsedSet used {used_bits, a.dev_hit_used + off, ev_hits <= kUsedBits};
    if (used.in_lds) {
      for (uint32_t w = lane; w < kUsedWords; w += kWave) used_bits[w] = 0;
    }
    else {
      for (uint32_t i = lane; i < ev_hits; i += kWave) used.bytes[i] = 0;
    }
    wave_sync();

    SearchState st {0, 0, 0, 0, 0, 0, VELO_TTF_MODULO};
    ModuleData md[6];
    // steps with first module 51, 49, ..., 5; the first has nothing to
    // forward (ProcessModules.cu:22-59 seeds it, then the pairs loop)
    uint32_t prev_tracklets = 0, cur_begin = 0, cur_n = 0;
    for (uint32_t first = VELO_N_MODULES - 1;; first -= 2) {
      load_modules(md, first, mt);
      const SeedPrefetch pf = prefetch_seed_windows(md, h0c, h2c);
      const int mode = stage_window(first);
      st.next_begin = (cur_begin + cur_n) % VELO_TTF_MODULO;
      st.next_n = 0;
      st.cap = VELO_TTF_MODULO - cur_n;
      uint32_t tl;
      if (mode == kWindowAll) {
        // an empty forwarding module still has its first slot probed
        // (VeloTools.cuh:6-26 inspects array[0] for size 0), which may lie
        // past the window: that step forwards from global memory
        if (md[4].num && md[5].num) forwarding<kWindowAll>(md, cur_begin, cur_n, H, used, ring, tracks, weak, st);
        else forwarding<kWindowNone>(md, cur_begin, cur_n, H, used, ring, tracks, weak, st);
        wave_sync(); // hit_used marks of forwarding are read by seeding
        tl = seeding<kWindowAll>(md, stage, H, h0c, h2c, used, ring, st, seed_scratch, pf);
      }
      else if (mode == kWindowNoH0) {
        if (md[4].num && md[5].num) forwarding<kWindowNoH0>(md, cur_begin, cur_n, H, used, ring, tracks, weak, st);
        else forwarding<kWindowNone>(md, cur_begin, cur_n, H, used, ring, tracks, weak, st);
        wave_sync();
        tl = seeding<kWindowNoH0>(md, stage, H, h0c, h2c, used, ring, st, seed_scratch, pf);
      }
      else {
        tl = step_global(md, cur_begin, cur_n, H, used, ring, tracks, weak, st, stage, h0c, h2c, seed_scratch, pf);
      }
      if (tl + prev_tracklets > VELO_TTF_MODULO) st.status |= VELO_STATUS_TRACKLETS_OVERFLOW;
      prev_tracklets = tl;
      cur_begin = st.next_begin;
      cur_n = st.next_n;
      ring.swap();
      wave_sync();
      if (first < 7) break;
    }

    // ProcessModules.cu:141-162: tracklets still followed become weak tracks
    for (uint32_t base = 0; base < cur_n; base += kWave) {
      const uint32_t idx = base + lane;
      uint2 ent = make_uint2(0, 0);
      if (idx < cur_n) ent = ring.entry(cur_begin, idx);
      const bool tl = idx < cur_n && (ent.y & kTtfTrackletFlag);
      append_weak(tl, ent.y & 0xFFFFu, ent.x & 0xFFFFu, ent.x >> 16, weak, st);
    }

    if (lane == 0) {
      a.dev_atomics_velo[e] = (int32_t) st.n_tracks;
      const uint32_t ip = N + e * (VELO_NUM_ATOMICS - 1u);
      a.dev_atomics_velo[ip + 0] = (int32_t) st.n_weak;
      a.dev_atomics_velo[ip + 1] = (int32_t) st.n_tracklets;
      a.dev_atomics_velo[ip + 2] = 0;
      a.dev_atomics_velo[ip + 3] = 0;
      if (a.dev_velo_event_status && st.status) atomicOr(a.dev_velo_event_status + e, st.status);
    }
    if (used.in_lds) {
      wave_sync();
      for (uint32_t i = lane; i < ev_hits; i += kWave) used.bytes[i] = (uint8_t) used.test((int) i);
    }
  });
}

// WeakTracksAdder.cu:8-86
template<typename Hits>
__device__ inline float weak_chi2(const Hits& H, uint32_t a0, uint32_t a1, uint32_t a2)
{
  float s0 = 0.0f, sx = 0.0f, sz = 0.0f, sxz = 0.0f, sz2 = 0.0f;
  float u0 = 0.0f, uy = 0.0f, uz = 0.0f, uyz = 0.0f, uz2 = 0.0f;
  const uint32_t hs[3] = {a0, a1, a2};
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    const uint32_t i = hs[h];
    const float x = H.x(i), y = H.y(i), z = H.z(i);
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
  const float x0 = (sx * sz2 - sxz * sz) * dens;
  const float denu = 1.0f / (uz2 * u0 - uz * uz);
  const float ty = (uyz * u0 - uy * uz) * denu;
  const float y0 = (uy * uz2 - uyz * uz) * denu;
  float ch = 0.0f;
  int ndof = -4;
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    const uint32_t i = hs[h];
    const float z = H.z(i);
    const float x = x0 + tx * z;
    const float y = y0 + ty * z;
    const float dx = x - H.x(i);
    const float dy = y - H.y(i);
    ch += dx * dx * kParamW + dy * dy * kParamW;
    ndof += 2;
  }
  return ch / (float) ndof;
}

__global__ void __launch_bounds__(64) weak_tracks_adder_kernel(velo_weak_tracks_adder_args a)
{
  const uint32_t N = a.n_events;
  for_each_event(a.tuning, a.n_events, [&](const uint32_t e) {
    const uint32_t lane = lane_id();
    const uint32_t total = a.dev_module_cluster_start[VELO_N_MODULES * N];
    const uint32_t off = a.dev_module_cluster_start[e * VELO_N_MODULES];
    GlobalHits H;
    H.gy = reinterpret_cast<const float*>(a.dev_velo_cluster_container + off);
    H.gz = reinterpret_cast<const float*>(a.dev_velo_cluster_container + total + off);
    H.gx = reinterpret_cast<const float*>(a.dev_velo_cluster_container + 5 * (size_t) total + off);
    const uint8_t* used = a.dev_hit_used + off;
    uint8_t* tracks = static_cast<uint8_t*>(a.dev_tracks) + (size_t) e * VELO_MAX_TRACKS * VELO_TRACK_HITS_BYTES;
    const uint8_t* weak =
      static_cast<const uint8_t*>(a.dev_weak_tracks) + (size_t) e * VELO_MAX_WEAK_TRACKS * VELO_TRACKLET_BYTES;
    const uint32_t ip = N + e * (VELO_NUM_ATOMICS - 1u);
    uint32_t n_weak = (uint32_t) a.dev_atomics_velo[ip];
    if (n_weak > VELO_MAX_WEAK_TRACKS) n_weak = VELO_MAX_WEAK_TRACKS;
    uint32_t n_tracks = (uint32_t) a.dev_atomics_velo[e];
    uint32_t status = 0;
    for (uint32_t base = 0; base < n_weak; base += kWave) {
      const uint32_t w = base + lane;
      bool accept = false;
      uint32_t t0 = 0, t1 = 0, t2 = 0;
      if (w < n_weak) {
        const uint16_t* t = reinterpret_cast<const uint16_t*>(weak + (size_t) w * VELO_TRACKLET_BYTES);
        t0 = t[0];
        t1 = t[1];
        t2 = t[2];
        const bool any_used = used[t0] || used[t1] || used[t2];
        const float chi2 = weak_chi2(H, t0, t1, t2);
        accept = !any_used && chi2 < kMaxChi2;
      }
      const uint64_t m = ballot(accept);
      const uint32_t slot = n_tracks + count_below(m);
      if (accept && slot < VELO_MAX_TRACKS) {
        uint16_t* d = reinterpret_cast<uint16_t*>(tracks + (size_t) slot * VELO_TRACK_HITS_BYTES);
        d[0] = 3;
        d[1] = (uint16_t) t0;
        d[2] = (uint16_t) t1;
        d[3] = (uint16_t) t2;
      }
      const uint32_t cnt = (uint32_t) __popcll(m);
      if (n_tracks + cnt > VELO_MAX_TRACKS) {
        status |= VELO_STATUS_TRACKS_OVERFLOW;
        n_tracks = VELO_MAX_TRACKS;
      }
      else {
        n_tracks += cnt;
      }
    }
    if (lane == 0) {
      a.dev_atomics_velo[e] = (int32_t) n_tracks;
      if (a.dev_velo_event_status && status) atomicOr(a.dev_velo_event_status + e, status);
    }
  });
}

} // namespace velo

using namespace velo;

extern "C" {

int velo_search_by_triplet_launch(const velo_search_by_triplet_args* a, void* stream)
{
  if (!a || !a->dev_velo_cluster_container || !a->dev_module_cluster_start || !a->dev_module_cluster_num ||
      !a->dev_tracks || !a->dev_tracks_to_follow || !a->dev_weak_tracks || !a->dev_hit_used ||
      !a->dev_atomics_velo || !a->dev_h0_candidates || !a->dev_h2_candidates || !a->dev_velo_module_zs)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  velo_search_by_triplet_args args = *a;
  args.tuning.lds_tile = window_hits(args.tuning.lds_tile);
  hipLaunchKernelGGL(search_by_triplet_kernel, dim3(event_grid(args.tuning, args.n_events)), dim3(kWave), 4 * sizeof(float) * args.tuning.lds_tile,
                     (hipStream_t) stream, args);
  return (int) hipGetLastError();
}

uint32_t velo_search_by_triplet_resident_workgroups(uint32_t lds_tile)
{
  const uint32_t tile = window_hits(lds_tile);
  int device = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&device) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, search_by_triplet_kernel, kWave,
                                                   4 * sizeof(float) * tile) != hipSuccess)
    return 0;
  return (uint32_t) (cus * per_cu);
}

int velo_weak_tracks_adder_launch(const velo_weak_tracks_adder_args* a, void* stream)
{
  if (!a || !a->dev_velo_cluster_container || !a->dev_module_cluster_start || !a->dev_tracks ||
      !a->dev_weak_tracks || !a->dev_hit_used || !a->dev_atomics_velo)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(weak_tracks_adder_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(kWave), 0, (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

} // extern "C"
