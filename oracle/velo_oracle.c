/*
 * velo_oracle.c -- single-threaded CPU restatement of the reference's VELO
 * path.  TEST INFRASTRUCTURE ONLY (see velo_oracle.h): the product never links
 * or calls this file.
 *
 * Compile with -ffp-contract=off: every float expression below is evaluated
 * left to right with one IEEE rounding per operation, exactly as the HIP
 * kernels evaluate it, so float outputs match bit for bit.
 *
 * Output order.  The reference allocates cluster, track and weak-track slots
 * with atomics, so its slot order is arbitrary (any order is a valid output).
 * The oracle fixes one "natural" order:
 *   clusters   per module: banks in order, SPs in bank order, then the SP's
 *              clusters (isolated SP: LUT cluster 0 then 1; otherwise the
 *              candidate from the low nibble then the high nibble);
 *   tracks     forwarding outputs in tracks-to-follow order, per step, then
 *              weak tracks accepted by the adder in weak-track order;
 *   seeding    h1 candidates of module_data[2] then [3], ascending index, and
 *              for equal scatter the lexicographically smallest (h0, h2).
 * The reference's result set equals the oracle's whenever no floating-point
 * tie occurs in seeding (TrackSeeding.cu:340-345 picks the first strict
 * minimum over its thread layout) and no capacity overflows (status bits).
 */
#include "velo_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* constants                                                                */
/* ------------------------------------------------------------------------ */

/* VeloDefinitions.cuh:45-77 */
static const float k_phi_extrapolation = 0.062f;
static const float k_forward_phi_tolerance = 0.052f;
static const float k_max_slope = 0.4f;
static const float k_tolerance = 0.6f;
static const float k_max_scatter_seeding = 0.004f;
static const float k_max_scatter_forwarding = 0.004f;
static const unsigned k_max_skipped_modules = 1u;
static const float k_param_w = 3966.94f;
static const float k_param_w_inverted = 0.000252083f; /* VeloDefinitions.cuh:74 */
static const float k_max_chi2 = 20.0f;
static const float k_pixel_size = 0.055f; /* VeloDefinitions.cuh:29 */

/* ClusteringDefinitions.cuh:14-21 */
#define MASK_BOTTOM 0xFFFEFFFFu
#define MASK_TOP 0xFFFF7FFFu
#define MASK_TOP_LEFT 0x7FFF7FFFu
#define MASK_BOTTOM_RIGHT 0xFFFEFFFEu
#define MASK_LTR_TOP_RIGHT 0x7FFF0000u
#define MASK_RTL_BOTTOM_LEFT 0x0000FFFEu
#define MAX_CLUSTERING_ITERATIONS 12

/* Constants.cu:36 */
static const uint8_t k_candidate_ks[9] = {0, 0, 1, 4, 4, 5, 5, 5, 5};

/* ------------------------------------------------------------------------ */
/* phi: atan2 from a minimax polynomial                                    */
/* ------------------------------------------------------------------------ */

/* atan(t) = t * P(t^2) on [0,1]; coefficients fitted by iteratively
 * reweighted least squares on the relative error (max 2.4e-9 in exact
 * arithmetic, <=1.5 ulp evaluated in float by Horner).  The reference calls
 * CUDA's atan2 (VeloTools.cu:3-10); this is a different, documented
 * function shared bit-identically by the oracle and the kernels. */
static const float k_atan_c[10] = {
  1.0f, -0.3333328366279602f, 0.1999826729297638f, -0.14261946082115173f,
  0.10942086577415466f, -0.0837746188044548f, 0.05755440890789032f,
  -0.03080674819648266f, 0.010728039778769016f, -0.001754164695739746f};
static const float k_pi_f = 3.14159274101257324f;   /* CUDART_PI_F */
static const float k_pi_2_f = 1.57079637050628662f; /* float(pi/2) */

float oracle_atan2f(float y, float x)
{
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
    float p = k_atan_c[9];
    for (int k = 8; k >= 0; --k) {
      p = p * s;
      p = p + k_atan_c[k];
    }
    a = t * p;
  }
  if (ay > ax) a = k_pi_2_f - a;
  if (signbit(x)) a = k_pi_f - a;
  if (signbit(y)) a = -a;
  return a;
}

/* VeloTools.cu:3-10 */
float oracle_phi_odd(float x, float y) { return oracle_atan2f(y, x); }

float oracle_phi_even(float x, float y)
{
  const float phi = oracle_atan2f(y, x);
  const float shift = phi < 0.0f ? 2.0f * k_pi_f : 0.0f;
  return phi + shift;
}

/* ------------------------------------------------------------------------ */
/* raw format helpers (ClusteringDefinitions.cu:3-21, cuda/velo/readme.md)  */
/* ------------------------------------------------------------------------ */

static inline uint32_t rd32(const char* p)
{
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

typedef struct {
  uint32_t n_banks;
  uint32_t n_words;    /* the event's size in 32-bit words */
  const char* base;
  const char* offsets; /* n_banks+1 u32, bytes relative to payload */
  const char* payload;
} raw_event;

typedef struct {
  uint32_t sensor_index;
  uint32_t sp_count;
  const char* sp_words;
  int fits; /* header and SP words lie inside the event */
} raw_bank;

static raw_event make_event(const char* ev, uint32_t n_words)
{
  raw_event r;
  r.n_words = n_words;
  r.n_banks = n_words ? rd32(ev) : 0u;
  r.base = ev;
  r.offsets = ev + 4;
  r.payload = ev + 4 + 4 * (r.n_banks + 1);
  return r;
}

/* Bank b starts at word n_banks + 2 + offset / 4 (the decoder's word
   addressing).  A bank whose offset table, header or SP words would reach
   past the event does not fit: it is flagged and skipped like a mislabelled
   one (the reference reads on past the event).  Neither does a bank of 2^22
   SPs or more (16 MB; VELO banks hold a few hundred), a limit of the device
   decoder, which sums the SP counts of an event's banks in 32 bits. */
static raw_bank make_bank(const raw_event* e, uint32_t b)
{
  raw_bank k;
  k.sensor_index = 0xFFFFFFFFu;
  k.sp_count = 0;
  k.sp_words = NULL;
  k.fits = 0;
  if ((uint64_t) e->n_banks + 2u > e->n_words) return k;
  const uint64_t h = (uint64_t) e->n_banks + 2u + rd32(e->offsets + 4 * b) / 4u;
  if (h + 2u > e->n_words) return k;
  const char* p = e->base + 4 * h;
  k.sensor_index = rd32(p);
  k.sp_count = rd32(p + 4);
  k.sp_words = p + 8;
  k.fits = h + 2u + k.sp_count <= e->n_words && k.sp_count < (1u << 22);
  if (!k.fits) k.sp_count = 0;
  return k;
}

static inline uint32_t sp_word_at(const raw_bank* b, uint32_t i) { return rd32(b->sp_words + 4 * i); }

/* ClusteringDefinitions.cu:55-61 */
static inline uint32_t channel_id(uint32_t sensor, uint32_t chip, uint32_t col, uint32_t row)
{
  return (sensor << 18) | (chip << 16) | (col << 8) | row;
}
static inline uint32_t lhcb_id(uint32_t cid) { return 0x80000000u + cid; }

static inline uint32_t popc(uint32_t v) { return (uint32_t) __builtin_popcount(v); }

/* ------------------------------------------------------------------------ */
/* a3: SP pattern tables -- ClusteringCommon.cpp:3-101                      */
/* ------------------------------------------------------------------------ */

void oracle_sp_lut(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  memset(patterns, 0, 256);
  memset(sizes, 0, 256);
  for (int i = 0; i < 512; ++i) fx[i] = fy[i] = 0.0f;
  /* 8-connected neighbour deltas, same visiting order as the reference */
  static const int dx[8] = {-1, 0, 1, -1, 0, 1, -1, 1};
  static const int dy[8] = {-1, -1, -1, 1, 1, 1, 0, 0};
  for (unsigned sp = 0; sp < 256; ++sp) {
    unsigned char buf[8];
    unsigned char idxs[8];
    unsigned n_idx = 0;
    for (unsigned b = 0; b < 8; ++b) {
      buf[b] = (unsigned char) (sp & (1u << b));
      if (buf[b]) idxs[n_idx++] = (unsigned char) b;
    }
    unsigned clu = 0;
    for (unsigned ip = 0; ip < n_idx; ++ip) {
      unsigned char seed = idxs[ip];
      if (!buf[seed]) continue;
      unsigned char stack[8];
      unsigned sp_ptr = 0;
      stack[sp_ptr++] = seed;
      buf[seed] = 0;
      unsigned char x = 0, y = 0, n = 0;
      while (sp_ptr) {
        const unsigned char idx = stack[--sp_ptr];
        const int row = idx % 4, col = idx / 4;
        x += (unsigned char) col;
        y += (unsigned char) row;
        ++n;
        for (int k = 0; k < 8; ++k) {
          const int nc = col + dx[k];
          const int nr = row + dy[k];
          if (nc < 0 || nc > 1 || nr < 0 || nr > 3) continue;
          const unsigned nidx = (unsigned) ((nc << 2) | nr);
          if (!buf[nidx]) continue;
          stack[sp_ptr++] = (unsigned char) nidx;
          buf[nidx] = 0;
        }
      }
      const uint32_t cx = x / n;
      const uint32_t cy = y / n;
      const float fxv = (float) x / (float) n - (float) cx;
      const float fyv = (float) y / (float) n - (float) cy;
      sizes[sp] |= (uint8_t) (n << (4 * clu));
      patterns[sp] |= (uint8_t) (((cx << 2) | cy) << (4 * clu));
      patterns[sp] |= (uint8_t) (clu << 3);
      fx[2 * sp + clu] = fxv;
      fy[2 * sp + clu] = fyv;
      ++clu;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* shared per-SP decode                                                     */
/* ------------------------------------------------------------------------ */

/* Number of clusters of an isolated SP: EstimateInputSize.cu:71-79. */
static inline uint32_t isolated_cluster_count(uint32_t sp)
{
  const int pattern_0 = (sp & 0x88u) && !(sp & 0x44u) && (sp & 0x33u);
  const int pattern_1 = (sp & 0xCCu) && !(sp & 0x22u) && (sp & 0x11u);
  return 1u + (uint32_t) (pattern_0 | pattern_1);
}

/* Candidate byte of a non-isolated SP: EstimateInputSize.cu:118-183.
 * All SPs of the bank take part in the neighbour mask, isolated or not
 * (EstimateInputSize.cu:126 tests the current word's flag, Q1). */
static uint8_t sp_candidates(const raw_bank* bank, uint32_t sp_word)
{
  const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
  const uint32_t sp = sp_word & 0xFFu;
  uint32_t pixels = (sp & 0x0Fu) | ((sp & 0xF0u) << 2);
  const uint32_t sp_row = sp_addr & 0x3Fu;
  const uint32_t sp_col = sp_addr >> 6;
  for (uint32_t k = 0; k < bank->sp_count; ++k) {
    const uint32_t o = sp_word_at(bank, k);
    const uint32_t o_addr = (o & 0x007FFF00u) >> 8;
    const uint32_t o_row = o_addr & 0x3Fu;
    const uint32_t o_col = o_addr >> 6;
    const uint32_t osp = o & 0xFFu;
    const int is_top = o_row == sp_row + 1u && o_col == sp_col;
    const int is_top_right = o_row == sp_row + 1u && o_col == sp_col + 1u;
    const int is_right = o_row == sp_row && o_col == sp_col + 1u;
    const int is_right_bottom = o_row == sp_row - 1u && o_col == sp_col + 1u;
    const int is_bottom = o_row == sp_row - 1u && o_col == sp_col;
    if (is_top) pixels |= ((osp & 0x01u) | ((osp & 0x10u) << 2)) << 4;
    if (is_top_right) pixels |= (osp & 0x01u) << 16;
    if (is_right) pixels |= (osp & 0x0Fu) << 12;
    if (is_right_bottom) pixels |= (osp & 0x08u) << 8;
    if (is_bottom) pixels |= (osp & 0x80u) >> 2;
  }
  const uint32_t inside = pixels & 0x3CFu;
  const uint32_t mask = (inside << 1) | (inside << 5) | (inside << 6) | (inside << 7);
  const uint32_t working = mask & ~pixels;
  const uint32_t cand = (working >> 1) & (working >> 5) & (working >> 6) & (working >> 7) & pixels;
  return (uint8_t) ((cand & 0x03u) | ((cand & 0xC0u) >> 4) | ((cand & 0x0Cu) << 2) | ((cand & 0x0300u) >> 2));
}

/* ------------------------------------------------------------------------ */
/* a4: estimate_input_size                                                  */
/* ------------------------------------------------------------------------ */

void oracle_estimate_input_size(
  const char* events, const uint32_t* offsets, uint32_t n_events, uint32_t* est,
  uint32_t* module_cluster_num, uint32_t* cand_num, uint32_t* candidates, uint32_t* status)
{
  for (uint32_t e = 0; e < n_events; ++e) {
    uint32_t* est_e = est + e * OR_N_MODULES;
    uint32_t* num_e = module_cluster_num + e * OR_N_MODULES;
    uint32_t* cand_e = candidates + e * OR_MAX_CANDIDATES_EVENT;
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) est_e[m] = num_e[m] = 0;
    uint32_t n_cand = 0;
    status[e] |= OR_STATUS_BANK_ORDER_CHECKED;
    const raw_event ev = make_event(events + offsets[e], (offsets[e + 1] - offsets[e]) / 4u);
    for (uint32_t b = 0; b < ev.n_banks && b < OR_N_SENSORS; ++b) {
      const raw_bank bank = make_bank(&ev, b);
      if (!bank.fits || bank.sensor_index != b) {
        status[e] |= OR_STATUS_BAD_BANK;
        continue;
      }
      if (bank.sp_count > 256u) status[e] |= OR_STATUS_SP_INDEX_TRUNCATED;
      for (uint32_t i = 1; i < bank.sp_count; ++i)
        if (((sp_word_at(&bank, i - 1) >> 8) & 0x7FFFu) >= ((sp_word_at(&bank, i) >> 8) & 0x7FFFu))
          status[e] |= OR_STATUS_UNSORTED_BANK;
      uint32_t* est_m = est_e + (bank.sensor_index >> 2);
      for (uint32_t i = 0; i < bank.sp_count; ++i) {
        const uint32_t w = sp_word_at(&bank, i);
        if (w & 0x80000000u) {
          *est_m += isolated_cluster_count(w & 0xFFu);
        }
        else {
          const uint8_t c = sp_candidates(&bank, w);
          uint32_t found = 0;
          if (c & 0x0Fu) {
            const uint32_t k = k_candidate_ks[c & 0x0Fu];
            if (n_cand < OR_MAX_CANDIDATES_EVENT) cand_e[n_cand] = (i << 11) | (b << 3) | k;
            ++n_cand;
            ++found;
          }
          if (c & 0xF0u) {
            const uint32_t k = k_candidate_ks[c >> 4] + 2u;
            if (n_cand < OR_MAX_CANDIDATES_EVENT) cand_e[n_cand] = (i << 11) | (b << 3) | k;
            ++n_cand;
            ++found;
          }
          *est_m += found;
        }
      }
    }
    if (n_cand > OR_MAX_CANDIDATES_EVENT) {
      status[e] |= OR_STATUS_CANDIDATES_OVERFLOW;
      n_cand = OR_MAX_CANDIDATES_EVENT;
    }
    cand_num[e] = n_cand;
  }
}

/* ------------------------------------------------------------------------ */
/* prefix sums: PrefixSumHandler.cuh:15-58 (exclusive, total at [n])        */
/* ------------------------------------------------------------------------ */

void oracle_prefix_sum(uint32_t* a, uint32_t n)
{
  uint32_t acc = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t v = a[i];
    a[i] = acc;
    acc += v;
  }
  a[n] = acc;
}

/* ------------------------------------------------------------------------ */
/* a6: masked clustering                                                    */
/* ------------------------------------------------------------------------ */

/* MaskedVeloClustering.cu:4-46 */
static inline uint32_t current_mask(uint32_t p)
{
  return ((p & MASK_TOP) << 1) | ((p & MASK_BOTTOM) >> 1) | ((p & MASK_BOTTOM_RIGHT) << 15) |
         ((p & MASK_TOP_LEFT) >> 15) | (p >> 16) | (p >> 17) | (p << 16) | (p << 17);
}
static inline uint32_t mask_ltr(uint32_t p) { return ((p & MASK_LTR_TOP_RIGHT) >> 15) | (p >> 16) | (p >> 17); }
static inline uint32_t mask_rtl(uint32_t p) { return ((p & MASK_RTL_BOTTOM_LEFT) << 15) | (p << 16) | (p << 17); }
static inline uint32_t cluster_current_mask(uint32_t p)
{
  return ((p & MASK_TOP) << 1) | ((p & MASK_BOTTOM_RIGHT) << 15) | (p << 16) | (p << 17);
}
static inline uint32_t cand_mask_rc(uint32_t p, uint32_t rp)
{
  return ((p & MASK_BOTTOM) >> 1) & (((p & MASK_TOP_LEFT) >> 15) | (rp << 17)) & ((p >> 16) | (rp << 16)) &
         ((p >> 17) | ((rp & MASK_RTL_BOTTOM_LEFT) << 15));
}

typedef struct {
  float x, y, z;
  uint32_t id;
} or_hit;

/* Local -> global transform, MaskedVeloClustering.cu:122-138 and :431-449.
 * local_y is formed in double exactly as the reference's expression
 * (cy + 0.5 + fy) * pixel_size promotes it. */
static or_hit make_hit(const char* geometry, uint32_t sensor, uint32_t cx, uint32_t cy, float fx, float fy)
{
  const float* local_x = (const float*) geometry;
  const float* x_pitch = local_x + 768;
  const float* ltg = x_pitch + 768 + 16 * sensor;
  const uint32_t chip = cx / 256u;
  const uint32_t cid = channel_id(sensor, chip, cx % 256u, cy);
  const float lx = local_x[cx] + fx * x_pitch[cx];
  const float ly = (float) (((double) cy + 0.5 + (double) fy) * (double) k_pixel_size);
  or_hit h;
  h.x = ltg[0] * lx + ltg[1] * ly + ltg[9];
  h.y = ltg[3] * lx + ltg[4] * ly + ltg[10];
  h.z = ltg[6] * lx + ltg[7] * ly + ltg[11];
  h.id = lhcb_id(cid);
  return h;
}

/* Cluster grown from candidate pixel k of the SP at sp_index
 * (MaskedVeloClustering.cu:174-461).  Returns 1 and fills *out when stored. */
static int cluster_candidate(
  const raw_bank* bank, uint32_t sp_index, uint32_t candidate_k, const char* geometry, or_hit* out)
{
  const uint32_t sp_word = sp_word_at(bank, sp_index);
  const uint32_t sp_addr = (sp_word & 0x007FFF00u) >> 8;
  const int32_t sp_row = (int32_t) (sp_addr & 0x3Fu);
  const int32_t sp_col = (int32_t) (sp_addr >> 6);
  const int32_t row_lo = sp_row - 2, row_hi = sp_row + 1;
  const int32_t col_lo = sp_col - 1, col_hi = sp_col + 1;
  const int32_t row_lower_limit = row_lo * 4;
  const int32_t col_lower_limit = col_lo * 2;

  uint32_t pa[3] = {0, 0, 0};
  for (uint32_t k = 0; k < bank->sp_count; ++k) {
    const uint32_t o = sp_word_at(bank, k);
    if (o & 0x80000000u) continue;
    const uint32_t o_addr = (o & 0x007FFF00u) >> 8;
    const int32_t o_row = (int32_t) (o_addr & 0x3Fu);
    const int32_t o_col = (int32_t) (o_addr >> 6);
    const uint32_t osp = o & 0xFFu;
    if (o_row >= row_lo && o_row <= row_hi && o_col >= col_lo && o_col <= col_hi) {
      const int rr = o_row - row_lo;
      const int rc = o_col - col_lo;
      pa[rc] |= ((osp & 0x0Fu) << (4 * rr)) | ((osp & 0xF0u) << (12 + 4 * rr));
    }
  }

  const uint32_t row = (uint32_t) (sp_row * 4) + (candidate_k % 4u);
  const uint32_t col = (uint32_t) (sp_col * 2) + (candidate_k >= 4u);
  const uint32_t start = (1u << (row - (uint32_t) row_lower_limit)) << (16u * (col % 2u));
  uint32_t cl[3] = {0, start, 0};
  uint32_t wc[3] = {0, start, 0};
  pa[1] &= ~cl[1];
  for (int it = 0; it < MAX_CLUSTERING_ITERATIONS; ++it) {
    const uint32_t m0 = current_mask(wc[0]) | mask_rtl(wc[1]);
    const uint32_t m1 = current_mask(wc[1]) | mask_rtl(wc[2]) | mask_ltr(wc[0]);
    const uint32_t m2 = current_mask(wc[2]) | mask_ltr(wc[1]);
    wc[0] = pa[0] & m0;
    wc[1] = pa[1] & m1;
    wc[2] = pa[2] & m2;
    if (!wc[0] && !wc[1] && !wc[2]) break;
    cl[0] |= wc[0];
    cl[1] |= wc[1];
    cl[2] |= wc[2];
    pa[0] &= ~cl[0];
    pa[1] &= ~cl[1];
    pa[2] &= ~cl[2];
  }
  if (cl[2]) return 0; /* :323-325 */

  const int n = (int) (popc(cl[0]) + popc(cl[1]));
  int do_store = 1;
  if (n >= 3) { /* :331-398 */
    const uint32_t w0 = cluster_current_mask(cl[0]) & ~cl[0];
    const uint32_t w1 = (cluster_current_mask(cl[1]) | mask_ltr(cl[0])) & ~cl[1];
    const uint32_t w2 = mask_ltr(cl[1]);
    uint32_t c0 = cand_mask_rc(w0, w1) & cl[0];
    uint32_t c1 = cand_mask_rc(w1, w2) & cl[1];
    c1 ^= start;
    if (c0 || c1) {
      const int32_t neg_mask = (int32_t) ~(0u - start);
      const int under = start < 4096u;
      uint32_t smaller_row = under ? (0xFFFu & (uint32_t) neg_mask) : (0xFFFu & (uint32_t) (neg_mask >> 16));
      smaller_row |= smaller_row << 16;
      const uint32_t current_pixel_mask = smaller_row | (under ? 0xFFFF0000u : 0u);
      do_store = ((c0 & smaller_row) | (c1 & current_pixel_mask)) == 0;
    }
  }
  if (!do_store) return 0;

  /* Centroid: sum of absolute pixel columns and rows (:401-429). */
  int x = 0, y = 0;
  for (int c = 0; c < 2; ++c) {
    for (int half = 0; half < 2; ++half) {
      const uint32_t colbits = (cl[c] >> (16 * half)) & 0xFFFFu;
      const int abs_col = col_lower_limit + 2 * c + half;
      x += (int) popc(colbits) * abs_col;
      for (int r = 0; r < 16; ++r)
        if (colbits & (1u << r)) y += row_lower_limit + r;
    }
  }
  const uint32_t cx = (uint32_t) (x / n);
  const uint32_t cy = (uint32_t) (y / n);
  const float fx = (float) x / (float) n - (float) cx;
  const float fy = (float) y / (float) n - (float) cy;
  *out = make_hit(geometry, bank->sensor_index, cx, cy, fx, fy);
  return 1;
}

static void store_hit(uint32_t* container, uint32_t total, uint32_t slot, or_hit h)
{
  memcpy(&container[slot], &h.x, 4);
  memcpy(&container[total + slot], &h.y, 4);
  memcpy(&container[2 * total + slot], &h.z, 4);
  container[3 * total + slot] = h.id;
}

void oracle_masked_clustering(
  const char* events, const uint32_t* offsets, uint32_t n_events, const uint32_t* module_cluster_start,
  uint32_t* module_cluster_num, const char* geometry, uint32_t* container, uint32_t* status)
{
  uint8_t patterns[256], sizes[256];
  float lut_fx[512], lut_fy[512];
  oracle_sp_lut(patterns, sizes, lut_fx, lut_fy);
  const uint32_t total = module_cluster_start[OR_N_MODULES * n_events];

  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t* start_e = module_cluster_start + e * OR_N_MODULES;
    uint32_t* num_e = module_cluster_num + e * OR_N_MODULES;
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) num_e[m] = 0;
    const raw_event ev = make_event(events + offsets[e], (offsets[e + 1] - offsets[e]) / 4u);
    for (uint32_t b = 0; b < ev.n_banks && b < OR_N_SENSORS; ++b) {
      const raw_bank bank = make_bank(&ev, b);
      if (!bank.fits || bank.sensor_index != b) continue; /* flagged by the estimate */
      const uint32_t module = b >> 2;
      for (uint32_t i = 0; i < bank.sp_count; ++i) {
        const uint32_t w = sp_word_at(&bank, i);
        const uint32_t sp_addr = (w & 0x007FFF00u) >> 8;
        const uint32_t sp_row = sp_addr & 0x3Fu;
        const uint32_t sp_col = sp_addr >> 6;
        const uint32_t sp = w & 0xFFu;
        if (w & 0x80000000u) {
          /* isolated SP, :98-167 */
          const uint32_t idx = patterns[sp];
          for (uint32_t c = 0; c < 1u + ((idx >> 3) & 1u); ++c) {
            const uint32_t code = (idx >> (4 * c)) & 0x7u;
            const uint32_t cx = sp_col * 2u + ((code >> 2) & 1u);
            const uint32_t cy = sp_row * 4u + (code & 3u);
            const or_hit h = make_hit(geometry, bank.sensor_index, cx, cy, lut_fx[2 * sp + c], lut_fy[2 * sp + c]);
            store_hit(container, total, start_e[module] + num_e[module], h);
            num_e[module]++;
          }
        }
        else {
          /* candidates of this SP, :174-461.  The reference decodes the
           * candidate's SP index through a uint8_t (:179, Q2). */
          const uint8_t c = sp_candidates(&bank, w);
          uint32_t ks[2];
          uint32_t nk = 0;
          if (c & 0x0Fu) ks[nk++] = k_candidate_ks[c & 0x0Fu];
          if (c & 0xF0u) ks[nk++] = k_candidate_ks[c >> 4] + 2u;
          for (uint32_t j = 0; j < nk; ++j) {
            or_hit h;
            if (cluster_candidate(&bank, i & 0xFFu, ks[j], geometry, &h)) {
              store_hit(container, total, start_e[module] + num_e[module], h);
              num_e[module]++;
            }
          }
        }
      }
    }
    uint32_t ev_hits = 0;
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) {
      if (num_e[m] > OR_MAX_NUMHITS_IN_MODULE) status[e] |= OR_STATUS_MODULE_HITS_OVERFLOW;
      ev_hits += num_e[m];
    }
    if (ev_hits >= OR_MAX_HITS_PER_EVENT) status[e] |= OR_STATUS_EVENT_HITS_OVERFLOW;
  }
}

/* ------------------------------------------------------------------------ */
/* a7: phi + stable per-module sort + permutation                           */
/* ------------------------------------------------------------------------ */

static inline float as_f(uint32_t u)
{
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static inline uint32_t as_u(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

void oracle_calculate_phi_and_sort(
  uint32_t n_events, const uint32_t* module_cluster_start, const uint32_t* module_cluster_num,
  uint32_t* container, uint32_t* perm, uint32_t* status)
{
  (void) status;
  const uint32_t total = module_cluster_start[OR_N_MODULES * n_events];
  uint32_t* X = container;
  uint32_t* Y = container + total;
  uint32_t* Z = container + 2 * total;
  uint32_t* ID = container + 3 * total;
  uint32_t* PHI = container + 4 * total;
  uint32_t* TMP = container + 5 * total;
  float phis[OR_MAX_HITS_PER_EVENT > 4096 ? 16384 : 4096];

  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t* start_e = module_cluster_start + e * OR_N_MODULES;
    const uint32_t* num_e = module_cluster_num + e * OR_N_MODULES;
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) {
      const uint32_t s = start_e[m], n = num_e[m];
      /* Estimated-but-unfilled slots ("holes") are defined as NaN in every
       * sorted slot so no search selects them; the reference leaves them
       * uninitialised (CalculatePhiAndSort.cu:38 zero permutation). */
      for (uint32_t p = s + n; p < start_e[m + 1]; ++p) X[p] = Y[p] = Z[p] = PHI[p] = TMP[p] = 0x7FC00000u;
      if (n > 16384u) continue;
      for (uint32_t i = 0; i < n; ++i) {
        const float x = as_f(X[s + i]), y = as_f(Y[s + i]);
        phis[i] = (m & 1u) ? oracle_phi_odd(x, y) : oracle_phi_even(x, y);
      }
      /* CalculatePhi.cu:47-59: rank with slot tie-break */
      for (uint32_t i = 0; i < n; ++i) {
        uint32_t pos = 0;
        for (uint32_t j = 0; j < n; ++j) pos += phis[i] > phis[j] || (phis[i] == phis[j] && i > j);
        perm[s + pos] = s + i;
        PHI[s + pos] = as_u(phis[i]);
      }
    }
    /* SortByPhi.cu:21-33: X->slot 5, Y->slot 0, Z->slot 1, ID->slot 2.
     * Only the module ranges are defined; holes are not written. */
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) {
      const uint32_t s = start_e[m], n = num_e[m];
      for (uint32_t i = 0; i < n; ++i) TMP[s + i] = X[perm[s + i]];
    }
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) {
      const uint32_t s = start_e[m], n = num_e[m];
      for (uint32_t i = 0; i < n; ++i) X[s + i] = Y[perm[s + i]];
    }
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) {
      const uint32_t s = start_e[m], n = num_e[m];
      for (uint32_t i = 0; i < n; ++i) Y[s + i] = Z[perm[s + i]];
    }
    for (uint32_t m = 0; m < OR_N_MODULES; ++m) {
      const uint32_t s = start_e[m], n = num_e[m];
      for (uint32_t i = 0; i < n; ++i) Z[s + i] = ID[perm[s + i]];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* a8: fill_candidates with the exact binary-search semantics               */
/* ------------------------------------------------------------------------ */

/* VeloTools.cuh:6-26.  array_size 0 still inspects array[0], as the
 * reference does when called from track forwarding. */
static int bs_first_candidate(const float* a, uint32_t size, float value, float margin)
{
  int found = 0;
  int l = 0;
  int r = (int) (size - 1u);
  while (l < r) {
    const int m = (l + r) / 2;
    const float el = a[m];
    found |= fabsf(value - el) < margin;
    if (value - margin > el) l = m + 1;
    else r = m;
  }
  found |= fabsf(value - a[l]) < margin;
  return found ? l : -1;
}

/* VeloTools.cuh:28-45 */
static int bs_second_candidate(const float* a, uint32_t size, float value, float margin)
{
  int l = 0;
  int r = (int) (size - 1u);
  while (l < r) {
    const int m = (l + r) / 2;
    if (value + margin > a[m]) l = m + 1;
    else r = m;
  }
  const int last_compatible = fabsf(value - a[l]) < margin;
  return last_compatible ? l + 1 : l;
}

void oracle_fill_candidates(
  uint32_t n_events, const uint32_t* container, const uint32_t* module_cluster_start,
  const uint32_t* module_cluster_num, int16_t* h0c, int16_t* h2c)
{
  const uint32_t total = module_cluster_start[OR_N_MODULES * n_events];
  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t* start_e = module_cluster_start + e * OR_N_MODULES;
    const uint32_t* num_e = module_cluster_num + e * OR_N_MODULES;
    const uint32_t off = start_e[0];
    const float* phis = (const float*) (container + 4 * total + off);
    int16_t* h0 = h0c + 2 * off;
    int16_t* h2 = h2c + 2 * off;
    for (uint32_t m = 2; m < 50; ++m) {
      const uint32_t m0s = start_e[m + 2] - off, m0n = num_e[m + 2];
      const uint32_t m2s = start_e[m - 2] - off, m2n = num_e[m - 2];
      for (uint32_t r = 0; r < num_e[m]; ++r) {
        const uint32_t h1 = start_e[m] + r - off;
        const float phi = phis[h1];
        int f0 = -1, l0 = -1, f2 = -1, l2 = -1;
        if (m0n > 0) {
          f0 = bs_first_candidate(phis + m0s, m0n, phi, k_phi_extrapolation);
          if (f0 != -1) {
            l0 = bs_second_candidate(phis + m0s + f0, m0n - (uint32_t) f0, phi, k_phi_extrapolation);
            f0 += (int) m0s;
            l0 = l0 == 0 ? f0 + 1 : f0 + l0;
          }
        }
        if (m2n > 0) {
          f2 = bs_first_candidate(phis + m2s, m2n, phi, k_phi_extrapolation);
          if (f2 != -1) {
            l2 = bs_second_candidate(phis + m2s + f2, m2n - (uint32_t) f2, phi, k_phi_extrapolation);
            f2 += (int) m2s;
            l2 = l2 == 0 ? f2 + 1 : f2 + l2;
          }
        }
        h0[2 * h1] = (int16_t) f0;
        h0[2 * h1 + 1] = (int16_t) l0;
        h2[2 * h1] = (int16_t) f2;
        h2[2 * h1 + 1] = (int16_t) l2;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* a9-a11: search_by_triplet                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t start; /* event-relative */
  uint32_t num;
  float z;
} or_module;

typedef struct {
  const float *x, *y, *z, *phi;
} or_hits_view;

/* One tracks-to-follow entry: the reference's encoding (ProcessModules.cu,
 * TrackForwarding.cu:63-66): bit 31 tracklet flag, bits 28-30 skipped
 * modules, bits 0-27 the tracklet or track number. */
typedef struct {
  uint32_t* items;
  uint32_t n;
} ttf_list;

typedef struct {
  or_tracklet_hits* tracklets; /* every tracklet ever created, by number */
  uint32_t n_tracklets;
  uint32_t cap_tracklets;
  uint32_t step_first_tracklet[32];
} tracklet_store;

/* TrackSeeding.cu:7-200 */
static void seeding(
  const or_module* md, const or_hits_view* H, const int16_t* h0c, const int16_t* h2c, const uint8_t* used,
  tracklet_store* ts, ttf_list* next, uint32_t next_cap, uint32_t* status)
{
  uint16_t h1s[2 * 16384];
  uint32_t n_h1 = 0;
  for (int side = 2; side <= 3; ++side) {
    for (uint32_t r = 0; r < md[side].num; ++r) {
      const uint32_t h1 = md[side].start + r;
      if (!used[h1] && h0c[2 * h1] != -1 && h2c[2 * h1] != -1 && n_h1 < 2 * 16384) h1s[n_h1++] = (uint16_t) h1;
    }
  }
  for (uint32_t i = 0; i < n_h1; ++i) {
    const uint32_t h1 = h1s[i];
    const float h1x = H->x[h1], h1y = H->y[h1], h1z = H->z[h1];
    const int f0 = h0c[2 * h1], l0 = h0c[2 * h1 + 1];
    const int f2 = h2c[2 * h1], l2 = h2c[2 * h1 + 1];
    float best = FLT_MAX;
    uint16_t b0 = 0, b2 = 0;
    for (int h0 = f0; h0 < l0; ++h0) {
      if (used[h0]) continue;
      const float h0x = H->x[h0], h0y = H->y[h0], h0z = H->z[h0];
      for (int h2 = f2; h2 < l2; ++h2) {
        if (used[h2]) continue;
        const float h2x = H->x[h2], h2y = H->y[h2], h2z = H->z[h2];
        const float dmax = k_max_slope * (h0z - h1z);
        const float dz21 = h2z - h1z;
        const float scatter_denom2 = 1.0f / (dz21 * dz21);
        const float z2_tz = (h2z - h0z) / (h1z - h0z);
        const float x = h0x + (h1x - h0x) * z2_tz;
        const float y = h0y + (h1y - h0y) * z2_tz;
        const float dx = x - h2x;
        const float dy = y - h2y;
        const float scatter_num = (dx * dx) + (dy * dy);
        const float scatter = scatter_num * scatter_denom2;
        const int cond = fabsf(h1x - h0x) < dmax && fabsf(h1y - h0y) < dmax && fabsf(dx) < k_tolerance &&
                         fabsf(dy) < k_tolerance && scatter < k_max_scatter_seeding && scatter < best;
        if (cond) {
          best = scatter;
          b0 = (uint16_t) h0;
          b2 = (uint16_t) h2;
        }
      }
    }
    if (best < FLT_MAX) {
      if (ts->n_tracklets >= ts->cap_tracklets) continue;
      const uint32_t tn = ts->n_tracklets++;
      ts->tracklets[tn].hits[0] = b0;
      ts->tracklets[tn].hits[1] = (uint16_t) h1;
      ts->tracklets[tn].hits[2] = b2;
      if (next->n < next_cap) next->items[next->n++] = 0x80000000u | tn;
      else *status |= OR_STATUS_TTF_OVERFLOW;
    }
  }
}

/* TrackForwarding.cuh:28-58 */
static void forward_candidates(
  const or_module* mod, float tx, float ty, const float* phis, float h0x, float h0y, float h0z, int odd,
  int* first, int* last)
{
  const float dz = mod->z - h0z;
  const float predx = tx * dz;
  const float predy = ty * dz;
  const float xp = h0x + predx;
  const float yp = h0y + predy;
  const float phi = odd ? oracle_phi_odd(xp, yp) : oracle_phi_even(xp, yp);
  int f = bs_first_candidate(phis + mod->start, mod->num, phi, k_forward_phi_tolerance);
  int l = -1;
  if (f != -1) {
    l = bs_second_candidate(phis + mod->start + f, mod->num - (uint32_t) f, phi, k_forward_phi_tolerance);
    f += (int) mod->start;
    l = l == 0 ? f + 1 : f + l;
  }
  *first = f;
  *last = l;
}

/* TrackForwarding.cu:41-182 */
static void forwarding(
  const or_module* md, const or_hits_view* H, uint8_t* used, const tracklet_store* ts, const ttf_list* cur,
  ttf_list* next, uint32_t next_cap, or_track_hits* tracks, uint32_t* n_tracks, or_tracklet_hits* weak,
  uint32_t* n_weak, uint32_t* status)
{
  for (uint32_t e = 0; e < cur->n; ++e) {
    const uint32_t full = cur->items[e];
    const int track_flag = (full & 0x80000000u) != 0;
    const uint32_t skipped = (full & 0x70000000u) >> 28;
    uint32_t trackno = full & 0x0FFFFFFFu;
    or_track_hits t;
    if (track_flag) {
      t.hitsNum = 3;
      t.hits[0] = ts->tracklets[trackno].hits[0];
      t.hits[1] = ts->tracklets[trackno].hits[1];
      t.hits[2] = ts->tracklets[trackno].hits[2];
    }
    else {
      t = tracks[trackno];
    }
    const uint32_t h0n = t.hits[t.hitsNum - 2];
    const uint32_t h1n = t.hits[t.hitsNum - 1];
    const float h0x = H->x[h0n], h0y = H->y[h0n], h0z = H->z[h0n];
    const float h1x = H->x[h1n], h1y = H->y[h1n], h1z = H->z[h1n];
    const float td = 1.0f / (h1z - h0z);
    const float txn = h1x - h0x;
    const float tyn = h1y - h0y;
    const float tx = txn * td;
    const float ty = tyn * td;

    int of, ol, ef, el;
    forward_candidates(&md[4], tx, ty, H->phi, h0x, h0y, h0z, 1, &of, &ol);
    forward_candidates(&md[5], tx, ty, H->phi, h0x, h0y, h0z, 0, &ef, &el);
    const int n_odd = ol - of;
    const int n_even = el - ef;
    float best = FLT_MAX;
    uint32_t best_h2 = 0;
    for (int j = 0; j < n_odd + n_even; ++j) {
      const int h2 = j < n_odd ? of + j : ef + j - n_odd;
      const float h2x = H->x[h2], h2y = H->y[h2], h2z = H->z[h2];
      const float dz = h2z - h0z;
      const float predx = tx * dz;
      const float predy = ty * dz;
      const float dz21 = h2z - h1z;
      const float scatter_denom2 = 1.0f / (dz21 * dz21);
      /* fit_hit_to_track, TrackForwarding.cu:12-36 */
      const float xp = h0x + predx;
      const float dx = fabsf(xp - h2x);
      const float yp = h0y + predy;
      const float dy = fabsf(yp - h2y);
      const float scatter_num = (dx * dx) + (dy * dy);
      const float scatter = scatter_num * scatter_denom2;
      const int cond = dx < k_tolerance && dy < k_tolerance && scatter < k_max_scatter_forwarding;
      const float fit = cond ? scatter : FLT_MAX;
      if (fit < best) {
        best = fit;
        best_h2 = (uint32_t) h2;
      }
    }
    if (best != FLT_MAX) {
      used[best_h2] = 1;
      t.hits[t.hitsNum++] = (uint16_t) best_h2;
      if (t.hitsNum <= 4) {
        used[t.hits[0]] = 1;
        used[t.hits[1]] = 1;
        used[t.hits[2]] = 1;
        if (*n_tracks >= OR_MAX_TRACKS) {
          *status |= OR_STATUS_TRACKS_OVERFLOW;
          continue;
        }
        trackno = (*n_tracks)++;
      }
      tracks[trackno] = t;
      if (next->n < next_cap) next->items[next->n++] = trackno;
      else *status |= OR_STATUS_TTF_OVERFLOW;
    }
    else if (skipped < k_max_skipped_modules) {
      if (next->n < next_cap) next->items[next->n++] = ((skipped + 1u) << 28) | (full & 0x8FFFFFFFu);
      else *status |= OR_STATUS_TTF_OVERFLOW;
    }
    else if (t.hitsNum == 3) {
      if (*n_weak < OR_MAX_WEAK_TRACKS) {
        weak[*n_weak].hits[0] = t.hits[0];
        weak[*n_weak].hits[1] = t.hits[1];
        weak[*n_weak].hits[2] = t.hits[2];
        (*n_weak)++;
      }
      else *status |= OR_STATUS_WEAK_OVERFLOW;
    }
  }
}

static void load_modules(
  or_module* md, uint32_t first_module, const uint32_t* start_e, const uint32_t* num_e, uint32_t off,
  const float* module_zs)
{
  for (uint32_t i = 0; i < 6; ++i) {
    const uint32_t m = first_module - i;
    md[i].start = start_e[m] - off;
    md[i].num = num_e[m];
    md[i].z = module_zs[m];
  }
}

void oracle_search_by_triplet(
  uint32_t n_events, const uint32_t* container, const uint32_t* module_cluster_start,
  const uint32_t* module_cluster_num, or_track_hits* tracks, or_tracklet_hits* weak_tracks, uint8_t* hit_used,
  int32_t* atomics, const int16_t* h0c, const int16_t* h2c, const float* module_zs, uint32_t* status)
{
  const uint32_t total = module_cluster_start[OR_N_MODULES * n_events];
  uint32_t* bufs = (uint32_t*) malloc(sizeof(uint32_t) * 2 * OR_TTF_MODULO);
  tracklet_store ts;
  ts.cap_tracklets = 65536;
  ts.tracklets = (or_tracklet_hits*) malloc(sizeof(or_tracklet_hits) * ts.cap_tracklets);

  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t* start_e = module_cluster_start + e * OR_N_MODULES;
    const uint32_t* num_e = module_cluster_num + e * OR_N_MODULES;
    const uint32_t off = start_e[0];
    const uint32_t ev_hits = start_e[OR_N_MODULES] - off;
    or_hits_view H;
    H.y = (const float*) (container + off);
    H.z = (const float*) (container + total + off);
    H.phi = (const float*) (container + 4 * total + off);
    H.x = (const float*) (container + 5 * total + off);
    uint8_t* used = hit_used + off;
    memset(used, 0, ev_hits);
    or_track_hits* trk = tracks + (size_t) e * OR_MAX_TRACKS;
    or_tracklet_hits* weak = weak_tracks + (size_t) e * OR_MAX_WEAK_TRACKS;
    uint32_t n_tracks = 0, n_weak = 0;
    ts.n_tracklets = 0;

    ttf_list cur = {bufs, 0}, nxt = {bufs + OR_TTF_MODULO, 0};
    or_module md[6];
    uint32_t step = 0;
    uint32_t first = OR_N_MODULES - 1;
    load_modules(md, first, start_e, num_e, off, module_zs);
    ts.step_first_tracklet[step] = 0;
    seeding(md, &H, h0c + 2 * off, h2c + 2 * off, used, &ts, &cur, OR_TTF_MODULO, &status[e]);
    first -= 2;
    while (first >= 4) {
      ++step;
      load_modules(md, first, start_e, num_e, off, module_zs);
      nxt.n = 0;
      /* new entries of this step share the ring with the unread ones */
      const uint32_t cap = OR_TTF_MODULO - cur.n;
      ts.step_first_tracklet[step] = ts.n_tracklets;
      forwarding(md, &H, used, &ts, &cur, &nxt, cap, trk, &n_tracks, weak, &n_weak, &status[e]);
      seeding(md, &H, h0c + 2 * off, h2c + 2 * off, used, &ts, &nxt, cap, &status[e]);
      /* a tracklet stays referenced for two steps (forward or skip) */
      if (ts.n_tracklets - ts.step_first_tracklet[step - 1] > OR_TTF_MODULO)
        status[e] |= OR_STATUS_TRACKLETS_OVERFLOW;
      ttf_list tmp = cur;
      cur = nxt;
      nxt = tmp;
      first -= 2;
    }
    /* ProcessModules.cu:141-162: leftover tracklets become weak tracks */
    for (uint32_t i = 0; i < cur.n; ++i) {
      const uint32_t full = cur.items[i];
      if (full & 0x80000000u) {
        const uint32_t tn = full & 0x0FFFFFFFu;
        if (n_weak < OR_MAX_WEAK_TRACKS) weak[n_weak++] = ts.tracklets[tn];
        else status[e] |= OR_STATUS_WEAK_OVERFLOW;
      }
    }
    atomics[e] = (int32_t) n_tracks;
    const uint32_t ip = n_events + e * (OR_NUM_ATOMICS - 1u);
    atomics[ip + 0] = (int32_t) n_weak;
    atomics[ip + 1] = (int32_t) ts.n_tracklets;
    atomics[ip + 2] = 0;
    atomics[ip + 3] = 0;
  }
  free(bufs);
  free(ts.tracklets);
}

/* ------------------------------------------------------------------------ */
/* a12: weak tracks adder                                                   */
/* ------------------------------------------------------------------------ */

/* WeakTracksAdder.cu:8-86 */
static float weak_chi2(const or_hits_view* H, const or_tracklet_hits* t)
{
  float s0 = 0.0f, sx = 0.0f, sz = 0.0f, sxz = 0.0f, sz2 = 0.0f;
  float u0 = 0.0f, uy = 0.0f, uz = 0.0f, uyz = 0.0f, uz2 = 0.0f;
  for (int h = 0; h < 3; ++h) {
    const uint32_t i = t->hits[h];
    const float x = H->x[i], y = H->y[i], z = H->z[i];
    const float wx = k_param_w;
    const float wx_t_x = wx * x;
    const float wx_t_z = wx * z;
    s0 += wx;
    sx += wx_t_x;
    sz += wx_t_z;
    sxz += wx_t_x * z;
    sz2 += wx_t_z * z;
    const float wy = k_param_w;
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
  for (int h = 0; h < 3; ++h) {
    const uint32_t i = t->hits[h];
    const float z = H->z[i];
    const float x = x0 + tx * z;
    const float y = y0 + ty * z;
    const float dx = x - H->x[i];
    const float dy = y - H->y[i];
    ch += dx * dx * k_param_w + dy * dy * k_param_w;
    ndof += 2;
  }
  return ch / (float) ndof;
}

void oracle_weak_tracks_adder(
  uint32_t n_events, const uint32_t* container, const uint32_t* module_cluster_start, or_track_hits* tracks,
  const or_tracklet_hits* weak_tracks, const uint8_t* hit_used, int32_t* atomics, uint32_t* status)
{
  const uint32_t total = module_cluster_start[OR_N_MODULES * n_events];
  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t off = module_cluster_start[e * OR_N_MODULES];
    or_hits_view H;
    H.y = (const float*) (container + off);
    H.z = (const float*) (container + total + off);
    H.phi = (const float*) (container + 4 * total + off);
    H.x = (const float*) (container + 5 * total + off);
    const uint8_t* used = hit_used + off;
    or_track_hits* trk = tracks + (size_t) e * OR_MAX_TRACKS;
    const or_tracklet_hits* weak = weak_tracks + (size_t) e * OR_MAX_WEAK_TRACKS;
    const uint32_t ip = n_events + e * (OR_NUM_ATOMICS - 1u);
    uint32_t n_weak = (uint32_t) atomics[ip];
    if (n_weak > OR_MAX_WEAK_TRACKS) n_weak = OR_MAX_WEAK_TRACKS;
    uint32_t n_tracks = (uint32_t) atomics[e];
    for (uint32_t w = 0; w < n_weak; ++w) {
      const or_tracklet_hits* t = &weak[w];
      const int any_used = used[t->hits[0]] || used[t->hits[1]] || used[t->hits[2]];
      const float chi2 = weak_chi2(&H, t);
      if (!any_used && chi2 < k_max_chi2) {
        if (n_tracks >= OR_MAX_TRACKS) {
          status[e] |= OR_STATUS_TRACKS_OVERFLOW;
          continue;
        }
        or_track_hits* d = &trk[n_tracks++];
        d->hitsNum = 3;
        d->hits[0] = t->hits[0];
        d->hits[1] = t->hits[1];
        d->hits[2] = t->hits[2];
      }
    }
    atomics[e] = (int32_t) n_tracks;
  }
}

/* ------------------------------------------------------------------------ */
/* a13/a14                                                                  */
/* ------------------------------------------------------------------------ */

void oracle_copy_and_prefix_sum(uint32_t n_events, uint32_t* atomics)
{
  for (uint32_t e = 0; e < n_events; ++e) atomics[n_events + e] = atomics[e];
  oracle_prefix_sum(atomics + n_events, n_events);
}

void oracle_copy_velo_track_hit_number(
  uint32_t n_events, const or_track_hits* tracks, const uint32_t* atomics, uint32_t* hit_number)
{
  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t acc = atomics[n_events + e];
    const uint32_t n = atomics[e];
    for (uint32_t i = 0; i < n; ++i) hit_number[acc + i] = tracks[(size_t) e * OR_MAX_TRACKS + i].hitsNum;
  }
}

/* ------------------------------------------------------------------------ */
/* a16: consolidation                                                       */
/* ------------------------------------------------------------------------ */

/* VeloKalmanFilter.cu:6-41: prediction to zhit, gain, update of the
 * (x, tx) state and its covariance, chi2 = r^2 R. */
float oracle_kalman_step(float z, float zhit, float xhit, float whit, float* x, float* tx, float* cov_xx,
                         float* cov_xtx, float* cov_txtx)
{
  const float dz = zhit - z;
  const float predx = *x + dz * *tx;
  const float dz_t_covtxtx = dz * *cov_txtx;
  const float predcov_xtx = *cov_xtx + dz_t_covtxtx;
  const float dx_t_covxtx = dz * *cov_xtx;
  const float predcov_xx = *cov_xx + 2.0f * dx_t_covxtx + dz * dz_t_covtxtx;
  const float predcov_txtx = *cov_txtx;
  const float r_gain = 1.0f / ((1.0f / whit) + predcov_xx);
  const float kx = predcov_xx * r_gain;
  const float ktx = predcov_xtx * r_gain;
  const float r = xhit - predx;
  *x = predx + kx * r;
  *tx = *tx + ktx * r;
  *cov_xx = (1.0f - kx) * predcov_xx;
  *cov_xtx = (1.0f - kx) * predcov_xtx;
  *cov_txtx = predcov_txtx - ktx * predcov_xtx;
  return r * r * r_gain;
}

/* VeloKalmanFilter.cuh:22-91 (simplified_fit<true>) for every track. */
void oracle_velo_kalman_fit(
  uint32_t n_events, const uint32_t* atomics, const uint32_t* hit_number, const char* track_hits,
  const char* states, char* out)
{
  const uint32_t T = atomics[2 * n_events];
  const uint32_t Ht = hit_number[T];
  const float* hx = (const float*) track_hits;
  const float* hy = hx + Ht;
  const float* hz = hy + Ht;
  const float* S[12];
  float* O[12];
  for (int k = 0; k < 12; ++k) {
    S[k] = (const float*) states + (size_t) k * T;
    O[k] = (float*) out + (size_t) k * T;
  }
  const uint8_t* bwd_in = (const uint8_t*) (states + 12 * sizeof(float) * (size_t) T);
  uint8_t* bwd_out = (uint8_t*) (out + 12 * sizeof(float) * (size_t) T);
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t h0 = hit_number[t];
    const int nh = (int) (hit_number[t + 1] - h0);
    const float* x = hx + h0;
    const float* y = hy + h0;
    const float* z = hz + h0;
    const float beam_tx = S[2][t], beam_ty = S[3][t], beam_z = S[11][t];
    const int backward = beam_z > z[0];
    const int direction = backward ? 1 : -1; /* upstream fit */
    const float slope2 = beam_tx * beam_tx + beam_ty * beam_ty;
    const float noise2 = (float) (1e-8 + 7e-6 * (double) slope2);
    int first = 0, last = nh - 1, step = 1;
    if ((z[last] - z[first]) * (float) direction < 0.0f) {
      first = nh - 1;
      last = 0;
      step = -1;
    }
    float sx = x[first], sy = y[first], sz = z[first], stx = beam_tx, sty = beam_ty;
    float c00 = k_param_w_inverted, c11 = k_param_w_inverted, c20 = 0.0f, c31 = 0.0f, c22 = 1.0f, c33 = 1.0f;
    float chi2 = 0.0f;
    for (int i = first + step; i != last + step; i += step) {
      c22 += noise2;
      c33 += noise2;
      chi2 += oracle_kalman_step(sz, z[i], x[i], k_param_w, &sx, &stx, &c00, &c20, &c22);
      chi2 += oracle_kalman_step(sz, z[i], y[i], k_param_w, &sy, &sty, &c11, &c31, &c33);
      sz = z[i];
    }
    c22 += noise2;
    c33 += noise2;
    const float v[12] = {sx, sy, stx, sty, c00, c20, c22, c11, c31, c33, chi2, sz};
    for (int k = 0; k < 12; ++k) O[k][t] = v[k];
    bwd_out[t] = bwd_in[t];
  }
}

void oracle_consolidate_velo_tracks(
  uint32_t n_events, const uint32_t* atomics, const or_track_hits* tracks, const uint32_t* hit_number,
  const uint32_t* container, const uint32_t* module_cluster_start, char* track_hits, char* states)
{
  const uint32_t total = module_cluster_start[OR_N_MODULES * n_events];
  const uint32_t total_tracks = atomics[2 * n_events];
  const uint32_t total_hits = hit_number[total_tracks];
  float* cx = (float*) track_hits;
  float* cy = cx + total_hits;
  float* cz = cy + total_hits;
  uint32_t* cid = (uint32_t*) (cz + total_hits);
  float* S[12];
  for (int k = 0; k < 12; ++k) S[k] = (float*) states + (size_t) k * total_tracks;
  uint8_t* backward = (uint8_t*) (states + 12 * sizeof(float) * (size_t) total_tracks);

  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t off = module_cluster_start[e * OR_N_MODULES];
    const float* Ys = (const float*) (container + off);
    const float* Zs = (const float*) (container + total + off);
    const uint32_t* IDs = container + 2 * total + off;
    const float* Xs = (const float*) (container + 5 * total + off);
    const uint32_t n = atomics[e];
    const uint32_t toff = atomics[n_events + e];
    for (uint32_t i = 0; i < n; ++i) {
      const or_track_hits* t = &tracks[(size_t) e * OR_MAX_TRACKS + i];
      const uint32_t ho = hit_number[toff + i];
      for (uint32_t h = 0; h < t->hitsNum; ++h) {
        const uint32_t hi = t->hits[h];
        cx[ho + h] = Xs[hi];
        cy[ho + h] = Ys[hi];
        cz[ho + h] = Zs[hi];
        cid[ho + h] = IDs[hi];
      }
      /* means_square_fit, ConsolidateVelo.cu:6-114 */
      float s0 = 0.0f, sx = 0.0f, sz = 0.0f, sxz = 0.0f, sz2 = 0.0f;
      float u0 = 0.0f, uy = 0.0f, uz = 0.0f, uyz = 0.0f, uz2 = 0.0f;
      for (uint32_t h = 0; h < t->hitsNum; ++h) {
        const uint32_t hi = t->hits[h];
        const float x = Xs[hi], y = Ys[hi], z = Zs[hi];
        const float wx = k_param_w;
        const float wx_t_x = wx * x;
        const float wx_t_z = wx * z;
        s0 += wx;
        sx += wx_t_x;
        sz += wx_t_z;
        sxz += wx_t_x * z;
        sz2 += wx_t_z * z;
        const float wy = k_param_w;
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
      const int bwd = zb > cz[ho];
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
      for (uint32_t h = 0; h < t->hitsNum; ++h) {
        const float z = cz[ho + h];
        const float x = x0 + tx * z;
        const float y = y0 + ty * z;
        const float dx = x - cx[ho + h];
        const float dy = y - cy[ho + h];
        ch += dx * dx * k_param_w + dy * dy * k_param_w;
        ndof += 2;
      }
      const float chi2 = ch / (float) ndof;
      x0 = x0 + tx * zb;
      y0 = y0 + ty * zb;
      const uint32_t ti = toff + i;
      S[0][ti] = x0;
      S[1][ti] = y0;
      S[2][ti] = tx;
      S[3][ti] = ty;
      S[4][ti] = m22 * den20;
      S[5][ti] = -m20 * den20;
      S[6][ti] = m00 * den20;
      S[7][ti] = m33 * den31;
      S[8][ti] = -m31 * den31;
      S[9][ti] = m11 * den31;
      S[10][ti] = chi2;
      S[11][ti] = zb;
      backward[ti] = (uint8_t) bwd;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* textbook 8-connected labelling (Clustering.cpp:140-270 semantics)        */
/* ------------------------------------------------------------------------ */

uint32_t oracle_ccl_event(
  const char* event, uint32_t* ids, uint32_t* modules, int32_t* sum_x, int32_t* sum_y, int32_t* npix,
  uint32_t capacity)
{
  enum { COLS = 768, ROWS = 256 };
  static uint8_t grid[COLS * ROWS];
  static uint32_t stack[COLS * ROWS];
  static uint32_t touched[COLS * ROWS];
  uint32_t n_out = 0;
  const raw_event ev = make_event(event, 0xFFFFFFFFu); /* a well-formed event (test helper) */
  for (uint32_t b = 0; b < ev.n_banks && b < OR_N_SENSORS; ++b) {
    const raw_bank bank = make_bank(&ev, b);
    uint32_t n_touched = 0;
    for (uint32_t i = 0; i < bank.sp_count; ++i) {
      const uint32_t w = sp_word_at(&bank, i);
      const uint32_t addr = (w & 0x007FFF00u) >> 8;
      const uint32_t row = addr & 0x3Fu, col = addr >> 6;
      for (uint32_t k = 0; k < 8; ++k) {
        if (!(w & (1u << k))) continue;
        const uint32_t c = col * 2u + k / 4u, r = row * 4u + k % 4u;
        if (c >= COLS || r >= ROWS) continue;
        const uint32_t idx = c * ROWS + r;
        if (!grid[idx]) {
          grid[idx] = 1;
          touched[n_touched++] = idx;
        }
      }
    }
    /* seed order: ascending column then row */
    for (uint32_t c = 0; c < COLS; ++c) {
      for (uint32_t r = 0; r < ROWS; ++r) {
        const uint32_t seed = c * ROWS + r;
        if (grid[seed] != 1) continue;
        uint32_t sp = 0;
        stack[sp++] = seed;
        grid[seed] = 2;
        int32_t sx = 0, sy = 0, n = 0;
        while (sp) {
          const uint32_t idx = stack[--sp];
          const int32_t cc = (int32_t) (idx / ROWS), rr = (int32_t) (idx % ROWS);
          sx += cc;
          sy += rr;
          ++n;
          for (int dc = -1; dc <= 1; ++dc) {
            for (int dr = -1; dr <= 1; ++dr) {
              const int32_t nc = cc + dc, nr = rr + dr;
              if ((dc == 0 && dr == 0) || nc < 0 || nc >= COLS || nr < 0 || nr >= ROWS) continue;
              const uint32_t ni = (uint32_t) nc * ROWS + (uint32_t) nr;
              if (grid[ni] == 1) {
                grid[ni] = 2;
                stack[sp++] = ni;
              }
            }
          }
        }
        if (n_out < capacity) {
          const uint32_t cx = (uint32_t) (sx / n), cy = (uint32_t) (sy / n);
          ids[n_out] = lhcb_id(channel_id(bank.sensor_index, cx / 256u, cx % 256u, cy));
          modules[n_out] = bank.sensor_index >> 2;
          sum_x[n_out] = sx;
          sum_y[n_out] = sy;
          npix[n_out] = n;
        }
        ++n_out;
      }
    }
    for (uint32_t t = 0; t < n_touched; ++t) grid[touched[t]] = 0;
  }
  return n_out;
}

/* ------------------------------------------------------------------------ */
/* whole pipeline                                                           */
/* ------------------------------------------------------------------------ */

int oracle_pipeline(
  const char* events, const uint32_t* offsets, uint32_t n, const char* geometry, const float* module_zs,
  or_pipeline_result* out)
{
  memset(out, 0, sizeof(*out));
  out->n_events = n;
  out->status = (uint32_t*) calloc(n ? n : 1, 4);
  out->module_cluster_start = (uint32_t*) calloc((size_t) n * OR_N_MODULES + 1, 4);
  out->module_cluster_num = (uint32_t*) calloc((size_t) n * OR_N_MODULES + 1, 4);
  uint32_t* cand_num = (uint32_t*) calloc(n ? n : 1, 4);
  uint32_t* cands = (uint32_t*) malloc((size_t) (n ? n : 1) * OR_MAX_CANDIDATES_EVENT * 4);
  oracle_estimate_input_size(
    events, offsets, n, out->module_cluster_start, out->module_cluster_num, cand_num, cands, out->status);
  oracle_prefix_sum(out->module_cluster_start, n * OR_N_MODULES);
  const uint32_t total = out->module_cluster_start[n * OR_N_MODULES];
  out->total_estimated_hits = total;
  out->cluster_container = (uint32_t*) calloc((size_t) 6 * total + 1, 4);
  oracle_masked_clustering(
    events, offsets, n, out->module_cluster_start, out->module_cluster_num, geometry, out->cluster_container,
    out->status);
  uint32_t* perm = (uint32_t*) calloc((size_t) total + 1, 4);
  oracle_calculate_phi_and_sort(
    n, out->module_cluster_start, out->module_cluster_num, out->cluster_container, perm, out->status);
  int16_t* h0c = (int16_t*) calloc((size_t) 2 * total + 2, 2);
  int16_t* h2c = (int16_t*) calloc((size_t) 2 * total + 2, 2);
  oracle_fill_candidates(n, out->cluster_container, out->module_cluster_start, out->module_cluster_num, h0c, h2c);
  or_track_hits* tracks = (or_track_hits*) calloc((size_t) (n ? n : 1) * OR_MAX_TRACKS, sizeof(or_track_hits));
  or_tracklet_hits* weak =
    (or_tracklet_hits*) calloc((size_t) (n ? n : 1) * OR_MAX_WEAK_TRACKS, sizeof(or_tracklet_hits));
  uint8_t* used = (uint8_t*) calloc((size_t) total + 1, 1);
  out->atomics_velo = (uint32_t*) calloc((size_t) n * OR_NUM_ATOMICS + 1, 4);
  oracle_search_by_triplet(
    n, out->cluster_container, out->module_cluster_start, out->module_cluster_num, tracks, weak, used,
    (int32_t*) out->atomics_velo, h0c, h2c, module_zs, out->status);
  oracle_weak_tracks_adder(
    n, out->cluster_container, out->module_cluster_start, tracks, weak, used, (int32_t*) out->atomics_velo,
    out->status);
  oracle_copy_and_prefix_sum(n, out->atomics_velo);
  const uint32_t total_tracks = out->atomics_velo[2 * n];
  out->total_tracks = total_tracks;
  out->velo_track_hit_number = (uint32_t*) calloc((size_t) total_tracks + 1, 4);
  oracle_copy_velo_track_hit_number(n, tracks, out->atomics_velo, out->velo_track_hit_number);
  oracle_prefix_sum(out->velo_track_hit_number, total_tracks);
  const uint32_t total_hits = out->velo_track_hit_number[total_tracks];
  out->total_track_hits = total_hits;
  out->velo_track_hits = (char*) calloc((size_t) 16 * total_hits + 16, 1);
  out->velo_states = (char*) calloc((size_t) 49 * total_tracks + 49, 1);
  oracle_consolidate_velo_tracks(
    n, out->atomics_velo, tracks, out->velo_track_hit_number, out->cluster_container, out->module_cluster_start,
    out->velo_track_hits, out->velo_states);
  free(cand_num);
  free(cands);
  free(perm);
  free(h0c);
  free(h2c);
  free(tracks);
  free(weak);
  free(used);
  return 0;
}

void oracle_pipeline_free(or_pipeline_result* r)
{
  free(r->status);
  free(r->module_cluster_start);
  free(r->module_cluster_num);
  free(r->cluster_container);
  free(r->atomics_velo);
  free(r->velo_track_hit_number);
  free(r->velo_track_hits);
  free(r->velo_states);
  memset(r, 0, sizeof(*r));
}

/* GlobalEventCut.cu:3-55, one event after the other.  The SciFi estimate
 * is the summed bank bytes without the 4-byte header word, halved, minus 2,
 * all in 32-bit unsigned arithmetic (an event with under 4 bytes of SciFi
 * data wraps and is cut).  UT banks are added in bank order, one valid
 * order of the reference's atomics, with its per-bank early exit. */
uint32_t oracle_global_event_cut(
  uint32_t n_events, const char* ut_raw, const uint32_t* ut_offsets, const char* scifi_raw,
  const uint32_t* scifi_offsets, uint32_t* event_list)
{
  uint32_t selected = 0;
  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t* sw = (const uint32_t*) (scifi_raw + scifi_offsets[e]);
    const uint32_t scifi_banks = sw[0];
    uint32_t n_scifi = 0;
    for (uint32_t i = 0; i < scifi_banks; ++i) n_scifi += sw[1 + i + 1] - sw[1 + i] - 4u;
    n_scifi = n_scifi / 2u - 2u;
    if (n_scifi >= OR_GEC_MAX_SCIFI_UT_CLUSTERS) continue;

    const uint32_t* uw = (const uint32_t*) (ut_raw + ut_offsets[e]);
    const uint32_t ut_banks = uw[0];
    const char* payload = (const char*) (uw + 2 + ut_banks);
    uint32_t n_ut = 0;
    int cut = 0;
    for (uint32_t i = 0; i < ut_banks && !cut; ++i) {
      uint32_t head;
      memcpy(&head, payload + uw[1 + i] + 4, 4);
      const uint32_t hits = head & 0xFFFFu;
      const uint32_t before = n_ut;
      n_ut += hits;
      if (before + hits >= OR_GEC_MAX_SCIFI_UT_CLUSTERS) cut = 1;
    }
    if (cut || n_ut + n_scifi >= OR_GEC_MAX_SCIFI_UT_CLUSTERS) continue;
    event_list[selected++] = e;
  }
  return selected;
}

/* ======================================================================
 * Beamline primary vertices (cuda/PV/beamlinePV/src/*.cu)
 * ====================================================================== */

#define OR_PV_ZMIN (-300.f)
#define OR_PV_ZMAX 300.f
#define OR_PV_DZ 0.25f
#define OR_PV_MAX_TRACK_Z0_ERR 1.5f
#define OR_PV_MIN_DENSITY 1.0f
#define OR_PV_MIN_DIP_DENSITY 2.0f
#define OR_PV_MIN_TRACKS_IN_SEED 2.5f
#define OR_PV_MAX_VERTEX_RHO2 0.09f
#define OR_PV_MAX_FIT_ITER 2u
#define OR_PV_MAX_DELTA_CHI2 9.f
#define OR_PV_MIN_WEIGHT 0.3f
#define OR_PV_MIN_TRACKS_PER_VERTEX 5u

/* e^x by Cody-Waite reduction (x = k ln2 + r, ln2 split in a 16-bit high
 * part and a low part), a degree-7 Taylor polynomial of e^r in Horner form
 * and ldexp by k; the beamline fit's exponential. */
float oracle_pv_exp(float x)
{
  if (!(x == x)) return x;
  if (x > 88.72284f) return INFINITY;
  if (x < -104.0f) return 0.0f;
  const float k = rintf(x * 1.44269504f);
  const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-06f;
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

/* pv_beamline_histo.cu:3-11 (polynomial order 2, a = sqrt(7)). */
float oracle_pv_gauss_integral(float x)
{
  const float a = sqrtf(7.0f);
  const float xi = x / a;
  const float eta = 1.0f - xi * xi;
  return 0.5f + xi * (0.5f + eta * (0.25f + eta * 0.1875f));
}

/* pv_beamline_extrapolate.cu:3-30 with the PVTrack constructor
 * (TrackBeamLineVertexFinder.cuh:12-25): closest approach to (0, 0). */
void oracle_pv_beamline_extrapolate(
  uint32_t n_events, const uint32_t* atomics_velo, const char* kalman_states, float* pvtracks)
{
  const size_t T = atomics_velo[2 * n_events];
  const float* S = (const float*) kalman_states;
  for (size_t t = 0; t < T; ++t) {
    const float x = S[t], y = S[T + t], tx = S[2 * T + t], ty = S[3 * T + t];
    const float c00 = S[4 * T + t], c20 = S[5 * T + t], c22 = S[6 * T + t];
    const float c11 = S[7 * T + t], c31 = S[8 * T + t], c33 = S[9 * T + t];
    const float z = S[11 * T + t];
    const float dz = (tx * (0.f - x) + ty * (0.f - y)) / (tx * tx + ty * ty);
    float tmp_c00 = c00, tmp_c11 = c11;
    const float dz2 = dz * dz;
    tmp_c00 += dz2 * c22 + 2.f * dz * c20;
    tmp_c11 += dz2 * c33 + 2.f * dz * c31;
    float* p = pvtracks + 7 * t;
    p[0] = z + dz;
    p[1] = x + dz * tx;
    p[2] = y + dz * ty;
    p[3] = tx;
    p[4] = ty;
    p[5] = 1.f / tmp_c00;
    p[6] = 1.f / tmp_c11;
  }
}

/* pv_beamline_histo.cu:13-76, tracks in order. */
void oracle_pv_beamline_histo(uint32_t n_events, const uint32_t* atomics_velo, const float* pvtracks, float* zhisto)
{
  const float a = sqrtf(7.0f);
  for (uint32_t e = 0; e < n_events; ++e) {
    float* h = zhisto + (size_t) e * OR_PV_NBINS;
    for (uint32_t i = 0; i < OR_PV_NBINS; ++i) h[i] = 0.f;
    const uint32_t n = atomics_velo[e], off = atomics_velo[n_events + e];
    for (uint32_t k = 0; k < n; ++k) {
      const float* trk = pvtracks + 7 * ((size_t) off + k);
      const float z = trk[0], tx = trk[3], ty = trk[4], w00 = trk[5], w11 = trk[6];
      if (!(OR_PV_ZMIN < z && z < OR_PV_ZMAX)) continue;
      const float zbin = (z - OR_PV_ZMIN) / OR_PV_DZ;
      const float zweight = tx * tx * w00 + ty * ty * w11;
      const float zerr = 1.f / sqrtf(zweight);
      if (!(zerr < OR_PV_MAX_TRACK_Z0_ERR)) continue;
      const float halfwindow = a * zerr / OR_PV_DZ;
      int minbin = (int) (zbin - halfwindow);
      if (minbin < 1) minbin = 1;
      int maxbin = (int) (zbin + halfwindow);
      if (maxbin > (int) OR_PV_NBINS - 2) maxbin = (int) OR_PV_NBINS - 2;
      if (maxbin < minbin) continue;
      float integral = 0;
      for (int i = minbin; i < maxbin; ++i) {
        const float relz = (OR_PV_ZMIN + (i + 1) * OR_PV_DZ - z) / zerr;
        const float thisintegral = oracle_pv_gauss_integral(relz);
        h[i] += thisintegral - integral;
        integral = thisintegral;
      }
      h[maxbin] += 1.f - integral;
    }
  }
}

/* pv_beamline_peak.cu:3-147 for one event; h[OR_PV_NBINS] reads as 0 (the
 * reference reads the next event's first bin there, always empty).  A push
 * beyond an array's size (200 edges, 32 extrema, 50 subclusters, 200
 * clusters, 32 peaks) is dropped. */
static uint32_t oracle_pv_peaks_event(const float* h, float* zpeaks)
{
  float zh[OR_PV_NBINS + 1];
  memcpy(zh, h, OR_PV_NBINS * sizeof(float));
  zh[OR_PV_NBINS] = 0.f;
  unsigned short edges[200];
  unsigned n_edges = 0;
  unsigned short cluster_peak[200];
  unsigned n_clusters = 0;
  {
    const float threshold = OR_PV_DZ * (1.f / (10.f * OR_PV_MAX_TRACK_Z0_ERR));
    int prevempty = 1;
    float integral = zh[0];
    for (int i = 1; i < (int) OR_PV_NBINS; ++i) {
      integral += zh[i];
      const int empty = zh[i] < threshold;
      if (empty != prevempty) {
        if (prevempty || integral > OR_PV_MIN_TRACKS_IN_SEED) {
          if (n_edges < 200) edges[n_edges++] = (unsigned short) i;
        }
        else {
          n_edges--;
        }
        prevempty = empty;
        integral = 0;
      }
    }
  }
  const unsigned n_proto = n_edges / 2;
  for (unsigned p = 0; p < n_proto; ++p) {
    const int ibegin = edges[2 * p], iend = edges[2 * p + 1];
    const float mindip = OR_PV_MIN_DIP_DENSITY * OR_PV_DZ;
    const float minpeak = OR_PV_MIN_DENSITY * OR_PV_DZ;
    unsigned short ex_index[32];
    float ex_value[32], ex_integral[32];
    int n_ex = 0;
#define OR_PV_PUSH(I, V, IN)                                                                                          \
  do {                                                                                                                \
    if (n_ex < 32) {                                                                                                  \
      ex_index[n_ex] = (unsigned short) (I);                                                                          \
      ex_value[n_ex] = (V);                                                                                           \
      ex_integral[n_ex] = (IN);                                                                                       \
      ++n_ex;                                                                                                         \
    }                                                                                                                 \
  } while (0)
    {
      int rising = 1;
      float integral = zh[ibegin];
      OR_PV_PUSH(ibegin, zh[ibegin], integral);
      for (int i = ibegin; i < iend; ++i) {
        const float value = zh[i];
        const int stillrising = zh[i + 1] > value;
        if (rising && !stillrising && value >= minpeak) {
          const int n = n_ex;
          if (n >= 2) {
            const float dv1 = ex_value[n - 2] - ex_value[n - 1];
            const float dv2 = value - ex_value[n - 1];
            if (dv1 > mindip && dv2 > mindip) {
              OR_PV_PUSH(i, value, integral + 0.5f * value);
            }
            else if (dv1 > dv2) {
              n_ex = n_ex - 1 < 0 ? 0 : n_ex - 1;
            }
            else {
              n_ex = n_ex - 2 < 0 ? 0 : n_ex - 2;
              OR_PV_PUSH(i, value, integral + 0.5f * value);
            }
          }
          else {
            OR_PV_PUSH(i, value, integral + 0.5f * value);
          }
        }
        else if (rising != stillrising) {
          OR_PV_PUSH(i, value, integral + 0.5f * value);
        }
        rising = stillrising;
        integral += value;
      }
      OR_PV_PUSH(iend, zh[iend], integral);
    }
#undef OR_PV_PUSH
    const int n = n_ex;
    unsigned short sub_peak[50];
    unsigned n_sub = 0;
    if (n > 3) {
      for (int i = 1; i < n / 2 + 1; ++i)
        if (ex_integral[2 * i] - ex_integral[2 * i - 2] > OR_PV_MIN_TRACKS_IN_SEED && n_sub < 50)
          sub_peak[n_sub++] = ex_index[2 * i - 1];
    }
    if (n_sub == 0) {
      if (ex_value[1] >= minpeak && n_clusters < 200) cluster_peak[n_clusters++] = ex_index[1];
    }
    else {
      for (unsigned i = 0; i < n_sub && n_clusters < 200; ++i) cluster_peak[n_clusters++] = sub_peak[i];
    }
  }
  uint32_t n_peaks = 0;
  for (unsigned i = 0; i < n_clusters && n_peaks < OR_PV_MAX_VERTICES; ++i) {
    const int izmax = cluster_peak[i];
    const float d1 = zh[izmax] - zh[izmax - 1];
    const float d2 = zh[izmax] - zh[izmax + 1];
    const float idz = d1 + d2 > 0 ? 0.5f * (d1 - d2) / (d1 + d2) : 0.0f;
    zpeaks[n_peaks++] = OR_PV_ZMIN + OR_PV_DZ * ((float) izmax + idz + 0.5f);
  }
  return n_peaks;
}

void oracle_pv_beamline_peak(uint32_t n_events, const float* zhisto, float* zpeaks, uint32_t* n_zpeaks)
{
  for (uint32_t e = 0; e < n_events; ++e)
    n_zpeaks[e] =
      oracle_pv_peaks_event(zhisto + (size_t) e * OR_PV_NBINS, zpeaks + (size_t) e * OR_PV_MAX_VERTICES);
}

/* pv_beamline_multi_fitter.cu:5-183, one seed after the other, each track's
 * weight denominator summed over the seeds as the reference does. */
void oracle_pv_beamline_multi_fitter(
  uint32_t n_events, const uint32_t* atomics_velo, const float* pvtracks, const float* zpeaks,
  const uint32_t* n_zpeaks, uint32_t* vertices, uint32_t* n_vertices)
{
  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t n = atomics_velo[e], off = atomics_velo[n_events + e];
    const float* seeds = zpeaks + (size_t) e * OR_PV_MAX_VERTICES;
    uint32_t S = n_zpeaks[e];
    if (S > OR_PV_MAX_VERTICES) S = OR_PV_MAX_VERTICES;
    uint32_t kept = 0;
    for (uint32_t s = 0; s < S; ++s) {
      int converged = 0;
      float cov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float vx = 0.f, vy = 0.f, vz = seeds[s];
      float chi2tot = 0.f;
      for (unsigned iter = 0; iter < OR_PV_MAX_FIT_ITER && !converged; ++iter) {
        float h00 = 0.f, h11 = 0.f, h20 = 0.f, h21 = 0.f, h22 = 0.f;
        float gx = 0.f, gy = 0.f, gz = 0.f;
        chi2tot = 0.f;
        unsigned nsel = 0;
        for (uint32_t k = 0; k < n; ++k) {
          const float* trk = pvtracks + 7 * ((size_t) off + k);
          const float z = trk[0], x = trk[1], y = trk[2], tx = trk[3], ty = trk[4], w00 = trk[5], w11 = trk[6];
          if (OR_PV_ZMIN > z || z > OR_PV_ZMAX) continue;
          const float dz = vz - z;
          const float rx = vx - (x + tx * dz);
          const float ry = vy - (y + ty * dz);
          const float chi2 = rx * rx * w00 + ry * ry * w11;
          if (!(chi2 < OR_PV_MAX_DELTA_CHI2)) continue;
          ++nsel;
          const float chi2_cut = 0.1f + 0.01f * (float) OR_PV_MAX_FIT_ITER / (float) (iter + 1);
          float weight = oracle_pv_exp(-chi2 * 0.5f);
          float denom = oracle_pv_exp(-chi2_cut * 0.5f);
          for (uint32_t o = 0; o < S; ++o) {
            const float dzo = seeds[o] - z;
            const float ox = 0.f - (x + tx * dzo);
            const float oy = 0.f - (y + ty * dzo);
            const float chi2o = ox * ox * w00 + oy * oy * w11;
            denom += oracle_pv_exp(-chi2o * 0.5f);
          }
          weight = weight / denom;
          if (weight < OR_PV_MIN_WEIGHT) continue;
          const float hx = rx * w00;
          const float hy = ry * w11;
          const float hz = -tx * rx * w00 - ty * ry * w11;
          gx = gx + hx * weight;
          gy = gy + hy * weight;
          gz = gz + hz * weight;
          h00 += weight * w00;
          h11 += weight * w11;
          h20 += weight * (-tx * w00);
          h21 += weight * (-ty * w11);
          h22 += weight * (-tx * -tx * w00 + -ty * -ty * w11);
          chi2tot += weight * chi2;
        }
        if (nsel >= 2) {
          const float det = h00 * (h22 * h11 - h21 * h21) + h20 * (-h11 * h20);
          const float inv_det = 1.f / det;
          cov[0] = (h22 * h11 - h21 * h21) * inv_det;
          cov[1] = -(-h20 * h21) * inv_det;
          cov[2] = (h22 * h00 - h20 * h20) * inv_det;
          cov[3] = (-h20 * h11) * inv_det;
          cov[4] = -(h21 * h00) * inv_det;
          cov[5] = (h11 * h00) * inv_det;
          const float dx = -1.f * (cov[0] * gx + cov[1] * gy + cov[3] * gz);
          const float dy = -1.f * (cov[1] * gx + cov[2] * gy + cov[4] * gz);
          const float dzv = -1.f * (cov[3] * gx + cov[4] * gy + cov[5] * gz);
          chi2tot += dx * gx + dy * gy + dzv * gz;
          vx = vx + dx;
          vy = vy + dy;
          vz = vz + dzv;
          converged = fabsf(dzv) < 0.001f;
        }
        else {
          break;
        }
      }
      const float rho2 = (vx - 0.f) * (vx - 0.f) + (vy - 0.f) * (vy - 0.f);
      if (n >= OR_PV_MIN_TRACKS_PER_VERTEX && rho2 < OR_PV_MAX_VERTEX_RHO2) {
        uint32_t* v = vertices + ((size_t) e * OR_PV_MAX_VERTICES + kept) * OR_PV_VERTEX_WORDS;
        const float f[4] = {vx, vy, vz, chi2tot};
        memcpy(v, f, sizeof f);
        v[4] = 0;
        v[5] = n;
        memcpy(v + 6, cov, sizeof cov);
        v[12] = 0;
        ++kept;
      }
    }
    n_vertices[e] = kept;
  }
}
