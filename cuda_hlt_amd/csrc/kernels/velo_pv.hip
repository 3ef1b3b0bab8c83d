// Beamline primary-vertex finding (cuda/PV/beamlinePV/src/*.cu) for gfx950.
//
// extrapolate: one lane per track, the SoA states read coalesced.
// histo: one wave per event with the 2400-bin histogram in LDS; tracks are
//   taken in order (staged 64 at a time) and the bins of one track's window
//   (at most ~33) are spread over the lanes, so every bin receives its
//   shares in track order without atomics.
// peak: one wave per event stages the histogram in LDS (coalesced) and
//   lane 0 runs the reference's sequential cluster search on it.
// multi_fitter: one wave per event, one lane per seed.  The weight
//   denominator of a track, exp(-chi2_cut/2) + sum over seeds of
//   exp(-chi2(seed)/2), does not depend on the vertex being fitted, so per
//   iteration each lane computes it for one track of a 64-track chunk (in
//   the reference's order) and all seed lanes read it from LDS.
#include "velo_device.h"
#include "velo_pv.h"
#include "velo_pv_math.h"

namespace velo {

// BeamlinePVConstants.cuh:3-17, PV_Definitions.cuh:10-22
constexpr float kPvZmin = VELO_PV_ZMIN;
constexpr float kPvZmax = VELO_PV_ZMAX;
constexpr float kPvDz = VELO_PV_DZ;
constexpr int kPvBins = (int) VELO_PV_NBINS;
constexpr float kMaxTrackZ0Err = 1.5f;
constexpr float kMinDensity = 1.0f;
constexpr float kMinDipDensity = 2.0f;
constexpr float kMinTracksInSeed = 2.5f;
constexpr float kMaxVertexRho2 = 0.09f;
constexpr unsigned kMaxFitIter = 2;
constexpr float kMaxDeltaChi2 = 9.f;
constexpr float kMinWeight = 0.3f;
constexpr unsigned kMinTracksPerVertex = 5;
constexpr unsigned kMaxPv = VELO_PV_MAX_VERTICES;
constexpr unsigned kMaxClusters = 200;
constexpr unsigned kMaxEdges = 200;
constexpr unsigned kMaxExtrema = 32;
constexpr unsigned kMaxSubclusters = 50;
constexpr unsigned kPvTrackFloats = 7;

struct PvEvent {
  uint32_t n, offset, total;
};

__device__ inline PvEvent pv_event(const int32_t* atomics, uint32_t N, uint32_t e)
{
  return {(uint32_t) atomics[e], (uint32_t) atomics[N + e], (uint32_t) atomics[2 * N]};
}

// ---- pv_beamline_extrapolate (pv_beamline_extrapolate.cu:3-30,
// TrackBeamLineVertexFinder.cuh:12-25) ----
__global__ void __launch_bounds__(1024) pv_extrapolate_kernel(velo_pv_beamline_extrapolate_args a)
{
  const uint32_t N = a.n_events;
  for_each_event(a.tuning, N, [&](const uint32_t e) {
    const PvEvent ev = pv_event(a.dev_atomics_velo, N, e);
    const float* S = reinterpret_cast<const float*>(a.dev_velo_kalman_beamline_states);
    const size_t T = ev.total;
    for (uint32_t i = threadIdx.x; i < ev.n; i += blockDim.x) {
      const size_t t = ev.offset + i;
      const float x = S[t], y = S[T + t], tx = S[2 * T + t], ty = S[3 * T + t];
      const float c00 = S[4 * T + t], c20 = S[5 * T + t], c22 = S[6 * T + t];
      const float c11 = S[7 * T + t], c31 = S[8 * T + t], c33 = S[9 * T + t];
      const float z = S[11 * T + t];
      const float dz = (tx * (0.f - x) + ty * (0.f - y)) / (tx * tx + ty * ty);
      const float dz2 = dz * dz;
      const float cxx = c00 + (dz2 * c22 + 2.f * dz * c20);
      const float cyy = c11 + (dz2 * c33 + 2.f * dz * c31);
      velo_pv_track p;
      p.z = z + dz;
      p.x = x + dz * tx;
      p.y = y + dz * ty;
      p.tx = tx;
      p.ty = ty;
      p.W_00 = 1.f / cxx;
      p.W_11 = 1.f / cyy;
      a.dev_pvtracks[t] = p;
    }
  });
}

// ---- pv_beamline_histo (pv_beamline_histo.cu:13-76) ----
__global__ void __launch_bounds__(64) pv_histo_kernel(velo_pv_beamline_histo_args a)
{
  __shared__ float h[kPvBins];
  __shared__ float trk[kWave][kPvTrackFloats];
  const uint32_t N = a.n_events;
  const uint32_t lane = threadIdx.x;
  for_each_event(a.tuning, N, [&](const uint32_t e) {
    const PvEvent ev = pv_event(a.dev_atomics_velo, N, e);
    for (int b = lane; b < kPvBins; b += kWave) h[b] = 0.f;
    const float wnd = sqrtf(7.0f);
    for (uint32_t base = 0; base < ev.n; base += kWave) {
      __syncthreads();
      if (base + lane < ev.n) {
        const float* p = reinterpret_cast<const float*>(a.dev_pvtracks + ev.offset + base + lane);
        for (uint32_t f = 0; f < kPvTrackFloats; ++f) trk[lane][f] = p[f];
      }
      __syncthreads();
      const uint32_t cnt = min((uint32_t) kWave, ev.n - base);
      for (uint32_t k = 0; k < cnt; ++k) {
        const float z = trk[k][0], tx = trk[k][3], ty = trk[k][4], w00 = trk[k][5], w11 = trk[k][6];
        if (!(kPvZmin < z && z < kPvZmax)) continue;
        const float zbin = (z - kPvZmin) / kPvDz;
        const float zweight = tx * tx * w00 + ty * ty * w11;
        const float zerr = 1.f / sqrtf(zweight);
        if (!(zerr < kMaxTrackZ0Err)) continue;
        const float halfwindow = wnd * zerr / kPvDz;
        const int minbin = max((int) (zbin - halfwindow), 1);
        const int maxbin = min((int) (zbin + halfwindow), kPvBins - 2);
        for (int i = minbin + (int) lane; i <= maxbin; i += kWave) {
          const float prev = i == minbin ? 0.f : pv_gauss_integral((kPvZmin + (float) i * kPvDz - z) / zerr);
          const float share =
            i < maxbin ? pv_gauss_integral((kPvZmin + (float) (i + 1) * kPvDz - z) / zerr) - prev : 1.f - prev;
          h[i] += share;
        }
      }
    }
    __syncthreads();
    float* out = a.dev_zhisto + (size_t) e * kPvBins;
    for (int b = lane; b < kPvBins; b += kWave) out[b] = h[b];
    __syncthreads();
  });
}

// ---- pv_beamline_peak (pv_beamline_peak.cu:3-147) ----
// Work arrays of the search (the reference's sizes).  Only a cluster's
// peak bin (izmax) reaches the output, so clusters keep just that.
struct PvPeakScratch {
  uint16_t edges[kMaxEdges];
  uint16_t cluster_peak[kMaxClusters];
  uint16_t sub_peak[kMaxSubclusters];
  uint16_t ex_index[kMaxExtrema];
  float ex_value[kMaxExtrema];
  float ex_integral[kMaxExtrema];
};

// The reference's sequential search over one event's histogram `zh`
// (with zh[kPvBins] = 0, the value the reference reads there: the next
// event's first bin, always empty).  A push beyond an array's size is
// dropped.
__device__ inline uint32_t pv_find_peaks(const float* zh, PvPeakScratch& w, float* zpeaks)
{
  uint32_t n_clusters = 0;
  uint32_t n_edges = 0;
  {
    const float threshold = kPvDz * (1.f / (10.f * kMaxTrackZ0Err));
    bool prevempty = true;
    float integral = zh[0];
    for (int i = 1; i < kPvBins; ++i) {
      integral += zh[i];
      const bool empty = zh[i] < threshold;
      if (empty != prevempty) {
        if (prevempty || integral > kMinTracksInSeed) {
          if (n_edges < kMaxEdges) w.edges[n_edges++] = (uint16_t) i;
        }
        else {
          --n_edges;
        }
        prevempty = empty;
        integral = 0;
      }
    }
  }
  const uint32_t n_proto = n_edges / 2;
  for (uint32_t p = 0; p < n_proto; ++p) {
    const int ibegin = w.edges[p * 2], iend = w.edges[p * 2 + 1];
    const float mindip = kMinDipDensity * kPvDz;
    const float minpeak = kMinDensity * kPvDz;
    int n_ex = 0;
    auto push = [&](int i, float v, float in) {
      if (n_ex < (int) kMaxExtrema) {
        w.ex_index[n_ex] = (uint16_t) i;
        w.ex_value[n_ex] = v;
        w.ex_integral[n_ex] = in;
        ++n_ex;
      }
    };
    {
      bool rising = true;
      float integral = zh[ibegin];
      push(ibegin, zh[ibegin], integral);
      for (int i = ibegin; i < iend; ++i) {
        const float value = zh[i];
        const bool stillrising = zh[i + 1] > value;
        if (rising && !stillrising && value >= minpeak) {
          const int n = n_ex;
          if (n >= 2) {
            const float dv1 = w.ex_value[n - 2] - w.ex_value[n - 1];
            const float dv2 = value - w.ex_value[n - 1];
            if (dv1 > mindip && dv2 > mindip) {
              push(i, value, integral + 0.5f * value);
            }
            else if (dv1 > dv2) {
              n_ex = max(n_ex - 1, 0);
            }
            else {
              n_ex = max(n_ex - 2, 0);
              push(i, value, integral + 0.5f * value);
            }
          }
          else {
            push(i, value, integral + 0.5f * value);
          }
        }
        else if (rising != stillrising) {
          push(i, value, integral + 0.5f * value);
        }
        rising = stillrising;
        integral += value;
      }
      push(iend, zh[iend], integral);
    }
    const int n = n_ex;
    uint32_t n_sub = 0;
    if (n > 3) {
      for (int i = 1; i < n / 2 + 1; ++i) {
        if (w.ex_integral[2 * i] - w.ex_integral[2 * i - 2] > kMinTracksInSeed && n_sub < kMaxSubclusters)
          w.sub_peak[n_sub++] = w.ex_index[2 * i - 1];
      }
    }
    if (n_sub == 0) {
      if (w.ex_value[1] >= minpeak && n_clusters < kMaxClusters) w.cluster_peak[n_clusters++] = w.ex_index[1];
    }
    else {
      // the reference's boundary adjustment of the first and last
      // subcluster leaves their peaks as they are
      for (uint32_t i = 0; i < n_sub && n_clusters < kMaxClusters; ++i) w.cluster_peak[n_clusters++] = w.sub_peak[i];
    }
  }
  uint32_t n_peaks = 0;
  for (uint32_t i = 0; i < n_clusters && n_peaks < kMaxPv; ++i) {
    const int izmax = w.cluster_peak[i];
    const float d1 = zh[izmax] - zh[izmax - 1];
    const float d2 = zh[izmax] - zh[izmax + 1];
    const float idz = d1 + d2 > 0 ? 0.5f * (d1 - d2) / (d1 + d2) : 0.0f;
    zpeaks[n_peaks++] = kPvZmin + kPvDz * ((float) izmax + idz + 0.5f);
  }
  return n_peaks;
}

// The same search with the work spread over the wave.  The edge pass of
// the reference decomposes exactly: every empty -> filled transition opens
// an edge, and the next filled -> empty transition keeps the pair when the
// sum of the bins after the opening transition up to and including the
// closing one (in bin order, from 0) exceeds kMinTracksInSeed, else drops
// it.  So each lane sums one pair's bins in that order, and each lane then
// runs the extrema and subcluster pass of one kept pair (proto-cluster);
// the peaks come out in proto-cluster order through a wave prefix sum.
// Every float operation is the sequential search's, in its order, so the
// results are identical.  When more than 99 pairs are kept (the sequential
// search would then reach its 200-edge limit) or transitions exceed the
// table, lane 0 runs the sequential search instead.
constexpr uint32_t kMaxTransitions = 512;
constexpr uint32_t kPeakLanes = 32; // lanes of the extrema pass

struct PvPeakParallel {
  uint16_t trans[kMaxTransitions];
  uint16_t kept[kMaxTransitions];  // kept pairs: begin, end, begin, end, ...
  uint16_t ex_index[kPeakLanes][kMaxExtrema];
  float ex_value[kPeakLanes][kMaxExtrema];
  float ex_integral[kPeakLanes][kMaxExtrema];
  uint32_t n_trans, n_kept;
};

// Extrema of the proto-cluster [ibegin, iend] into lane slot `w` (the
// reference's loop, pv_beamline_peak.cu:43-96); returns their number.
__device__ inline int pv_extrema(const float* zh, int ibegin, int iend, uint16_t* ex_index, float* ex_value,
                                 float* ex_integral)
{
  const float mindip = kMinDipDensity * kPvDz;
  const float minpeak = kMinDensity * kPvDz;
  int n_ex = 0;
  auto push = [&](int i, float v, float in) {
    if (n_ex < (int) kMaxExtrema) {
      ex_index[n_ex] = (uint16_t) i;
      ex_value[n_ex] = v;
      ex_integral[n_ex] = in;
      ++n_ex;
    }
  };
  bool rising = true;
  float integral = zh[ibegin];
  push(ibegin, zh[ibegin], integral);
  for (int i = ibegin; i < iend; ++i) {
    const float value = zh[i];
    const bool stillrising = zh[i + 1] > value;
    if (rising && !stillrising && value >= minpeak) {
      const int n = n_ex;
      if (n >= 2) {
        const float dv1 = ex_value[n - 2] - ex_value[n - 1];
        const float dv2 = value - ex_value[n - 1];
        if (dv1 > mindip && dv2 > mindip) {
          push(i, value, integral + 0.5f * value);
        }
        else if (dv1 > dv2) {
          n_ex = max(n_ex - 1, 0);
        }
        else {
          n_ex = max(n_ex - 2, 0);
          push(i, value, integral + 0.5f * value);
        }
      }
      else {
        push(i, value, integral + 0.5f * value);
      }
    }
    else if (rising != stillrising) {
      push(i, value, integral + 0.5f * value);
    }
    rising = stillrising;
    integral += value;
  }
  push(iend, zh[iend], integral);
  return n_ex;
}

__device__ inline float pv_cluster_mean(const float* zh, int izmax)
{
  const float d1 = zh[izmax] - zh[izmax - 1];
  const float d2 = zh[izmax] - zh[izmax + 1];
  const float idz = d1 + d2 > 0 ? 0.5f * (d1 - d2) / (d1 + d2) : 0.0f;
  return kPvZmin + kPvDz * ((float) izmax + idz + 0.5f);
}

__global__ void __launch_bounds__(64) pv_peak_kernel(velo_pv_beamline_peak_args a)
{
  __shared__ float zh[kPvBins + 1];
  __shared__ PvPeakParallel w;
  __shared__ PvPeakScratch fallback;
  const uint32_t N = a.n_events;
  const uint32_t lane = threadIdx.x;
  const float threshold = kPvDz * (1.f / (10.f * kMaxTrackZ0Err));
  for_each_event(a.tuning, N, [&](const uint32_t e) {
    const float* in = a.dev_zhisto + (size_t) e * kPvBins;
    for (int b = lane; b < kPvBins; b += kWave) zh[b] = in[b];
    if (lane == 0) zh[kPvBins] = 0.f;
    __syncthreads();

    // transitions of the empty flag (the flag before bin 1 counts as empty)
    uint32_t n_trans = 0;
    for (int base = 1; base < kPvBins; base += kWave) {
      const int i = base + (int) lane;
      bool t = false;
      if (i < kPvBins) {
        const bool prev = i == 1 ? true : zh[i - 1] < threshold;
        t = (zh[i] < threshold) != prev;
      }
      const uint64_t m = ballot(t);
      const uint32_t at = n_trans + count_below(m);
      if (t && at < kMaxTransitions) w.trans[at] = (uint16_t) i;
      n_trans += (uint32_t) __popcll(m);
    }
    __syncthreads();

    // pairs (opening, closing transition) kept by their bin sums
    const uint32_t n_pairs = min(n_trans, kMaxTransitions) / 2;
    uint32_t n_kept = 0;
    for (uint32_t base = 0; base < n_pairs; base += kWave) {
      const uint32_t k = base + lane;
      bool keep = false;
      int t0 = 0, t1 = 0;
      if (k < n_pairs) {
        t0 = w.trans[2 * k];
        t1 = w.trans[2 * k + 1];
        float integral = 0;
        for (int i = t0 + 1; i <= t1; ++i) integral += zh[i];
        keep = integral > kMinTracksInSeed;
      }
      const uint64_t m = ballot(keep);
      const uint32_t at = n_kept + count_below(m);
      if (keep) {
        w.kept[2 * at] = (uint16_t) t0;
        w.kept[2 * at + 1] = (uint16_t) t1;
      }
      n_kept += (uint32_t) __popcll(m);
    }
    __syncthreads();

    float* out = a.dev_zpeaks + (size_t) e * kMaxPv;
    if (n_trans > kMaxTransitions || n_kept > 99) {
      if (lane == 0) a.dev_number_of_zpeaks[e] = pv_find_peaks(zh, fallback, out);
      __syncthreads();
      return;
    }
    // extrema pass, one lane per kept pair; peaks in pair order
    uint32_t n_peaks = 0;
    const float minpeak = kMinDensity * kPvDz;
    for (uint32_t base = 0; base < n_kept; base += kPeakLanes) {
      const uint32_t p = base + lane;
      uint32_t count = 0;
      int n = 0;
      if (lane < kPeakLanes && p < n_kept) {
        n = pv_extrema(zh, w.kept[2 * p], w.kept[2 * p + 1], w.ex_index[lane], w.ex_value[lane],
                       w.ex_integral[lane]);
        if (n > 3)
          for (int i = 1; i < n / 2 + 1; ++i)
            count += w.ex_integral[lane][2 * i] - w.ex_integral[lane][2 * i - 2] > kMinTracksInSeed ? 1u : 0u;
        count = min(count, kMaxSubclusters);
        if (count == 0 && w.ex_value[lane][1] >= minpeak) count = 1;
      }
      const uint32_t incl = wave_inclusive_scan(count);
      uint32_t at = n_peaks + incl - count;
      if (count) {
        uint32_t sub = 0;
        if (n > 3)
          for (int i = 1; i < n / 2 + 1 && sub < kMaxSubclusters; ++i)
            if (w.ex_integral[lane][2 * i] - w.ex_integral[lane][2 * i - 2] > kMinTracksInSeed) {
              if (at < kMaxPv) out[at] = pv_cluster_mean(zh, w.ex_index[lane][2 * i - 1]);
              ++at;
              ++sub;
            }
        if (sub == 0 && at < kMaxPv) out[at] = pv_cluster_mean(zh, w.ex_index[lane][1]);
      }
      n_peaks += __shfl(incl, kWave - 1, kWave);
    }
    if (lane == 0) a.dev_number_of_zpeaks[e] = min(n_peaks, kMaxPv);
    __syncthreads();
  });
}

// ---- pv_beamline_multi_fitter (pv_beamline_multi_fitter.cu:5-183) ----
__global__ void __launch_bounds__(64) pv_multi_fitter_kernel(velo_pv_beamline_multi_fitter_args a)
{
  __shared__ float trk[kWave][kPvTrackFloats];
  __shared__ float denom[kWave];
  __shared__ float seeds[kMaxPv];
  const uint32_t N = a.n_events;
  const uint32_t lane = threadIdx.x;
  for_each_event(a.tuning, N, [&](const uint32_t e) {
    const PvEvent ev = pv_event(a.dev_atomics_velo, N, e);
    const uint32_t S = min(a.dev_number_of_zpeaks[e], kMaxPv);
    if (lane < S) seeds[lane] = a.dev_zpeaks[(size_t) e * kMaxPv + lane];
    __syncthreads();
    const bool mine = lane < S;
    float vx = 0.f, vy = 0.f, vz = mine ? seeds[lane] : 0.f;
    float cov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float chi2tot = 0.f;
    bool active = mine;
    for (unsigned iter = 0; iter < kMaxFitIter; ++iter) {
      if (!__any(active)) break;
      float h00 = 0.f, h11 = 0.f, h20 = 0.f, h21 = 0.f, h22 = 0.f;
      float gx = 0.f, gy = 0.f, gz = 0.f;
      uint32_t nsel = 0;
      if (active) chi2tot = 0.f;
      const float chi2_cut = 0.1f + 0.01f * (float) kMaxFitIter / (float) (iter + 1);
      for (uint32_t base = 0; base < ev.n; base += kWave) {
        __syncthreads();
        if (base + lane < ev.n) {
          const float* p = reinterpret_cast<const float*>(a.dev_pvtracks + ev.offset + base + lane);
          float f[kPvTrackFloats];
          for (uint32_t i = 0; i < kPvTrackFloats; ++i) f[i] = trk[lane][i] = p[i];
          float d = pv_exp(-chi2_cut * 0.5f);
          for (uint32_t s = 0; s < S; ++s) {
            const float dz = seeds[s] - f[0];
            const float rx = 0.f - (f[1] + f[3] * dz);
            const float ry = 0.f - (f[2] + f[4] * dz);
            const float c = rx * rx * f[5] + ry * ry * f[6];
            d += pv_exp(-c * 0.5f);
          }
          denom[lane] = d;
        }
        __syncthreads();
        if (active) {
          const uint32_t cnt = min((uint32_t) kWave, ev.n - base);
          for (uint32_t k = 0; k < cnt; ++k) {
            const float z = trk[k][0];
            if (kPvZmin > z || z > kPvZmax) continue;
            const float x = trk[k][1], y = trk[k][2], tx = trk[k][3], ty = trk[k][4];
            const float w00 = trk[k][5], w11 = trk[k][6];
            const float dz = vz - z;
            const float rx = vx - (x + tx * dz);
            const float ry = vy - (y + ty * dz);
            const float chi2 = rx * rx * w00 + ry * ry * w11;
            if (!(chi2 < kMaxDeltaChi2)) continue;
            ++nsel;
            const float w = pv_exp(-chi2 * 0.5f) / denom[k];
            if (w < kMinWeight) continue;
            const float hx = rx * w00;
            const float hy = ry * w11;
            const float hz = -tx * rx * w00 - ty * ry * w11;
            gx = gx + hx * w;
            gy = gy + hy * w;
            gz = gz + hz * w;
            const float hw20 = -tx * w00;
            const float hw21 = -ty * w11;
            const float hw22 = -tx * -tx * w00 + -ty * -ty * w11;
            h00 += w * w00;
            h11 += w * w11;
            h20 += w * hw20;
            h21 += w * hw21;
            h22 += w * hw22;
            chi2tot += w * chi2;
          }
        }
      }
      if (active) {
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
          if (fabsf(dzv) < 0.001f) active = false;
        }
        else {
          active = false;
        }
      }
    }
    // a seed's vertex is kept with >= 5 event tracks within 0.3 mm of the beam
    const float rho2 = (vx - 0.f) * (vx - 0.f) + (vy - 0.f) * (vy - 0.f);
    const bool keep = mine && ev.n >= kMinTracksPerVertex && rho2 < kMaxVertexRho2;
    const uint64_t ballot = __ballot(keep);
    if (keep) {
      velo_pv_vertex v;
      v.x = vx;
      v.y = vy;
      v.z = vz;
      v.chi2 = chi2tot;
      v.ndof = 0;
      v.n_tracks = ev.n;
      v.cov00 = cov[0];
      v.cov10 = cov[1];
      v.cov11 = cov[2];
      v.cov20 = cov[3];
      v.cov21 = cov[4];
      v.cov22 = cov[5];
      v.nTracks = 0;
      a.dev_multi_fit_vertices[(size_t) e * kMaxPv + __popcll(ballot & __lanemask_lt())] = v;
    }
    if (lane == 0) a.dev_number_of_multi_fit_vertices[e] = (uint32_t) __popcll(ballot);
    __syncthreads();
  });
}

} // namespace velo

using namespace velo;

extern "C" {

float velo_pv_exp(float x) { return pv_exp(x); }

int velo_pv_beamline_extrapolate_launch(const velo_pv_beamline_extrapolate_args* a, void* stream)
{
  if (!a || !a->dev_velo_kalman_beamline_states || !a->dev_atomics_velo || !a->dev_pvtracks)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  const uint32_t block = a->tuning.block_dim ? a->tuning.block_dim : 128;
  if (block % kWave || block > 1024) return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(pv_extrapolate_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(block), 0,
                     (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

int velo_pv_beamline_histo_launch(const velo_pv_beamline_histo_args* a, void* stream)
{
  if (!a || !a->dev_atomics_velo || !a->dev_pvtracks || !a->dev_zhisto) return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  if (a->tuning.block_dim && a->tuning.block_dim != kWave) return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(pv_histo_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(kWave), 0,
                     (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

int velo_pv_beamline_peak_launch(const velo_pv_beamline_peak_args* a, void* stream)
{
  if (!a || !a->dev_zhisto || !a->dev_zpeaks || !a->dev_number_of_zpeaks) return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  if (a->tuning.block_dim && a->tuning.block_dim != kWave) return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(pv_peak_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(kWave), 0,
                     (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

int velo_pv_beamline_multi_fitter_launch(const velo_pv_beamline_multi_fitter_args* a, void* stream)
{
  if (!a || !a->dev_atomics_velo || !a->dev_pvtracks || !a->dev_zpeaks || !a->dev_number_of_zpeaks ||
      !a->dev_multi_fit_vertices || !a->dev_number_of_multi_fit_vertices)
    return (int) hipErrorInvalidValue;
  if (a->n_events == 0) return 0;
  if (!schedule_ok(a->tuning)) return (int) hipErrorInvalidValue;
  if (a->tuning.block_dim && a->tuning.block_dim != kWave) return (int) hipErrorInvalidValue;
  hipLaunchKernelGGL(pv_multi_fitter_kernel, dim3(event_grid(a->tuning, a->n_events)), dim3(kWave), 0,
                     (hipStream_t) stream, *a);
  return (int) hipGetLastError();
}

} // extern "C"
