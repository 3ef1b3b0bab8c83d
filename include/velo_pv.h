/* velo_pv.h -- beamline primary-vertex finding on the VELO tracks.
 *
 * The reference's four pv_beamline algorithms, run after velo_kalman_fit on
 * its beamline states:
 *   pv_beamline_extrapolate   pv_beamline_extrapolate.cu:3-30
 *   pv_beamline_histo         pv_beamline_histo.cu:3-76
 *   pv_beamline_peak          pv_beamline_peak.cu:3-147
 *   pv_beamline_multi_fitter  pv_beamline_multi_fitter.cu:5-183
 * with the constants of BeamlinePVConstants.cuh:3-25 and PV_Definitions.cuh
 * and the layouts PVTrack (TrackBeamLineVertexFinder.cuh:10-31) and
 * PV::Vertex (PV_Definitions.cuh:76-108).
 *
 * Each launcher takes the reference kernel's device pointers in its order,
 * the event count and a velo_tuning block (velo_hlt.h); it validates its
 * pointers, never allocates or synchronises, and returns a hipError_t.
 * Float results equal the oracle's bit for bit: sums run in a fixed order
 * (the reference's float atomics have none) and the exponential is one
 * function shared in definition with the oracle (velo_pv_exp). */
#ifndef VELO_PV_H
#define VELO_PV_H

#include <stddef.h>
#include <stdint.h>

#include "velo_hlt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VELO_PV_NBINS 2400u       /* (zmax - zmin) / dz */
#define VELO_PV_ZMIN (-300.0f)    /* mm */
#define VELO_PV_ZMAX 300.0f
#define VELO_PV_DZ 0.25f
#define VELO_PV_MAX_VERTICES 32u  /* peaks and vertices per event */

/* PVTrack: the track at its closest approach to the beamline. 28 B. */
typedef struct {
  float z;
  float x, y;
  float tx, ty;
  float W_00, W_11; /* inverse x / y variances there */
} velo_pv_track;

/* PV::Vertex. 52 B.  n_tracks is the event's VELO track count, as in the
 * reference (its count runs over fresh copies whose weight is always 1);
 * ndof and nTracks are 0 (the reference leaves ndof unset). */
typedef struct {
  float x, y, z;
  float chi2;
  int32_t ndof;
  uint32_t n_tracks;
  float cov00, cov10, cov11, cov20, cov21, cov22;
  int32_t nTracks;
} velo_pv_vertex;

/* ---- pv_beamline_extrapolate: every track of every event to its point of
 * closest approach to the beamline (0, 0) from its velo_kalman_fit state. */
typedef struct {
  const char* dev_velo_kalman_beamline_states;
  const int32_t* dev_atomics_velo;
  const uint32_t* dev_velo_track_hit_number;
  velo_pv_track* dev_pvtracks;          /* total tracks */
  uint32_t n_events;
  velo_tuning tuning;
} velo_pv_beamline_extrapolate_args;
int velo_pv_beamline_extrapolate_launch(const velo_pv_beamline_extrapolate_args* a, void* stream);

/* ---- pv_beamline_histo: per event, the z histogram (2400 bins of 0.25 mm
 * over -300..300 mm) of the tracks' z0 resolution functions.  Tracks are
 * added in track order, each bin's share once. */
typedef struct {
  const int32_t* dev_atomics_velo;
  const uint32_t* dev_velo_track_hit_number;
  const velo_pv_track* dev_pvtracks;
  float* dev_zhisto;                    /* N * 2400 */
  uint32_t n_events;
  velo_tuning tuning;
} velo_pv_beamline_histo_args;
int velo_pv_beamline_histo_launch(const velo_pv_beamline_histo_args* a, void* stream);

/* ---- pv_beamline_peak: per event, the z seeds (cluster peaks) of the
 * histogram, at most 32 (the reference writes past its 32 slots when an
 * event has more; here the first 32 are kept). */
typedef struct {
  const float* dev_zhisto;
  float* dev_zpeaks;                    /* N * 32 */
  uint32_t* dev_number_of_zpeaks;       /* N */
  uint32_t n_events;
  velo_tuning tuning;
} velo_pv_beamline_peak_args;
int velo_pv_beamline_peak_launch(const velo_pv_beamline_peak_args* a, void* stream);

/* ---- pv_beamline_multi_fitter: per seed, two iterations of the adaptive
 * multi-vertex fit; vertices with at least 5 event tracks and within
 * 0.3 mm of the beamline are kept, in seed order (the reference appends
 * them in atomic-arrival order). */
typedef struct {
  const int32_t* dev_atomics_velo;
  const uint32_t* dev_velo_track_hit_number;
  const velo_pv_track* dev_pvtracks;
  const float* dev_zpeaks;
  const uint32_t* dev_number_of_zpeaks;
  velo_pv_vertex* dev_multi_fit_vertices;       /* N * 32 */
  uint32_t* dev_number_of_multi_fit_vertices;   /* N */
  uint32_t n_events;
  velo_tuning tuning;
} velo_pv_beamline_multi_fitter_args;
int velo_pv_beamline_multi_fitter_launch(const velo_pv_beamline_multi_fitter_args* a, void* stream);

/* The exponential the fit uses, on the host (the kernels evaluate the same
 * operations): Cody-Waite reduction x = k ln2 + r, |r| <= ln2 / 2, a
 * degree-7 Taylor polynomial of e^r in Horner form, and ldexp by k. */
float velo_pv_exp(float x);

#ifdef __cplusplus
}
#endif

#endif
