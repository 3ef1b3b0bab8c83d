/*
 * velo_oracle.h -- CPU oracle for the VELO reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded C restatement of the
 * reference's VELO path (raw SuperPixel decoding, masked clustering, phi sort,
 * fill_candidates, search_by_triplet, weak tracks, prefix sums, consolidation).
 * It is the checker the HIP kernels are compared against; nothing in the
 * product links, loads or calls it.  Only tests/, the smoke check and the
 * bench's CPU-baseline leg use it.
 *
 * Buffer layouts are the reference's (so one input feeds both the oracle and
 * the GPU launcher):
 *   raw events      ClusteringDefinitions.cu:3-21, cuda/velo/readme.md:9-33
 *   module counters EstimateInputSize.cu:20-23 (N*52+1 estimates, N*52 nums)
 *   container       MaskedVeloClustering.cu:75-79 (SoA x,y,z,id,phi,temp)
 *   tracks          VeloEventModel.cuh:43-69 (TrackHits 54 B, TrackletHits 6 B)
 *   consolidated    VeloConsolidated.cuh:12-144, ConsolidatedTypes.cuh:11-67
 *
 * Where the reference allocates output slots with atomics (clusters, tracks,
 * weak tracks), the oracle emits them in a fixed "natural" order; any order is
 * a valid reference output, and the kernels reproduce this one exactly.
 */
#ifndef VELO_ORACLE_H
#define VELO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* VeloDefinitions.cuh:12-77, ClusteringDefinitions.cuh:12 */
#define OR_N_MODULES 52u
#define OR_N_SENSORS 208u
#define OR_MAX_CANDIDATES_EVENT 3000u
#define OR_MAX_NUMHITS_IN_MODULE 500u
#define OR_MAX_HITS_PER_EVENT 9500u
#define OR_MAX_TRACKS 1200u
#define OR_MAX_TRACK_SIZE 26u
#define OR_TTF_MODULO 2000u
#define OR_MAX_WEAK_TRACKS 500u
#define OR_NUM_ATOMICS 5u
#define OR_GEOMETRY_BYTES 19456u

/* Per-event status bits (no reference counterpart: the reference only
 * asserts these capacities in debug builds, VeloDefinitions.cuh:18-69). */
#define OR_STATUS_SP_INDEX_TRUNCATED 0x001u /* a bank holds >256 SPs (Q2)      */
#define OR_STATUS_CANDIDATES_OVERFLOW 0x002u /* >3000 cluster candidates        */
#define OR_STATUS_MODULE_HITS_OVERFLOW 0x004u /* >500 clusters in one module     */
#define OR_STATUS_EVENT_HITS_OVERFLOW 0x008u /* >=9500 clusters in the event    */
#define OR_STATUS_TTF_OVERFLOW 0x010u        /* >2000 live tracks-to-follow     */
#define OR_STATUS_TRACKLETS_OVERFLOW 0x020u  /* >2000 tracklets in 2 steps      */
#define OR_STATUS_WEAK_OVERFLOW 0x040u       /* >500 weak tracks                */
#define OR_STATUS_TRACKS_OVERFLOW 0x080u     /* >1200 tracks                    */
#define OR_STATUS_BAD_BANK 0x100u            /* bank index != sensor index      */
#define OR_STATUS_UNSORTED_BANK 0x200u       /* SP addresses not increasing     */
#define OR_STATUS_BANK_ORDER_CHECKED 0x80000000u /* estimate ran on the event (not an error) */
#define OR_STATUS_ERROR_MASK 0x3FFu

#pragma pack(push, 1)
typedef struct {
  uint16_t hitsNum;
  uint16_t hits[OR_MAX_TRACK_SIZE];
} or_track_hits; /* 54 B, VeloEventModel.cuh:58-69 */
#pragma pack(pop)

typedef struct {
  uint16_t hits[3];
} or_tracklet_hits; /* 6 B, VeloEventModel.cuh:43-55 */

/* ---- math shared with the kernels only through identical restatement ---- */
float oracle_atan2f(float y, float x);
float oracle_phi_odd(float x, float y);
float oracle_phi_even(float x, float y);

/* ---- a3: SP pattern tables (ClusteringCommon.cpp:3-101) ---- */
void oracle_sp_lut(uint8_t* patterns256, uint8_t* sizes256, float* fx512, float* fy512);

/* ---- a4: estimate_input_size (EstimateInputSize.cu:3-255) ---- */
void oracle_estimate_input_size(
  const char* events, const uint32_t* offsets, uint32_t n_events,
  uint32_t* estimated_input_size, /* N*52 (+1 left untouched) */
  uint32_t* module_cluster_num,   /* N*52, zeroed */
  uint32_t* event_candidate_num,  /* N */
  uint32_t* cluster_candidates,   /* N*3000 */
  uint32_t* status);              /* N, OR-ed */

/* ---- a5/a13/a15: exclusive scan in place, total written at [n] ---- */
void oracle_prefix_sum(uint32_t* array, uint32_t n);

/* ---- a6: masked_velo_clustering (MaskedVeloClustering.cu:48-464) ---- */
void oracle_masked_clustering(
  const char* events, const uint32_t* offsets, uint32_t n_events,
  const uint32_t* module_cluster_start, /* N*52+1 after the scan */
  uint32_t* module_cluster_num,         /* N*52 */
  const char* geometry,
  uint32_t* cluster_container,          /* 6*total */
  uint32_t* status);

/* ---- a7: calculate_phi_and_sort (CalculatePhiAndSort.cu:6-52) ---- */
void oracle_calculate_phi_and_sort(
  uint32_t n_events, const uint32_t* module_cluster_start, const uint32_t* module_cluster_num,
  uint32_t* cluster_container, uint32_t* hit_permutation, uint32_t* status);

/* ---- a8: fill_candidates (FillCandidates.cu:8-97) ---- */
void oracle_fill_candidates(
  uint32_t n_events, const uint32_t* cluster_container, const uint32_t* module_cluster_start,
  const uint32_t* module_cluster_num, int16_t* h0_candidates, int16_t* h2_candidates);

/* ---- a9..a11: search_by_triplet (SearchByTriplet.cu, ProcessModules.cu,
 *      TrackSeeding.cu, TrackForwarding.cu) ---- */
void oracle_search_by_triplet(
  uint32_t n_events, const uint32_t* cluster_container, const uint32_t* module_cluster_start,
  const uint32_t* module_cluster_num, or_track_hits* tracks, or_tracklet_hits* weak_tracks,
  uint8_t* hit_used, int32_t* atomics_velo, const int16_t* h0_candidates,
  const int16_t* h2_candidates, const float* module_zs, uint32_t* status);

/* ---- a12: weak_tracks_adder (WeakTracksAdder.cu:8-157) ---- */
void oracle_weak_tracks_adder(
  uint32_t n_events, const uint32_t* cluster_container, const uint32_t* module_cluster_start,
  or_track_hits* tracks, const or_tracklet_hits* weak_tracks, const uint8_t* hit_used,
  int32_t* atomics_velo, uint32_t* status);

/* ---- a13: copy_and_prefix_sum_single_block (PrefixSumSingleBlock.cu:130-151) ---- */
void oracle_copy_and_prefix_sum(uint32_t n_events, uint32_t* atomics_velo);

/* ---- a14: copy_velo_track_hit_number (PrefixSumSingleBlock.cu:153-171) ---- */
void oracle_copy_velo_track_hit_number(
  uint32_t n_events, const or_track_hits* tracks, const uint32_t* atomics_velo,
  uint32_t* velo_track_hit_number);

/* ---- a16: consolidate_velo_tracks (ConsolidateVelo.cu:6-170) ---- */
void oracle_consolidate_velo_tracks(
  uint32_t n_events, const uint32_t* atomics_velo, const or_track_hits* tracks,
  const uint32_t* velo_track_hit_number, const uint32_t* cluster_container,
  const uint32_t* module_cluster_start, char* velo_track_hits, char* velo_states);

/* ---- velo_kalman_fit (VeloKalmanFilter.cu:43-73, VeloKalmanFilter.cuh:22-91):
 *      simplified Kalman fit of every consolidated track from its first
 *      hit towards the beamline; output in the consolidated state layout.
 *      The reference leaves the output's backward flag uninitialised; here
 *      it is the input state's flag. ---- */
void oracle_velo_kalman_fit(
  uint32_t n_events, const uint32_t* atomics_velo, const uint32_t* velo_track_hit_number,
  const char* velo_track_hits, const char* velo_states, char* kalman_states);

/* One filter step of one projection (VeloKalmanFilter.cu:6-41); returns
 * the chi2 contribution.  Exposed for the closed-form tests. */
float oracle_kalman_step(float z, float zhit, float xhit, float whit, float* x, float* tx, float* cov_xx,
                         float* cov_xtx, float* cov_txtx);

/* ---- global_event_cut (GlobalEventCut.cu:3-55, GlobalEventCut.cuh:9)
 *      UT and SciFi events in the raw-event layout (u32 n_banks, u32 byte
 *      offsets[n_banks+1], per bank the source id and its data words).
 *      Writes the kept events to event_list in increasing order (the
 *      reference appends them in atomic-arrival order) and returns their
 *      count. ---- */
#define OR_GEC_MAX_SCIFI_UT_CLUSTERS 9750u
uint32_t oracle_global_event_cut(
  uint32_t n_events, const char* ut_raw, const uint32_t* ut_offsets, const char* scifi_raw,
  const uint32_t* scifi_offsets, uint32_t* event_list);

/* ---- beamline primary vertices (cuda/PV/beamlinePV/src/*.cu,
 *      BeamlinePVConstants.cuh:3-17).  PV tracks: 7 floats per track
 *      (z, x, y, tx, ty, W_00, W_11; TrackBeamLineVertexFinder.cuh:10-31);
 *      histograms: 2400 floats per event; peaks: 32 per event; vertices:
 *      13 words per vertex, 32 per event (PV::Vertex: x, y, z, chi2, ndof,
 *      n_tracks, cov00, cov10, cov11, cov20, cov21, cov22, nTracks).
 *      Histogram shares are added in track order, vertices kept in seed
 *      order (the reference's float and index atomics have no order). ---- */
#define OR_PV_NBINS 2400u
#define OR_PV_MAX_VERTICES 32u
#define OR_PV_VERTEX_WORDS 13u
float oracle_pv_exp(float x);
float oracle_pv_gauss_integral(float x);
void oracle_pv_beamline_extrapolate(
  uint32_t n_events, const uint32_t* atomics_velo, const char* kalman_states, float* pvtracks);
void oracle_pv_beamline_histo(uint32_t n_events, const uint32_t* atomics_velo, const float* pvtracks, float* zhisto);
void oracle_pv_beamline_peak(uint32_t n_events, const float* zhisto, float* zpeaks, uint32_t* n_zpeaks);
void oracle_pv_beamline_multi_fitter(
  uint32_t n_events, const uint32_t* atomics_velo, const float* pvtracks, const float* zpeaks,
  const uint32_t* n_zpeaks, uint32_t* vertices, uint32_t* n_vertices);

/* ---- textbook 8-connected component labelling (Clustering.cpp:3-280
 *      semantics): LHCbIDs of all clusters of one event, plus centroid
 *      sums so coordinates can be formed; used to pin a6 ---- */
uint32_t oracle_ccl_event(
  const char* event, uint32_t* lhcb_ids, uint32_t* modules, int32_t* sum_x, int32_t* sum_y,
  int32_t* npix, uint32_t capacity);

/* ---- whole plain pipeline on host buffers (the CPU baseline) ---- */
typedef struct {
  uint32_t n_events;
  uint32_t total_estimated_hits; /* scan total of the estimates */
  uint32_t total_tracks;
  uint32_t total_track_hits;
  /* pointers owned by the oracle; valid until oracle_pipeline_free */
  uint32_t* module_cluster_start;  /* N*52+1 */
  uint32_t* module_cluster_num;    /* N*52 */
  uint32_t* cluster_container;     /* 6*total_estimated_hits (sorted slots) */
  uint32_t* atomics_velo;          /* 5N (+1) after consolidation scans */
  uint32_t* velo_track_hit_number; /* total_tracks+1 */
  char* velo_track_hits;           /* 16*total_track_hits */
  char* velo_states;               /* 49*total_tracks */
  uint32_t* status;                /* N */
} or_pipeline_result;

int oracle_pipeline(
  const char* events, const uint32_t* offsets, uint32_t n_events, const char* geometry,
  const float* module_zs, or_pipeline_result* out);
void oracle_pipeline_free(or_pipeline_result* r);

#ifdef __cplusplus
}
#endif
#endif
