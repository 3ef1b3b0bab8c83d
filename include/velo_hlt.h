/*
 * velo_hlt.h -- C ABI of the MI355X (gfx950) VELO reconstruction kernels.
 *
 * One launcher per algorithm of the reference's Velo sequence
 * (configuration/sequences/Velo.h:5-18).  Each launcher takes a POD struct
 * holding the same device pointers and scalars as the reference kernel's
 * signature, plus the event count and a tuning block, and an opaque stream
 * (a hipStream_t, NULL = default stream).  Launchers are asynchronous: they
 * never allocate device memory and never synchronise.  They return a
 * hipError_t value (0 = success); argument errors return
 * hipErrorInvalidValue (1).
 *
 * Output order.  Where the reference allocates slots with atomics (clusters,
 * tracks, weak tracks) these kernels emit a fixed, deterministic order (see
 * each launcher); any slot order is a valid output of the reference.
 *
 * Per-event status bits (dev_velo_event_status, may be NULL) report the
 * capacity overflows the reference only asserts in debug builds
 * (VeloDefinitions.cuh:18-69, ClusteringDefinitions.cuh:12).
 */
#ifndef VELO_HLT_H
#define VELO_HLT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- capacities and layout constants (VeloDefinitions.cuh:7-77) ---- */
#define VELO_N_MODULES 52u
#define VELO_N_SENSORS 208u
#define VELO_MAX_CANDIDATES_EVENT 3000u /* ClusteringDefinitions.cuh:12 */
#define VELO_MAX_NUMHITS_IN_MODULE 500u
#define VELO_MAX_HITS_PER_EVENT 9500u
#define VELO_MAX_TRACKS 1200u
#define VELO_MAX_TRACK_SIZE 26u
#define VELO_TTF_MODULO 2000u
#define VELO_MAX_WEAK_TRACKS 500u
#define VELO_NUM_ATOMICS 5u
#define VELO_GEOMETRY_BYTES 19456u
#define VELO_TRACK_HITS_BYTES 54u /* sizeof(Velo::TrackHits), VeloEventModel.cuh:58 */
#define VELO_TRACKLET_BYTES 6u    /* sizeof(Velo::TrackletHits) */
#define VELO_STATE_BYTES 49u      /* 12 floats + bool per track, VeloConsolidated.cuh:68-101 */
#define VELO_HIT_BYTES 16u        /* sizeof(Velo::Hit) */
/* Tracks-to-follow entries are 8 bytes here (last two hits carried along),
 * so dev_tracks_to_follow holds 2 u32 per entry. */
#define VELO_TTF_ENTRY_WORDS 2u

/* ---- per-event status bits ---- */
#define VELO_STATUS_SP_INDEX_TRUNCATED 0x001u /* bank with >256 SPs; clustering reads SP index mod 256 as the reference does */
#define VELO_STATUS_CANDIDATES_OVERFLOW 0x002u /* >3000 candidates; list truncated */
#define VELO_STATUS_MODULE_HITS_OVERFLOW 0x004u /* >500 clusters in a module */
#define VELO_STATUS_EVENT_HITS_OVERFLOW 0x008u /* >=9500 clusters in the event */
#define VELO_STATUS_TTF_OVERFLOW 0x010u        /* >2000 live tracks to follow; extra entries dropped */
#define VELO_STATUS_TRACKLETS_OVERFLOW 0x020u  /* >2000 tracklets within two steps */
#define VELO_STATUS_WEAK_OVERFLOW 0x040u       /* >500 weak tracks; extra dropped */
#define VELO_STATUS_TRACKS_OVERFLOW 0x080u     /* >1200 tracks; extra dropped */
#define VELO_STATUS_BAD_BANK 0x100u            /* bank index != sensor index, the bank reaches past its event, or it holds 2^22 SPs or more; bank skipped */
#define VELO_STATUS_UNSORTED_BANK 0x200u       /* a bank's SP addresses do not strictly increase (set by estimate_input_size) */
/* Not an error: estimate_input_size sets this bit on every event it
 * processes, after checking the order of the event's banks.  Masked
 * clustering skips its own bank-order check only in events whose status
 * word carries this bit and not UNSORTED_BANK; a status word estimate did
 * not write (zeroed, or a caller's own) makes clustering check every bank. */
#define VELO_STATUS_BANK_ORDER_CHECKED 0x80000000u
/* The bits above that report a problem with the event. */
#define VELO_STATUS_ERROR_MASK 0x3FFu

/* ---- tuning block ----
 * Launch shape of one launcher call.  With grid_dim = 0 a per-event kernel
 * runs one workgroup per event (workgroup e handles event e).  With
 * grid_dim > 0 the launch is persistent: grid_dim workgroups take event
 * positions k = 0, 1, ... from dev_work_counter and handle event
 * dev_work_order[k] (k itself when dev_work_order is NULL), so a
 * largest-first order balances events of very uneven size.
 * dev_work_counter points to two u32 that must be zero before the first
 * launch; the last workgroup to finish sets them back to zero, so the same
 * pair serves the next launch (also inside a captured graph) as long as
 * launches that share a pair do not overlap.  Outputs do not depend on
 * the launch shape.  Scans (velo_prefix_sum_launch) use dev_count: when
 * non-NULL the element count is read there on the device and array_size
 * is only the capacity. */
typedef struct {
  uint32_t block_dim; /* threads per workgroup, multiple of 64; 0 = default */
  uint32_t lds_tile;  /* LDS staging capacity in elements (event words, hits or phis, per kernel); 0 = default */
  uint32_t grid_dim;  /* 0: one workgroup per event; >0: persistent workgroups */
  uint32_t reserved;  /* must be 0 */
  uint32_t* dev_work_counter;     /* persistent launches: 2 u32, zero at the first launch */
  const uint32_t* dev_work_order; /* persistent launches: position -> event, or NULL */
  const uint32_t* dev_count;      /* scans: element count on the device, or NULL */
} velo_tuning;

/* ---- device constants (Constants.cu:22-53, :239-250) ---- */
typedef struct {
  uint8_t* dev_velo_sp_patterns;  /* 256 */
  float* dev_velo_sp_fx;          /* 512 */
  float* dev_velo_sp_fy;          /* 512 */
  uint8_t* dev_velo_candidate_ks; /* 9 */
  float* dev_velo_module_zs;      /* 52 */
  char* dev_velo_geometry;        /* VELO_GEOMETRY_BYTES */
} velo_constants;

/* Replaces Constants::initialize_constants (Constants.cu:123-154) for the
 * VELO part and initialize_geometry_constants (Constants.cu:239-250):
 * allocates and fills the device tables.  module_zs may be NULL (the
 * reference's 52 z positions are used). */
int velo_constants_init(const void* geometry, size_t geometry_bytes, const float* module_zs, velo_constants* out);
int velo_constants_free(velo_constants* c);
/* Host copies of the SP pattern tables the device holds. */
void velo_sp_tables(uint8_t* patterns256, uint8_t* sizes256, float* fx512, float* fy512);

/* ---- K0: init_event_list (InitEventList.cu:7-11) ---- */
typedef struct {
  uint32_t* dev_event_list;
  uint32_t n_events;
  velo_tuning tuning;
} velo_init_event_list_args;
int velo_init_event_list_launch(const velo_init_event_list_args* a, void* stream);

/* ---- event_order: no reference counterpart; fills the reference's
 * dev_event_order argument (ArgumentsCommon.cuh:16, declared but unused
 * there).  Position k of dev_event_order receives a selected-event index,
 * largest raw event first (in buckets of 64 bytes; order within a bucket
 * unspecified); persistent launches take events in this order.  Outputs do
 * not depend on it.  dev_event_list NULL = identity. */
typedef struct {
  const uint32_t* dev_raw_input_offsets;
  const uint32_t* dev_event_list;
  uint32_t* dev_event_order; /* N */
  uint32_t n_events;
  velo_tuning tuning;
} velo_event_order_args;
int velo_event_order_launch(const velo_event_order_args* a, void* stream);

/* ---- global_event_cut (GlobalEventCut.cu:3-55, GlobalEventCut.cuh:9) ----
 * Keeps an event when its SciFi cluster estimate (bank bytes without the
 * 4-byte source word, halved, minus 2, in 32-bit unsigned arithmetic) and
 * that estimate plus its UT hit count (low 16 bits of each UT bank's first
 * data word) both stay below VELO_GEC_MAX_SCIFI_UT_CLUSTERS.  UT and SciFi
 * events use the raw-event layout (u32 n_banks, u32 byte offsets[n_banks+1],
 * per bank the source id and its data words).  dev_event_list receives the
 * kept events in increasing order (the reference appends them in
 * atomic-arrival order; the set and the count are the same) and
 * dev_number_of_selected_events their count.  An event whose banks reach
 * past ut_events_size / scifi_events_size is not kept (the reference has no
 * sizes and reads on).  dev_event_list holds n_events words; it is used as
 * scratch for per-event flags before the compaction. */
#define VELO_GEC_MAX_SCIFI_UT_CLUSTERS 9750u
typedef struct {
  const char* dev_ut_raw_input;
  const uint32_t* dev_ut_raw_input_offsets;
  const char* dev_scifi_raw_input;
  const uint32_t* dev_scifi_raw_input_offsets;
  uint32_t* dev_number_of_selected_events;
  uint32_t* dev_event_list;
  size_t ut_events_size;
  size_t scifi_events_size;
  uint32_t n_events;
  velo_tuning tuning;
} velo_global_event_cut_args;
int velo_global_event_cut_launch(const velo_global_event_cut_args* a, void* stream);

/* ---- K1: estimate_input_size (EstimateInputSize.cu:3-255) ----
 * Per module: clusters of isolated SPs plus candidates of the others.
 * Candidates are listed per event in bank, SP, low-then-high-nibble order. */
typedef struct {
  const char* dev_raw_input;
  const uint32_t* dev_raw_input_offsets;
  uint32_t* dev_estimated_input_size; /* N*52 (+1 for the scan total) */
  uint32_t* dev_module_cluster_num;   /* N*52, zeroed */
  uint32_t* dev_event_candidate_num;  /* N */
  uint32_t* dev_cluster_candidates;   /* N*3000 */
  const uint32_t* dev_event_list;     /* N, NULL = identity */
  uint32_t* dev_event_order;          /* unused, kept for signature parity */
  const uint8_t* dev_velo_candidate_ks;
  uint32_t* dev_velo_event_status;    /* N or NULL */
  uint32_t n_events;
  velo_tuning tuning;
} velo_estimate_input_size_args;
int velo_estimate_input_size_launch(const velo_estimate_input_size_args* a, void* stream);

/* ---- K2-K4 / K12-K14: PREFIX_SUM_ALGORITHM (PrefixSumHandler.cuh:15-58) ----
 * Exclusive scan of dev_array[0, array_size) in place; the total is written
 * to dev_array[array_size].  dev_auxiliary_array holds
 * velo_prefix_sum_aux_size(array_size) words. */
typedef struct {
  uint32_t* dev_array;
  uint32_t* dev_auxiliary_array;
  uint32_t array_size;
  velo_tuning tuning;
} velo_prefix_sum_args;
uint32_t velo_prefix_sum_aux_size(uint32_t array_size);
int velo_prefix_sum_launch(const velo_prefix_sum_args* a, void* stream);

/* Allocation-free variant of velo_prefix_sum_launch.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-tile sums live in a device array inside the library
 * (256 KiB on each GPU it runs on, kept until the process ends) instead of
 * a caller's dev_auxiliary_array; nothing is allocated and nothing needs
 * to be zeroed.  Same scan as velo_prefix_sum_launch: dev_array[0,
 * array_size) scanned in place, the total written to dev_array[array_size];
 * dev_count (may be NULL) gives the element count on the device, array_size
 * being then the capacity.  array_size above
 * VELO_PREFIX_SUM_NOALLOC_MAX_ELEMENTS returns hipErrorInvalidValue.  Every
 * call on one GPU uses that GPU's one array: calls that overlap on one GPU
 * (on different streams, or from different threads without ordering) race
 * and give wrong sums; order them on one stream. */
#define VELO_PREFIX_SUM_NOALLOC_MAX_ELEMENTS 134217728u /* 65536 tiles of 2048 */
int velo_prefix_sum_noalloc_launch(uint32_t* dev_array, uint32_t array_size, const uint32_t* dev_count, void* stream);

/* ---- K5: masked_velo_clustering (MaskedVeloClustering.cu:48-464) ----
 * Clusters of module m are written to [start_m, start_m + num_m) in bank,
 * SP, cluster order (isolated SP: LUT cluster 0 then 1; otherwise the
 * low-nibble candidate then the high-nibble one).  Candidates are read from
 * dev_cluster_candidates / dev_event_candidate_num (estimate_input_size's
 * list, as the reference does) when both are given and the event's list is
 * below VELO_MAX_CANDIDATES_EVENT and fits in LDS beside the event; else
 * (NULL, a full list, a large event) they are derived from the raw banks,
 * with the same result. */
typedef struct {
  const char* dev_raw_input;
  const uint32_t* dev_raw_input_offsets;
  const uint32_t* dev_module_cluster_start; /* scanned estimates, N*52+1 */
  uint32_t* dev_module_cluster_num;         /* N*52 */
  const uint32_t* dev_event_candidate_num;
  const uint32_t* dev_cluster_candidates;
  uint32_t* dev_velo_cluster_container;     /* 6 * total */
  const uint32_t* dev_event_list;           /* NULL = identity */
  uint32_t* dev_event_order;
  const char* dev_velo_geometry;
  const uint8_t* dev_velo_sp_patterns;
  const float* dev_velo_sp_fx;
  const float* dev_velo_sp_fy;
  uint32_t* dev_velo_event_status;          /* N or NULL; as estimate_input_size left it: bank order is
                                               checked only in events with VELO_STATUS_UNSORTED_BANK
                                               (NULL: in every event) */
  uint32_t n_events;
  velo_tuning tuning;
} velo_masked_clustering_args;
int velo_masked_clustering_launch(const velo_masked_clustering_args* a, void* stream);

/* ---- K6: calculate_phi_and_sort (CalculatePhiAndSort.cu:6-52) ----
 * Slots after the sort: Y=0, Z=1, ID=2, unsorted ID=3, phi=4, X=5.  Only
 * [start, start+num) of each module holds hits; its holes (estimated but
 * unfilled slots) read as NaN in slots 0, 1, 2, 4 and 5 and are left out of
 * dev_hit_permutation.  tuning.lds_tile is not used. */
typedef struct {
  const uint32_t* dev_module_cluster_start;
  const uint32_t* dev_module_cluster_num;
  uint32_t* dev_velo_cluster_container;
  uint32_t* dev_hit_permutation;
  uint32_t n_events;
  velo_tuning tuning;
} velo_calculate_phi_and_sort_args;
int velo_calculate_phi_and_sort_launch(const velo_calculate_phi_and_sort_args* a, void* stream);

/* ---- K7: fill_candidates (FillCandidates.cu:8-97) ---- */
typedef struct {
  const uint32_t* dev_velo_cluster_container;
  const uint32_t* dev_module_cluster_start;
  const uint32_t* dev_module_cluster_num;
  int16_t* dev_h0_candidates; /* 2 * total */
  int16_t* dev_h2_candidates; /* 2 * total */
  uint32_t n_events;
  velo_tuning tuning;
} velo_fill_candidates_args;
int velo_fill_candidates_launch(const velo_fill_candidates_args* a, void* stream);

/* ---- K8: search_by_triplet (SearchByTriplet.cu:7-111) ----
 * Tracks are numbered in forwarding order; seeding takes h1 candidates of
 * module_data[2] then [3] in index order and, for equal scatter, the
 * lexicographically smallest (h0, h2).  dev_tracks_to_follow holds
 * N * 2000 * VELO_TTF_ENTRY_WORDS words (a step's first 128 entries stay
 * in LDS and are not written there); dev_tracklets and dev_rel_indices are
 * not used (tracklet hits travel inside the tracks-to-follow entries).
 * tuning.lds_tile: hits of the LDS window, rounded down to a power of two
 * (0 = 512). */
typedef struct {
  const uint32_t* dev_velo_cluster_container;
  const uint32_t* dev_module_cluster_start;
  const uint32_t* dev_module_cluster_num;
  void* dev_tracks;             /* N*1200 TrackHits (54 B) */
  void* dev_tracklets;          /* unused */
  uint32_t* dev_tracks_to_follow;
  void* dev_weak_tracks;        /* N*500 TrackletHits (6 B) */
  uint8_t* dev_hit_used;        /* total */
  int32_t* dev_atomics_velo;    /* N*5 */
  const int16_t* dev_h0_candidates;
  const int16_t* dev_h2_candidates;
  uint16_t* dev_rel_indices;    /* unused */
  const float* dev_velo_module_zs;
  uint32_t* dev_velo_event_status;
  uint32_t n_events;
  velo_tuning tuning;
} velo_search_by_triplet_args;
int velo_search_by_triplet_launch(const velo_search_by_triplet_args* a, void* stream);
/* Workgroups of search_by_triplet that fit on the current device at once
 * with an LDS window of lds_tile hits (0 = the default): the persistent
 * grid the sequence uses when no search shape is given.  0 on error. */
uint32_t velo_search_by_triplet_resident_workgroups(uint32_t lds_tile);

/* ---- K9: weak_tracks_adder (WeakTracksAdder.cu:117-157) ----
 * Accepted weak tracks are appended in weak-track order. */
typedef struct {
  const uint32_t* dev_velo_cluster_container;
  const uint32_t* dev_module_cluster_start;
  void* dev_tracks;
  const void* dev_weak_tracks;
  const uint8_t* dev_hit_used;
  int32_t* dev_atomics_velo;
  uint32_t* dev_velo_event_status;
  uint32_t n_events;
  velo_tuning tuning;
} velo_weak_tracks_adder_args;
int velo_weak_tracks_adder_launch(const velo_weak_tracks_adder_args* a, void* stream);

/* ---- K10: copy_and_prefix_sum_single_block (PrefixSumSingleBlock.cu:130-151) ----
 * dev_output_array[i] = dev_input_array[i] scanned exclusively, total at
 * *dev_total_sum. */
typedef struct {
  uint32_t* dev_total_sum;
  const uint32_t* dev_input_array;
  uint32_t* dev_output_array;
  uint32_t array_size;
  velo_tuning tuning;
} velo_copy_and_prefix_sum_single_block_args;
int velo_copy_and_prefix_sum_single_block_launch(const velo_copy_and_prefix_sum_single_block_args* a, void* stream);

/* ---- K11: copy_velo_track_hit_number (PrefixSumSingleBlock.cu:153-171) ---- */
typedef struct {
  const void* dev_tracks;
  const int32_t* dev_atomics_storage;
  uint32_t* dev_velo_track_hit_number;
  uint32_t n_events;
  velo_tuning tuning;
} velo_copy_velo_track_hit_number_args;
int velo_copy_velo_track_hit_number_launch(const velo_copy_velo_track_hit_number_args* a, void* stream);

/* ---- velo_kalman_fit (VeloKalmanFilter.cu:43-73, VeloKalmanFilter.cuh:22-91)
 * Simplified Kalman fit of every consolidated track from the hit farthest
 * from the beamline towards it (scattering noise 1e-8 + 7e-6 (tx^2 + ty^2)
 * per layer); the states go to dev_velo_kalman_beamline_states in the
 * consolidated state layout (12 float arrays of total_tracks, then one
 * byte per track).  The reference leaves the backward byte of these states
 * uninitialised; here it is the input state's. */
typedef struct {
  const int32_t* dev_atomics_velo;
  const uint32_t* dev_velo_track_hit_number;
  const char* dev_velo_track_hits;
  const char* dev_velo_states;
  char* dev_velo_kalman_beamline_states;
  uint32_t n_events;
  velo_tuning tuning;
} velo_kalman_fit_args;
int velo_kalman_fit_launch(const velo_kalman_fit_args* a, void* stream);

/* ---- K15: consolidate_velo_tracks (ConsolidateVelo.cu:116-170) ----
 * Also sets dev_accepted_velo_tracks[0, total_tracks) to 1 (the memset of
 * ConsolidateVeloVisitor.cu:42-43) when non-NULL.  With dev_ghost_model
 * (a velo_ghost_model, velo_ghost.h; no reference counterpart) the track
 * is scored inline and accepted only when p(ghost) < the model's
 * threshold; dev_ghost_score, when non-NULL, receives p(ghost) per track.
 * Hits and states are the same with or without the model. */
typedef struct {
  const int32_t* dev_atomics_velo;
  const void* dev_tracks;
  const uint32_t* dev_velo_track_hit_number;
  const uint32_t* dev_velo_cluster_container;
  const uint32_t* dev_module_cluster_start;
  const uint32_t* dev_module_cluster_num;
  char* dev_velo_track_hits;
  char* dev_velo_states;
  uint8_t* dev_accepted_velo_tracks;
  const void* dev_ghost_model;
  float* dev_ghost_score;
  uint32_t n_events;
  velo_tuning tuning;
} velo_consolidate_velo_tracks_args;
int velo_consolidate_velo_tracks_launch(const velo_consolidate_velo_tracks_args* a, void* stream);

/* ---- library information ---- */
const char* velo_hlt_version(void);
const char* velo_error_string(int code);
/* sizeof of a struct declared in this library's headers (velo_hlt.h,
 * velo_pv.h, velo_ghost.h, velo_checker.h, velo_mdf.h, velo_sequence.h) as
 * the library was built, by its typedef name; 0 for an unknown name.  A
 * binding compares its own mirror of every struct the library reads or
 * writes through a pointer and refuses a library that differs, so that a
 * library built from a newer header never writes past a caller's struct.
 * No reference counterpart. */
size_t velo_abi_struct_size(const char* name);

#ifdef __cplusplus
}
#endif
#endif
