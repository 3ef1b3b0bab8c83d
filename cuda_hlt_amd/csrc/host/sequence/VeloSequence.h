// The algorithms of the VELO sequence and the configured sequence itself
// (configuration/sequences/Velo.h:5-18; the global event cut runs when the
// caller supplies UT and SciFi banks and asks for it).  Each ALGORITHM binds
// a C-ABI launcher of include/velo_hlt.h with the reference's dependency
// list (plus the status word).
#pragma once

#include <tuple>

#include "../gear/ArgumentManager.h"
#include "../gear/Handler.h"
#include "ArgumentsVelo.h"
#include "velo_pv.h"

ALGORITHM(
  velo_init_event_list_launch,
  init_event_list_t,
  ARGUMENTS(
    dev_velo_raw_input,
    dev_velo_raw_input_offsets,
    dev_ut_raw_input,
    dev_ut_raw_input_offsets,
    dev_scifi_raw_input,
    dev_scifi_raw_input_offsets,
    dev_number_of_selected_events,
    dev_event_list))

ALGORITHM(
  velo_global_event_cut_launch,
  global_event_cut_t,
  ARGUMENTS(
    dev_ut_raw_input,
    dev_ut_raw_input_offsets,
    dev_scifi_raw_input,
    dev_scifi_raw_input_offsets,
    dev_number_of_selected_events,
    dev_event_list))

// no reference counterpart: fills dev_event_order (declared, unused there)
ALGORITHM(
  velo_event_order_launch,
  velo_event_order_t,
  ARGUMENTS(dev_velo_raw_input_offsets, dev_event_list, dev_event_order))

ALGORITHM(
  velo_estimate_input_size_launch,
  velo_estimate_input_size_t,
  ARGUMENTS(
    dev_velo_raw_input,
    dev_velo_raw_input_offsets,
    dev_estimated_input_size,
    dev_module_cluster_num,
    dev_module_candidate_num,
    dev_cluster_candidates,
    dev_event_list,
    dev_event_order,
    dev_velo_event_status))

PREFIX_SUM_ALGORITHM(prefix_sum_velo_clusters_t, ARGUMENTS(dev_estimated_input_size, dev_cluster_offset))

ALGORITHM(
  velo_masked_clustering_launch,
  velo_masked_clustering_t,
  ARGUMENTS(
    dev_velo_raw_input,
    dev_velo_raw_input_offsets,
    dev_estimated_input_size,
    dev_module_cluster_num,
    dev_module_candidate_num,
    dev_cluster_candidates,
    dev_velo_cluster_container,
    dev_event_list,
    dev_event_order,
    dev_velo_event_status))

ALGORITHM(
  velo_calculate_phi_and_sort_launch,
  velo_calculate_phi_and_sort_t,
  ARGUMENTS(dev_estimated_input_size, dev_module_cluster_num, dev_velo_cluster_container, dev_hit_permutation))

ALGORITHM(
  velo_fill_candidates_launch,
  velo_fill_candidates_t,
  ARGUMENTS(
    dev_velo_cluster_container,
    dev_estimated_input_size,
    dev_module_cluster_num,
    dev_h0_candidates,
    dev_h2_candidates))

ALGORITHM(
  velo_search_by_triplet_launch,
  velo_search_by_triplet_t,
  ARGUMENTS(
    dev_velo_cluster_container,
    dev_estimated_input_size,
    dev_module_cluster_num,
    dev_tracks,
    dev_tracklets,
    dev_tracks_to_follow,
    dev_weak_tracks,
    dev_hit_used,
    dev_atomics_velo,
    dev_h0_candidates,
    dev_h2_candidates,
    dev_rel_indices,
    dev_velo_event_status,
    dev_event_order))

ALGORITHM(
  velo_weak_tracks_adder_launch,
  velo_weak_tracks_adder_t,
  ARGUMENTS(
    dev_velo_cluster_container,
    dev_estimated_input_size,
    dev_tracks,
    dev_weak_tracks,
    dev_hit_used,
    dev_atomics_velo,
    dev_velo_event_status))

ALGORITHM(
  velo_copy_and_prefix_sum_single_block_launch,
  copy_and_prefix_sum_single_block_velo_t,
  ARGUMENTS(dev_atomics_velo))

ALGORITHM(
  velo_copy_velo_track_hit_number_launch,
  copy_velo_track_hit_number_t,
  ARGUMENTS(dev_tracks, dev_atomics_velo, dev_velo_track_hit_number))

PREFIX_SUM_ALGORITHM(
  prefix_sum_velo_track_hit_number_t,
  // dev_atomics_velo: batched runs read the track total on the device
  ARGUMENTS(dev_velo_track_hit_number, dev_prefix_sum_auxiliary_array_2, dev_atomics_velo))

ALGORITHM(
  velo_consolidate_velo_tracks_launch,
  consolidate_velo_tracks_t,
  ARGUMENTS(
    dev_atomics_velo,
    dev_tracks,
    dev_velo_track_hit_number,
    dev_velo_cluster_container,
    dev_estimated_input_size,
    dev_module_cluster_num,
    dev_velo_track_hits,
    dev_velo_states,
    dev_accepted_velo_tracks))

// VeloKalmanFilter.cuh:94-102; run when RuntimeOptions::run_velo_kalman_fit
// is set (the reference's Velo sequence ends with consolidation, its
// default sequence continues with this fit)
ALGORITHM(
  velo_kalman_fit_launch,
  velo_kalman_fit_t,
  ARGUMENTS(
    dev_atomics_velo,
    dev_velo_track_hit_number,
    dev_velo_track_hits,
    dev_velo_states,
    dev_velo_kalman_beamline_states))

// pv_beamline_*.cuh (cuda/PV/beamlinePV/include)
ALGORITHM(
  velo_pv_beamline_extrapolate_launch,
  pv_beamline_extrapolate_t,
  ARGUMENTS(dev_velo_kalman_beamline_states, dev_atomics_velo, dev_velo_track_hit_number, dev_pvtracks))

ALGORITHM(
  velo_pv_beamline_histo_launch,
  pv_beamline_histo_t,
  ARGUMENTS(dev_atomics_velo, dev_velo_track_hit_number, dev_pvtracks, dev_zhisto))

ALGORITHM(velo_pv_beamline_peak_launch, pv_beamline_peak_t, ARGUMENTS(dev_zhisto, dev_zpeaks, dev_number_of_zpeaks))

ALGORITHM(
  velo_pv_beamline_multi_fitter_launch,
  pv_beamline_multi_fitter_t,
  ARGUMENTS(
    dev_atomics_velo,
    dev_velo_track_hit_number,
    dev_pvtracks,
    dev_zpeaks,
    dev_number_of_zpeaks,
    dev_multi_fit_vertices,
    dev_number_of_multi_fit_vertices))

#define SEQUENCE_T(...) typedef std::tuple<__VA_ARGS__> configured_sequence_t;

SEQUENCE_T(
  init_event_list_t,
  global_event_cut_t,
  velo_event_order_t,
  velo_estimate_input_size_t,
  prefix_sum_velo_clusters_t,
  velo_masked_clustering_t,
  velo_calculate_phi_and_sort_t,
  velo_fill_candidates_t,
  velo_search_by_triplet_t,
  velo_weak_tracks_adder_t,
  copy_and_prefix_sum_single_block_velo_t,
  copy_velo_track_hit_number_t,
  prefix_sum_velo_track_hit_number_t,
  consolidate_velo_tracks_t,
  velo_kalman_fit_t,
  pv_beamline_extrapolate_t,
  pv_beamline_histo_t,
  pv_beamline_peak_t,
  pv_beamline_multi_fitter_t)

// Kept allocated from first use to the end of the sequence
// (OutputArguments.cuh:16-34, VELO part).
typedef std::tuple<
  dev_atomics_velo,
  dev_velo_track_hit_number,
  dev_velo_track_hits,
  dev_velo_states,
  dev_accepted_velo_tracks,
  dev_velo_kalman_beamline_states,
  dev_multi_fit_vertices,
  dev_number_of_multi_fit_vertices,
  dev_event_list,
  dev_number_of_selected_events,
  dev_velo_event_status>
  output_arguments_t;
