// Arguments of the VELO sequence, with the reference's names and element
// types (ArgumentsVelo.cuh:10-30, ArgumentsCommon.cuh:9-19).
#pragma once

#include <cstdint>

#include "../gear/Argument.h"
#include "VeloEventModel.h"
#include "velo_pv.h"

using uint = unsigned int;

ARGUMENT(dev_velo_raw_input, char)
ARGUMENT(dev_velo_raw_input_offsets, uint)
ARGUMENT(dev_ut_raw_input, char)
ARGUMENT(dev_ut_raw_input_offsets, uint)
ARGUMENT(dev_scifi_raw_input, char)
ARGUMENT(dev_scifi_raw_input_offsets, uint)
ARGUMENT(dev_event_list, uint)
ARGUMENT(dev_event_order, uint)
ARGUMENT(dev_number_of_selected_events, uint)
ARGUMENT(dev_accepted_velo_tracks, uint8_t)
ARGUMENT(dev_estimated_input_size, uint)
ARGUMENT(dev_module_cluster_num, uint)
ARGUMENT(dev_module_candidate_num, uint)
ARGUMENT(dev_cluster_offset, uint)
ARGUMENT(dev_cluster_candidates, uint)
ARGUMENT(dev_velo_cluster_container, uint)
ARGUMENT(dev_tracks, Velo::TrackHits)
ARGUMENT(dev_tracks_to_follow, uint)
ARGUMENT(dev_hit_used, uint8_t)
ARGUMENT(dev_atomics_velo, int)
ARGUMENT(dev_tracklets, Velo::TrackletHits)
ARGUMENT(dev_weak_tracks, Velo::TrackletHits)
ARGUMENT(dev_h0_candidates, short)
ARGUMENT(dev_h2_candidates, short)
ARGUMENT(dev_rel_indices, unsigned short)
ARGUMENT(dev_hit_permutation, uint)
ARGUMENT(dev_velo_track_hit_number, uint)
ARGUMENT(dev_prefix_sum_auxiliary_array_2, uint)
ARGUMENT(dev_velo_track_hits, char)
ARGUMENT(dev_velo_states, char)
ARGUMENT(dev_velo_kalman_beamline_states, char)
// ArgumentsPV.cuh:11-16
ARGUMENT(dev_pvtracks, velo_pv_track)
ARGUMENT(dev_zhisto, float)
ARGUMENT(dev_zpeaks, float)
ARGUMENT(dev_number_of_zpeaks, uint)
ARGUMENT(dev_multi_fit_vertices, velo_pv_vertex)
ARGUMENT(dev_number_of_multi_fit_vertices, uint)
// Per-event capacity-overflow bits (no reference counterpart).
ARGUMENT(dev_velo_event_status, uint)
