// Per-algorithm host glue: buffer sizing from host counters and the launch
// (reference interface: stream/sequence/include/SequenceVisitor.cuh:11-58).
// The specialisations for the VELO sequence are declared here and defined in
// VeloVisitors.cpp.
#pragma once

#include <hip/hip_runtime_api.h>

#include "CheckerInvoker.h"
#include "Constants.h"
#include "HostBuffers.h"
#include "RuntimeOptions.h"
#include "VeloSequence.h"

struct SequenceVisitor {
  template<typename T>
  void set_arguments_size(
    typename T::arguments_t arguments,
    const RuntimeOptions& runtime_options,
    const Constants& constants,
    const HostBuffers& host_buffers);

  template<typename T>
  void visit(
    T& state,
    const typename T::arguments_t& arguments,
    const RuntimeOptions& runtime_options,
    const Constants& constants,
    HostBuffers& host_buffers,
    hipStream_t& stream,
    hipEvent_t& event);

  // Checks the output of step T against MC truth after a run (reference:
  // SequenceVisitor.cuh:42-48); steps without a checker do nothing.
  template<typename T>
  void check(
    const unsigned& start_event_offset,
    const unsigned& number_of_events_requested,
    const HostBuffers& host_buffers,
    const Constants& constants,
    CheckerInvoker& checker_invoker) const
  {}
};

#define DECLARE_VISITOR(_TYPE)                                                                          \
  template<>                                                                                            \
  void SequenceVisitor::set_arguments_size<_TYPE>(                                                      \
    typename _TYPE::arguments_t, const RuntimeOptions&, const Constants&, const HostBuffers&);          \
  template<>                                                                                            \
  void SequenceVisitor::visit<_TYPE>(                                                                   \
    _TYPE&, const typename _TYPE::arguments_t&, const RuntimeOptions&, const Constants&, HostBuffers&,  \
    hipStream_t&, hipEvent_t&);

#define DEFINE_EMPTY_SET_ARGUMENTS_SIZE(_TYPE)                                                          \
  template<>                                                                                            \
  void SequenceVisitor::set_arguments_size<_TYPE>(                                                      \
    typename _TYPE::arguments_t, const RuntimeOptions&, const Constants&, const HostBuffers&)           \
  {}

DECLARE_VISITOR(init_event_list_t)
DECLARE_VISITOR(global_event_cut_t)
DECLARE_VISITOR(velo_event_order_t)
DECLARE_VISITOR(velo_estimate_input_size_t)
DECLARE_VISITOR(prefix_sum_velo_clusters_t)
DECLARE_VISITOR(velo_masked_clustering_t)
DECLARE_VISITOR(velo_calculate_phi_and_sort_t)
DECLARE_VISITOR(velo_fill_candidates_t)
DECLARE_VISITOR(velo_search_by_triplet_t)
DECLARE_VISITOR(velo_weak_tracks_adder_t)
DECLARE_VISITOR(copy_and_prefix_sum_single_block_velo_t)
DECLARE_VISITOR(copy_velo_track_hit_number_t)
DECLARE_VISITOR(prefix_sum_velo_track_hit_number_t)
DECLARE_VISITOR(consolidate_velo_tracks_t)
DECLARE_VISITOR(velo_kalman_fit_t)
DECLARE_VISITOR(pv_beamline_extrapolate_t)
DECLARE_VISITOR(pv_beamline_histo_t)
DECLARE_VISITOR(pv_beamline_peak_t)
DECLARE_VISITOR(pv_beamline_multi_fitter_t)

// VeloSequenceCheckers_impl.cuh:8-24: the VELO tracks after consolidation.
template<>
void SequenceVisitor::check<consolidate_velo_tracks_t>(
  const unsigned& start_event_offset,
  const unsigned& number_of_events_requested,
  const HostBuffers& host_buffers,
  const Constants& constants,
  CheckerInvoker& checker_invoker) const;
