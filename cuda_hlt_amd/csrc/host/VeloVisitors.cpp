// SequenceVisitor specialisations of the VELO sequence.  Sizing and the
// three host reads of totals follow the reference's visitors
// (stream/visitors/velo/src/*.cu, InitEventListVisitor.cu:4-89); launches go
// through the C ABI and are checked.
#include <cstring>

#include "sequence/SequenceVisitor.h"

namespace {
// Counter pairs of the persistent steps (two words each, see velo_tuning).
enum WorkSlot : unsigned {
  kSlotEstimate,
  kSlotClustering,
  kSlotSort,
  kSlotFill,
  kSlotSearch,
  kSlotConsolidate,
  kNumWorkSlots
};

template<typename S>
void tune(S& state, const velo_tuning& t, const RuntimeOptions& ro, WorkSlot slot)
{
  velo_tuning x = t;
  if (x.grid_dim) {
    if (!ro.dev_work_counters) throw std::invalid_argument("persistent launch without work counters");
    x.dev_work_counter = ro.dev_work_counters + 2 * slot;
    // a caller's work order lists input events; after the cut, slots are
    // selected events, so only an order computed on the device applies
    if (!x.dev_work_order && !ro.run_global_event_cut) x.dev_work_order = ro.dev_work_order;
  }
  state.handler.set_tuning(x);
}

// Host read of one total between steps (the plain sequence's three
// synchronisation points); batched runs size by bounds instead.
void read_total(const HostBuffers& hb, unsigned* host, const unsigned* dev, hipStream_t& stream, hipEvent_t& event)
{
  hb.copy_to_host(host, dev, sizeof(unsigned), stream, "a total");
  hipCheck(hipEventRecord(event, stream));
  hipCheck(hipEventSynchronize(event));
}
static_assert(2 * kNumWorkSlots <= VELO_WORK_COUNTER_WORDS, "work counter words");
} // namespace

// ---- init_event_list_t ------------------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<init_event_list_t>(
  init_event_list_t::arguments_t arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  const HostBuffers&)
{
  if (runtime_options.dev_velopix_events) {
    arguments.set_external<dev_velo_raw_input>(
      const_cast<char*>(runtime_options.dev_velopix_events), runtime_options.host_velopix_events_size);
    arguments.set_external<dev_velo_raw_input_offsets>(
      const_cast<unsigned*>(runtime_options.dev_velopix_event_offsets),
      (runtime_options.number_of_events + 1) * sizeof(unsigned));
  }
  else {
    arguments.clear_external<dev_velo_raw_input>();
    arguments.clear_external<dev_velo_raw_input_offsets>();
    arguments.set_size<dev_velo_raw_input>(runtime_options.host_velopix_events_size);
    arguments.set_size<dev_velo_raw_input_offsets>(runtime_options.host_velopix_event_offsets_size);
  }
  // UT and SciFi banks travel only when the global event cut runs
  const bool gec = runtime_options.run_global_event_cut;
  const size_t n1 = (size_t) runtime_options.number_of_events + 1;
  if (gec && runtime_options.dev_ut_events) {
    arguments.set_external<dev_ut_raw_input>(
      const_cast<char*>(runtime_options.dev_ut_events), runtime_options.host_ut_events_size);
    arguments.set_external<dev_ut_raw_input_offsets>(
      const_cast<unsigned*>(runtime_options.dev_ut_event_offsets), n1 * sizeof(unsigned));
  }
  else {
    arguments.clear_external<dev_ut_raw_input>();
    arguments.clear_external<dev_ut_raw_input_offsets>();
    arguments.set_size<dev_ut_raw_input>(gec ? runtime_options.host_ut_events_size : 0);
    arguments.set_size<dev_ut_raw_input_offsets>(gec ? n1 : 0);
  }
  if (gec && runtime_options.dev_scifi_events) {
    arguments.set_external<dev_scifi_raw_input>(
      const_cast<char*>(runtime_options.dev_scifi_events), runtime_options.host_scifi_events_size);
    arguments.set_external<dev_scifi_raw_input_offsets>(
      const_cast<unsigned*>(runtime_options.dev_scifi_event_offsets), n1 * sizeof(unsigned));
  }
  else {
    arguments.clear_external<dev_scifi_raw_input>();
    arguments.clear_external<dev_scifi_raw_input_offsets>();
    arguments.set_size<dev_scifi_raw_input>(gec ? runtime_options.host_scifi_events_size : 0);
    arguments.set_size<dev_scifi_raw_input_offsets>(gec ? n1 : 0);
  }
  arguments.set_size<dev_event_list>(runtime_options.number_of_events);
  arguments.set_size<dev_number_of_selected_events>(1);
}

template<>
void SequenceVisitor::visit<init_event_list_t>(
  init_event_list_t& state,
  const init_event_list_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers&,
  hipStream_t& stream,
  hipEvent_t&)
{
  if (!runtime_options.dev_velopix_events) {
    hipCheck(hipMemcpyAsync(
      arguments.offset<dev_velo_raw_input>(), runtime_options.host_velopix_events,
      arguments.size<dev_velo_raw_input>(), hipMemcpyHostToDevice, stream));
    hipCheck(hipMemcpyAsync(
      arguments.offset<dev_velo_raw_input_offsets>(), runtime_options.host_velopix_event_offsets,
      arguments.size<dev_velo_raw_input_offsets>(), hipMemcpyHostToDevice, stream));
  }
  if (runtime_options.run_global_event_cut) {
    const size_t n1 = (size_t) runtime_options.number_of_events + 1;
    if (!runtime_options.dev_ut_events) {
      hipCheck(hipMemcpyAsync(
        arguments.offset<dev_ut_raw_input>(), runtime_options.host_ut_events, runtime_options.host_ut_events_size,
        hipMemcpyHostToDevice, stream));
      hipCheck(hipMemcpyAsync(
        arguments.offset<dev_ut_raw_input_offsets>(), runtime_options.host_ut_event_offsets, n1 * sizeof(unsigned),
        hipMemcpyHostToDevice, stream));
    }
    if (!runtime_options.dev_scifi_events) {
      hipCheck(hipMemcpyAsync(
        arguments.offset<dev_scifi_raw_input>(), runtime_options.host_scifi_events,
        runtime_options.host_scifi_events_size, hipMemcpyHostToDevice, stream));
      hipCheck(hipMemcpyAsync(
        arguments.offset<dev_scifi_raw_input_offsets>(), runtime_options.host_scifi_event_offsets,
        n1 * sizeof(unsigned), hipMemcpyHostToDevice, stream));
    }
  }
  state.set_opts(dim3(runtime_options.number_of_events), dim3(256), stream);
  state.set_arguments(arguments.offset<dev_event_list>());
  state.invoke();
}

// ---- global_event_cut_t (GlobalEventCutVisitor.cu:4-55) ------------------------
DEFINE_EMPTY_SET_ARGUMENTS_SIZE(global_event_cut_t)

template<>
void SequenceVisitor::visit<global_event_cut_t>(
  global_event_cut_t& state,
  const global_event_cut_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t& event)
{
  if (!runtime_options.run_global_event_cut) {
    for (unsigned i = 0; i < runtime_options.number_of_events; ++i) host_buffers.host_event_list[i] = i;
    return;
  }
  // the launcher writes the count itself: no host-to-device reset
  state.set_opts(dim3(runtime_options.number_of_events), dim3(256), stream);
  state.set_arguments(
    arguments.offset<dev_ut_raw_input>(),
    arguments.offset<dev_ut_raw_input_offsets>(),
    arguments.offset<dev_scifi_raw_input>(),
    arguments.offset<dev_scifi_raw_input_offsets>(),
    arguments.offset<dev_number_of_selected_events>(),
    arguments.offset<dev_event_list>(),
    arguments.size<dev_ut_raw_input>(),
    arguments.size<dev_scifi_raw_input>());
  state.invoke();
  host_buffers.copy_to_host(host_buffers.host_event_list, arguments.offset<dev_event_list>(),
                            (size_t) runtime_options.number_of_events * sizeof(unsigned), stream, "host_event_list");
  read_total(host_buffers, host_buffers.host_number_of_selected_events, arguments.offset<dev_number_of_selected_events>(), stream,
             event);
}

// ---- velo_event_order_t (no reference counterpart) --------------------------------
template<>
void SequenceVisitor::set_arguments_size<velo_event_order_t>(
  velo_event_order_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  arguments.set_size<dev_event_order>(host_buffers.host_number_of_selected_events[0]);
}

template<>
void SequenceVisitor::visit<velo_event_order_t>(
  velo_event_order_t& state,
  const velo_event_order_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(1024), stream);
  state.set_arguments(
    arguments.offset<dev_velo_raw_input_offsets>(),
    runtime_options.run_global_event_cut ? arguments.offset<dev_event_list>() : nullptr,
    arguments.offset<dev_event_order>());
  state.invoke();
}

// ---- velo_estimate_input_size_t ----------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<velo_estimate_input_size_t>(
  velo_estimate_input_size_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t n = host_buffers.host_number_of_selected_events[0];
  arguments.set_size<dev_estimated_input_size>(n * VELO_N_MODULES + 1);
  arguments.set_size<dev_module_cluster_num>(n * VELO_N_MODULES);
  arguments.set_size<dev_module_candidate_num>(n);
  arguments.set_size<dev_cluster_candidates>(n * VELO_MAX_CANDIDATES_EVENT);
  arguments.set_size<dev_velo_event_status>(n);
}

template<>
void SequenceVisitor::visit<velo_estimate_input_size_t>(
  velo_estimate_input_size_t& state,
  const velo_estimate_input_size_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants& constants,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  // 256 threads: four waves share the event's staged words and flat bank
  // table; small workgroups keep more events in flight per CU
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(256), stream);
  state.set_arguments(
    arguments.offset<dev_velo_raw_input>(),
    arguments.offset<dev_velo_raw_input_offsets>(),
    arguments.offset<dev_estimated_input_size>(),
    arguments.offset<dev_module_cluster_num>(),
    arguments.offset<dev_module_candidate_num>(),
    arguments.offset<dev_cluster_candidates>(),
    arguments.offset<dev_event_list>(),
    arguments.offset<dev_event_order>(),
    constants.dev_velo_candidate_ks,
    arguments.offset<dev_velo_event_status>());
  tune(state, runtime_options.tuning.estimate_input_size, runtime_options, kSlotEstimate);
  state.invoke();
}

// ---- prefix_sum_velo_clusters_t (host sync #1) --------------------------------
template<>
void SequenceVisitor::set_arguments_size<prefix_sum_velo_clusters_t>(
  prefix_sum_velo_clusters_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  arguments.set_size<dev_cluster_offset>(
    prefix_sum_velo_clusters_t::aux_array_size(host_buffers.host_number_of_selected_events[0] * VELO_N_MODULES));
}

template<>
void SequenceVisitor::visit<prefix_sum_velo_clusters_t>(
  prefix_sum_velo_clusters_t& state,
  const prefix_sum_velo_clusters_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t& event)
{
  const size_t n = host_buffers.host_number_of_selected_events[0];
  state.set_size(n * VELO_N_MODULES);
  state.set_opts(stream);
  state.set_arguments(arguments.offset<dev_estimated_input_size>(), arguments.offset<dev_cluster_offset>());
  state.invoke();
  if (!runtime_options.batched)
    read_total(host_buffers, host_buffers.host_total_number_of_velo_clusters,
               arguments.offset<dev_estimated_input_size>() + n * VELO_N_MODULES, stream, event);
}

// ---- velo_masked_clustering_t -------------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<velo_masked_clustering_t>(
  velo_masked_clustering_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  arguments.set_size<dev_velo_cluster_container>(6 * (size_t) host_buffers.host_total_number_of_velo_clusters[0]);
}

template<>
void SequenceVisitor::visit<velo_masked_clustering_t>(
  velo_masked_clustering_t& state,
  const velo_masked_clustering_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants& constants,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  // 256 threads: four waves per event keep more events in flight per CU
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(256), stream);
  state.set_arguments(
    arguments.offset<dev_velo_raw_input>(),
    arguments.offset<dev_velo_raw_input_offsets>(),
    arguments.offset<dev_estimated_input_size>(),
    arguments.offset<dev_module_cluster_num>(),
    arguments.offset<dev_module_candidate_num>(),
    arguments.offset<dev_cluster_candidates>(),
    arguments.offset<dev_velo_cluster_container>(),
    arguments.offset<dev_event_list>(),
    arguments.offset<dev_event_order>(),
    constants.dev_velo_geometry,
    constants.dev_velo_sp_patterns,
    constants.dev_velo_sp_fx,
    constants.dev_velo_sp_fy,
    arguments.offset<dev_velo_event_status>());
  tune(state, runtime_options.tuning.masked_clustering, runtime_options, kSlotClustering);
  state.invoke();
}

// ---- velo_calculate_phi_and_sort_t --------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<velo_calculate_phi_and_sort_t>(
  velo_calculate_phi_and_sort_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  arguments.set_size<dev_hit_permutation>(host_buffers.host_total_number_of_velo_clusters[0]);
}

template<>
void SequenceVisitor::visit<velo_calculate_phi_and_sort_t>(
  velo_calculate_phi_and_sort_t& state,
  const velo_calculate_phi_and_sort_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(256), stream);
  state.set_arguments(
    arguments.offset<dev_estimated_input_size>(),
    arguments.offset<dev_module_cluster_num>(),
    arguments.offset<dev_velo_cluster_container>(),
    arguments.offset<dev_hit_permutation>());
  tune(state, runtime_options.tuning.calculate_phi_and_sort, runtime_options, kSlotSort);
  state.invoke();
}

// ---- velo_fill_candidates_t ----------------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<velo_fill_candidates_t>(
  velo_fill_candidates_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  arguments.set_size<dev_h0_candidates>(2 * (size_t) host_buffers.host_total_number_of_velo_clusters[0]);
  arguments.set_size<dev_h2_candidates>(2 * (size_t) host_buffers.host_total_number_of_velo_clusters[0]);
}

template<>
void SequenceVisitor::visit<velo_fill_candidates_t>(
  velo_fill_candidates_t& state,
  const velo_fill_candidates_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(256), stream);
  state.set_arguments(
    arguments.offset<dev_velo_cluster_container>(),
    arguments.offset<dev_estimated_input_size>(),
    arguments.offset<dev_module_cluster_num>(),
    arguments.offset<dev_h0_candidates>(),
    arguments.offset<dev_h2_candidates>());
  tune(state, runtime_options.tuning.fill_candidates, runtime_options, kSlotFill);
  state.invoke();
}

// ---- velo_search_by_triplet_t --------------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<velo_search_by_triplet_t>(
  velo_search_by_triplet_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t n = host_buffers.host_number_of_selected_events[0];
  arguments.set_size<dev_tracks>(n * VELO_MAX_TRACKS);
  arguments.set_size<dev_tracklets>(0); // tracklet hits travel in the tracks-to-follow entries
  arguments.set_size<dev_tracks_to_follow>(n * VELO_TTF_MODULO * VELO_TTF_ENTRY_WORDS);
  arguments.set_size<dev_weak_tracks>(n * VELO_MAX_WEAK_TRACKS);
  arguments.set_size<dev_hit_used>(host_buffers.host_total_number_of_velo_clusters[0]);
  arguments.set_size<dev_atomics_velo>(n * VELO_NUM_ATOMICS + 1);
  arguments.set_size<dev_rel_indices>(0); // h1 candidates are staged in LDS
}

template<>
void SequenceVisitor::visit<velo_search_by_triplet_t>(
  velo_search_by_triplet_t& state,
  const velo_search_by_triplet_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants& constants,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(64), stream);
  state.set_arguments(
    arguments.offset<dev_velo_cluster_container>(),
    arguments.offset<dev_estimated_input_size>(),
    arguments.offset<dev_module_cluster_num>(),
    arguments.offset<dev_tracks>(),
    arguments.offset<dev_tracklets>(),
    arguments.offset<dev_tracks_to_follow>(),
    arguments.offset<dev_weak_tracks>(),
    arguments.offset<dev_hit_used>(),
    arguments.offset<dev_atomics_velo>(),
    arguments.offset<dev_h0_candidates>(),
    arguments.offset<dev_h2_candidates>(),
    arguments.offset<dev_rel_indices>(),
    constants.dev_velo_module_zs,
    arguments.offset<dev_velo_event_status>());
  // No search shape given: persistent over the largest-first order, as many
  // workgroups as fit on the device at once (the tail of one workgroup per
  // event otherwise leaves most of the last round idle).
  velo_tuning t = runtime_options.tuning.search_by_triplet;
  const velo_tuning none {};
  if (std::memcmp(&t, &none, sizeof(t)) == 0) {
    static thread_local uint32_t resident = 0;
    if (!resident) resident = velo_search_by_triplet_resident_workgroups(0);
    t.grid_dim = resident;
    t.dev_work_order = arguments.offset<dev_event_order>();
  }
  tune(state, t, runtime_options, kSlotSearch);
  state.invoke();
}

// ---- velo_weak_tracks_adder_t ---------------------------------------------------
DEFINE_EMPTY_SET_ARGUMENTS_SIZE(velo_weak_tracks_adder_t)

template<>
void SequenceVisitor::visit<velo_weak_tracks_adder_t>(
  velo_weak_tracks_adder_t& state,
  const velo_weak_tracks_adder_t::arguments_t& arguments,
  const RuntimeOptions&,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(64), stream);
  state.set_arguments(
    arguments.offset<dev_velo_cluster_container>(),
    arguments.offset<dev_estimated_input_size>(),
    arguments.offset<dev_tracks>(),
    arguments.offset<dev_weak_tracks>(),
    arguments.offset<dev_hit_used>(),
    arguments.offset<dev_atomics_velo>(),
    arguments.offset<dev_velo_event_status>());
  state.invoke();
}

// ---- copy_and_prefix_sum_single_block_velo_t (host sync #2) ----------------------
DEFINE_EMPTY_SET_ARGUMENTS_SIZE(copy_and_prefix_sum_single_block_velo_t)

template<>
void SequenceVisitor::visit<copy_and_prefix_sum_single_block_velo_t>(
  copy_and_prefix_sum_single_block_velo_t& state,
  const copy_and_prefix_sum_single_block_velo_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t& event)
{
  const unsigned n = host_buffers.host_number_of_selected_events[0];
  unsigned* atomics = reinterpret_cast<unsigned*>(arguments.offset<dev_atomics_velo>());
  state.set_opts(dim3(1), dim3(1024), stream);
  state.set_arguments(atomics + 2 * n, atomics, atomics + n, n);
  state.invoke();
  if (!runtime_options.batched)
    read_total(host_buffers, host_buffers.host_number_of_reconstructed_velo_tracks, atomics + 2 * n, stream, event);
}

// ---- copy_velo_track_hit_number_t ------------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<copy_velo_track_hit_number_t>(
  copy_velo_track_hit_number_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  arguments.set_size<dev_velo_track_hit_number>(host_buffers.velo_track_hit_number_size());
}

template<>
void SequenceVisitor::visit<copy_velo_track_hit_number_t>(
  copy_velo_track_hit_number_t& state,
  const copy_velo_track_hit_number_t::arguments_t& arguments,
  const RuntimeOptions&,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(256), stream);
  state.set_arguments(
    arguments.offset<dev_tracks>(), arguments.offset<dev_atomics_velo>(), arguments.offset<dev_velo_track_hit_number>());
  state.invoke();
}

// ---- prefix_sum_velo_track_hit_number_t (host sync #3) ----------------------------
template<>
void SequenceVisitor::set_arguments_size<prefix_sum_velo_track_hit_number_t>(
  prefix_sum_velo_track_hit_number_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  arguments.set_size<dev_prefix_sum_auxiliary_array_2>(
    prefix_sum_velo_track_hit_number_t::aux_array_size(host_buffers.host_number_of_reconstructed_velo_tracks[0]));
}

template<>
void SequenceVisitor::visit<prefix_sum_velo_track_hit_number_t>(
  prefix_sum_velo_track_hit_number_t& state,
  const prefix_sum_velo_track_hit_number_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t& event)
{
  const unsigned t = host_buffers.host_number_of_reconstructed_velo_tracks[0];
  state.set_size(t);
  state.set_opts(stream);
  state.set_arguments(arguments.offset<dev_velo_track_hit_number>(), arguments.offset<dev_prefix_sum_auxiliary_array_2>());
  // batched: t is the capacity and the track total sits at dev_atomics_velo[2N];
  // plain: t is the total (the handler outlives the run, so always set)
  const unsigned n = host_buffers.host_number_of_selected_events[0];
  state.set_device_count(
    runtime_options.batched ? reinterpret_cast<const uint32_t*>(arguments.offset<dev_atomics_velo>()) + 2 * n
                            : nullptr);
  state.invoke();
  if (!runtime_options.batched)
    read_total(host_buffers, host_buffers.host_accumulated_number_of_hits_in_velo_tracks,
               arguments.offset<dev_velo_track_hit_number>() + t, stream, event);
}

// ---- consolidate_velo_tracks_t ------------------------------------------------------
template<>
void SequenceVisitor::set_arguments_size<consolidate_velo_tracks_t>(
  consolidate_velo_tracks_t::arguments_t arguments,
  const RuntimeOptions&,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t t = host_buffers.host_number_of_reconstructed_velo_tracks[0];
  arguments.set_size<dev_velo_track_hits>(host_buffers.host_accumulated_number_of_hits_in_velo_tracks[0] * sizeof(Velo::Hit));
  arguments.set_size<dev_velo_states>(t * sizeof(VeloState));
  arguments.set_size<dev_accepted_velo_tracks>(t);
}

template<>
void SequenceVisitor::visit<consolidate_velo_tracks_t>(
  consolidate_velo_tracks_t& state,
  const consolidate_velo_tracks_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  const unsigned n = host_buffers.host_number_of_selected_events[0];
  state.set_opts(dim3(n), dim3(512), stream);
  state.set_arguments(
    arguments.offset<dev_atomics_velo>(),
    arguments.offset<dev_tracks>(),
    arguments.offset<dev_velo_track_hit_number>(),
    arguments.offset<dev_velo_cluster_container>(),
    arguments.offset<dev_estimated_input_size>(),
    arguments.offset<dev_module_cluster_num>(),
    arguments.offset<dev_velo_track_hits>(),
    arguments.offset<dev_velo_states>(),
    arguments.offset<dev_accepted_velo_tracks>(),
    runtime_options.dev_ghost_model,
    static_cast<float*>(nullptr));
  tune(state, runtime_options.tuning.consolidate, runtime_options, kSlotConsolidate);
  state.invoke();
  if (!runtime_options.copy_outputs_to_host || runtime_options.batched) return; // batched: copied after the graph
  // ConsolidateVeloVisitor.cu:47-66: device to host
  const size_t t = host_buffers.host_number_of_reconstructed_velo_tracks[0];
  const size_t h = host_buffers.host_accumulated_number_of_hits_in_velo_tracks[0];
  host_buffers.ensure_outputs(t, h);
  host_buffers.copy_to_host(host_buffers.host_atomics_velo, arguments.offset<dev_atomics_velo>(),
                            (2 * (size_t) n + 1) * sizeof(unsigned), stream, "host_atomics_velo");
  host_buffers.copy_to_host(host_buffers.host_velo_track_hit_number, arguments.offset<dev_velo_track_hit_number>(),
                            (t + 1) * sizeof(unsigned), stream, "host_velo_track_hit_number");
  host_buffers.copy_to_host(host_buffers.host_velo_track_hits, arguments.offset<dev_velo_track_hits>(),
                            h * sizeof(Velo::Hit), stream, "host_velo_track_hits");
  host_buffers.copy_to_host(host_buffers.host_velo_states, arguments.offset<dev_velo_states>(),
                            t * VELO_STATE_BYTES, stream, "host_velo_states");
  host_buffers.copy_to_host(host_buffers.host_accepted_velo_tracks, arguments.offset<dev_accepted_velo_tracks>(),
                            t, stream, "host_accepted_velo_tracks");
}

// ---- velo_kalman_fit_t (KalmanFilterVisitor.cu:4-44) -------------------------------
template<>
void SequenceVisitor::set_arguments_size<velo_kalman_fit_t>(
  velo_kalman_fit_t::arguments_t arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t t = runtime_options.run_velo_kalman_fit ? host_buffers.host_number_of_reconstructed_velo_tracks[0] : 0;
  arguments.set_size<dev_velo_kalman_beamline_states>(t * VELO_STATE_BYTES);
}

template<>
void SequenceVisitor::visit<velo_kalman_fit_t>(
  velo_kalman_fit_t& state,
  const velo_kalman_fit_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  if (!runtime_options.run_velo_kalman_fit) return;
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(256), stream);
  state.set_arguments(
    arguments.offset<dev_atomics_velo>(),
    arguments.offset<dev_velo_track_hit_number>(),
    arguments.offset<dev_velo_track_hits>(),
    arguments.offset<dev_velo_states>(),
    arguments.offset<dev_velo_kalman_beamline_states>());
  state.invoke();
  if (!runtime_options.copy_outputs_to_host || runtime_options.batched) return;
  const size_t t = host_buffers.host_number_of_reconstructed_velo_tracks[0];
  host_buffers.ensure_kalman(t);
  host_buffers.copy_to_host(host_buffers.host_kalmanvelo_states, arguments.offset<dev_velo_kalman_beamline_states>(),
                            t * VELO_STATE_BYTES, stream, "host_kalmanvelo_states");
}

// ---- pv_beamline_extrapolate_t (ExtrapolateVisitor.cu:5-33) --------------------
template<>
void SequenceVisitor::set_arguments_size<pv_beamline_extrapolate_t>(
  pv_beamline_extrapolate_t::arguments_t arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t t = runtime_options.run_pv_beamline ? host_buffers.host_number_of_reconstructed_velo_tracks[0] : 0;
  arguments.set_size<dev_pvtracks>(t);
}

template<>
void SequenceVisitor::visit<pv_beamline_extrapolate_t>(
  pv_beamline_extrapolate_t& state,
  const pv_beamline_extrapolate_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  if (!runtime_options.run_pv_beamline) return;
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(128), stream);
  state.set_arguments(
    arguments.offset<dev_velo_kalman_beamline_states>(),
    arguments.offset<dev_atomics_velo>(),
    arguments.offset<dev_velo_track_hit_number>(),
    arguments.offset<dev_pvtracks>());
  state.invoke();
}

// ---- pv_beamline_histo_t (HistoVisitor.cu:5-33) ---------------------------------
template<>
void SequenceVisitor::set_arguments_size<pv_beamline_histo_t>(
  pv_beamline_histo_t::arguments_t arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t n = runtime_options.run_pv_beamline ? host_buffers.host_number_of_selected_events[0] : 0;
  arguments.set_size<dev_zhisto>(n * VELO_PV_NBINS);
}

template<>
void SequenceVisitor::visit<pv_beamline_histo_t>(
  pv_beamline_histo_t& state,
  const pv_beamline_histo_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  if (!runtime_options.run_pv_beamline) return;
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(64), stream);
  state.set_arguments(
    arguments.offset<dev_atomics_velo>(),
    arguments.offset<dev_velo_track_hit_number>(),
    arguments.offset<dev_pvtracks>(),
    arguments.offset<dev_zhisto>());
  state.invoke();
}

// ---- pv_beamline_peak_t (PeakVisitor.cu:5-40) -----------------------------------
template<>
void SequenceVisitor::set_arguments_size<pv_beamline_peak_t>(
  pv_beamline_peak_t::arguments_t arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t n = runtime_options.run_pv_beamline ? host_buffers.host_number_of_selected_events[0] : 0;
  arguments.set_size<dev_zpeaks>(n * VELO_PV_MAX_VERTICES);
  arguments.set_size<dev_number_of_zpeaks>(n);
}

template<>
void SequenceVisitor::visit<pv_beamline_peak_t>(
  pv_beamline_peak_t& state,
  const pv_beamline_peak_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  if (!runtime_options.run_pv_beamline) return;
  state.set_opts(dim3(host_buffers.host_number_of_selected_events[0]), dim3(64), stream);
  state.set_arguments(
    arguments.offset<dev_zhisto>(), arguments.offset<dev_zpeaks>(), arguments.offset<dev_number_of_zpeaks>());
  state.invoke();
}

// ---- pv_beamline_multi_fitter_t (MultiFitterVisitor.cu:5-60) ---------------------
template<>
void SequenceVisitor::set_arguments_size<pv_beamline_multi_fitter_t>(
  pv_beamline_multi_fitter_t::arguments_t arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  const HostBuffers& host_buffers)
{
  const size_t n = runtime_options.run_pv_beamline ? host_buffers.host_number_of_selected_events[0] : 0;
  arguments.set_size<dev_multi_fit_vertices>(n * VELO_PV_MAX_VERTICES);
  arguments.set_size<dev_number_of_multi_fit_vertices>(n);
}

template<>
void SequenceVisitor::visit<pv_beamline_multi_fitter_t>(
  pv_beamline_multi_fitter_t& state,
  const pv_beamline_multi_fitter_t::arguments_t& arguments,
  const RuntimeOptions& runtime_options,
  const Constants&,
  HostBuffers& host_buffers,
  hipStream_t& stream,
  hipEvent_t&)
{
  if (!runtime_options.run_pv_beamline) return;
  const unsigned n = host_buffers.host_number_of_selected_events[0];
  // the kernel writes every event's vertex count: no memset first
  state.set_opts(dim3(n), dim3(64), stream);
  state.set_arguments(
    arguments.offset<dev_atomics_velo>(),
    arguments.offset<dev_velo_track_hit_number>(),
    arguments.offset<dev_pvtracks>(),
    arguments.offset<dev_zpeaks>(),
    arguments.offset<dev_number_of_zpeaks>(),
    arguments.offset<dev_multi_fit_vertices>(),
    arguments.offset<dev_number_of_multi_fit_vertices>());
  state.invoke();
  if (!runtime_options.copy_outputs_to_host || runtime_options.batched) return;
  host_buffers.copy_to_host(host_buffers.host_reconstructed_multi_pvs, arguments.offset<dev_multi_fit_vertices>(),
                            (size_t) n * VELO_PV_MAX_VERTICES * sizeof(velo_pv_vertex), stream, "host_reconstructed_multi_pvs");
  host_buffers.copy_to_host(host_buffers.host_number_of_multivertex, arguments.offset<dev_number_of_multi_fit_vertices>(),
                            (size_t) n * sizeof(unsigned), stream, "host_number_of_multivertex");
}
