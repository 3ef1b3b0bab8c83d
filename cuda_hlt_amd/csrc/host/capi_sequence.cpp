// C ABI of include/velo_sequence.h over the C++ Stream.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>
#include <exception>
#include <memory>

#include "sequence/Stream.h"
#include "velo_sequence.h"

struct velo_sequence {
  Constants constants;
  std::unique_ptr<Stream> stream;
  RuntimeOptions last {};
};

namespace {
void set_error(char* err, size_t len, const char* msg)
{
  if (err && len) {
    std::snprintf(err, len, "%s", msg);
  }
}
} // namespace

extern "C" {

velo_sequence* velo_sequence_create(
  const velo_sequence_config* cfg, const void* geometry, size_t geometry_bytes, const float* module_zs, char* err,
  size_t err_len)
{
  if (!cfg || !geometry) {
    set_error(err, err_len, "velo_sequence_create: null argument");
    return nullptr;
  }
  auto s = std::make_unique<velo_sequence>();
  try {
    s->constants.reserve_and_initialize(geometry, geometry_bytes, module_zs);
    s->stream = std::make_unique<Stream>();
    s->stream->initialize(cfg->max_events, cfg->print_memory_manager != 0, 0, cfg->reserve_mb, 0, s->constants);
  }
  catch (const std::exception& e) {
    set_error(err, err_len, e.what());
    s->stream.reset();
    s->constants.release();
    return nullptr;
  }
  return s.release();
}

void velo_sequence_destroy(velo_sequence* s)
{
  if (!s) return;
  s->stream.reset();
  s->constants.release();
  delete s;
}

int velo_sequence_run(velo_sequence* s, const velo_sequence_input* in, char* err, size_t err_len)
{
  if (!s || !in || (!in->dev_events && !in->host_events) || (in->dev_events && !in->dev_offsets) ||
      (!in->dev_events && !in->host_offsets) || in->n_events == 0) {
    set_error(err, err_len, "velo_sequence_run: invalid input");
    return (int) hipErrorInvalidValue;
  }
  RuntimeOptions ro;
  ro.host_velopix_events = in->host_events;
  ro.host_velopix_event_offsets = in->host_offsets;
  ro.host_velopix_events_size = in->events_size;
  ro.host_velopix_event_offsets_size = in->n_events + 1;
  ro.dev_velopix_events = in->dev_events;
  ro.dev_velopix_event_offsets = in->dev_offsets;
  ro.number_of_events = in->n_events;
  ro.number_of_repetitions = in->repetitions ? in->repetitions : 1;
  ro.copy_outputs_to_host = in->copy_outputs_to_host != 0;
  ro.tuning.estimate_input_size = in->tuning[0];
  ro.tuning.masked_clustering = in->tuning[1];
  ro.tuning.calculate_phi_and_sort = in->tuning[2];
  ro.tuning.fill_candidates = in->tuning[3];
  ro.tuning.search_by_triplet = in->tuning[4];
  ro.tuning.consolidate = in->tuning[5];
  ro.batched = in->batched != 0;
  ro.dev_work_order = in->dev_work_order;
  ro.dev_ghost_model = in->dev_ghost_model;
  ro.run_pv_beamline = in->run_pv_beamline != 0;
  ro.run_velo_kalman_fit = in->run_velo_kalman_fit != 0 || ro.run_pv_beamline;
  ro.run_global_event_cut = in->run_global_event_cut != 0;
  if (ro.run_global_event_cut) {
    const bool ut_ok = in->dev_ut_events ? in->dev_ut_offsets != nullptr : in->host_ut_events && in->host_ut_offsets;
    const bool scifi_ok =
      in->dev_scifi_events ? in->dev_scifi_offsets != nullptr : in->host_scifi_events && in->host_scifi_offsets;
    if (!ut_ok || !scifi_ok || ro.batched) {
      set_error(err, err_len,
                ro.batched ? "velo_sequence_run: the global event cut runs in plain runs only"
                           : "velo_sequence_run: the global event cut needs UT and SciFi events with offsets");
      return (int) hipErrorInvalidValue;
    }
    ro.host_ut_events = in->host_ut_events;
    ro.host_ut_event_offsets = in->host_ut_offsets;
    ro.host_ut_events_size = in->ut_events_size;
    ro.host_scifi_events = in->host_scifi_events;
    ro.host_scifi_event_offsets = in->host_scifi_offsets;
    ro.host_scifi_events_size = in->scifi_events_size;
    ro.dev_ut_events = in->dev_ut_events;
    ro.dev_ut_event_offsets = in->dev_ut_offsets;
    ro.dev_scifi_events = in->dev_scifi_events;
    ro.dev_scifi_event_offsets = in->dev_scifi_offsets;
  }
  try {
    s->stream->run_sequence(ro);
  }
  catch (const std::exception& e) {
    set_error(err, err_len, e.what());
    return (int) hipErrorUnknown;
  }
  s->last = ro;
  return 0;
}

int velo_sequence_output_get(const velo_sequence* s, velo_sequence_output* out)
{
  if (!s || !out || !s->stream) return (int) hipErrorInvalidValue;
  const Stream& st = *s->stream;
  const HostBuffers& hb = st.host_buffers;
  std::memset(out, 0, sizeof(*out));
  out->n_events = hb.host_number_of_selected_events[0];
  out->total_estimated_hits = hb.host_total_number_of_velo_clusters[0];
  out->total_tracks = hb.host_number_of_reconstructed_velo_tracks[0];
  out->total_track_hits = hb.host_accumulated_number_of_hits_in_velo_tracks[0];
  if (s->last.copy_outputs_to_host) {
    out->host_atomics_velo = hb.host_atomics_velo;
    out->host_track_hit_number = hb.host_velo_track_hit_number;
    out->host_track_hits = hb.host_velo_track_hits;
    out->host_states = hb.host_velo_states;
    out->host_accepted = hb.host_accepted_velo_tracks;
    if (s->last.run_velo_kalman_fit) out->host_kalman_states = hb.host_kalmanvelo_states;
  }
  out->dev_atomics_velo = st.output<dev_atomics_velo>();
  out->dev_track_hit_number = st.output<dev_velo_track_hit_number>();
  out->dev_track_hits = st.output<dev_velo_track_hits>();
  out->dev_states = st.output<dev_velo_states>();
  out->dev_accepted = st.output<dev_accepted_velo_tracks>();
  out->dev_kalman_states = st.output<dev_velo_kalman_beamline_states>();
  out->dev_event_status = st.output<dev_velo_event_status>();
  out->dev_module_cluster_start = st.output<dev_estimated_input_size>();
  out->dev_module_cluster_num = st.output<dev_module_cluster_num>();
  out->dev_cluster_container = st.output<dev_velo_cluster_container>();
  out->memory_required = st.scheduler.memory_manager.total_memory_required;
  out->host_event_list = hb.host_event_list;
  out->dev_event_list = st.output<dev_event_list>();
  if (s->last.run_pv_beamline) {
    if (s->last.copy_outputs_to_host) {
      out->host_pv_vertices = reinterpret_cast<const velo_pv_vertex*>(hb.host_reconstructed_multi_pvs);
      out->host_pv_counts = hb.host_number_of_multivertex;
    }
    out->dev_pv_vertices = st.output<dev_multi_fit_vertices>();
    out->dev_pv_counts = st.output<dev_number_of_multi_fit_vertices>();
  }
  return 0;
}

size_t velo_sequence_struct_size(int which)
{
  switch (which) {
  case 0: return sizeof(velo_sequence_config);
  case 1: return sizeof(velo_sequence_input);
  case 2: return sizeof(velo_sequence_output);
  default: return 0;
  }
}

size_t velo_workspace_bytes(uint32_t n_events, size_t events_bytes, int run_velo_kalman_fit, int run_pv_beamline)
{
  try {
    return Stream::batched_arena_bytes(n_events, events_bytes, run_velo_kalman_fit != 0, run_pv_beamline != 0);
  }
  catch (const std::exception&) {
    return 0;
  }
}

int velo_status(const velo_sequence* s, uint32_t* host_status, uint32_t max_events, uint32_t* combined)
{
  if (!s || !s->stream) return -1;
  const Stream& st = *s->stream;
  const uint32_t n = st.host_buffers.host_number_of_selected_events[0];
  const uint32_t* dev = st.output<dev_velo_event_status>();
  if (combined) *combined = 0;
  if (n == 0) return 0;
  if (!dev) return -1;
  std::vector<uint32_t> all(n);
  if (hipMemcpyAsync(all.data(), dev, (size_t) n * sizeof(uint32_t), hipMemcpyDeviceToHost, st.stream) != hipSuccess ||
      hipStreamSynchronize(st.stream) != hipSuccess)
    return -1;
  if (host_status) std::memcpy(host_status, all.data(), (size_t) std::min(n, max_events) * sizeof(uint32_t));
  if (combined)
    for (uint32_t v : all) *combined |= v & VELO_STATUS_ERROR_MASK;
  return (int) n;
}

int velo_sequence_check(velo_sequence* s, const velo_mc_truth* truth, uint32_t start_event,
                        velo_checker_result* result, char* err, size_t err_len)
{
  if (!s || !s->stream || !truth || !result) {
    set_error(err, err_len, "velo_sequence_check: null argument");
    return -1;
  }
  if (!s->last.copy_outputs_to_host) {
    set_error(err, err_len, "velo_sequence_check: the last run did not copy its outputs to the host");
    return -1;
  }
  CheckerInvoker invoker;
  invoker.mcps = truth->mcps;
  invoker.n_mcps = truth->n_mcps;
  invoker.mc_hits = truth->mc_hits;
  invoker.n_mc_events = truth->n_events;
  invoker.velo_result = result;
  try {
    s->stream->run_monte_carlo_test(invoker, start_event);
  }
  catch (const std::exception& e) {
    set_error(err, err_len, e.what());
    return -1;
  }
  if (!invoker.velo_checked) {
    set_error(err, err_len, "velo_sequence_check: no step checked the tracks");
    return -1;
  }
  return 0;
}

void* velo_sequence_stream(const velo_sequence* s) { return s && s->stream ? (void*) s->stream->stream : nullptr; }

int velo_sequence_enable_step_timing(velo_sequence* s, int enable)
{
  if (!s || !s->stream) return (int) hipErrorInvalidValue;
  s->stream->time_steps = enable != 0;
  return 0;
}

int velo_sequence_step_times(const velo_sequence* s, float* ms, const char** names, int max_steps)
{
  if (!s || !s->stream) return -1;
  const Stream& st = *s->stream;
  const int n = (int) st.step_ms.size();
  for (int i = 0; i < n && i < max_steps; ++i) {
    if (ms) ms[i] = st.step_ms[i];
    if (names) names[i] = st.step_names[i];
  }
  return n;
}

} // extern "C"
