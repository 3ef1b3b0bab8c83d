// Stream::initialize / run_sequence (Stream.cu:16-83).
#include "sequence/Stream.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

hipError_t Stream::initialize(
  unsigned max_number_of_events,
  bool print_memory_manager,
  unsigned param_start_event_offset,
  size_t reserve_mb,
  unsigned param_stream_number,
  const Constants& param_constants)
{
  // Default priority; non-blocking so that synchronous copies other threads
  // issue on the legacy stream never wait on (or join a capture of) it.
  hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  hipCheck(hipEventCreateWithFlags(&event, hipEventDisableTiming));
  stream_number = param_stream_number;
  do_print_memory_manager = print_memory_manager;
  start_event_offset = param_start_event_offset;
  constants = param_constants;
  host_buffers.reserve(max_number_of_events);
  reserved_bytes = reserve_mb * 1024 * 1024;
  hipCheck(hipMalloc((void**) &dev_base_pointer, reserved_bytes));
  scheduler.initialize(do_print_memory_manager, reserved_bytes, dev_base_pointer);
  hipCheck(hipMalloc((void**) &dev_work_counters, VELO_WORK_COUNTER_WORDS * sizeof(unsigned)));
  hipCheck(hipMemset(dev_work_counters, 0, VELO_WORK_COUNTER_WORDS * sizeof(unsigned)));
  hipCheck(hipEventCreate(&graph_start));
  hipCheck(hipEventCreate(&graph_end));
  return hipSuccess;
}

bool Stream::GraphKey::operator==(const GraphKey& o) const
{
  return dev_events == o.dev_events && dev_offsets == o.dev_offsets && dev_work_order == o.dev_work_order &&
         events_size == o.events_size && n_events == o.n_events && dev_ghost_model == o.dev_ghost_model &&
         run_velo_kalman_fit == o.run_velo_kalman_fit && run_pv_beamline == o.run_pv_beamline &&
         std::memcmp(&tuning, &o.tuning, sizeof(tuning)) == 0;
}

void Stream::batched_capacities(const RuntimeOptions& ro, HostBuffers& hb)
{
  const size_t n = ro.number_of_events;
  // Every SP word yields at most two clusters (estimate_input_size counts at
  // most two per SP), and the input holds fewer SP words than bytes / 4.
  const size_t clusters = ro.host_velopix_events_size / 2 + 1;
  const size_t tracks = n * VELO_MAX_TRACKS;
  const size_t track_hits = tracks * VELO_MAX_TRACK_SIZE;
  if (clusters > 0xFFFFFFFFull || track_hits > 0xFFFFFFFFull)
    throw std::invalid_argument("batched run: batch too large for 32-bit hit indices");
  hb.host_number_of_selected_events[0] = (unsigned) n;
  hb.host_total_number_of_velo_clusters[0] = (unsigned) clusters;
  hb.host_number_of_reconstructed_velo_tracks[0] = (unsigned) tracks;
  hb.host_accumulated_number_of_hits_in_velo_tracks[0] = (unsigned) track_hits;
}

size_t Stream::batched_arena_bytes(unsigned n_events, size_t events_bytes, bool run_velo_kalman_fit,
                                   bool run_pv_beamline)
{
  RuntimeOptions ro;
  ro.number_of_events = n_events;
  ro.host_velopix_events_size = events_bytes;
  ro.host_velopix_event_offsets_size = (size_t) n_events + 1;
  // the input is bound in place in a batched run: any non-null address
  // makes the sizing take it as external (nothing is dereferenced)
  static const char resident[4] = {};
  ro.dev_velopix_events = resident;
  ro.dev_velopix_event_offsets = reinterpret_cast<const unsigned*>(resident);
  ro.batched = true;
  ro.run_pv_beamline = run_pv_beamline;
  ro.run_velo_kalman_fit = run_velo_kalman_fit || run_pv_beamline;
  unsigned counters[4] = {};
  HostBuffers hb;
  hb.host_number_of_selected_events = counters;
  hb.host_total_number_of_velo_clusters = counters + 1;
  hb.host_number_of_reconstructed_velo_tracks = counters + 2;
  hb.host_accumulated_number_of_hits_in_velo_tracks = counters + 3;
  batched_capacities(ro, hb);
  scheduler_t sched;
  sched.initialize(false, (size_t) 1 << 62, nullptr);
  SequenceVisitor visitor;
  Constants constants;
  Sch::size_sequence(sched, visitor, std::forward_as_tuple(static_cast<const RuntimeOptions&>(ro),
                                                           static_cast<const Constants&>(constants),
                                                           static_cast<const HostBuffers&>(hb)));
  const size_t bytes = sched.memory_manager.total_memory_required;
  hb.host_number_of_selected_events = nullptr; // not pinned memory: nothing for release() to free
  hb.host_total_number_of_velo_clusters = hb.host_number_of_reconstructed_velo_tracks = nullptr;
  hb.host_accumulated_number_of_hits_in_velo_tracks = nullptr;
  return bytes;
}

void Stream::run_monte_carlo_test(CheckerInvoker& invoker, unsigned start_event_offset)
{
  const unsigned n = host_buffers.host_number_of_selected_events[0];
  Sch::check_sequence<configured_sequence_t>(
    sequence_visitor, start_event_offset, n, static_cast<const HostBuffers&>(host_buffers),
    static_cast<const Constants&>(constants), invoker);
}

void Stream::drop_graphs()
{
  for (auto& g : graphs) {
    if (g.exec) (void) hipGraphExecDestroy(g.exec);
    if (g.graph) (void) hipGraphDestroy(g.graph);
  }
  graphs.clear();
}

Stream::GraphEntry& Stream::capture(const RuntimeOptions& ro, const GraphKey& key)
{
  if (graphs.size() >= max_graphs) { // evict the least recently used
    auto lru = std::min_element(graphs.begin(), graphs.end(),
                                [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
    if (lru->exec) (void) hipGraphExecDestroy(lru->exec);
    if (lru->graph) (void) hipGraphDestroy(lru->graph);
    graphs.erase(lru);
  }
  batched_capacities(ro, host_buffers);
  scheduler.reset();
  hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  constexpr size_t n_steps = std::tuple_size_v<configured_sequence_t>;
  step_names.assign(n_steps + 1, "");
  auto hook = [&](size_t i, const char* name) { step_names[i] = name; };
  GraphEntry entry;
  entry.key = key;
  try {
    Sch::run_sequence(
      scheduler,
      sequence_visitor,
      hook,
      std::forward_as_tuple(ro, static_cast<const Constants&>(constants), static_cast<const HostBuffers&>(host_buffers)),
      std::forward_as_tuple(ro, static_cast<const Constants&>(constants), host_buffers, stream, event));
  } catch (...) {
    hipGraph_t partial = nullptr;
    (void) hipStreamEndCapture(stream, &partial);
    if (partial) (void) hipGraphDestroy(partial);
    throw;
  }
  hipCheck(hipStreamEndCapture(stream, &entry.graph));
  const hipError_t rc = hipGraphInstantiate(&entry.exec, entry.graph, nullptr, nullptr, 0);
  if (rc != hipSuccess) {
    (void) hipGraphDestroy(entry.graph);
    hipCheck(rc);
  }
  entry.layout = scheduler.argument_manager.arguments_tuple;
  entry.memory_required = scheduler.memory_manager.total_memory_required;
  graphs.push_back(entry);
  return graphs.back();
}

hipError_t Stream::run_batched(const RuntimeOptions& ro)
{
  if (!ro.dev_velopix_events || !ro.dev_velopix_event_offsets)
    throw std::invalid_argument("batched run: the input must be device-resident");
  if (ro.run_global_event_cut) throw std::invalid_argument("batched run: the global event cut runs in plain runs only");
  GraphKey key;
  key.dev_events = ro.dev_velopix_events;
  key.dev_offsets = ro.dev_velopix_event_offsets;
  key.dev_work_order = ro.dev_work_order;
  key.dev_ghost_model = ro.dev_ghost_model;
  key.run_velo_kalman_fit = ro.run_velo_kalman_fit;
  key.run_pv_beamline = ro.run_pv_beamline;
  key.events_size = ro.host_velopix_events_size;
  key.n_events = ro.number_of_events;
  key.tuning = ro.tuning;
  auto it = std::find_if(graphs.begin(), graphs.end(), [&](const GraphEntry& g) { return g.key == key; });
  GraphEntry& g = it != graphs.end() ? *it : capture(ro, key);
  g.last_use = ++graph_clock;
  scheduler.argument_manager.arguments_tuple = g.layout;
  const unsigned n = ro.number_of_events;
  hipCheck(hipEventRecord(graph_start, stream));
  for (unsigned r = 0; r < ro.number_of_repetitions; ++r) hipCheck(hipGraphLaunch(g.exec, stream));
  hipCheck(hipEventRecord(graph_end, stream));
  hipCheck(hipEventSynchronize(graph_end));
  hipCheck(hipEventElapsedTime(&graph_ms, graph_start, graph_end));
  if (time_steps) {
    step_ms.assign(1, graph_ms / (float) ro.number_of_repetitions);
    step_names.assign(2, "");
    step_names[0] = "velo_batched_graph";
    step_names[1] = "end";
  }

  // totals (the plain sequence reads them between steps), on this stream
  unsigned* hb = host_buffers.host_number_of_selected_events;
  hb[0] = n;
  auto d2h = [&](void* dst, const void* src, size_t bytes, const char* what) {
    host_buffers.copy_to_host(dst, src, bytes, stream, what);
  };
  auto wait = [&]() {
    hipCheck(hipEventRecord(event, stream));
    hipCheck(hipEventSynchronize(event));
  };
  d2h(hb + 1, output<dev_estimated_input_size>() + (size_t) n * VELO_N_MODULES, sizeof(unsigned), "cluster total");
  d2h(hb + 2, reinterpret_cast<unsigned*>(output<dev_atomics_velo>()) + 2 * (size_t) n, sizeof(unsigned),
      "track total");
  wait();
  // the totals come from the device: they index and size host copies only
  // within the bounds the graph was captured for
  const size_t max_tracks = (size_t) n * VELO_MAX_TRACKS;
  if (hb[2] > max_tracks)
    throw std::runtime_error("batched run: track total " + std::to_string(hb[2]) + " exceeds its bound " +
                             std::to_string(max_tracks));
  d2h(hb + 3, output<dev_velo_track_hit_number>() + hb[2], sizeof(unsigned), "track hit total");
  wait();
  if ((size_t) hb[3] > (size_t) hb[2] * VELO_MAX_TRACK_SIZE)
    throw std::runtime_error("batched run: track hit total " + std::to_string(hb[3]) + " exceeds its bound");
  if (ro.copy_outputs_to_host) {
    const size_t t = hb[2], h = hb[3];
    host_buffers.ensure_outputs(t, h);
    d2h(host_buffers.host_atomics_velo, output<dev_atomics_velo>(), (2 * (size_t) n + 1) * sizeof(unsigned),
        "atomics");
    d2h(host_buffers.host_velo_track_hit_number, output<dev_velo_track_hit_number>(), (t + 1) * sizeof(unsigned),
        "track hit number");
    d2h(host_buffers.host_velo_track_hits, output<dev_velo_track_hits>(), h * 16, "track hits");
    d2h(host_buffers.host_velo_states, output<dev_velo_states>(), t * VELO_STATE_BYTES, "states");
    d2h(host_buffers.host_accepted_velo_tracks, output<dev_accepted_velo_tracks>(), t, "accepted");
    if (ro.run_velo_kalman_fit) {
      host_buffers.ensure_kalman(t);
      d2h(host_buffers.host_kalmanvelo_states, output<dev_velo_kalman_beamline_states>(), t * VELO_STATE_BYTES,
          "kalman states");
    }
    if (ro.run_pv_beamline) {
      d2h(host_buffers.host_reconstructed_multi_pvs, output<dev_multi_fit_vertices>(),
          (size_t) n * VELO_PV_MAX_VERTICES * sizeof(velo_pv_vertex), "multi-fit vertices");
      d2h(host_buffers.host_number_of_multivertex, output<dev_number_of_multi_fit_vertices>(), n * sizeof(unsigned),
          "vertex counts");
    }
    wait();
  }
  host_buffers.check_guards();
  return hipSuccess;
}

hipError_t Stream::run_sequence(const RuntimeOptions& options)
{
  if (options.number_of_events > host_buffers.host_max_number_of_events)
    throw std::invalid_argument("run_sequence: more events than the stream was initialised for");
  RuntimeOptions runtime_options = options;
  runtime_options.dev_work_counters = dev_work_counters;
  if (runtime_options.batched) return run_batched(runtime_options);
  // the graphs stay valid: each restores its own layout of the arena
  for (unsigned repetition = 0; repetition < runtime_options.number_of_repetitions; ++repetition) {
    host_buffers.host_number_of_selected_events[0] = runtime_options.number_of_events;
    scheduler.reset();
    constexpr size_t n_steps = std::tuple_size_v<configured_sequence_t>;
    if (time_steps && step_events.size() != n_steps + 1) {
      for (auto e : step_events) (void) hipEventDestroy(e);
      step_events.assign(n_steps + 1, nullptr);
      for (auto& e : step_events) hipCheck(hipEventCreate(&e));
    }
    step_names.assign(n_steps + 1, "");
    auto hook = [&](size_t i, const char* name) {
      step_names[i] = name;
      if (time_steps) hipCheck(hipEventRecord(step_events[i], stream));
    };
    Sch::run_sequence(
      scheduler,
      sequence_visitor,
      hook,
      std::forward_as_tuple(runtime_options, static_cast<const Constants&>(constants), static_cast<const HostBuffers&>(host_buffers)),
      std::forward_as_tuple(runtime_options, static_cast<const Constants&>(constants), host_buffers, stream, event));
    hipCheck(hipEventRecord(event, stream));
    hipCheck(hipEventSynchronize(event));
    if (time_steps) {
      step_ms.assign(n_steps, 0.f);
      for (size_t i = 0; i < n_steps; ++i) hipCheck(hipEventElapsedTime(&step_ms[i], step_events[i], step_events[i + 1]));
    }
    host_buffers.check_guards();
  }
  return hipSuccess;
}

void Stream::release()
{
  drop_graphs();
  if (dev_work_counters) (void) hipFree(dev_work_counters);
  dev_work_counters = nullptr;
  if (graph_start) (void) hipEventDestroy(graph_start);
  if (graph_end) (void) hipEventDestroy(graph_end);
  graph_start = graph_end = nullptr;
  for (auto e : step_events)
    if (e) (void) hipEventDestroy(e);
  step_events.clear();
  if (dev_base_pointer) (void) hipFree(dev_base_pointer);
  dev_base_pointer = nullptr;
  host_buffers.release();
  if (event) (void) hipEventDestroy(event);
  if (stream) (void) hipStreamDestroy(stream);
  event = nullptr;
  stream = nullptr;
}
