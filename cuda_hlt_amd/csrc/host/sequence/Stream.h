// One stream of the sequence: a HIP stream, a private device arena managed
// by the scheduler, pinned host buffers and the shared constants
// (reference interface: stream/sequence/include/Stream.cuh,
// Stream.cu:16-83).
#pragma once

#include <hip/hip_runtime_api.h>

#include <vector>

#include "../gear/Scheduler.h"
#include "SequenceVisitor.h"

struct Stream {
  using scheduler_t = Scheduler<configured_sequence_t, output_arguments_t>;

  scheduler_t scheduler;
  hipStream_t stream = nullptr;
  hipEvent_t event = nullptr;
  unsigned stream_number = 0;
  bool do_print_memory_manager = false;
  HostBuffers host_buffers;
  unsigned start_event_offset = 0;
  char* dev_base_pointer = nullptr;
  size_t reserved_bytes = 0;
  Constants constants;
  SequenceVisitor sequence_visitor;

  // Work counters of persistent launches (zeroed once; every persistent
  // launch leaves its pair at zero).
  unsigned* dev_work_counters = nullptr;

  // Batched runs: the captured sequence and what it was captured for.
  struct GraphKey {
    const char* dev_events = nullptr;
    const unsigned* dev_offsets = nullptr;
    const unsigned* dev_work_order = nullptr;
    const void* dev_ghost_model = nullptr;
    bool run_velo_kalman_fit = false;
    bool run_pv_beamline = false;
    size_t events_size = 0;
    unsigned n_events = 0;
    VeloTuningOptions tuning {};
    bool operator==(const GraphKey& o) const;
  };
  // One captured graph per key (batch shape), with the argument layout it
  // was captured on; replaying a graph restores its layout so output<A>()
  // points at that run's buffers.  All graphs share the arena: one runs at
  // a time on this stream.
  struct GraphEntry {
    GraphKey key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    typename scheduler_t::arguments_tuple_t layout;
    size_t memory_required = 0;
    unsigned long long last_use = 0;
  };
  static constexpr size_t max_graphs = 16;
  std::vector<GraphEntry> graphs;
  unsigned long long graph_clock = 0;
  hipEvent_t graph_start = nullptr, graph_end = nullptr;
  float graph_ms = 0.f;

  // Optional per-step timing with HIP events recorded on `stream` before
  // every step (for the benchmark and profiling); times of the last run.
  bool time_steps = false;
  std::vector<hipEvent_t> step_events;
  std::vector<float> step_ms;
  std::vector<const char*> step_names;

  Stream() = default;
  Stream(const Stream&) = delete;
  ~Stream() { release(); }

  hipError_t initialize(
    unsigned max_number_of_events,
    bool print_memory_manager,
    unsigned param_start_event_offset,
    size_t reserve_mb,
    unsigned param_stream_number,
    const Constants& param_constants);

  hipError_t run_sequence(const RuntimeOptions& runtime_options);

  // Capacities a batched run sizes its buffers by (no host reads).
  static void batched_capacities(const RuntimeOptions& ro, HostBuffers& hb);

  // Arena bytes a batched run of n_events events holding events_bytes raw
  // bytes needs: the scheduler lays the sequence out on the batched bounds
  // (set_arguments_size and setup of every step, nothing launched, no
  // device call) and reports its high-water mark.
  static size_t batched_arena_bytes(unsigned n_events, size_t events_bytes, bool run_velo_kalman_fit,
                                    bool run_pv_beamline);

  // Checks the last run's outputs (copied to the host) against MC truth
  // with every step's SequenceVisitor::check (Stream.cu:85-105).
  void run_monte_carlo_test(CheckerInvoker& invoker, unsigned start_event_offset);

  template<typename A>
  typename A::type* output() const
  {
    return const_cast<scheduler_t&>(scheduler).argument_manager.template offset<A>();
  }

  void release();

private:
  hipError_t run_batched(const RuntimeOptions& runtime_options);
  GraphEntry& capture(const RuntimeOptions& runtime_options, const GraphKey& key);
  void drop_graphs();
};
