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

  template<typename A>
  typename A::type* output() const
  {
    return const_cast<scheduler_t&>(scheduler).argument_manager.template offset<A>();
  }

  void release();
};
