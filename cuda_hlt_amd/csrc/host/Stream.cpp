// Stream::initialize / run_sequence (Stream.cu:16-83).
#include "sequence/Stream.h"

hipError_t Stream::initialize(
  unsigned max_number_of_events,
  bool print_memory_manager,
  unsigned param_start_event_offset,
  size_t reserve_mb,
  unsigned param_stream_number,
  const Constants& param_constants)
{
  hipCheck(hipStreamCreate(&stream)); // default priority
  hipCheck(hipEventCreateWithFlags(&event, hipEventDisableTiming));
  stream_number = param_stream_number;
  do_print_memory_manager = print_memory_manager;
  start_event_offset = param_start_event_offset;
  constants = param_constants;
  host_buffers.reserve(max_number_of_events);
  reserved_bytes = reserve_mb * 1024 * 1024;
  hipCheck(hipMalloc((void**) &dev_base_pointer, reserved_bytes));
  scheduler.initialize(do_print_memory_manager, reserved_bytes, dev_base_pointer);
  return hipSuccess;
}

hipError_t Stream::run_sequence(const RuntimeOptions& runtime_options)
{
  if (runtime_options.number_of_events > host_buffers.host_max_number_of_events)
    throw std::invalid_argument("run_sequence: more events than the stream was initialised for");
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
  }
  return hipSuccess;
}

void Stream::release()
{
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
