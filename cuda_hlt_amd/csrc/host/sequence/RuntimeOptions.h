// Inputs of one run of the sequence (main/include/RuntimeOptions.h, VELO
// part).  Raw events come either from host memory (copied to the device by
// init_event_list, as the reference does) or from device memory the caller
// keeps resident (dev_velopix_events); the second skips the copy.
#pragma once

#include <cstddef>
#include <cstdint>

#include "velo_hlt.h"

struct VeloTuningOptions {
  velo_tuning estimate_input_size {};
  velo_tuning masked_clustering {};
  velo_tuning calculate_phi_and_sort {};
  velo_tuning fill_candidates {};
  velo_tuning search_by_triplet {};
  velo_tuning consolidate {};
};

struct RuntimeOptions {
  const char* host_velopix_events = nullptr;
  const unsigned* host_velopix_event_offsets = nullptr;
  size_t host_velopix_events_size = 0;
  size_t host_velopix_event_offsets_size = 0; // elements (N+1)
  const char* dev_velopix_events = nullptr;
  const unsigned* dev_velopix_event_offsets = nullptr;
  unsigned number_of_events = 0;
  unsigned number_of_repetitions = 1;
  bool copy_outputs_to_host = true;
  VeloTuningOptions tuning {};
};
