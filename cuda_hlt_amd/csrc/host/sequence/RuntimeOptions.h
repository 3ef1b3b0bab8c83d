// Inputs of one run of the sequence (main/include/RuntimeOptions.h, VELO
// part).  Raw events come either from host memory (copied to the device by
// init_event_list, as the reference does) or from device memory the caller
// keeps resident (dev_velopix_events); the second skips the copy.
#pragma once

#include <cstddef>
#include <cstdint>

#include "velo_hlt.h"

// Device words the Stream keeps for the persistent steps' work counters.
constexpr unsigned VELO_WORK_COUNTER_WORDS = 16;

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
  // Batched runs: no host reads between steps (buffers are sized by upper
  // bounds from the input size and the per-event capacities), the whole
  // sequence is captured once in a hipGraph and replayed; needs
  // device-resident input.
  bool batched = false;
  // Persistent launches (tuning.grid_dim > 0): processing order of the
  // events (largest first balances uneven events), NULL = identity.
  const unsigned* dev_work_order = nullptr;
  // UT and SciFi banks for the global event cut (main/include/
  // RuntimeOptions.h:14-21): host input copied by init_event_list, or
  // device-resident input used in place.  Raw-event layout (u32 n_banks,
  // byte offsets, per bank the source id and its words).
  bool run_global_event_cut = false;
  const char* host_ut_events = nullptr;
  const unsigned* host_ut_event_offsets = nullptr;
  size_t host_ut_events_size = 0;
  const char* host_scifi_events = nullptr;
  const unsigned* host_scifi_event_offsets = nullptr;
  size_t host_scifi_events_size = 0;
  const char* dev_ut_events = nullptr;
  const unsigned* dev_ut_event_offsets = nullptr;
  const char* dev_scifi_events = nullptr;
  const unsigned* dev_scifi_event_offsets = nullptr;
  // Run velo_kalman_fit after consolidation (KalmanFilterVisitor.cu:4-44).
  bool run_velo_kalman_fit = false;
  // Run the four pv_beamline steps after it (DefaultSequence.h:19-23);
  // implies run_velo_kalman_fit.
  bool run_pv_beamline = false;
  // Ghost filter run inline by consolidation (a device velo_ghost_model),
  // NULL = every track accepted.
  const void* dev_ghost_model = nullptr;
  // Set by the Stream: two counter words per persistent step.
  unsigned* dev_work_counters = nullptr;
};
