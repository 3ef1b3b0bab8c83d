// Pinned host buffers of one stream (HostBuffers.cuh / HostBuffers.cu:5-83,
// VELO part): the counters the visitors size buffers from, and the
// consolidated output copied back at the end of the sequence.  Output
// buffers grow on demand.
#pragma once

#include <cstdint>

#include "../gear/Common.h"
#include "velo_pv.h"

struct HostBuffers {
  unsigned* host_number_of_selected_events = nullptr;
  unsigned* host_total_number_of_velo_clusters = nullptr;
  unsigned* host_number_of_reconstructed_velo_tracks = nullptr;
  unsigned* host_accumulated_number_of_hits_in_velo_tracks = nullptr;
  unsigned host_max_number_of_events = 0;

  unsigned* host_event_list = nullptr;      // max_events
  unsigned* host_atomics_velo = nullptr;    // 2 * max_events + 1
  unsigned* host_velo_event_status = nullptr; // max_events
  unsigned* host_velo_track_hit_number = nullptr;
  char* host_velo_track_hits = nullptr;
  char* host_velo_states = nullptr;
  uint8_t* host_accepted_velo_tracks = nullptr;
  char* host_kalmanvelo_states = nullptr;
  char* host_reconstructed_multi_pvs = nullptr;   // max_events * 32 vertices
  unsigned* host_number_of_multivertex = nullptr;  // max_events
  size_t cap_track_hit_number = 0, cap_track_hits = 0, cap_states = 0, cap_accepted = 0, cap_kalman = 0;

  void reserve(unsigned max_number_of_events)
  {
    host_max_number_of_events = max_number_of_events;
    hipCheck(hipHostMalloc((void**) &host_number_of_selected_events, 4 * sizeof(unsigned), hipHostMallocDefault));
    host_total_number_of_velo_clusters = host_number_of_selected_events + 1;
    host_number_of_reconstructed_velo_tracks = host_number_of_selected_events + 2;
    host_accumulated_number_of_hits_in_velo_tracks = host_number_of_selected_events + 3;
    for (int i = 0; i < 4; ++i) host_number_of_selected_events[i] = 0;
    hipCheck(hipHostMalloc((void**) &host_event_list, max_number_of_events * sizeof(unsigned) + 4, hipHostMallocDefault));
    hipCheck(hipHostMalloc(
      (void**) &host_atomics_velo, (2 * (size_t) max_number_of_events + 1) * sizeof(unsigned), hipHostMallocDefault));
    hipCheck(
      hipHostMalloc((void**) &host_velo_event_status, max_number_of_events * sizeof(unsigned) + 4, hipHostMallocDefault));
    hipCheck(hipHostMalloc((void**) &host_reconstructed_multi_pvs,
                           (size_t) max_number_of_events * VELO_PV_MAX_VERTICES * sizeof(velo_pv_vertex) + 64,
                           hipHostMallocDefault));
    hipCheck(hipHostMalloc(
      (void**) &host_number_of_multivertex, max_number_of_events * sizeof(unsigned) + 4, hipHostMallocDefault));
  }

  static void grow(char** p, size_t* cap, size_t need)
  {
    if (need <= *cap) return;
    if (*p) hipCheck(hipHostFree(*p));
    const size_t n = need + need / 2 + 64;
    hipCheck(hipHostMalloc((void**) p, n, hipHostMallocDefault));
    *cap = n;
  }

  void ensure_outputs(size_t tracks, size_t hits)
  {
    grow(reinterpret_cast<char**>(&host_velo_track_hit_number), &cap_track_hit_number, (tracks + 1) * sizeof(unsigned));
    grow(&host_velo_track_hits, &cap_track_hits, hits * 16 + 16);
    grow(&host_velo_states, &cap_states, tracks * VELO_STATE_BYTES + 64);
    grow(reinterpret_cast<char**>(&host_accepted_velo_tracks), &cap_accepted, tracks + 64);
  }

  void ensure_kalman(size_t tracks) { grow(&host_kalmanvelo_states, &cap_kalman, tracks * VELO_STATE_BYTES + 64); }

  size_t velo_track_hit_number_size() const { return host_number_of_reconstructed_velo_tracks[0] + 1; }

  void release()
  {
    for (void* p : {(void*) host_number_of_selected_events, (void*) host_event_list, (void*) host_atomics_velo,
                    (void*) host_velo_event_status, (void*) host_velo_track_hit_number, (void*) host_velo_track_hits,
                    (void*) host_velo_states, (void*) host_accepted_velo_tracks, (void*) host_kalmanvelo_states,
                    (void*) host_reconstructed_multi_pvs, (void*) host_number_of_multivertex})
      if (p) (void) hipHostFree(p);
    host_number_of_selected_events = nullptr;
    host_event_list = nullptr;
    host_atomics_velo = nullptr;
    host_velo_event_status = nullptr;
    host_velo_track_hit_number = nullptr;
    host_velo_track_hits = nullptr;
    host_velo_states = nullptr;
    host_accepted_velo_tracks = nullptr;
    host_kalmanvelo_states = nullptr;
    host_reconstructed_multi_pvs = nullptr;
    host_number_of_multivertex = nullptr;
    cap_track_hit_number = cap_track_hits = cap_states = cap_accepted = cap_kalman = 0;
  }
};
