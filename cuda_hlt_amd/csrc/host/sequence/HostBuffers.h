// Pinned host buffers of one stream (HostBuffers.cuh / HostBuffers.cu:5-83,
// VELO part): the counters the visitors size buffers from, and the
// consolidated output copied back at the end of the sequence.  Output
// buffers grow on demand.
//
// Every pinned buffer carries its usable capacity and a 64-byte guard of a
// fixed pattern behind it.  All device-to-host copies of the sequence go
// through copy_to_host, which refuses a copy larger than its destination,
// and the Stream checks every guard after each run, so a copy that ever
// wrote past a buffer is reported as an error instead of corrupting memory
// silently.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

#include "../gear/Common.h"
#include "velo_pv.h"

struct HostBuffers {
  static constexpr size_t guard_bytes = 64;
  static constexpr uint8_t guard_value = 0xA5;

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

  // usable bytes of each buffer above (same order)
  size_t cap_counters = 0, cap_event_list = 0, cap_atomics = 0, cap_status = 0, cap_track_hit_number = 0,
         cap_track_hits = 0, cap_states = 0, cap_accepted = 0, cap_kalman = 0, cap_multi_pvs = 0,
         cap_multivertex = 0;

  static void alloc(void** p, size_t* cap, size_t bytes)
  {
    hipCheck(hipHostMalloc(p, bytes + guard_bytes, hipHostMallocDefault));
    std::memset(static_cast<char*>(*p) + bytes, guard_value, guard_bytes);
    *cap = bytes;
  }

  template<typename T>
  static void alloc(T** p, size_t* cap, size_t bytes)
  {
    void* q = nullptr;
    alloc(&q, cap, bytes);
    *p = static_cast<T*>(q);
  }

  void reserve(unsigned max_number_of_events)
  {
    host_max_number_of_events = max_number_of_events;
    const size_t n = max_number_of_events;
    alloc(&host_number_of_selected_events, &cap_counters, 4 * sizeof(unsigned));
    host_total_number_of_velo_clusters = host_number_of_selected_events + 1;
    host_number_of_reconstructed_velo_tracks = host_number_of_selected_events + 2;
    host_accumulated_number_of_hits_in_velo_tracks = host_number_of_selected_events + 3;
    for (int i = 0; i < 4; ++i) host_number_of_selected_events[i] = 0;
    alloc(&host_event_list, &cap_event_list, n * sizeof(unsigned) + 4);
    alloc(&host_atomics_velo, &cap_atomics, (2 * n + 1) * sizeof(unsigned));
    alloc(&host_velo_event_status, &cap_status, n * sizeof(unsigned) + 4);
    alloc(&host_reconstructed_multi_pvs, &cap_multi_pvs, n * VELO_PV_MAX_VERTICES * sizeof(velo_pv_vertex) + 64);
    alloc(&host_number_of_multivertex, &cap_multivertex, n * sizeof(unsigned) + 4);
  }

  template<typename T>
  static void grow(T** p, size_t* cap, size_t need)
  {
    if (need <= *cap && *p) return;
    if (*p) hipCheck(hipHostFree(*p));
    *p = nullptr;
    alloc(p, cap, need + need / 2 + 64);
  }

  void ensure_outputs(size_t tracks, size_t hits)
  {
    grow(&host_velo_track_hit_number, &cap_track_hit_number, (tracks + 1) * sizeof(unsigned));
    grow(&host_velo_track_hits, &cap_track_hits, hits * 16 + 16);
    grow(&host_velo_states, &cap_states, tracks * VELO_STATE_BYTES + 64);
    grow(&host_accepted_velo_tracks, &cap_accepted, tracks + 64);
  }

  void ensure_kalman(size_t tracks) { grow(&host_kalmanvelo_states, &cap_kalman, tracks * VELO_STATE_BYTES + 64); }

  size_t velo_track_hit_number_size() const { return host_number_of_reconstructed_velo_tracks[0] + 1; }

  // Capacity of the pinned buffer that holds [dst, dst + 1), 0 if none does.
  size_t room_at(const void* dst) const
  {
    const char* d = static_cast<const char*>(dst);
    for (const auto& b : buffers()) {
      const char* base = static_cast<const char*>(b.first);
      if (base && d >= base && d < base + b.second) return b.second - static_cast<size_t>(d - base);
    }
    return 0;
  }

  // Asynchronous device-to-host copy into one of these buffers; throws when
  // the copy would not fit its destination.
  void copy_to_host(void* dst, const void* src, size_t bytes, hipStream_t stream, const char* what) const
  {
    if (bytes == 0) return;
    const size_t room = room_at(dst);
    if (bytes > room)
      throw std::invalid_argument(std::string("device-to-host copy of ") + what + ": " + std::to_string(bytes) +
                                  " bytes into a pinned buffer with " + std::to_string(room) + " bytes left");
    hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
  }

  // Throws if any guard behind a buffer changed.
  void check_guards() const
  {
    static const char* names[] = {"counters", "event list", "atomics", "status", "track hit number", "track hits",
                                  "states", "accepted", "kalman states", "multi-fit vertices", "vertex counts"};
    int i = 0;
    for (const auto& b : buffers()) {
      if (b.first) {
        const uint8_t* g = static_cast<const uint8_t*>(b.first) + b.second;
        for (size_t k = 0; k < guard_bytes; ++k)
          if (g[k] != guard_value)
            throw std::runtime_error(std::string("pinned host buffer '") + names[i] + "' was written past its " +
                                     std::to_string(b.second) + " bytes");
      }
      ++i;
    }
  }

  void release()
  {
    for (const auto& b : buffers())
      if (b.first) (void) hipHostFree(const_cast<void*>(b.first));
    host_number_of_selected_events = nullptr;
    host_total_number_of_velo_clusters = nullptr;
    host_number_of_reconstructed_velo_tracks = nullptr;
    host_accumulated_number_of_hits_in_velo_tracks = nullptr;
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
    cap_counters = cap_event_list = cap_atomics = cap_status = cap_track_hit_number = cap_track_hits = cap_states =
      cap_accepted = cap_kalman = cap_multi_pvs = cap_multivertex = 0;
  }

private:
  struct BufferList {
    std::pair<const void*, size_t> v[11];
    const std::pair<const void*, size_t>* begin() const { return v; }
    const std::pair<const void*, size_t>* end() const { return v + 11; }
  };
  BufferList buffers() const
  {
    return BufferList {{{host_number_of_selected_events, cap_counters},
                        {host_event_list, cap_event_list},
                        {host_atomics_velo, cap_atomics},
                        {host_velo_event_status, cap_status},
                        {host_velo_track_hit_number, cap_track_hit_number},
                        {host_velo_track_hits, cap_track_hits},
                        {host_velo_states, cap_states},
                        {host_accepted_velo_tracks, cap_accepted},
                        {host_kalmanvelo_states, cap_kalman},
                        {host_reconstructed_multi_pvs, cap_multi_pvs},
                        {host_number_of_multivertex, cap_multivertex}}};
  }
};
