/*
 * velo_sequence.h -- C ABI of the VELO sequence (host side).
 *
 * Wraps one Stream of the C++ sequence layer: Stream::initialize and
 * Stream::run_sequence (Stream.cu:16-83) over the configured Velo sequence
 * (configuration/sequences/Velo.h:5-18 without the global event cut), with
 * Constants::reserve_and_initialize (Constants.cu:123-154) done at creation.
 * Errors are reported as a non-zero return and a message; no exception
 * crosses the ABI.
 */
#ifndef VELO_SEQUENCE_H
#define VELO_SEQUENCE_H

#include <stddef.h>
#include <stdint.h>

#include "velo_checker.h"
#include "velo_hlt.h"
#include "velo_pv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct velo_sequence velo_sequence;

typedef struct {
  uint32_t max_events;      /* events per run (host buffers are sized for it) */
  size_t reserve_mb;        /* device arena, MiB (the reference's -m) */
  int print_memory_manager; /* print the arena after every step (-p) */
} velo_sequence_config;

typedef struct {
  /* host input (copied to the device by init_event_list) ... */
  const char* host_events;
  const uint32_t* host_offsets; /* n_events+1 byte offsets */
  size_t events_size;           /* bytes */
  /* ... or device-resident input (used in place when non-NULL) */
  const char* dev_events;
  const uint32_t* dev_offsets;
  uint32_t n_events;
  uint32_t repetitions;        /* 0 = 1 */
  int copy_outputs_to_host;    /* copy the consolidated tracks back */
  velo_tuning tuning[6];       /* estimate, clustering, sort, fill, search, consolidate */
  /* Batched run (no reference counterpart): no host reads between steps,
   * buffers sized by bounds (clusters <= events_size / 2, tracks <= 1200 and
   * track hits <= 1200 * 26 per event), the sequence captured once in a
   * hipGraph and replayed while the input pointers, n_events, events_size,
   * dev_work_order and tuning stay the same.  Needs device-resident input.
   * Outputs are identical to the plain run's. */
  int batched;
  const uint32_t* dev_work_order; /* persistent steps: event order (largest first), NULL = identity */
  /* Ghost filter run inline by consolidation (device velo_ghost_model,
   * velo_ghost.h), NULL = every track accepted.  No reference counterpart. */
  const void* dev_ghost_model;
  /* Also run velo_kalman_fit after consolidation (KalmanFilterVisitor.cu:4-44). */
  int run_velo_kalman_fit;
  /* Run global_event_cut (GlobalEventCut.cu:3-55) on UT and SciFi banks in
   * the raw-event layout (velo_mdf.h), from host memory (copied by
   * init_event_list) or device memory (used in place when non-NULL); the
   * VELO steps then run on the kept events only, in increasing order
   * (output host_event_list / dev_event_list map each slot to its input
   * event).  Plain runs only (batched = 0). */
  int run_global_event_cut;
  const char* host_ut_events;
  const uint32_t* host_ut_offsets; /* n_events+1 */
  size_t ut_events_size;
  const char* host_scifi_events;
  const uint32_t* host_scifi_offsets;
  size_t scifi_events_size;
  const char* dev_ut_events;
  const uint32_t* dev_ut_offsets;
  const char* dev_scifi_events;
  const uint32_t* dev_scifi_offsets;
  /* Also run the four pv_beamline steps after velo_kalman_fit
   * (DefaultSequence.h:19-23; velo_pv.h); implies run_velo_kalman_fit. */
  int run_pv_beamline;
} velo_sequence_input;

typedef struct {
  uint32_t n_events;
  uint32_t total_estimated_hits;
  uint32_t total_tracks;
  uint32_t total_track_hits;
  /* host copies (valid when copy_outputs_to_host was set) */
  const uint32_t* host_atomics_velo;       /* 2N+1 */
  const uint32_t* host_track_hit_number;   /* T+1 */
  const char* host_track_hits;             /* 16*Ht: SoA x, y, z, LHCbID */
  const char* host_states;                 /* 49*T: SoA 12 floats + bool */
  const uint8_t* host_accepted;            /* T: dev_accepted_velo_tracks */
  const char* host_kalman_states;          /* 49*T when velo_kalman_fit ran */
  /* device outputs (valid until the next run) */
  const int32_t* dev_atomics_velo;
  const uint32_t* dev_track_hit_number;
  const char* dev_track_hits;
  const char* dev_states;
  const uint8_t* dev_accepted;
  const char* dev_kalman_states;
  const uint32_t* dev_event_status;
  const uint32_t* dev_module_cluster_start;
  const uint32_t* dev_module_cluster_num;
  const uint32_t* dev_cluster_container;
  size_t memory_required; /* arena high-water mark, bytes */
  /* input event of each output slot (n_events entries; identity without
   * the global event cut) */
  const uint32_t* host_event_list;
  const uint32_t* dev_event_list;
  /* primary vertices when the pv_beamline steps ran: 32 velo_pv_vertex
   * slots and a count per event */
  const velo_pv_vertex* host_pv_vertices;
  const uint32_t* host_pv_counts;
  const velo_pv_vertex* dev_pv_vertices;
  const uint32_t* dev_pv_counts;
} velo_sequence_output;

velo_sequence* velo_sequence_create(
  const velo_sequence_config* cfg, const void* geometry, size_t geometry_bytes, const float* module_zs, char* err,
  size_t err_len);
void velo_sequence_destroy(velo_sequence* s);
int velo_sequence_run(velo_sequence* s, const velo_sequence_input* in, char* err, size_t err_len);
int velo_sequence_output_get(const velo_sequence* s, velo_sequence_output* out);
/* sizeof of this header's structs as the library was built (0: config,
 * 1: input, 2: output, else 0), so that a binding can refuse a library
 * whose structs differ from its own. */
size_t velo_sequence_struct_size(int which);
void* velo_sequence_stream(const velo_sequence* s); /* the hipStream_t of the sequence */

/* Per-step timing (no reference counterpart): when enabled, HIP events are
 * recorded on the sequence's stream before every step; after a run,
 * velo_sequence_step_times writes up to max_steps durations in ms and the
 * step names, and returns the number of steps. */
int velo_sequence_enable_step_timing(velo_sequence* s, int enable);
int velo_sequence_step_times(const velo_sequence* s, float* ms, const char** names, int max_steps);

/* Device arena bytes (velo_sequence_config.reserve_mb covers it) that a
 * batched run of n_events events holding events_bytes raw bytes needs: the
 * scheduler lays the sequence out on the batched run's bounds (the
 * set_arguments_size of every step, SequenceVisitor.cuh:11-22, and the
 * memory manager's first-fit placement, MemoryManager.cuh:8-170) without
 * launching anything; the result is its high-water mark.  A plain run of
 * the same events needs less (it sizes by totals read between steps).  The
 * reference has no such query: its arena is a fixed command-line
 * reservation.  No device call, no allocation. */
size_t velo_workspace_bytes(uint32_t n_events, size_t events_bytes, int run_velo_kalman_fit, int run_pv_beamline);

/* The per-event status words (VELO_STATUS_* of velo_hlt.h) of the last
 * run: up to max_events of them copied to host_status (may be NULL), and
 * the OR of the error bits (VELO_STATUS_ERROR_MASK) over all events to
 * *combined (may be NULL).  Returns the number of events of the last run,
 * or -1.  Waits for the sequence's stream.  No reference counterpart: the
 * reference only asserts on its capacities in debug builds. */
int velo_status(const velo_sequence* s, uint32_t* host_status, uint32_t max_events, uint32_t* combined);

/* Checks the last run's VELO tracks (copy_outputs_to_host set) against MC
 * truth: every step's SequenceVisitor::check in sequence order
 * (Stream::run_monte_carlo_test, Stream.cu:85-105; the VELO checker,
 * VeloSequenceCheckers_impl.cuh:8-24).  The truth covers input events
 * start_event, start_event + 1, ...; output slots map to input events
 * through the event list (the global event cut may have dropped some).
 * Returns 0, or -1 with a message in err. */
int velo_sequence_check(velo_sequence* s, const velo_mc_truth* truth, uint32_t start_event,
                        velo_checker_result* result, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif
#endif
