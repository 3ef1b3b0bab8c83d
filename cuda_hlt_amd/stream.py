"""ctypes binding of include/velo_sequence.h: the C++ VELO sequence
(Stream::run_sequence over the configured Velo sequence)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as nat

P = C.c_void_p


class SequenceConfig(C.Structure):
    _fields_ = [("max_events", C.c_uint32), ("reserve_mb", C.c_size_t), ("print_memory_manager", C.c_int)]


class SequenceInput(C.Structure):
    _fields_ = [("host_events", P), ("host_offsets", P), ("events_size", C.c_size_t), ("dev_events", P),
                ("dev_offsets", P), ("n_events", C.c_uint32), ("repetitions", C.c_uint32),
                ("copy_outputs_to_host", C.c_int), ("tuning", nat.Tuning * 6), ("batched", C.c_int),
                ("dev_work_order", P), ("dev_ghost_model", P), ("run_velo_kalman_fit", C.c_int),
                ("run_global_event_cut", C.c_int), ("host_ut_events", P), ("host_ut_offsets", P),
                ("ut_events_size", C.c_size_t), ("host_scifi_events", P), ("host_scifi_offsets", P),
                ("scifi_events_size", C.c_size_t), ("dev_ut_events", P), ("dev_ut_offsets", P),
                ("dev_scifi_events", P), ("dev_scifi_offsets", P), ("run_pv_beamline", C.c_int)]


class SequenceOutput(C.Structure):
    _fields_ = [("n_events", C.c_uint32), ("total_estimated_hits", C.c_uint32), ("total_tracks", C.c_uint32),
                ("total_track_hits", C.c_uint32), ("host_atomics_velo", P), ("host_track_hit_number", P),
                ("host_track_hits", P), ("host_states", P), ("host_accepted", P), ("host_kalman_states", P),
                ("dev_atomics_velo", P), ("dev_track_hit_number", P),
                ("dev_track_hits", P), ("dev_states", P), ("dev_accepted", P), ("dev_kalman_states", P),
                ("dev_event_status", P),
                ("dev_module_cluster_start", P), ("dev_module_cluster_num", P), ("dev_cluster_container", P),
                ("memory_required", C.c_size_t), ("host_event_list", P), ("dev_event_list", P),
                ("host_pv_vertices", P), ("host_pv_counts", P), ("dev_pv_vertices", P), ("dev_pv_counts", P)]


PV_VERTEX_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("chi2", "<f4"), ("ndof", "<i4"),
                            ("n_tracks", "<u4"), ("cov00", "<f4"), ("cov10", "<f4"), ("cov11", "<f4"),
                            ("cov20", "<f4"), ("cov21", "<f4"), ("cov22", "<f4"), ("nTracks", "<i4")])

TUNING_SLOTS = ("estimate_input_size", "masked_clustering", "calculate_phi_and_sort", "fill_candidates",
                "search_by_triplet", "consolidate")

_bound = False


def _lib():
    global _bound
    L = nat.lib()
    if not _bound:
        L.velo_sequence_create.argtypes = [C.POINTER(SequenceConfig), C.c_char_p, C.c_size_t, P, C.c_char_p, C.c_size_t]
        L.velo_sequence_create.restype = P
        L.velo_sequence_destroy.argtypes = [P]
        L.velo_sequence_run.argtypes = [P, C.POINTER(SequenceInput), C.c_char_p, C.c_size_t]
        L.velo_sequence_run.restype = C.c_int
        L.velo_sequence_output_get.argtypes = [P, C.POINTER(SequenceOutput)]
        L.velo_sequence_output_get.restype = C.c_int
        L.velo_sequence_stream.argtypes = [P]
        L.velo_sequence_stream.restype = P
        L.velo_sequence_enable_step_timing.argtypes = [P, C.c_int]
        L.velo_sequence_enable_step_timing.restype = C.c_int
        L.velo_sequence_step_times.argtypes = [P, C.POINTER(C.c_float), C.POINTER(C.c_char_p), C.c_int]
        L.velo_sequence_step_times.restype = C.c_int
        L.velo_workspace_bytes.argtypes = [C.c_uint32, C.c_size_t, C.c_int, C.c_int]
        L.velo_workspace_bytes.restype = C.c_size_t
        L.velo_status.argtypes = [P, P, C.c_uint32, C.POINTER(C.c_uint32)]
        L.velo_status.restype = C.c_int
        L.velo_sequence_check.argtypes = [P, P, C.c_uint32, P, C.c_char_p, C.c_size_t]
        L.velo_sequence_check.restype = C.c_int
        L.velo_sequence_struct_size.argtypes = [C.c_int]
        L.velo_sequence_struct_size.restype = C.c_size_t
        nat.check_mirrors(L.velo_abi_struct_size, {
            SequenceConfig: "velo_sequence_config", SequenceInput: "velo_sequence_input",
            SequenceOutput: "velo_sequence_output", nat.Tuning: "velo_tuning"}, nat.LIB_PATH)
        _bound = True
    return L


def workspace_bytes(n_events: int, events_bytes: int, velo_kalman_fit=False, pv_beamline=False) -> int:
    """Arena bytes of a batched run (velo_workspace_bytes); no GPU needed."""
    return int(_lib().velo_workspace_bytes(n_events, events_bytes, int(velo_kalman_fit), int(pv_beamline)))


@dataclass
class HostTracks:
    """Consolidated output copied to the host (reference layouts)."""
    n_events: int
    atomics: np.ndarray       # uint32[2N+1]
    hit_number: np.ndarray    # uint32[T+1]
    track_hits: np.ndarray    # uint8[16 Ht]
    states: np.ndarray        # uint8[49 T]
    accepted: np.ndarray      # uint8[T]: dev_accepted_velo_tracks
    kalman: np.ndarray | None = None  # uint8[49 T]: velo_kalman_fit states, when run
    event_list: np.ndarray | None = None  # uint32[N]: input event of each slot
    pv_vertices: np.ndarray | None = None  # PV_VERTEX_DTYPE[N, 32], when the pv_beamline steps ran
    pv_counts: np.ndarray | None = None    # uint32[N]


class VeloSequence:
    """One C++ Stream of the VELO sequence with its own device arena."""

    def __init__(self, geometry: bytes, module_zs=None, max_events=10000, reserve_mb=4096, print_memory=False):
        self._zs = None if module_zs is None else np.ascontiguousarray(module_zs, np.float32)
        err = C.create_string_buffer(512)
        cfg = SequenceConfig(max_events, reserve_mb, int(print_memory))
        self.h = _lib().velo_sequence_create(C.byref(cfg), geometry, len(geometry),
                                             None if self._zs is None else self._zs.ctypes.data, err, 512)
        if not self.h:
            raise nat.VeloError(f"velo_sequence_create: {err.value.decode()}")
        self.max_events = max_events

    def close(self):
        if getattr(self, "h", None):
            _lib().velo_sequence_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return _lib().velo_sequence_stream(self.h)

    def run(self, n_events, host_events=None, host_offsets=None, dev_events=None, dev_offsets=None,
            copy_outputs=True, repetitions=1, tuning=None, events_size=0, batched=False,
            dev_work_order=None, dev_ghost_model=None, velo_kalman_fit=False, gec=None,
            pv_beamline=False) -> SequenceOutput:
        """Run the sequence on host input (copied by init_event_list) or on
        device-resident input (dev_events/dev_offsets; events_size = their
        bytes, needed by batched runs to bound the buffers).  batched=True
        captures the sequence in a hipGraph on the first call and replays it
        while inputs, sizes and tuning stay the same.  tuning maps a step
        name to (block_dim, lds_tile) or (block_dim, lds_tile, grid_dim);
        grid_dim > 0 runs that step persistent in dev_work_order.
        dev_ghost_model (a device velo_ghost_model) makes consolidation
        score every track and clear the accepted flag of likely ghosts.
        gec runs the global event cut first: host arrays (ut_events,
        ut_offsets, scifi_events, scifi_offsets), e.g. a gen.GecBanks, or
        a dict of device pointers with the same keys plus ut_events_size and
        scifi_events_size; the output then covers the kept events, listed
        in HostTracks.event_list.  pv_beamline also runs velo_kalman_fit and
        the four beamline PV steps (HostTracks.pv_vertices / pv_counts)."""
        inp = SequenceInput()
        if dev_events is not None:
            inp.dev_events, inp.dev_offsets = int(dev_events), int(dev_offsets)
            inp.events_size = int(events_size)
        else:
            self._keep = (np.ascontiguousarray(host_events), np.ascontiguousarray(host_offsets, np.uint32))
            inp.host_events = self._keep[0].ctypes.data
            inp.host_offsets = self._keep[1].ctypes.data
            inp.events_size = self._keep[0].nbytes
        inp.n_events = n_events
        inp.repetitions = repetitions
        inp.copy_outputs_to_host = int(copy_outputs)
        for k, shape in (tuning or {}).items():
            inp.tuning[TUNING_SLOTS.index(k)] = nat.Tuning(*shape)
        inp.batched = int(batched)
        inp.dev_work_order = None if dev_work_order is None else int(dev_work_order)
        inp.dev_ghost_model = None if dev_ghost_model is None else int(dev_ghost_model)
        inp.run_velo_kalman_fit = int(velo_kalman_fit)
        inp.run_pv_beamline = int(pv_beamline)
        if gec is not None:
            inp.run_global_event_cut = 1
            if isinstance(gec, dict):
                inp.dev_ut_events, inp.dev_ut_offsets = int(gec["ut_events"]), int(gec["ut_offsets"])
                inp.dev_scifi_events, inp.dev_scifi_offsets = int(gec["scifi_events"]), int(gec["scifi_offsets"])
                inp.ut_events_size, inp.scifi_events_size = int(gec["ut_events_size"]), int(gec["scifi_events_size"])
            else:
                self._keep_gec = [np.ascontiguousarray(a) for a in (gec.ut_events, gec.scifi_events)] + \
                    [np.ascontiguousarray(a, np.uint32) for a in (gec.ut_offsets, gec.scifi_offsets)]
                ue, se, uo, so = self._keep_gec
                inp.host_ut_events, inp.host_ut_offsets, inp.ut_events_size = ue.ctypes.data, uo.ctypes.data, ue.nbytes
                inp.host_scifi_events, inp.host_scifi_offsets = se.ctypes.data, so.ctypes.data
                inp.scifi_events_size = se.nbytes
        err = C.create_string_buffer(1024)
        rc = _lib().velo_sequence_run(self.h, C.byref(inp), err, 1024)
        if rc != 0:
            raise nat.VeloError(f"velo_sequence_run: {err.value.decode()}")
        return self.output()

    def status(self):
        """(per-event status words of the last run, OR of their error bits)."""
        n = int(_lib().velo_status(self.h, None, 0, None))
        if n < 0:
            raise nat.VeloError("velo_status failed")
        out = np.zeros(max(n, 1), np.uint32)
        combined = C.c_uint32(0)
        if _lib().velo_status(self.h, out.ctypes.data, n, C.byref(combined)) != n:
            raise nat.VeloError("velo_status failed")
        return out[:n], int(combined.value)

    def check(self, mcps, mc_hits, start_event=0):
        """The last run's tracks against MC truth (MCP_DTYPE particles and
        their LHCbIDs, events numbered from start_event) through the
        sequence's checkers; returns a checker.CheckerResult."""
        from . import checker
        from .gen import MCP_DTYPE
        m = np.ascontiguousarray(mcps, MCP_DTYPE)
        h = np.ascontiguousarray(mc_hits, np.uint32)
        truth = checker._McTruth(m.ctypes.data if len(m) else None, len(m), h.ctypes.data if len(h) else None,
                                 len(h), int(m["event"].max()) + 1 if len(m) else 0)
        res = checker._Result()
        err = C.create_string_buffer(512)
        checker._lib()
        if _lib().velo_sequence_check(self.h, C.byref(truth), start_event, C.byref(res), err, 512) != 0:
            raise nat.VeloError(err.value.decode())
        return checker.result_of(res)

    def enable_step_timing(self, enable=True):
        nat.check(_lib().velo_sequence_enable_step_timing(self.h, int(enable)), "enable_step_timing")

    def step_times(self):
        """[(step name, ms)] of the last run (step timing must be enabled)."""
        ms = (C.c_float * 64)()
        names = (C.c_char_p * 64)()
        n = _lib().velo_sequence_step_times(self.h, ms, names, 64)
        return [(names[i].decode(), float(ms[i])) for i in range(max(n, 0))]

    def output(self) -> SequenceOutput:
        out = SequenceOutput()
        nat.check(_lib().velo_sequence_output_get(self.h, C.byref(out)), "velo_sequence_output_get")
        return out

    def host_tracks(self) -> HostTracks:
        o = self.output()
        if not o.host_atomics_velo:
            raise nat.VeloError("outputs were not copied to the host")
        n, T, Ht = o.n_events, o.total_tracks, o.total_track_hits

        def arr(ptr, count, dtype):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), (max(count, 1) * np.dtype(dtype).itemsize,)
                                         ).view(dtype)[:count].copy()
        return HostTracks(n, arr(o.host_atomics_velo, 2 * n + 1, np.uint32),
                          arr(o.host_track_hit_number, T + 1, np.uint32),
                          arr(o.host_track_hits, 16 * Ht, np.uint8), arr(o.host_states, 49 * T, np.uint8),
                          arr(o.host_accepted, T, np.uint8),
                          arr(o.host_kalman_states, 49 * T, np.uint8) if o.host_kalman_states else None,
                          arr(o.host_event_list, n, np.uint32),
                          arr(o.host_pv_vertices, 32 * n, PV_VERTEX_DTYPE).reshape(n, 32) if o.host_pv_vertices
                          else None,
                          arr(o.host_pv_counts, n, np.uint32) if o.host_pv_counts else None)
