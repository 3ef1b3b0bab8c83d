"""ctypes binding of the C ABI in include/velo_hlt.h (libvelo_hlt.so).

The structures below mirror the header field for field.  Device pointers
are passed as integers (e.g. torch.Tensor.data_ptr()), streams as the raw
hipStream_t value (torch.cuda.Stream.cuda_stream) or None for the default
stream.  Every launcher call is checked; a non-zero hipError_t raises
VeloError.  There is no CPU fallback: if the library is missing, loading
raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libvelo_hlt.so"

N_MODULES = 52
N_SENSORS = 208
MAX_CANDIDATES_EVENT = 3000
MAX_NUMHITS_IN_MODULE = 500
MAX_HITS_PER_EVENT = 9500
MAX_TRACKS = 1200
MAX_TRACK_SIZE = 26
TTF_MODULO = 2000
MAX_WEAK_TRACKS = 500
NUM_ATOMICS = 5
GEOMETRY_BYTES = 19456
TRACK_HITS_BYTES = 54
TRACKLET_BYTES = 6
STATE_BYTES = 49
HIT_BYTES = 16
TTF_ENTRY_WORDS = 2

STATUS = {
    "SP_INDEX_TRUNCATED": 0x001, "CANDIDATES_OVERFLOW": 0x002, "MODULE_HITS_OVERFLOW": 0x004,
    "EVENT_HITS_OVERFLOW": 0x008, "TTF_OVERFLOW": 0x010, "TRACKLETS_OVERFLOW": 0x020,
    "WEAK_OVERFLOW": 0x040, "TRACKS_OVERFLOW": 0x080, "BAD_BANK": 0x100, "UNSORTED_BANK": 0x200,
    "BANK_ORDER_CHECKED": 0x80000000,  # not an error: estimate_input_size checked the event's bank order
}
STATUS_ERROR_MASK = 0x3FF

P = C.c_void_p
U32 = C.c_uint32


class VeloError(RuntimeError):
    pass


class Tuning(C.Structure):
    """velo_tuning: launch shape; grid_dim > 0 makes a per-event launch
    persistent (workgroups take events from dev_work_counter, in
    dev_work_order); dev_count gives a scan its element count on the device."""
    _fields_ = [("block_dim", U32), ("lds_tile", U32), ("grid_dim", U32), ("reserved", U32),
                ("dev_work_counter", P), ("dev_work_order", P), ("dev_count", P)]


class Constants(C.Structure):
    _fields_ = [("dev_velo_sp_patterns", P), ("dev_velo_sp_fx", P), ("dev_velo_sp_fy", P),
                ("dev_velo_candidate_ks", P), ("dev_velo_module_zs", P), ("dev_velo_geometry", P)]


def _args(name, fields):
    return type(name, (C.Structure,), {"_fields_": [(f, P) for f in fields] + [("n_events", U32), ("tuning", Tuning)]})


InitEventListArgs = type("InitEventListArgs", (C.Structure,),
                         {"_fields_": [("dev_event_list", P), ("n_events", U32), ("tuning", Tuning)]})
EstimateInputSizeArgs = _args("EstimateInputSizeArgs", [
    "dev_raw_input", "dev_raw_input_offsets", "dev_estimated_input_size", "dev_module_cluster_num",
    "dev_event_candidate_num", "dev_cluster_candidates", "dev_event_list", "dev_event_order",
    "dev_velo_candidate_ks", "dev_velo_event_status"])
PrefixSumArgs = type("PrefixSumArgs", (C.Structure,), {"_fields_": [
    ("dev_array", P), ("dev_auxiliary_array", P), ("array_size", U32), ("tuning", Tuning)]})
MaskedClusteringArgs = _args("MaskedClusteringArgs", [
    "dev_raw_input", "dev_raw_input_offsets", "dev_module_cluster_start", "dev_module_cluster_num",
    "dev_event_candidate_num", "dev_cluster_candidates", "dev_velo_cluster_container", "dev_event_list",
    "dev_event_order", "dev_velo_geometry", "dev_velo_sp_patterns", "dev_velo_sp_fx", "dev_velo_sp_fy",
    "dev_velo_event_status"])
CalculatePhiAndSortArgs = _args("CalculatePhiAndSortArgs", [
    "dev_module_cluster_start", "dev_module_cluster_num", "dev_velo_cluster_container", "dev_hit_permutation"])
FillCandidatesArgs = _args("FillCandidatesArgs", [
    "dev_velo_cluster_container", "dev_module_cluster_start", "dev_module_cluster_num", "dev_h0_candidates",
    "dev_h2_candidates"])
SearchByTripletArgs = _args("SearchByTripletArgs", [
    "dev_velo_cluster_container", "dev_module_cluster_start", "dev_module_cluster_num", "dev_tracks",
    "dev_tracklets", "dev_tracks_to_follow", "dev_weak_tracks", "dev_hit_used", "dev_atomics_velo",
    "dev_h0_candidates", "dev_h2_candidates", "dev_rel_indices", "dev_velo_module_zs", "dev_velo_event_status"])
WeakTracksAdderArgs = _args("WeakTracksAdderArgs", [
    "dev_velo_cluster_container", "dev_module_cluster_start", "dev_tracks", "dev_weak_tracks", "dev_hit_used",
    "dev_atomics_velo", "dev_velo_event_status"])
CopyAndPrefixSumArgs = type("CopyAndPrefixSumArgs", (C.Structure,), {"_fields_": [
    ("dev_total_sum", P), ("dev_input_array", P), ("dev_output_array", P), ("array_size", U32),
    ("tuning", Tuning)]})
CopyTrackHitNumberArgs = _args("CopyTrackHitNumberArgs", [
    "dev_tracks", "dev_atomics_storage", "dev_velo_track_hit_number"])
ConsolidateArgs = _args("ConsolidateArgs", [
    "dev_atomics_velo", "dev_tracks", "dev_velo_track_hit_number", "dev_velo_cluster_container",
    "dev_module_cluster_start", "dev_module_cluster_num", "dev_velo_track_hits", "dev_velo_states",
    "dev_accepted_velo_tracks", "dev_ghost_model", "dev_ghost_score"])

KalmanFitArgs = _args("KalmanFitArgs", [
    "dev_atomics_velo", "dev_velo_track_hit_number", "dev_velo_track_hits", "dev_velo_states",
    "dev_velo_kalman_beamline_states"])

EventOrderArgs = _args("EventOrderArgs", ["dev_raw_input_offsets", "dev_event_list", "dev_event_order"])
GlobalEventCutArgs = type("GlobalEventCutArgs", (C.Structure,), {"_fields_": [
    ("dev_ut_raw_input", P), ("dev_ut_raw_input_offsets", P), ("dev_scifi_raw_input", P),
    ("dev_scifi_raw_input_offsets", P), ("dev_number_of_selected_events", P), ("dev_event_list", P),
    ("ut_events_size", C.c_size_t), ("scifi_events_size", C.c_size_t), ("n_events", U32), ("tuning", Tuning)]})

PvExtrapolateArgs = _args("PvExtrapolateArgs", [
    "dev_velo_kalman_beamline_states", "dev_atomics_velo", "dev_velo_track_hit_number", "dev_pvtracks"])
PvHistoArgs = _args("PvHistoArgs", ["dev_atomics_velo", "dev_velo_track_hit_number", "dev_pvtracks", "dev_zhisto"])
PvPeakArgs = _args("PvPeakArgs", ["dev_zhisto", "dev_zpeaks", "dev_number_of_zpeaks"])
PvMultiFitterArgs = _args("PvMultiFitterArgs", [
    "dev_atomics_velo", "dev_velo_track_hit_number", "dev_pvtracks", "dev_zpeaks", "dev_number_of_zpeaks",
    "dev_multi_fit_vertices", "dev_number_of_multi_fit_vertices"])

LAUNCHERS = {
    "velo_init_event_list_launch": InitEventListArgs,
    "velo_global_event_cut_launch": GlobalEventCutArgs,
    "velo_event_order_launch": EventOrderArgs,
    "velo_estimate_input_size_launch": EstimateInputSizeArgs,
    "velo_prefix_sum_launch": PrefixSumArgs,
    "velo_masked_clustering_launch": MaskedClusteringArgs,
    "velo_calculate_phi_and_sort_launch": CalculatePhiAndSortArgs,
    "velo_fill_candidates_launch": FillCandidatesArgs,
    "velo_search_by_triplet_launch": SearchByTripletArgs,
    "velo_weak_tracks_adder_launch": WeakTracksAdderArgs,
    "velo_copy_and_prefix_sum_single_block_launch": CopyAndPrefixSumArgs,
    "velo_copy_velo_track_hit_number_launch": CopyTrackHitNumberArgs,
    "velo_consolidate_velo_tracks_launch": ConsolidateArgs,
    "velo_kalman_fit_launch": KalmanFitArgs,
    "velo_pv_beamline_extrapolate_launch": PvExtrapolateArgs,
    "velo_pv_beamline_histo_launch": PvHistoArgs,
    "velo_pv_beamline_peak_launch": PvPeakArgs,
    "velo_pv_beamline_multi_fitter_launch": PvMultiFitterArgs,
}

_lib = None

# C typedef name of every ctypes mirror in this module.
_C_NAMES = {
    Tuning: "velo_tuning", Constants: "velo_constants", InitEventListArgs: "velo_init_event_list_args",
    EventOrderArgs: "velo_event_order_args", GlobalEventCutArgs: "velo_global_event_cut_args",
    EstimateInputSizeArgs: "velo_estimate_input_size_args", PrefixSumArgs: "velo_prefix_sum_args",
    MaskedClusteringArgs: "velo_masked_clustering_args", CalculatePhiAndSortArgs: "velo_calculate_phi_and_sort_args",
    FillCandidatesArgs: "velo_fill_candidates_args", SearchByTripletArgs: "velo_search_by_triplet_args",
    WeakTracksAdderArgs: "velo_weak_tracks_adder_args",
    CopyAndPrefixSumArgs: "velo_copy_and_prefix_sum_single_block_args",
    CopyTrackHitNumberArgs: "velo_copy_velo_track_hit_number_args", KalmanFitArgs: "velo_kalman_fit_args",
    ConsolidateArgs: "velo_consolidate_velo_tracks_args", PvExtrapolateArgs: "velo_pv_beamline_extrapolate_args",
    PvHistoArgs: "velo_pv_beamline_histo_args", PvPeakArgs: "velo_pv_beamline_peak_args",
    PvMultiFitterArgs: "velo_pv_beamline_multi_fitter_args",
}


def check_mirrors(size_fn, mirrors: dict, path) -> None:
    """Refuse a library whose structs differ in size from these ctypes
    mirrors ({ctypes type: C typedef name}); size_fn(name) is the library's
    sizeof query.  A library built from a newer header would otherwise
    read or write past a mirror."""
    bad = []
    for st, name in mirrors.items():
        n = int(size_fn(name.encode()))
        if n != C.sizeof(st):
            bad.append(f"{name}: {n} bytes in the library, {C.sizeof(st)} in {st.__name__}")
    if bad:
        raise VeloError(f"{path}: struct layout differs from the binding (rebuild it): " + "; ".join(bad))


def lib():
    """Load libvelo_hlt.so (raises when it is not built: no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise VeloError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # One HIP runtime per process: PyTorch ships its own libamdhip64 and
        # loads it by file name, so it has to be loaded before this library
        # (whose libamdhip64.so.7 dependency then resolves to the same copy).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, st in LAUNCHERS.items():
            f = getattr(L, name)
            f.argtypes = [C.POINTER(st), P]
            f.restype = C.c_int
        L.velo_prefix_sum_noalloc_launch.argtypes = [P, U32, P, P]
        L.velo_prefix_sum_noalloc_launch.restype = C.c_int
        L.velo_prefix_sum_aux_size.argtypes = [U32]
        L.velo_prefix_sum_aux_size.restype = U32
        L.velo_constants_init.argtypes = [C.c_char_p, C.c_size_t, P, C.POINTER(Constants)]
        L.velo_constants_init.restype = C.c_int
        L.velo_constants_free.argtypes = [C.POINTER(Constants)]
        L.velo_constants_free.restype = C.c_int
        L.velo_sp_tables.argtypes = [P, P, P, P]
        L.velo_hlt_version.restype = C.c_char_p
        L.velo_search_by_triplet_resident_workgroups.argtypes = [U32]
        L.velo_search_by_triplet_resident_workgroups.restype = U32
        L.velo_pv_exp.argtypes = [C.c_float]
        L.velo_pv_exp.restype = C.c_float
        L.velo_error_string.argtypes = [C.c_int]
        L.velo_error_string.restype = C.c_char_p
        L.velo_abi_struct_size.argtypes = [C.c_char_p]
        L.velo_abi_struct_size.restype = C.c_size_t
        check_mirrors(L.velo_abi_struct_size, _C_NAMES, LIB_PATH)
        _lib = L
    return _lib


def check(code: int, what: str) -> None:
    if code != 0:
        msg = lib().velo_error_string(code).decode()
        raise VeloError(f"{what} failed: hipError {code} ({msg})")


def launch(name: str, args, stream=None) -> None:
    """Call one C-ABI launcher with a filled argument structure."""
    check(getattr(lib(), name)(C.byref(args), stream), name)


_hip = None


def d2h(device_ptr: int, nbytes: int):
    """Copy nbytes from a device pointer to a new numpy uint8 array."""
    import numpy as np
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [P, P, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    out = np.empty(max(int(nbytes), 1), np.uint8)
    check(_hip.hipMemcpy(out.ctypes.data, device_ptr, int(nbytes), 2), "hipMemcpy")  # 2 = DeviceToHost
    return out[: int(nbytes)]


def prefix_sum_aux_size(n: int) -> int:
    return int(lib().velo_prefix_sum_aux_size(n))


def sp_tables():
    import numpy as np
    p = np.zeros(256, np.uint8); s = np.zeros(256, np.uint8)
    fx = np.zeros(512, np.float32); fy = np.zeros(512, np.float32)
    lib().velo_sp_tables(p.ctypes.data, s.ctypes.data, fx.ctypes.data, fy.ctypes.data)
    return p, s, fx, fy


class DeviceConstants:
    """Device tables (SP patterns, candidate ks, module z, geometry)."""

    def __init__(self, geometry: bytes, module_zs=None):
        import numpy as np
        self.c = Constants()
        zs = None
        if module_zs is not None:
            self._zs = np.ascontiguousarray(module_zs, np.float32)
            zs = self._zs.ctypes.data
        check(lib().velo_constants_init(geometry, len(geometry), zs, C.byref(self.c)), "velo_constants_init")

    def free(self):
        if self.c.dev_velo_geometry:
            lib().velo_constants_free(C.byref(self.c))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
