"""ctypes binding of the CPU oracle (oracle/velo_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the bench's
CPU-baseline leg as the checker.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libvelo_oracle.so"
REF_SP_PATH = HERE / "_ref" / "libref_sp.so"

N_MODULES = 52
MAX_CANDIDATES_EVENT = 3000
MAX_TRACKS = 1200
MAX_WEAK_TRACKS = 500
NUM_ATOMICS = 5

TRACK_DTYPE = np.dtype([("hitsNum", "<u2"), ("hits", "<u2", (26,))])  # 54 B, VeloEventModel.cuh:58-69
TRACKLET_DTYPE = np.dtype([("hits", "<u2", (3,))])                   # 6 B

STATUS_BITS = {
    "SP_INDEX_TRUNCATED": 0x001, "CANDIDATES_OVERFLOW": 0x002, "MODULE_HITS_OVERFLOW": 0x004,
    "EVENT_HITS_OVERFLOW": 0x008, "TTF_OVERFLOW": 0x010, "TRACKLETS_OVERFLOW": 0x020,
    "WEAK_OVERFLOW": 0x040, "TRACKS_OVERFLOW": 0x080, "BAD_BANK": 0x100, "UNSORTED_BANK": 0x200,
    "BANK_ORDER_CHECKED": 0x80000000,  # not an error: estimate_input_size checked the event's bank order
}
STATUS_ERROR_MASK = 0x3FF

_lib = None
_P = C.c_void_p


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        L.oracle_atan2f.argtypes = [C.c_float, C.c_float]
        L.oracle_atan2f.restype = C.c_float
        L.oracle_phi_odd.argtypes = [C.c_float, C.c_float]
        L.oracle_phi_odd.restype = C.c_float
        L.oracle_phi_even.argtypes = [C.c_float, C.c_float]
        L.oracle_phi_even.restype = C.c_float
        L.oracle_sp_lut.argtypes = [_P] * 4
        L.oracle_estimate_input_size.argtypes = [_P, _P, C.c_uint32, _P, _P, _P, _P, _P]
        L.oracle_prefix_sum.argtypes = [_P, C.c_uint32]
        L.oracle_masked_clustering.argtypes = [_P, _P, C.c_uint32, _P, _P, _P, _P, _P]
        L.oracle_calculate_phi_and_sort.argtypes = [C.c_uint32, _P, _P, _P, _P, _P]
        L.oracle_fill_candidates.argtypes = [C.c_uint32, _P, _P, _P, _P, _P]
        L.oracle_search_by_triplet.argtypes = [C.c_uint32] + [_P] * 11
        L.oracle_weak_tracks_adder.argtypes = [C.c_uint32] + [_P] * 7
        L.oracle_copy_and_prefix_sum.argtypes = [C.c_uint32, _P]
        L.oracle_copy_velo_track_hit_number.argtypes = [C.c_uint32, _P, _P, _P]
        L.oracle_consolidate_velo_tracks.argtypes = [C.c_uint32] + [_P] * 7
        L.oracle_ccl_event.argtypes = [_P, _P, _P, _P, _P, _P, C.c_uint32]
        L.oracle_velo_kalman_fit.argtypes = [C.c_uint32] + [_P] * 5
        L.oracle_global_event_cut.argtypes = [C.c_uint32] + [_P] * 5
        L.oracle_global_event_cut.restype = C.c_uint32
        L.oracle_pv_exp.argtypes = [C.c_float]
        L.oracle_pv_exp.restype = C.c_float
        L.oracle_pv_gauss_integral.argtypes = [C.c_float]
        L.oracle_pv_gauss_integral.restype = C.c_float
        L.oracle_pv_beamline_extrapolate.argtypes = [C.c_uint32, _P, _P, _P]
        L.oracle_pv_beamline_histo.argtypes = [C.c_uint32, _P, _P, _P]
        L.oracle_pv_beamline_peak.argtypes = [C.c_uint32, _P, _P, _P]
        L.oracle_pv_beamline_multi_fitter.argtypes = [C.c_uint32] + [_P] * 6
        _F = C.c_float
        _PF = C.POINTER(C.c_float)
        L.oracle_kalman_step.argtypes = [_F, _F, _F, _F, _PF, _PF, _PF, _PF, _PF]
        L.oracle_kalman_step.restype = C.c_float
        L.oracle_ccl_event.restype = C.c_uint32
        _lib = L
    return _lib


def ref_sp_lib():
    """The reference's own cache_sp_patterns, or None when not built."""
    if not REF_SP_PATH.exists():
        return None
    L = C.CDLL(str(REF_SP_PATH))
    L.ref_cache_sp_patterns.argtypes = [_P] * 4
    return L


# --------------------------------------------------------------------------
# stage wrappers (numpy in, numpy out; buffers use the reference layouts)
# --------------------------------------------------------------------------

def sp_lut():
    p = np.zeros(256, np.uint8); s = np.zeros(256, np.uint8)
    fx = np.zeros(512, np.float32); fy = np.zeros(512, np.float32)
    lib().oracle_sp_lut(_ptr(p), _ptr(s), _ptr(fx), _ptr(fy))
    return p, s, fx, fy


def ref_sp_lut():
    L = ref_sp_lib()
    if L is None:
        return None
    p = np.zeros(256, np.uint8); s = np.zeros(256, np.uint8)
    fx = np.zeros(512, np.float32); fy = np.zeros(512, np.float32)
    L.ref_cache_sp_patterns(_ptr(p), _ptr(s), _ptr(fx), _ptr(fy))
    return p, s, fx, fy


def estimate_input_size(events: np.ndarray, offsets: np.ndarray):
    n = len(offsets) - 1
    est = np.zeros(n * N_MODULES + 1, np.uint32)
    num = np.zeros(n * N_MODULES, np.uint32)
    cnum = np.zeros(n, np.uint32)
    cand = np.zeros(n * MAX_CANDIDATES_EVENT, np.uint32)
    status = np.zeros(n, np.uint32)
    lib().oracle_estimate_input_size(_ptr(events), _ptr(offsets), n, _ptr(est), _ptr(num), _ptr(cnum),
                                     _ptr(cand), _ptr(status))
    return est, num, cnum, cand, status


def prefix_sum(a: np.ndarray) -> np.ndarray:
    """Exclusive scan of a[:-1] in place semantics; returns a copy with total at [-1]."""
    out = np.ascontiguousarray(a, dtype=np.uint32).copy()
    lib().oracle_prefix_sum(_ptr(out), len(out) - 1)
    return out


def masked_clustering(events, offsets, starts, geometry: bytes, status=None):
    n = len(offsets) - 1
    total = int(starts[n * N_MODULES])
    num = np.zeros(n * N_MODULES, np.uint32)
    cont = np.zeros(6 * total + 1, np.uint32)
    st = np.zeros(n, np.uint32) if status is None else status
    g = np.frombuffer(geometry, np.uint8).copy()
    lib().oracle_masked_clustering(_ptr(events), _ptr(offsets), n, _ptr(starts), _ptr(num), _ptr(g),
                                   _ptr(cont), _ptr(st))
    return num, cont, st


def calculate_phi_and_sort(n, starts, nums, container):
    cont = container.copy()
    perm = np.zeros(max(1, int(starts[n * N_MODULES])), np.uint32)
    st = np.zeros(n, np.uint32)
    lib().oracle_calculate_phi_and_sort(n, _ptr(starts), _ptr(nums), _ptr(cont), _ptr(perm), _ptr(st))
    return cont, perm


def fill_candidates(n, container, starts, nums):
    total = int(starts[n * N_MODULES])
    h0 = np.full(2 * total + 2, -7, np.int16)
    h2 = np.full(2 * total + 2, -7, np.int16)
    lib().oracle_fill_candidates(n, _ptr(container), _ptr(starts), _ptr(nums), _ptr(h0), _ptr(h2))
    return h0, h2


def search_by_triplet(n, container, starts, nums, h0, h2, module_zs):
    total = int(starts[n * N_MODULES])
    tracks = np.zeros(n * MAX_TRACKS, TRACK_DTYPE)
    weak = np.zeros(n * MAX_WEAK_TRACKS, TRACKLET_DTYPE)
    used = np.zeros(total + 1, np.uint8)
    atomics = np.zeros(n * NUM_ATOMICS + 1, np.int32)
    st = np.zeros(n, np.uint32)
    zs = np.ascontiguousarray(module_zs, np.float32)
    lib().oracle_search_by_triplet(n, _ptr(container), _ptr(starts), _ptr(nums), _ptr(tracks), _ptr(weak),
                                   _ptr(used), _ptr(atomics), _ptr(h0), _ptr(h2), _ptr(zs), _ptr(st))
    return tracks, weak, used, atomics, st


def weak_tracks_adder(n, container, starts, tracks, weak, used, atomics):
    tracks = tracks.copy(); atomics = atomics.copy()
    st = np.zeros(n, np.uint32)
    lib().oracle_weak_tracks_adder(n, _ptr(container), _ptr(starts), _ptr(tracks), _ptr(weak), _ptr(used),
                                   _ptr(atomics), _ptr(st))
    return tracks, atomics, st


def copy_and_prefix_sum(n, atomics):
    a = atomics.astype(np.uint32).copy()
    lib().oracle_copy_and_prefix_sum(n, _ptr(a))
    return a


def copy_velo_track_hit_number(n, tracks, atomics):
    total = int(atomics[2 * n])
    hn = np.zeros(total + 1, np.uint32)
    lib().oracle_copy_velo_track_hit_number(n, _ptr(tracks), _ptr(atomics), _ptr(hn))
    return hn


def consolidate(n, atomics, tracks, hit_number, container, starts):
    total_tracks = int(atomics[2 * n])
    total_hits = int(hit_number[total_tracks])
    hits = np.zeros(16 * total_hits + 16, np.uint8)
    states = np.zeros(49 * total_tracks + 49, np.uint8)
    lib().oracle_consolidate_velo_tracks(n, _ptr(atomics), _ptr(tracks), _ptr(hit_number), _ptr(container),
                                         _ptr(starts), _ptr(hits), _ptr(states))
    return hits, states


PV_NBINS = 2400
PV_MAX_VERTICES = 32
PV_VERTEX_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("chi2", "<f4"), ("ndof", "<i4"),
                            ("n_tracks", "<u4"), ("cov00", "<f4"), ("cov10", "<f4"), ("cov11", "<f4"),
                            ("cov20", "<f4"), ("cov21", "<f4"), ("cov22", "<f4"), ("nTracks", "<i4")])


def pv_exp(x: float) -> float:
    return float(lib().oracle_pv_exp(x))


def pv_gauss_integral(x: float) -> float:
    return float(lib().oracle_pv_gauss_integral(x))


def pv_beamline_extrapolate(n, atomics, kalman_states):
    """PV tracks (T x 7 float32: z, x, y, tx, ty, W_00, W_11)."""
    atomics = np.ascontiguousarray(atomics, np.uint32)
    T = int(atomics[2 * n])
    out = np.zeros((max(T, 1), 7), np.float32)
    lib().oracle_pv_beamline_extrapolate(n, _ptr(atomics), _ptr(np.ascontiguousarray(kalman_states, np.uint8)),
                                         _ptr(out))
    return out[:T]


def pv_beamline_histo(n, atomics, pvtracks):
    out = np.zeros((max(n, 1), PV_NBINS), np.float32)
    lib().oracle_pv_beamline_histo(n, _ptr(np.ascontiguousarray(atomics, np.uint32)),
                                   _ptr(np.ascontiguousarray(pvtracks, np.float32)), _ptr(out))
    return out[:n]


def pv_beamline_peak(n, zhisto):
    """(zpeaks N x 32, counts N)."""
    zp = np.zeros((max(n, 1), PV_MAX_VERTICES), np.float32)
    cnt = np.zeros(max(n, 1), np.uint32)
    lib().oracle_pv_beamline_peak(n, _ptr(np.ascontiguousarray(zhisto, np.float32)), _ptr(zp), _ptr(cnt))
    return zp[:n], cnt[:n]


def pv_beamline_multi_fitter(n, atomics, pvtracks, zpeaks, n_zpeaks):
    """(vertices N x 32 PV_VERTEX_DTYPE, counts N)."""
    v = np.zeros((max(n, 1), PV_MAX_VERTICES), PV_VERTEX_DTYPE)
    cnt = np.zeros(max(n, 1), np.uint32)
    lib().oracle_pv_beamline_multi_fitter(
        n, _ptr(np.ascontiguousarray(atomics, np.uint32)), _ptr(np.ascontiguousarray(pvtracks, np.float32)),
        _ptr(np.ascontiguousarray(zpeaks, np.float32)), _ptr(np.ascontiguousarray(n_zpeaks, np.uint32)), _ptr(v),
        _ptr(cnt))
    return v[:n], cnt[:n]


def pv_beamline(n, atomics, kalman_states):
    """The four steps: dict of pvtracks, zhisto, zpeaks, n_zpeaks, vertices, n_vertices."""
    tr = pv_beamline_extrapolate(n, atomics, kalman_states)
    h = pv_beamline_histo(n, atomics, tr)
    zp, nz = pv_beamline_peak(n, h)
    v, nv = pv_beamline_multi_fitter(n, atomics, tr, zp, nz)
    return {"pvtracks": tr, "zhisto": h, "zpeaks": zp, "n_zpeaks": nz, "vertices": v, "n_vertices": nv}


def global_event_cut(ut_events, ut_offsets, scifi_events, scifi_offsets):
    """Kept events (increasing) of the global event cut."""
    n = len(ut_offsets) - 1
    assert len(scifi_offsets) - 1 == n
    ut = np.ascontiguousarray(ut_events, np.uint8)
    sf = np.ascontiguousarray(scifi_events, np.uint8)
    lst = np.zeros(max(n, 1), np.uint32)
    k = lib().oracle_global_event_cut(n, _ptr(ut), _ptr(np.ascontiguousarray(ut_offsets, np.uint32)), _ptr(sf),
                                      _ptr(np.ascontiguousarray(scifi_offsets, np.uint32)), _ptr(lst))
    return lst[:k].copy()


def velo_kalman_fit(n, atomics, hit_number, track_hits, states):
    """Kalman-fitted beamline states of the consolidated tracks (49 B each, SoA)."""
    T = int(atomics[2 * n])
    out = np.zeros(49 * T + 49, np.uint8)
    lib().oracle_velo_kalman_fit(n, _ptr(np.ascontiguousarray(atomics, np.uint32)),
                                 _ptr(np.ascontiguousarray(hit_number, np.uint32)),
                                 _ptr(np.ascontiguousarray(track_hits, np.uint8)),
                                 _ptr(np.ascontiguousarray(states, np.uint8)), _ptr(out))
    return out


def kalman_step(z, zhit, xhit, whit, x, tx, cxx, cxtx, ctxtx):
    """One filter step; returns (chi2, x, tx, cxx, cxtx, ctxtx) as float32."""
    v = [C.c_float(x), C.c_float(tx), C.c_float(cxx), C.c_float(cxtx), C.c_float(ctxtx)]
    chi2 = lib().oracle_kalman_step(z, zhit, xhit, whit, *[C.byref(a) for a in v])
    return (np.float32(chi2),) + tuple(np.float32(a.value) for a in v)


def ccl_event(event: np.ndarray, capacity: int = 200000):
    ids = np.zeros(capacity, np.uint32); mods = np.zeros(capacity, np.uint32)
    sx = np.zeros(capacity, np.int32); sy = np.zeros(capacity, np.int32); npx = np.zeros(capacity, np.int32)
    k = lib().oracle_ccl_event(_ptr(event), _ptr(ids), _ptr(mods), _ptr(sx), _ptr(sy), _ptr(npx), capacity)
    k = min(k, capacity)
    return ids[:k], mods[:k], sx[:k], sy[:k], npx[:k]


# --------------------------------------------------------------------------
# whole pipeline
# --------------------------------------------------------------------------

@dataclass
class PipelineResult:
    n_events: int
    starts: np.ndarray           # N*52+1 module cluster starts (scan of estimates)
    nums: np.ndarray             # N*52 clusters per module
    container: np.ndarray        # 6*total sorted cluster container
    status: np.ndarray
    tracks: np.ndarray           # N*1200 TrackHits (after the weak adder)
    atomics: np.ndarray          # after copy_and_prefix_sum
    hit_number: np.ndarray       # T+1 exclusive offsets
    track_hits: np.ndarray       # SoA x,y,z,id bytes
    states: np.ndarray           # SoA state bytes


class _CPipeline(C.Structure):
    _fields_ = [("n_events", C.c_uint32), ("total_estimated_hits", C.c_uint32), ("total_tracks", C.c_uint32),
                ("total_track_hits", C.c_uint32)] + [(f, C.c_void_p) for f in (
                    "module_cluster_start", "module_cluster_num", "cluster_container", "atomics_velo",
                    "velo_track_hit_number", "velo_track_hits", "velo_states", "status")]


def pipeline_c(events: np.ndarray, offsets: np.ndarray, geometry: bytes, module_zs: np.ndarray):
    """Whole oracle pipeline in one C call (the CPU baseline); returns
    (total tracks, total hits on tracks)."""
    L = lib()
    L.oracle_pipeline.argtypes = [_P, _P, C.c_uint32, C.c_char_p, _P, C.POINTER(_CPipeline)]
    L.oracle_pipeline.restype = C.c_int
    L.oracle_pipeline_free.argtypes = [C.POINTER(_CPipeline)]
    r = _CPipeline()
    zs = np.ascontiguousarray(module_zs, np.float32)
    L.oracle_pipeline(_ptr(events), _ptr(offsets), len(offsets) - 1, geometry, _ptr(zs), C.byref(r))
    out = (int(r.total_tracks), int(r.total_track_hits))
    L.oracle_pipeline_free(C.byref(r))
    return out


def pipeline(events: np.ndarray, offsets: np.ndarray, geometry: bytes, module_zs: np.ndarray) -> PipelineResult:
    n = len(offsets) - 1
    est, _, _, _, st0 = estimate_input_size(events, offsets)
    starts = prefix_sum(est)
    nums, cont, st1 = masked_clustering(events, offsets, starts, geometry)
    cont, _ = calculate_phi_and_sort(n, starts, nums, cont)
    h0, h2 = fill_candidates(n, cont, starts, nums)
    tracks, weak, used, atomics, st2 = search_by_triplet(n, cont, starts, nums, h0, h2, module_zs)
    tracks, atomics, st3 = weak_tracks_adder(n, cont, starts, tracks, weak, used, atomics)
    atomics = copy_and_prefix_sum(n, atomics)
    hn = copy_velo_track_hit_number(n, tracks, atomics)
    hn = prefix_sum(hn)
    hits, states = consolidate(n, atomics, tracks, hn, cont, starts)
    return PipelineResult(n, starts, nums, cont, st0 | st1 | st2 | st3, tracks, atomics, hn, hits, states)
