"""VELO track checker: efficiency per MC category, clone and ghost rates.

Python side of velo_checker_run (include/velo_checker.h), the host C++
restatement of the reference's VELO checker (MCAssociator.cpp:13-60,
TrackChecker.cpp:93-146 and :349-405, CategoriesTrackChecker.cpp:5-230).
It takes consolidated tracks in the reference layout (scanned track counts,
scanned hit counts, SoA hits whose last quarter is the LHCbIDs), so the
same call checks the GPU output, the oracle's and a filtered subset.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native as nat
from .gen import MCP_DTYPE, Events

N_CATEGORIES = 40
MIN_WEIGHT = 0.7
HEADLINE = "Not electron Velo"


class _Category(C.Structure):
    _fields_ = [("accepted", C.c_uint64), ("found", C.c_uint64), ("clones", C.c_uint64),
                ("eff_per_event", C.c_double), ("hit_eff", C.c_double), ("hit_purity", C.c_double)]


class _McTruth(C.Structure):
    _fields_ = [("mcps", C.c_void_p), ("n_mcps", C.c_size_t), ("mc_hits", C.c_void_p), ("n_mc_hits", C.c_size_t),
                ("n_events", C.c_uint32)]


class _ClusteringResult(C.Structure):
    _fields_ = [("reconstructible", C.c_uint64), ("reconstructed", C.c_uint64), ("correct", C.c_uint64),
                ("clones", C.c_uint64), ("ghosts", C.c_uint64), ("efficiency", C.c_float),
                ("clone_fraction", C.c_float), ("ghost_fraction", C.c_float)]


class _Result(C.Structure):
    _fields_ = [("events", C.c_uint64), ("tracks", C.c_uint64), ("ghosts", C.c_uint64),
                ("ghost_per_event", C.c_double), ("categories", _Category * N_CATEGORIES)]


_bound = False


def _lib():
    global _bound
    L = nat.lib()
    if not _bound:
        L.velo_checker_category_name.argtypes = [C.c_uint32]
        L.velo_checker_category_name.restype = C.c_char_p
        L.velo_checker_run.argtypes = [C.c_uint32] + [C.c_void_p] * 4 + [C.c_void_p, C.c_size_t, C.c_void_p,
                                                                         C.POINTER(_Result), C.c_void_p, C.c_void_p]
        L.velo_checker_run.restype = C.c_int
        nat.check_mirrors(L.velo_abi_struct_size, {_Category: "velo_checker_category",
                                                   _Result: "velo_checker_result",
                                                   _McTruth: "velo_mc_truth",
                                                   _ClusteringResult: "velo_clustering_check_result"}, nat.LIB_PATH)
        L.velo_x86_clustering.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t,
                                          C.POINTER(C.c_size_t)]
        L.velo_x86_clustering.restype = C.c_int
        L.velo_clustering_check.argtypes = [C.c_uint32] + [C.c_void_p] * 4 + [C.c_int, C.c_float,
                                                                             C.POINTER(_ClusteringResult)]
        L.velo_clustering_check.restype = C.c_int
        L.velo_mc_read_folder.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(_McTruth), C.c_char_p,
                                          C.c_size_t]
        L.velo_mc_read_folder.restype = C.c_int
        L.velo_mc_free.argtypes = [C.POINTER(_McTruth)]
        L.velo_mc_free.restype = None
        _bound = True
    return L


def category_names() -> list[str]:
    L = _lib()
    return [L.velo_checker_category_name(i).decode() for i in range(N_CATEGORIES)]


@dataclass
class CheckerResult:
    events: int
    tracks: int
    ghosts: int
    ghost_per_event: float
    categories: dict = field(default_factory=dict)
    track_mcp: np.ndarray | None = None     # int32 per track: matched particle (index into mcps) or -1
    track_weight: np.ndarray | None = None  # float32 per track: best weight (0 when unmatched)

    @property
    def ghost_rate(self) -> float:
        return self.ghosts / self.tracks if self.tracks else 0.0

    def efficiency(self, name: str = HEADLINE) -> float:
        c = self.categories[name]
        return c["found"] / c["accepted"] if c["accepted"] else 0.0

    @property
    def ghost_labels(self) -> np.ndarray:
        """Per track: True when the track is a ghost (no particle holds at
        least 70 % of its matched hits)."""
        return self.track_mcp < 0

    def report(self) -> str:
        """The reference's printout (TrackChecker.cpp:24-30, :126-145)."""
        lines = []
        for name, c in self.categories.items():
            if not c["accepted"]:
                continue
            lines.append(f"{name:<50s}: {c['found']:9d}/{c['accepted']:9d} {100 * self.efficiency(name):6.2f}% "
                         f"({100 * c['eff_per_event']:6.2f}%), {c['clones']:9d} ({100 * c['clone_rate']:6.2f}%) "
                         f"clones, hit eff {100 * c['hit_eff']:6.2f}% pur {100 * c['hit_purity']:6.2f}%")
        lines.append(f"{'TrackChecker output':<50s}: {self.ghosts:9d}/{self.tracks:9d} "
                     f"{100 * self.ghost_rate:6.2f}% ({100 * self.ghost_per_event:6.2f}%) ghosts")
        return "\n".join(lines)


def check_arrays(n_events: int, event_track_offsets, track_hit_offsets, hit_ids, mcps, mc_hits,
                 accepted=None) -> CheckerResult:
    """Check tracks given as raw arrays (see velo_checker_run)."""
    L = _lib()
    eto = np.ascontiguousarray(event_track_offsets, np.uint32)
    tho = np.ascontiguousarray(track_hit_offsets, np.uint32)
    ids = np.ascontiguousarray(hit_ids, np.uint32)
    mcps = np.ascontiguousarray(mcps, MCP_DTYPE)
    mch = np.ascontiguousarray(mc_hits, np.uint32)
    T = int(eto[n_events])
    acc = None if accepted is None else np.ascontiguousarray(accepted, np.uint8)
    if acc is not None and len(acc) < T:
        raise ValueError("accepted holds fewer entries than tracks")
    track_mcp = np.empty(max(T, 1), np.int32)
    track_weight = np.empty(max(T, 1), np.float32)
    res = _Result()
    rc = L.velo_checker_run(n_events, eto.ctypes.data, tho.ctypes.data, ids.ctypes.data if len(ids) else None,
                            acc.ctypes.data if acc is not None else None,
                            mcps.ctypes.data if len(mcps) else None, len(mcps),
                            mch.ctypes.data if len(mch) else None, C.byref(res),
                            track_mcp.ctypes.data, track_weight.ctypes.data)
    if rc != 0:
        raise ValueError("velo_checker_run rejected its input (null pointer or MC particles out of event order)")
    return result_of(res, track_mcp[:T], track_weight[:T])


def result_of(res, track_mcp=None, track_weight=None) -> CheckerResult:
    """CheckerResult of a velo_checker_result."""
    cats = {}
    for i, name in enumerate(category_names()):
        c = res.categories[i]
        cats[name] = {"accepted": int(c.accepted), "found": int(c.found), "clones": int(c.clones),
                      "eff_per_event": c.eff_per_event, "hit_eff": c.hit_eff, "hit_purity": c.hit_purity,
                      "clone_rate": c.clones / (2 * c.found) if c.found else 0.0}
    return CheckerResult(int(res.events), int(res.tracks), int(res.ghosts), res.ghost_per_event, cats,
                         track_mcp, track_weight)


def check(ev: Events, out, accepted=None) -> CheckerResult:
    """Check consolidated tracks `out` (anything with n_events, atomics,
    hit_number and track_hits in the reference layout) against ev's MC truth."""
    n = out.n_events
    T = int(out.atomics[2 * n])
    Ht = int(out.hit_number[T])
    ids = np.asarray(out.track_hits)[12 * Ht: 16 * Ht].view(np.uint32)
    return check_arrays(n, np.asarray(out.atomics)[n: 2 * n + 1], np.asarray(out.hit_number)[: T + 1], ids,
                        ev.mcps, ev.mc_hits, accepted)


def read_mc_folder(folder, n_events: int = 0, start_event: int = 0):
    """MC truth of a folder of per-event MC files (the reference's
    MC_info/tracks layout, MCEvent.cpp:16-103): (mcps, mc_hits) in the
    generator's MCP_DTYPE layout, particles with VELO hits only."""
    L = _lib()
    out = _McTruth()
    err = C.create_string_buffer(512)
    if L.velo_mc_read_folder(str(folder).encode(), n_events, start_event, C.byref(out), err, 512) != 0:
        raise ValueError(err.value.decode())
    try:
        mcps = np.frombuffer(C.string_at(out.mcps, out.n_mcps * MCP_DTYPE.itemsize), MCP_DTYPE).copy()
        hits = np.frombuffer(C.string_at(out.mc_hits, out.n_mc_hits * 4), np.uint32).copy()
    finally:
        L.velo_mc_free(C.byref(out))
    return mcps, hits


# ---- clustering checker (ClusteringChecker.cpp:32-135) ----------------------

@dataclass
class ClusteringResult:
    reconstructible: int
    reconstructed: int
    correct: int
    clones: int
    ghosts: int
    efficiency: float
    clone_fraction: float
    ghost_fraction: float

    def report(self) -> str:
        return (f"Clustering: efficiency {100 * self.efficiency:.3f}% ({self.correct}/{self.reconstructible}), "
                f"clones {100 * self.clone_fraction:.3f}%, ghosts {100 * self.ghost_fraction:.3f}%")


def x86_clustering(events, offsets):
    """The reference's x86 clustering of raw VP events (velo_x86_clustering):
    (ids, id_offsets), the cluster LHCb IDs of event e at
    ids[id_offsets[e]:id_offsets[e + 1]]."""
    L = _lib()
    ev = np.ascontiguousarray(events)
    off = np.ascontiguousarray(offsets, np.uint32)
    n = len(off) - 1
    id_off = np.zeros(n + 1, np.uint32)
    cap = max(ev.nbytes // 2, 1)  # at most two clusters per 4-byte SP word
    ids = np.empty(cap, np.uint32)
    total = C.c_size_t(0)
    rc = L.velo_x86_clustering(ev.ctypes.data, off.ctypes.data, n, ids.ctypes.data, id_off.ctypes.data, cap,
                               C.byref(total))
    if rc != 0:
        raise ValueError(f"velo_x86_clustering failed ({rc}, {total.value} clusters)")
    return ids[: total.value].copy(), id_off


def check_clustering(found_offsets, found_ids, expected_offsets, expected_ids, just_check_ids=True,
                     allowed_distance_error=1.0) -> ClusteringResult:
    """The reference's clustering report of found vs expected cluster IDs."""
    L = _lib()
    fo = np.ascontiguousarray(found_offsets, np.uint32)
    fi = np.ascontiguousarray(found_ids, np.uint32)
    xo = np.ascontiguousarray(expected_offsets, np.uint32)
    xi = np.ascontiguousarray(expected_ids, np.uint32)
    if len(fo) != len(xo):
        raise ValueError("found and expected clusters cover different numbers of events")
    r = _ClusteringResult()
    rc = L.velo_clustering_check(len(fo) - 1, fo.ctypes.data, fi.ctypes.data if len(fi) else None, xo.ctypes.data,
                                 xi.ctypes.data if len(xi) else None, int(just_check_ids),
                                 float(allowed_distance_error), C.byref(r))
    if rc != 0:
        raise ValueError("velo_clustering_check rejected its input")
    return ClusteringResult(int(r.reconstructible), int(r.reconstructed), int(r.correct), int(r.clones),
                            int(r.ghosts), float(r.efficiency), float(r.clone_fraction), float(r.ghost_fraction))


def container_ids(container, starts, nums, n_events, slot=3):
    """Cluster LHCb IDs per event from a cluster container (u32[6 * total];
    slot 3: the IDs in clustering order, slot 2: in phi order) and the
    module starts / counts of masked clustering: (ids, id_offsets)."""
    c = np.asarray(container).view(np.uint32)
    st = np.asarray(starts).view(np.uint32)
    nu = np.asarray(nums).view(np.uint32)
    total = int(st[52 * n_events])
    col = c[slot * total: (slot + 1) * total]
    out, off = [], [0]
    for e in range(n_events):
        for m in range(52):
            s0 = int(st[52 * e + m])
            out.append(col[s0: s0 + int(nu[52 * e + m])])
        off.append(off[-1] + int(nu[52 * e: 52 * e + 52].sum()))
    return (np.concatenate(out) if out else np.zeros(0, np.uint32)), np.asarray(off, np.uint32)
