"""Seeded synthetic VELO events (ctypes binding of include/velo_gen.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libvelo_gen.so"
_lib = None

GEOMETRY_BYTES = 19456
MCP_HAS_VELO, MCP_IS_LONG, MCP_IS_ELECTRON, MCP_BACKWARD = 0x1, 0x2, 0x4, 0x8


class GenConfig(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("n_events", C.c_uint32), ("n_threads", C.c_uint32),
        ("nu", C.c_double), ("nu_min", C.c_double), ("nu_max", C.c_double),
        ("tracks_per_pv", C.c_double), ("eta_min", C.c_double), ("eta_max", C.c_double),
        ("backward_fraction", C.c_double), ("hit_efficiency", C.c_double),
        ("noise_fraction", C.c_double),
    ]


MCP_DTYPE = np.dtype([
    ("event", "<u4"), ("key", "<u4"), ("pid", "<i4"), ("p", "<f4"), ("pt", "<f4"),
    ("eta", "<f4"), ("phi", "<f4"), ("flags", "<u4"), ("n_pv", "<u4"),
    ("hit_offset", "<u4"), ("n_hits", "<u4"),
])


def _load():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(_LIB_PATH))
        lib.velo_gen_default_config.argtypes = [C.POINTER(GenConfig)]
        lib.velo_gen_geometry.argtypes = [C.c_char_p, C.c_size_t]
        lib.velo_gen_module_zs.argtypes = [C.POINTER(C.c_float)]
        lib.velo_gen_run.argtypes = [C.POINTER(GenConfig)]
        lib.velo_gen_run.restype = C.c_void_p
        for name, rt in [("velo_gen_n_events", C.c_uint32), ("velo_gen_events_size", C.c_size_t),
                         ("velo_gen_events", C.c_void_p), ("velo_gen_offsets", C.c_void_p),
                         ("velo_gen_n_mcps", C.c_uint32), ("velo_gen_mcps", C.c_void_p),
                         ("velo_gen_n_mc_hits", C.c_uint32), ("velo_gen_mc_hits", C.c_void_p)]:
            f = getattr(lib, name)
            f.argtypes = [C.c_void_p]
            f.restype = rt
        lib.velo_gen_free.argtypes = [C.c_void_p]
        _lib = lib
    return _lib


@dataclass
class Events:
    """Raw VP events in the reference's layout plus MC truth."""
    events: np.ndarray      # uint8, concatenated raw events
    offsets: np.ndarray     # uint32[n+1], byte offsets
    mcps: np.ndarray        # MCP_DTYPE
    mc_hits: np.ndarray     # uint32 LHCbIDs, indexed by mcps['hit_offset']

    @property
    def n_events(self) -> int:
        return len(self.offsets) - 1

    def subset(self, start: int, stop: int) -> "Events":
        lo, hi = int(self.offsets[start]), int(self.offsets[stop])
        sel = (self.mcps["event"] >= start) & (self.mcps["event"] < stop)
        m = self.mcps[sel].copy()
        m["event"] -= start
        return Events(self.events[lo:hi].copy(), (self.offsets[start:stop + 1] - lo).astype(np.uint32),
                      m, self.mc_hits)

    def sp_counts(self) -> np.ndarray:
        """SuperPixel words per event (raw size = 2504 + 4 S bytes)."""
        sizes = np.diff(self.offsets.astype(np.int64))
        return ((sizes - 2504) // 4).astype(np.int64)


def default_config(**kw) -> GenConfig:
    cfg = GenConfig()
    _load().velo_gen_default_config(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def geometry() -> bytes:
    buf = C.create_string_buffer(GEOMETRY_BYTES)
    _load().velo_gen_geometry(buf, GEOMETRY_BYTES)
    return buf.raw


def module_zs() -> np.ndarray:
    out = np.zeros(52, np.float32)
    _load().velo_gen_module_zs(out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def generate(n_events: int, seed: int = 0, **kw) -> Events:
    lib = _load()
    cfg = default_config(n_events=n_events, seed=seed, **kw)
    h = lib.velo_gen_run(C.byref(cfg))
    if not h:
        raise RuntimeError("velo_gen_run failed")
    try:
        n = lib.velo_gen_n_events(h)
        size = lib.velo_gen_events_size(h)
        ev = np.ctypeslib.as_array(C.cast(lib.velo_gen_events(h), C.POINTER(C.c_uint8)), (max(size, 1),))[:size].copy()
        off = np.ctypeslib.as_array(C.cast(lib.velo_gen_offsets(h), C.POINTER(C.c_uint32)), (n + 1,)).copy()
        nm = lib.velo_gen_n_mcps(h)
        if nm:
            raw = C.string_at(lib.velo_gen_mcps(h), nm * MCP_DTYPE.itemsize)
            mcps = np.frombuffer(raw, MCP_DTYPE).copy()
        else:
            mcps = np.zeros(0, MCP_DTYPE)
        nh = lib.velo_gen_n_mc_hits(h)
        hits = (np.ctypeslib.as_array(C.cast(lib.velo_gen_mc_hits(h), C.POINTER(C.c_uint32)), (nh,)).copy()
                if nh else np.zeros(0, np.uint32))
    finally:
        lib.velo_gen_free(h)
    return Events(ev, off, mcps, hits)


# Workloads named by what they are (event count, pile-up, seed).
def minbias(n_events: int, seed: int = 0) -> Events:
    """nu = 7.6 minimum-bias events."""
    return generate(n_events, seed=seed)


def uneven(n_events: int, seed: int = 0) -> Events:
    """Per-event nu ~ U(2, 11): very uneven occupancy."""
    return generate(n_events, seed=seed, nu_min=2.0, nu_max=11.0)
