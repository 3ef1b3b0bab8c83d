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
        lib.velo_gen_struct_size.argtypes = [C.c_char_p]
        lib.velo_gen_struct_size.restype = C.c_size_t
        for st, name in ((GenConfig, b"velo_gen_config"), (MCP_DTYPE, b"velo_gen_mcp")):
            size = C.sizeof(st) if isinstance(st, type) else st.itemsize
            if lib.velo_gen_struct_size(name) != size:
                raise RuntimeError(f"{_LIB_PATH}: {name.decode()} is {lib.velo_gen_struct_size(name)} bytes in the "
                                   f"library, {size} in the binding (rebuild it)")
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


# --- UT and SciFi banks for the global event cut ---------------------------------

UT_BANKS = 180          # banks per event in the reference's UT sample
SCIFI_BANKS = 240       # banks per event in the reference's FTCluster sample
GEC_SCALE = 0.72        # UT + SciFi clusters per VP raw byte
UT_FRACTION = 0.21      # UT share of the two counts (the reference sample's)


@dataclass
class GecBanks:
    """UT and SciFi raw events (u32 n_banks, u32 byte offsets, per bank the
    source id and its data words) whose only content the global event cut
    reads is right: UT hit counts in the low 16 bits of each bank's first
    data word, SciFi bank sizes."""
    ut_events: np.ndarray
    ut_offsets: np.ndarray
    scifi_events: np.ndarray
    scifi_offsets: np.ndarray
    ut_hits: np.ndarray          # per event, as the cut counts them
    scifi_clusters: np.ndarray   # per event: data bytes / 2 - 2


def _raw_event(words_per_bank: np.ndarray, first_words: np.ndarray | None, rng) -> np.ndarray:
    nb = len(words_per_bank)
    sizes = 4 * (1 + words_per_bank.astype(np.int64))
    offs = np.zeros(nb + 1, np.uint32)
    offs[1:] = np.cumsum(sizes)
    payload = rng.integers(0, 2**32, size=int(offs[-1]) // 4, dtype=np.uint32)
    starts = offs[:-1] // 4
    payload[starts] = np.arange(nb, dtype=np.uint32)          # source ids
    if first_words is not None:
        payload[starts + 1] = first_words
    return np.concatenate([np.array([nb], np.uint32), offs, payload]).view(np.uint8)


def gec_banks(events: Events, seed: int = 0, scale: float = GEC_SCALE) -> GecBanks:
    """UT and SciFi events matching VP events: UT + SciFi clusters
    ~ scale x VP raw bytes (8 % spread), scaled so that about a tenth of
    nu = 7.6 minimum-bias events fail the cut, the removal the reference's
    threshold aims at (GlobalEventCut.cuh:9)."""
    rng = np.random.default_rng(seed)
    vp = np.diff(events.offsets.astype(np.int64))
    total = np.maximum(0, vp * scale * (1 + 0.08 * rng.standard_normal(len(vp)))).astype(np.int64)
    ut_total = (total * UT_FRACTION).astype(np.int64)
    scifi_total = 2 * ((total - ut_total) // 2)                # the estimate is always even
    uts, sfs = [], []
    for e in range(len(vp)):
        hits = rng.multinomial(int(ut_total[e]), np.full(UT_BANKS, 1.0 / UT_BANKS))
        uts.append(_raw_event(1 + (hits + 1) // 2, (0x0080 << 16) | hits.astype(np.uint32), rng))
        # data words = (estimate + 2) / 2 (4 bytes hold 2 clusters)
        words = rng.multinomial(int(scifi_total[e]) // 2 + 1, np.full(SCIFI_BANKS, 1.0 / SCIFI_BANKS))
        sfs.append(_raw_event(words, None, rng))

    def cat(parts):
        offs = np.zeros(len(parts) + 1, np.uint32)
        offs[1:] = np.cumsum([p.nbytes for p in parts])
        return (np.concatenate(parts) if parts else np.zeros(0, np.uint8)), offs

    ue, uo = cat(uts)
    se, so = cat(sfs)
    return GecBanks(ue, uo, se, so, ut_total.astype(np.uint32), scifi_total.astype(np.uint32))
