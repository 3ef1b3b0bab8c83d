"""VP raw banks from MDF files (include/velo_mdf.h).

`read_vp` binds velo_mdf_read_vp, the host C++ reader
(csrc/host/mdf/MdfReader.cpp) that replaces the reference's MDF input for
the VP banks (mdf/src/read_mdf.cpp:28-337, main/src/MDFReader.cpp:15-63):
it verifies record checksums, inflates zlib-compressed records and writes
the VELO event layout the sequence reads (u32 n_banks, u32 byte offsets,
per bank the sensor id and the SuperPixel words).  `check_vp_events`
restates check_velopix_events (main/src/CheckVP.cpp:5-52).  `write_vp`
writes events of that layout back as MDF records, one VP bank per sensor,
so that synthetic events can travel in the on-disk format.

    python -m cuda_hlt_amd.mdf FILE.mdf [...] [-n N] [--threads T]
"""
from __future__ import annotations

import argparse
import ctypes as C
import struct
import sys
import time
import zlib
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

from . import native as nat

BANK_MUON = 13
BANK_VP = 63
BANK_FT_CLUSTER = 64
BANK_UT = 66
BANK_VP_VERSION = 2
BANK_MAGIC = 0xCBCB
HEADER_VERSION = 3
SUBHEADER_WORDS = 7           # trigger mask[4], run, orbit, bunch
DATA_START = 19 + 4 * SUBHEADER_WORDS + 1
ZL_BLOCK_MAX = 0xFFFFFF       # 3-byte sizes in a "ZL" block header


class _Events(C.Structure):
    _fields_ = [("events", C.c_void_p), ("events_size", C.c_size_t), ("offsets", C.POINTER(C.c_uint32)),
                ("n_events", C.c_uint32), ("n_records", C.c_uint32)]


_bound = False


def _lib():
    global _bound
    L = nat.lib()
    if not _bound:
        L.velo_mdf_read_vp.argtypes = [C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                       C.c_uint32, C.POINTER(_Events), C.c_char_p, C.c_size_t]
        L.velo_mdf_read_vp.restype = C.c_int
        L.velo_mdf_read_banks.argtypes = [C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                          C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(_Events),
                                          C.c_char_p, C.c_size_t]
        L.velo_mdf_read_banks.restype = C.c_int
        L.velo_mdf_free.argtypes = [C.POINTER(_Events)]
        L.velo_mdf_free.restype = None
        L.velo_mdf_checksum.argtypes = [C.c_void_p, C.c_size_t]
        L.velo_mdf_checksum.restype = C.c_uint32
        L.velo_read_folder.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(_Events), C.c_char_p,
                                       C.c_size_t]
        L.velo_read_folder.restype = C.c_int
        L.velo_check_vp_events.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
        L.velo_check_vp_events.restype = C.c_int64
        nat.check_mirrors(L.velo_abi_struct_size, {_Events: "velo_mdf_events"}, nat.LIB_PATH)
        _bound = True
    return L


class MdfError(RuntimeError):
    pass


@dataclass
class MdfEvents:
    events: np.ndarray    # uint8, concatenated VP events
    offsets: np.ndarray   # uint32[n + 1], byte offsets
    n_records: int        # records read, skipped ones included

    @property
    def n_events(self) -> int:
        return len(self.offsets) - 1

    def event(self, i: int) -> bytes:
        return self.events[int(self.offsets[i]):int(self.offsets[i + 1])].tobytes()


def _take(L, out: _Events) -> MdfEvents:
    try:
        events = (np.ctypeslib.as_array((C.c_uint8 * out.events_size).from_address(out.events)).copy()
                  if out.events_size else np.zeros(0, np.uint8))
        offsets = np.ctypeslib.as_array(out.offsets, (out.n_events + 1,)).copy()
        return MdfEvents(events, offsets, int(out.n_records))
    finally:
        L.velo_mdf_free(C.byref(out))


def _names(paths):
    if isinstance(paths, (str, Path)):
        paths = [paths]
    return (C.c_char_p * len(paths))(*[str(p).encode() for p in paths]), len(paths)


def read_vp(paths: str | Path | Sequence[str | Path], max_events: int = 0, start_event: int = 0,
            verify_checksum: bool = True, threads: int = 1) -> MdfEvents:
    """VP events of the MDF files, in file and record order."""
    L = _lib()
    names, n = _names(paths)
    out = _Events()
    err = C.create_string_buffer(512)
    rc = L.velo_mdf_read_vp(names, n, max_events, start_event, int(verify_checksum), threads, C.byref(out), err, 512)
    if rc != 0:
        raise MdfError(err.value.decode())
    return _take(L, out)


def read_banks(paths: str | Path | Sequence[str | Path], bank_types: Sequence[int], max_events: int = 0,
               start_event: int = 0, verify_checksum: bool = True, threads: int = 1) -> dict[int, MdfEvents]:
    """Events of each bank type in one pass over the files (as the
    reference's MDF::read_events fills one buffer per type)."""
    L = _lib()
    names, n = _names(paths)
    types = (C.c_uint32 * len(bank_types))(*bank_types)
    outs = (_Events * len(bank_types))()
    err = C.create_string_buffer(512)
    rc = L.velo_mdf_read_banks(names, n, max_events, start_event, int(verify_checksum), threads, types,
                               len(bank_types), outs, err, 512)
    if rc != 0:
        raise MdfError(err.value.decode())
    return {int(t): _take(L, outs[k]) for k, t in enumerate(bank_types)}


def read_folder(folder: str | Path, n_events: int = 0, start_event: int = 0) -> MdfEvents:
    """Per-event files of a folder ("N.bin" / "N_M.bin", natural order,
    round-robin past the last file), as the reference's read_folder
    (main/src/InputTools.cpp:178-211); n_records = files in the folder."""
    L = _lib()
    out = _Events()
    err = C.create_string_buffer(512)
    if L.velo_read_folder(str(folder).encode(), n_events, start_event, C.byref(out), err, 512) != 0:
        raise MdfError(err.value.decode())
    return _take(L, out)


def count_bad_banks(events: np.ndarray, offsets: np.ndarray) -> int:
    """Banks whose sensor id is not their position (check_velopix_events,
    CheckVP.cpp:5-52), counted by the library; raises on a bank table that
    reaches past its event."""
    ev = np.ascontiguousarray(events, np.uint8)
    off = np.ascontiguousarray(offsets, np.uint32)
    n = int(_lib().velo_check_vp_events(ev.ctypes.data, off.ctypes.data, len(off) - 1))
    if n < 0:
        raise MdfError("a bank table or bank offset reaches past its event")
    return n


def checksum(data: bytes) -> int:
    """The MDF record checksum (raw_helpers.cpp:8-21) of data."""
    buf = C.create_string_buffer(bytes(data), len(data))
    return int(_lib().velo_mdf_checksum(buf, len(data)))


def check_vp_events(events: np.ndarray, offsets: np.ndarray) -> list[tuple[int, int, int]]:
    """(event, bank, sensor id) of every bank whose sensor id is not its
    position, as check_velopix_events reports (CheckVP.cpp:24-31)."""
    w = events.view(np.uint32)
    bad = []
    for e in range(len(offsets) - 1):
        base = int(offsets[e]) // 4
        n_banks = int(w[base])
        bank_offsets = w[base + 1:base + 2 + n_banks]
        payload = base + 2 + n_banks
        sensors = w[payload + bank_offsets[:-1] // 4]
        for b in np.flatnonzero(sensors != np.arange(n_banks, dtype=np.uint32)):
            bad.append((e, int(b), int(sensors[b])))
    return bad


# --- writing -----------------------------------------------------------------

def _vp_banks(event: bytes) -> bytes:
    """The raw banks (magic, length, type, version, source id, data) of one
    VP event blob."""
    w = np.frombuffer(event, np.uint32)
    n_banks = int(w[0])
    offs = w[1:2 + n_banks]
    payload = 2 + n_banks
    out = bytearray()
    for b in range(n_banks):
        lo, hi = payload + int(offs[b]) // 4, payload + int(offs[b + 1]) // 4
        source, data = int(w[lo]), w[lo + 1:hi].tobytes()
        out += struct.pack("<HHBBH", BANK_MAGIC, 8 + len(data), BANK_VP, BANK_VP_VERSION, source) + data
    return bytes(out)


def _zl_blocks(raw: bytes, level: int) -> bytes:
    out = bytearray()
    for at in range(0, len(raw), ZL_BLOCK_MAX):
        chunk = raw[at:at + ZL_BLOCK_MAX]
        z = zlib.compress(chunk, level)
        if len(z) > ZL_BLOCK_MAX:
            raise MdfError("compressed block larger than a ZL header can state")
        out += b"ZL" + bytes([zlib.DEFLATED]) + len(z).to_bytes(3, "little") + len(chunk).to_bytes(3, "little") + z
    return bytes(out)


def record(banks: bytes, compress: bool = True, level: int = 6, run: int = 0, extra_banks: bytes = b"",
           with_checksum: bool = True) -> bytes:
    """One MDF record (header version 3, 7-word subheader) holding banks."""
    body = banks + extra_banks
    if compress:
        body = _zl_blocks(body, level)
    size = DATA_START + len(body)
    head = struct.pack("<4I", size, size, size, 0) + bytes([1 if compress else 0, (HEADER_VERSION << 4) |
                                                             SUBHEADER_WORDS, 1])
    sub = struct.pack("<7I", 0, 0, 0, 0, run, 0, 0)
    rec = bytearray(head + sub + b"\0" + body)
    if with_checksum:
        rec[12:16] = struct.pack("<I", checksum(bytes(rec[16:])))
    return bytes(rec)


def write_vp(path: str | Path, events: np.ndarray, offsets: np.ndarray, compress: bool = True,
             level: int = 6) -> int:
    """Writes each event as one MDF record of VP banks; returns bytes written."""
    n = 0
    with open(path, "wb") as f:
        for e in range(len(offsets) - 1):
            ev = events[int(offsets[e]):int(offsets[e + 1])].tobytes()
            n += f.write(record(_vp_banks(ev), compress=compress, level=level))
    return n


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Read the VP banks of MDF files")
    ap.add_argument("files", nargs="+")
    ap.add_argument("-n", "--events", type=int, default=0)
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--no-checksum", action="store_true")
    ap.add_argument("--output", help="write the events (.npz: events, offsets)")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    ev = read_vp(a.files, a.events, a.start, not a.no_checksum, a.threads)
    dt = time.perf_counter() - t0
    bad = check_vp_events(ev.events, ev.offsets)
    print(f"{ev.n_events} events ({ev.n_records} records) {ev.events.nbytes} B of VP banks in {dt * 1e3:.1f} ms "
          f"({ev.n_events / dt:.0f} events/s); sensor-order errors: {len(bad)}")
    if a.output:
        np.savez(a.output, events=ev.events, offsets=ev.offsets)
    return 0 if not bad else 1


if __name__ == "__main__":
    sys.exit(main())
