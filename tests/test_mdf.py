"""MDF VP input (include/velo_mdf.h, cuda_hlt_amd/mdf.py).

Pinned by: the reference's own sample, whose MDF file and per-event VP
bank files (input/minbias/banks/{mdf,VP}) hold the same 10 events in both
formats, so the decoded MDF must equal the .bin files byte for byte
(read_mdf.cpp:160-183 writes that layout); the stored record checksums of
that file; the published test vectors of the one-at-a-time hash; and
round trips of synthetic events through an independent writer (Python's
zlib, struct packing) with and without compression, several compression
blocks, foreign and odd-sized banks.
"""
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from cuda_hlt_amd import mdf

REF_BANKS = Path("/root/reference/input/minbias/banks")
REF_MDF = REF_BANKS / "mdf" / "upgrade_mc_minbias_scifi_v5.mdf"


def _events_equal(got: mdf.MdfEvents, events, offsets):
    assert got.n_events == len(offsets) - 1
    np.testing.assert_array_equal(got.offsets, offsets)
    np.testing.assert_array_equal(got.events, events)


def test_checksum_published_vectors():
    # Bob Jenkins' one-at-a-time hash; bytes below 0x80 are the same signed or not
    assert mdf.checksum(b"a") == 0xCA2E9442
    assert mdf.checksum(b"The quick brown fox jumps over the lazy dog") == 0x519E91F5
    assert mdf.checksum(b"") == 0


def test_checksum_signed_bytes():
    """Bytes are added as signed char: 0xFF adds -1, not 255."""
    def oaat(data, signed):
        h = 0
        for b in data:
            v = b - 256 if signed and b >= 128 else b
            h = (h + v) & 0xFFFFFFFF
            h = (h + (h << 10)) & 0xFFFFFFFF
            h ^= h >> 6
        h = (h + (h << 3)) & 0xFFFFFFFF
        h ^= h >> 11
        return (h + (h << 15)) & 0xFFFFFFFF
    data = bytes([0xFF, 0x80, 0x01, 0xC3])
    assert oaat(data, False) != oaat(data, True)
    assert mdf.checksum(data) == oaat(data, True)


@pytest.mark.skipif(not REF_MDF.exists(), reason="reference sample absent")
def test_reference_mdf_equals_reference_vp_banks():
    bins = sorted((REF_BANKS / "VP").glob("*.bin"))
    got = mdf.read_vp(REF_MDF, threads=4)
    assert got.n_events == got.n_records == len(bins) == 10
    for i, b in enumerate(bins):
        assert got.event(i) == b.read_bytes(), b.name
    assert mdf.check_vp_events(got.events, got.offsets) == []
    # every record carries a checksum, so the read above verified all ten
    raw = REF_MDF.read_bytes()
    at, n = 0, 0
    while at < len(raw):
        size, _, _, chk = struct.unpack_from("<4I", raw, at)
        assert chk != 0 and mdf.checksum(raw[at + 16:at + size]) == chk
        assert raw[at + 16] & 0xF == 1          # zlib-compressed
        at += size
        n += 1
    assert n == 10


@pytest.mark.skipif(not REF_MDF.exists(), reason="reference sample absent")
def test_reference_mdf_slices_and_corruption(tmp_path):
    full = mdf.read_vp(REF_MDF)
    part = mdf.read_vp(REF_MDF, max_events=3, start_event=4)
    assert part.n_events == 3 and part.n_records == 7
    for i in range(3):
        assert part.event(i) == full.event(4 + i)
    # two files are read in order
    both = mdf.read_vp([REF_MDF, REF_MDF], threads=3)
    assert both.n_events == 20 and both.event(13) == full.event(3)
    # one flipped byte in record 2's compressed payload fails its checksum
    raw = bytearray(REF_MDF.read_bytes())
    at = 0
    for _ in range(2):
        at += struct.unpack_from("<I", raw, at)[0]
    raw[at + 100] ^= 0x10
    bad = tmp_path / "bad.mdf"
    bad.write_bytes(bytes(raw))
    with pytest.raises(mdf.MdfError, match="record 2: checksum mismatch"):
        mdf.read_vp(bad, threads=4)
    assert mdf.read_vp(bad, max_events=2).n_events == 2


@pytest.mark.parametrize("compress", [True, False])
def test_synthetic_round_trip(tmp_path, minbias_200, compress):
    ev = minbias_200.subset(0, 60)
    p = tmp_path / "ev.mdf"
    mdf.write_vp(p, ev.events, ev.offsets, compress=compress)
    for threads in (1, 5):
        _events_equal(mdf.read_vp(p, threads=threads), ev.events, ev.offsets)
    assert mdf.check_vp_events(ev.events, ev.offsets) == []
    sl = mdf.read_vp(p, max_events=7, start_event=50)
    sub = ev.subset(50, 57)
    _events_equal(sl, sub.events, sub.offsets)
    # start beyond the end: no events, all records counted
    none = mdf.read_vp(p, start_event=100)
    assert none.n_events == 0 and none.n_records == 60


def _vp_bank(source, words):
    data = np.asarray(words, np.uint32).tobytes()
    return struct.pack("<HHBBH", 0xCBCB, 8 + len(data), 63, 2, source) + data


def _foreign_bank(btype, nbytes):
    data = bytes(range(nbytes))
    pad = (-(8 + nbytes)) % 4
    return struct.pack("<HHBBH", 0xCBCB, 8 + nbytes, btype, 0, 7) + data + b"\xAA" * pad


def _blob(banks):
    """Expected VP event blob of (source, words) banks."""
    offs, payload = [0], []
    for s, w in banks:
        payload += [s] + list(w)
        offs.append(4 * len(payload))
    return np.array([len(banks)] + offs + payload, np.uint32).tobytes()


def _header(size, compress, chk=0):
    return struct.pack("<4I", size, size, size, chk) + bytes([1 if compress else 0, 0x37, 1]) + bytes(29)


def test_foreign_and_odd_sized_banks_are_skipped(tmp_path):
    vp = [(0, [2, 0x80000101, 0x00000203]), (1, [0]), (2, [1, 0x80001F0F])]
    body = (_foreign_bank(16, 5) + _vp_bank(*vp[0]) + _foreign_bank(40, 1) + _vp_bank(*vp[1])
            + _foreign_bank(63 + 1, 10) + _vp_bank(*vp[2]) + _foreign_bank(5, 3))
    rec_plain = mdf.record(b"", compress=False, extra_banks=body)
    rec_zip = mdf.record(b"", compress=True, extra_banks=body)
    p = tmp_path / "mixed.mdf"
    p.write_bytes(rec_plain + rec_zip + mdf.record(b"", compress=False))
    got = mdf.read_vp(p)
    assert got.n_events == 3
    assert got.event(0) == got.event(1) == _blob(vp)
    assert got.event(2) == _blob([])           # an event without VP banks


def test_several_zlib_blocks_in_one_record(tmp_path):
    vp = [(s, [3, s, s + 1, s + 2]) for s in range(5)]
    banks = b"".join(_vp_bank(*b) for b in vp)
    cut = 3 * 24
    blocks = b""
    for part in (banks[:cut], banks[cut:]):
        z = zlib.compress(part)
        blocks += b"ZL\x08" + len(z).to_bytes(3, "little") + len(part).to_bytes(3, "little") + z
    size = mdf.DATA_START + len(blocks)
    rec = bytearray(_header(size, True) + blocks)
    rec[12:16] = struct.pack("<I", mdf.checksum(bytes(rec[16:])))
    p = tmp_path / "blocks.mdf"
    p.write_bytes(bytes(rec))
    assert mdf.read_vp(p).event(0) == _blob(vp)


def test_checksum_off_reads_corrupted_data(tmp_path, minbias_200):
    ev = minbias_200.subset(0, 4)
    p = tmp_path / "ev.mdf"
    mdf.write_vp(p, ev.events, ev.offsets, compress=False)
    raw = bytearray(p.read_bytes())
    size0 = struct.unpack_from("<I", raw, 0)[0]
    at = size0 + size0 // 2 & ~3                    # a data word of record 1
    raw[at] ^= 0x01
    p.write_bytes(bytes(raw))
    with pytest.raises(mdf.MdfError, match="record 1: checksum mismatch"):
        mdf.read_vp(p)
    got = mdf.read_vp(p, verify_checksum=False)
    diff = np.flatnonzero(got.events != ev.events)
    assert len(diff) == 1


def test_zero_checksum_is_not_verified(tmp_path):
    vp = [(0, [1, 0x80000101])]
    rec = mdf.record(_vp_bank(*vp[0]), compress=False, with_checksum=False)
    p = tmp_path / "nochk.mdf"
    p.write_bytes(rec)
    assert mdf.read_vp(p).event(0) == _blob(vp)


@pytest.mark.parametrize("damage", ["truncated", "magic", "version", "overrun", "block"])
def test_malformed_input_is_an_error(tmp_path, damage):
    vp = [(0, [1, 0x80000101]), (1, [1, 0x80000202])]
    rec = bytearray(mdf.record(b"".join(_vp_bank(*b) for b in vp), compress=damage == "block",
                               with_checksum=False))
    if damage == "truncated":
        rec = rec[:-3]
    elif damage == "magic":
        rec[mdf.DATA_START] ^= 0xFF
    elif damage == "version":
        rec[17] = 0x27
    elif damage == "overrun":
        struct.pack_into("<H", rec, mdf.DATA_START + 2, 400)
    elif damage == "block":
        rec[mdf.DATA_START + 1] = ord("Z")      # "ZZ": not a zlib block
    p = tmp_path / "bad.mdf"
    p.write_bytes(bytes(rec))
    with pytest.raises(mdf.MdfError):
        mdf.read_vp(p)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(mdf.MdfError, match="cannot open"):
        mdf.read_vp(tmp_path / "nope.mdf")


def test_check_vp_events_reports_sensor_order(minbias_200):
    ev = minbias_200.subset(0, 3)
    events = ev.events.copy()
    w = events.view(np.uint32)
    n_banks = int(w[0])
    w[2 + n_banks + int(w[1 + 5]) // 4] = 77   # bank 5 of event 0 claims sensor 77
    assert mdf.check_vp_events(events, ev.offsets) == [(0, 5, 77)]
