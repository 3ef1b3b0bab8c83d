"""Per-event input folders and MC truth (CPU only).

* The folder reader's order and round-robin (InputTools.cpp:118-211) and
  the MC truth parser (MCEvent.cpp:16-103) on files written here.
* When the reference's own sample is present (its 10 minimum-bias VP
  events, their VELO geometry and MC truth), the oracle runs on those real
  events: every bank in place, no status bit, every textbook connected
  component of three pixels or fewer found, and the checker's efficiency
  and ghost rate in the range the reference's VELO tracking reaches on
  minimum bias (long tracks above 97 %, ghosts below 4 %).  Nothing of the
  sample is copied into the repository.
"""
import struct
from pathlib import Path

import numpy as np
import pytest

REF = Path("/root/reference/input")
VP = REF / "minbias" / "banks" / "VP"
MC = REF / "minbias" / "MC_info" / "tracks"
GEOMETRY = REF / "detector_configuration" / "velo_geometry.bin"


def test_folder_natural_order_and_round_robin(tmp_path):
    from cuda_hlt_amd import mdf
    names = {"1_10.bin": b"D" * 8, "1_2.bin": b"C" * 12, "0_5.bin": b"B" * 4, "3.bin": b"A" * 16}
    for k, v in names.items():
        (tmp_path / k).write_bytes(v)
    ev = mdf.read_folder(tmp_path)
    assert ev.n_events == 4 and ev.n_records == 4
    assert [ev.event(i) for i in range(4)] == [b"A" * 16, b"B" * 4, b"C" * 12, b"D" * 8]
    ev = mdf.read_folder(tmp_path, n_events=6, start_event=3)
    assert [ev.event(i)[:1] for i in range(6)] == [b"D", b"A", b"B", b"C", b"D", b"A"]


def test_folder_errors(tmp_path):
    from cuda_hlt_amd import mdf
    with pytest.raises(mdf.MdfError, match="could not be opened"):
        mdf.read_folder(tmp_path / "missing")
    with pytest.raises(mdf.MdfError, match="no .bin files"):
        mdf.read_folder(tmp_path)
    (tmp_path / "1.bin").write_bytes(b"x" * 4)
    (tmp_path / "notes.txt").write_bytes(b"x")
    with pytest.raises(mdf.MdfError, match="notes.txt"):
        mdf.read_folder(tmp_path)


def _mc_file(particles):
    out = struct.pack("<I", len(particles))
    for p in particles:
        out += struct.pack("<Iiffff8bI", p["key"], p["pid"], p["p"], p["pt"], p["eta"], p["phi"], *p["flags"], p["npv"])
        for det in ("velo", "ut", "scifi"):
            out += struct.pack("<I", len(p[det])) + struct.pack(f"<{len(p[det])}I", *p[det])
    return out


def test_mc_truth_parser(tmp_path):
    from cuda_hlt_amd import checker
    from cuda_hlt_amd.gen import MCP_HAS_VELO, MCP_IS_ELECTRON, MCP_IS_LONG, MCP_BACKWARD
    base = dict(p=5000.0, pt=600.0, phi=0.5, npv=3, ut=[7, 8], scifi=[9])
    a = dict(base, key=1, pid=211, eta=3.0, flags=(1, 0, 1, 1, 1, 0, 0, 0), velo=[10, 11, 12])
    b = dict(base, key=2, pid=-11, eta=-2.5, flags=(0, 0, 1, 0, 0, 1, 0, 1), velo=[20, 21, 22, 23])
    c = dict(base, key=3, pid=13, eta=2.2, flags=(0, 1, 0, 1, 1, 0, 0, 0), velo=[])   # no VELO hits: dropped
    (tmp_path / "5_1.bin").write_bytes(_mc_file([a, c]))
    (tmp_path / "5_2.bin").write_bytes(_mc_file([b]))
    mcps, hits = checker.read_mc_folder(tmp_path)
    assert list(mcps["key"]) == [1, 2] and list(mcps["event"]) == [0, 1]
    assert mcps[0]["flags"] == MCP_HAS_VELO | MCP_IS_LONG
    assert mcps[1]["flags"] == MCP_HAS_VELO | MCP_IS_ELECTRON | MCP_BACKWARD | 0x10 | 0x40
    assert [list(hits[m["hit_offset"]: m["hit_offset"] + m["n_hits"]]) for m in mcps] == [[10, 11, 12],
                                                                                           [20, 21, 22, 23]]
    (tmp_path / "5_3.bin").write_bytes(_mc_file([a]) + b"\0")
    with pytest.raises(ValueError, match="size mismatch"):
        checker.read_mc_folder(tmp_path)
    with pytest.raises(ValueError, match="only 3 MC files"):
        checker.read_mc_folder(tmp_path, n_events=3, start_event=1)


needs_sample = pytest.mark.skipif(not (VP.is_dir() and MC.is_dir() and GEOMETRY.is_file()),
                                  reason="the reference's minimum-bias sample is not present")


@pytest.fixture(scope="module")
def real():
    from cuda_hlt_amd import gen, mdf
    from oracle import pyoracle as po
    ev = mdf.read_folder(VP)
    geometry = GEOMETRY.read_bytes()
    return ev, geometry, po.pipeline(ev.events, ev.offsets, geometry, gen.module_zs())


@needs_sample
def test_real_sample_banks(real):
    from cuda_hlt_amd import mdf
    ev, _, _ = real
    assert ev.n_events == 10
    assert mdf.count_bad_banks(ev.events, ev.offsets) == 0
    assert mdf.check_vp_events(ev.events, ev.offsets) == []


@needs_sample
def test_real_sample_oracle_invariants(real):
    from oracle import pyoracle as po
    ev, _, r = real
    n = ev.n_events
    assert not (r.status & po.STATUS_ERROR_MASK).any()
    assert np.all(r.status & po.STATUS_BITS["BANK_ORDER_CHECKED"])
    T = int(r.atomics[2 * n])
    counts = np.diff(r.hit_number[: T + 1].astype(np.int64))
    assert T > 1000 and counts.min() >= 3 and counts.max() <= 26
    Ht = int(r.hit_number[T])
    ids = r.track_hits[12 * Ht: 16 * Ht].view(np.uint32)
    for t in range(T):
        tid = ids[r.hit_number[t]: r.hit_number[t + 1]]
        assert len(set(tid.tolist())) == len(tid)


@needs_sample
def test_real_sample_small_clusters_match_connected_components(real):
    """Every 8-connected component of at most three pixels (textbook
    labelling of the raw pixels) is a cluster of the oracle, with the same
    LHCbID."""
    from oracle import pyoracle as po
    ev, _, r = real
    n = ev.n_events
    total = int(r.starts[n * po.N_MODULES])
    found = set(r.container[2 * total: 3 * total].tolist())
    small = missing = 0
    for e in range(n):
        ids, _, _, _, npx = po.ccl_event(ev.events[ev.offsets[e]: ev.offsets[e + 1]])
        for i, k in zip(ids, npx):
            if k <= 3:
                small += 1
                missing += int(i) not in found
    assert small > 10000 and missing == 0


@needs_sample
def test_real_sample_efficiency_and_ghosts(real):
    from cuda_hlt_amd import checker, gen
    ev, _, r = real
    mcps, hits = checker.read_mc_folder(MC)
    assert np.unique(mcps["event"]).size == ev.n_events
    res = checker.check(gen.Events(ev.events, ev.offsets, mcps, hits), r)
    assert res.efficiency("Not electron long eta25") > 0.97
    assert res.efficiency("Not electron Velo") > 0.85
    assert res.ghost_rate < 0.04
