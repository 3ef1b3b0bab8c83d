"""The clustering checker: the x86 clustering it takes as the expected
clusters (Clustering.cpp:3-270) and the efficiency / clone / ghost report
(ClusteringChecker.cpp:32-135).

The x86 clustering (a row-scan flood fill seeded in pixel order, isolated
SPs through the pattern table) is pinned against the oracle's textbook
8-connected labelling (seeded in column order, a plain stack): the same
cluster IDs on every event, generated and hand-placed.  The report is
pinned by hand-built ID lists whose counts follow from the definitions.
"""
from pathlib import Path

import numpy as np
import pytest

from cuda_hlt_amd import checker
from oracle import pyoracle as po
from tests import rawevent

REF_VP = Path("/root/reference/input/minbias/banks/VP")


def ccl_ids(events, offsets):
    out = []
    for e in range(len(offsets) - 1):
        ids, *_ = po.ccl_event(np.ascontiguousarray(events[offsets[e]: offsets[e + 1]]))
        out.append(np.sort(ids))
    return out


def assert_same_clusters(events, offsets):
    ids, off = checker.x86_clustering(events, offsets)
    assert len(off) == len(offsets)
    for e, want in enumerate(ccl_ids(events, offsets)):
        np.testing.assert_array_equal(np.sort(ids[off[e]: off[e + 1]]), want, err_msg=f"event {e}")
    return ids, off


@pytest.mark.parametrize("kind", ["minbias", "uneven"])
def test_x86_clustering_equals_textbook_labelling(kind):
    from cuda_hlt_amd import gen
    ev = gen.minbias(150, seed=4) if kind == "minbias" else gen.uneven(150, seed=6)
    ids, off = assert_same_clusters(ev.events, ev.offsets)
    assert off[-1] > 100 * 150


def test_x86_clustering_hand_placed_shapes():
    pix = []
    # an L across four SPs (8-connected through a diagonal step)
    pix += [(3, 9, 7), (3, 10, 8), (3, 10, 9), (3, 11, 9), (3, 12, 9)]
    # two pixels touching only at a corner of two SPs, and a lone pixel beside them
    pix += [(7, 101, 3), (7, 102, 4), (7, 104, 4)]
    # an isolated SP holding two clusters (pattern table), in another sensor
    pix += [(12, 40, 20), (12, 41, 23)]
    # a long row run through several SPs, with a pixel above its far end
    pix += [(100, c, 50) for c in range(200, 215)] + [(100, 214, 51)]
    # pixels on the sensor's last column and row
    pix += [(207, 767, 255), (207, 766, 254)]
    ev, off = rawevent.encode(pix)
    ids, _ = assert_same_clusters(ev, off)
    assert len(ids) == 1 + 2 + 2 + 1 + 1


def test_x86_clustering_skips_empty_super_pixels():
    ev, off = rawevent.encode([(5, 20, 20)])
    words = ev.view(np.uint32).copy()
    n_banks = int(words[0])
    bank = 1 + n_banks + 1 + int(words[1 + 5]) // 4  # header word of bank 5
    assert words[bank] == 5 and words[bank + 1] == 1
    words[bank + 2] &= 0xFFFFFF00  # the SP keeps its address, loses its pixels
    ids, off = checker.x86_clustering(words.view(np.uint8), off)
    assert off[-1] == 0


@pytest.mark.skipif(not REF_VP.exists(), reason="needs the reference's sample events")
def test_x86_clustering_on_the_reference_events():
    from cuda_hlt_amd import mdf
    ev = mdf.read_folder(REF_VP)
    events = np.frombuffer(b"".join(ev.event(i) for i in range(ev.n_events)), np.uint8)
    sizes = [len(ev.event(i)) for i in range(ev.n_events)]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    ids, off = assert_same_clusters(events, offsets)
    assert off[-1] > 1000 * ev.n_events


def test_report_counts():
    # expected: 3 clusters; found: one exact, one twice (a clone), one ghost
    xo, xi = [0, 3], [0x80010203, 0x80010510, 0x80020001]
    fo, fi = [0, 4], [0x80010203, 0x80010510, 0x80010510, 0x80030000]
    r = checker.check_clustering(fo, fi, xo, xi)
    assert (r.reconstructible, r.reconstructed, r.correct, r.clones, r.ghosts) == (3, 4, 3, 1, 1)
    assert r.efficiency == pytest.approx(1.0)
    assert r.clone_fraction == pytest.approx(0.25) and r.ghost_fraction == pytest.approx(0.25)


def test_report_distance_tolerance():
    # same sensor and chip, one pixel apart (row 0x03 vs 0x04): a match only
    # when positions are compared, and only within the allowed distance
    xo, xi = [0, 1], [0x80010203]
    fo, fi = [0, 1], [0x80010204]
    assert checker.check_clustering(fo, fi, xo, xi).ghosts == 1
    r = checker.check_clustering(fo, fi, xo, xi, just_check_ids=False, allowed_distance_error=1.5)
    assert (r.correct, r.ghosts) == (1, 0)
    assert checker.check_clustering(fo, fi, xo, xi, just_check_ids=False, allowed_distance_error=1.0).ghosts == 1
    # a row below the expected one (unsigned difference in the reference) is as close
    r = checker.check_clustering(fo, [0x80010202], xo, xi, just_check_ids=False, allowed_distance_error=1.5)
    assert r.correct == 1
    # another chip (bits 16-17) never matches by position
    r = checker.check_clustering(fo, [0x80020203], xo, xi, just_check_ids=False, allowed_distance_error=100.0)
    assert r.ghosts == 1
    # near-duplicates count as clones when positions are compared
    r = checker.check_clustering([0, 2], [0x80010203, 0x80010204], xo, xi, just_check_ids=False,
                                 allowed_distance_error=1.5)
    assert r.clones == 2 and r.correct == 2


def test_report_empty_and_mismatched():
    r = checker.check_clustering([0, 0], [], [0, 0], [])
    assert r.efficiency == 0.0 and r.clone_fraction == 0.0
    with pytest.raises(ValueError):
        checker.check_clustering([0, 0, 0], [], [0, 0], [])


def test_oracle_masked_clustering_report():
    """The masked clustering (the GPU path's definition, through the oracle)
    against the x86 clustering on generated events: the same clusters up to
    a handful of merged-track shapes."""
    from cuda_hlt_amd import gen
    ev = gen.minbias(100, seed=8)
    r = po.pipeline(ev.events, ev.offsets, gen.geometry(), gen.module_zs())
    fi, fo = checker.container_ids(r.container, r.starts, r.nums, ev.n_events)
    xi, xo = checker.x86_clustering(ev.events, ev.offsets)
    res = checker.check_clustering(fo, fi, xo, xi)
    assert res.efficiency > 0.9995 and res.ghost_fraction < 5e-4 and res.clones == 0
