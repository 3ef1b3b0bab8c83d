"""The product's VELO checker (velo_checker_run) against closed-form cases
and the test-side restatement in tests/checker.py."""
import numpy as np
import pytest

from cuda_hlt_amd import checker as ck
from cuda_hlt_amd.gen import MCP_DTYPE, MCP_HAS_VELO, MCP_IS_ELECTRON


def mcps_of(spec):
    """spec: list of (event, flags, eta, hits)."""
    m = np.zeros(len(spec), MCP_DTYPE)
    hits, off = [], 0
    for i, (e, flags, eta, h) in enumerate(spec):
        m[i]["event"], m[i]["key"], m[i]["flags"], m[i]["eta"] = e, i, flags, eta
        m[i]["p"], m[i]["pt"] = 1e4, 1e3
        m[i]["hit_offset"], m[i]["n_hits"] = off, len(h)
        hits += list(h)
        off += len(h)
    return m, np.array(hits, np.uint32)


def run(n_events, tracks_per_event, mcps, mc_hits, accepted=None):
    counts = [len(t) for t in tracks_per_event]
    eto = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    flat = [t for ev in tracks_per_event for t in ev]
    tho = np.concatenate([[0], np.cumsum([len(t) for t in flat])]).astype(np.uint32)
    ids = np.concatenate([np.array(t, np.uint32) for t in flat]) if flat else np.zeros(0, np.uint32)
    return ck.check_arrays(n_events, eto, tho, ids, mcps, mc_hits, accepted)


def test_category_names_match_reference_labels():
    names = ck.category_names()
    assert len(names) == 40 and names[36] == "Not electron Velo" and names[16] == "Electrons Velo"


def test_weight_threshold_ghost_and_clone():
    V = MCP_HAS_VELO
    mcps, hits = mcps_of([(0, V, 3.0, [1, 2, 3, 4, 5]), (0, V, -3.0, [10, 11, 12]), (0, V | MCP_IS_ELECTRON, 3.0, [20, 21, 22])])
    tracks = [[
        [1, 2, 3, 4, 5],          # particle 0, weight 1
        [1, 2, 3, 10],            # particle 0, weight 0.75 -> clone
        [10, 11, 1, 2],           # weight 0.5 -> ghost
        [99, 98, 97],             # no match -> ghost
        [10, 11, 12, 99],         # particle 1 (unmatched hit ignored), weight 1
        [20, 21, 22],             # electron: found only in electron categories
    ]]
    r = run(1, tracks, mcps, hits)
    assert r.tracks == 6 and r.ghosts == 2
    c = r.categories["Not electron Velo"]
    assert (c["accepted"], c["found"], c["clones"]) == (2, 2, 1)
    assert r.categories["Not electron Velo backward"]["found"] == 1
    assert r.categories["Electrons Velo"]["found"] == 1
    np.testing.assert_array_equal(r.track_mcp, [0, 0, -1, -1, 1, 2])
    np.testing.assert_allclose(r.track_weight, [1.0, 0.75, 0.5, 0.0, 1.0, 1.0])
    # hit purity: mean of weights over matched tracks of the category (1, .75, 1)
    assert c["hit_purity"] == pytest.approx((1 + 0.75 + 1) / 3)
    # hit efficiency: n_matched * weight / mc hits = 5*1/5, 4*.75/5, 3*1/3
    assert c["hit_eff"] == pytest.approx((1 + 0.6 + 1) / 3)
    assert c["clone_rate"] == pytest.approx(1 / 4)  # reference formula clones / (2 found)


def test_threshold_is_inclusive_at_0_7():
    mcps, hits = mcps_of([(0, MCP_HAS_VELO, 3.0, list(range(1, 8))), (0, MCP_HAS_VELO, 3.0, [50, 51, 52])])
    r = run(1, [[[1, 2, 3, 4, 5, 6, 7, 50, 51, 52]]], mcps, hits)  # 7/10 = 0.7 -> matched
    assert r.ghosts == 0 and r.track_mcp[0] == 0
    r = run(1, [[[1, 2, 3, 4, 5, 6, 50, 51, 52]]], mcps, hits)  # 6/9 < 0.7 -> ghost
    assert r.ghosts == 1


def test_tie_goes_to_lowest_particle_and_shared_id_to_first_listed():
    mcps, hits = mcps_of([(0, MCP_HAS_VELO, 3.0, [1, 2, 7]), (0, MCP_HAS_VELO, 3.0, [3, 4, 7])])
    r = run(1, [[[1, 2, 3, 4]]], mcps, hits)  # 2 vs 2: weight .5, ghost, best is particle 0
    assert r.ghosts == 1 and r.track_weight[0] == pytest.approx(0.5)
    r = run(1, [[[3, 4, 7]]], mcps, hits)  # id 7 belongs to particle 0 (listed first): 2/3 for particle 1
    assert r.track_weight[0] == pytest.approx(2 / 3)


def test_accepted_mask_drops_tracks_and_events_stay_separate():
    mcps, hits = mcps_of([(0, MCP_HAS_VELO, 3.0, [1, 2, 3]), (1, MCP_HAS_VELO, 3.0, [1, 2, 3])])
    tracks = [[[1, 2, 3], [1, 2, 3]], [[1, 2, 3]]]
    r = run(2, tracks, mcps, hits)
    c = r.categories["Not electron Velo"]
    assert (c["found"], c["clones"], r.tracks) == (2, 1, 3)
    np.testing.assert_array_equal(r.track_mcp, [0, 0, 1])
    r = run(2, tracks, mcps, hits, accepted=np.array([1, 0, 1], np.uint8))
    c = r.categories["Not electron Velo"]
    assert (c["found"], c["clones"], r.tracks) == (2, 0, 2)
    assert "TrackChecker output" in r.report()


def test_rejects_particles_out_of_event_order():
    mcps, hits = mcps_of([(1, MCP_HAS_VELO, 3.0, [1, 2, 3]), (0, MCP_HAS_VELO, 3.0, [4, 5, 6])])
    with pytest.raises(ValueError):
        run(2, [[], []], mcps, hits)


def test_matches_test_restatement_on_oracle_output(minbias_200, geometry, module_zs):
    from oracle import pyoracle as po
    from tests.checker import velo_efficiency
    ev = minbias_200
    r = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    eff, ghost_rate, flags = velo_efficiency(ev, r)
    res = ck.check(ev, r)
    assert res.efficiency() == pytest.approx(eff)
    assert res.ghost_rate == pytest.approx(ghost_rate)
    np.testing.assert_array_equal(res.ghost_labels, flags)
    assert res.efficiency() > 0.95 and res.ghost_rate < 0.03
