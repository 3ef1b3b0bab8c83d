"""The autotuner: its selection and checking logic on the CPU (a stand-in
sequence with known step times), the batched arena bound, and on the GPU
that tuned launch shapes reproduce the default output bit for bit."""
import json

import numpy as np
import pytest

from cuda_hlt_amd import autotune


class FakeSequence:
    """Records every run; step times come from cost(alg, shape) for the
    tuned step and 1.0 for the others; the output is a function of the
    tuning for algorithms listed in `breaks`."""

    def __init__(self, cost, breaks=()):
        self.cost = cost
        self.breaks = set(breaks)
        self.runs = []
        self.tuning = {}

    def enable_step_timing(self, on=True):
        self.timing = on

    def run(self, n_events, tuning=None, **kw):
        self.runs.append((dict(tuning or {}), kw))
        self.tuning = dict(tuning or {})

    def step_times(self):
        out = []
        for alg, step in autotune.STEP_OF.items():
            shape = self.tuning.get(alg)
            out.append((step, self.cost(alg, shape) if shape else 1.0))
        return out


def fake_outputs(seq, n):
    # the default output unless a "breaking" algorithm runs a tuned shape
    changed = sorted(a for a in seq.breaks if a in seq.tuning and seq.tuning[a] != (0, 0, 0))
    return ("out", n, tuple(changed))


def cost_table(best):
    """Cost with a unique minimum at best[alg] (block, tile, grid)."""
    def cost(alg, shape):
        b, t, g = shape
        bb, bt, bg = best[alg]
        return 1.0 + abs(np.log2(b) - np.log2(bb)) + 0.01 * abs(t - bt) / 1024 + 0.001 * abs(g - bg) / 1024
    return cost


def test_tune_picks_the_fastest_shape_then_grid():
    want = {"estimate_input_size": (512, 4096, 2048), "masked_clustering": (256, 6144, 0),
            "calculate_phi_and_sort": (128, 0, 8192), "fill_candidates": (512, 2048, 1024),
            "search_by_triplet": (64, 512, 4096), "consolidate": (256, 2048, 0)}
    seq = FakeSequence(cost_table(want))
    best, table = autotune.tune(seq, 100, 0, 0, repeats=1, log=lambda *a: None, outputs=fake_outputs)
    assert best == want
    for alg, space in autotune.SPACE.items():
        # every (block, tile) of the space at grid 0, then every grid at the best shape
        keys = set(table[alg])
        assert {f"{b}x{t}x0" for b, t in space} <= keys
        b, t, _ = want[alg]
        assert {f"{b}x{t}x{g}" for g in autotune.BATCHED_GRIDS[1:]} <= keys
        assert len(keys) == len(space) + len(autotune.BATCHED_GRIDS) - 1
    # the final run uses the tuned shapes of every step at once
    assert seq.runs[-1][0] == best


def test_tune_times_one_step_at_a_time_and_takes_the_median():
    calls = []

    def cost(alg, shape):
        calls.append(alg)
        # three repeats: 5, 1, 3 -> median 3 for every shape except (64, 1)
        k = len(calls) % 3
        base = {0: 5.0, 1: 1.0, 2: 3.0}[k]
        return base - (2.5 if alg == "search_by_triplet" and shape[:2] == (64, 1) else 0.0)

    seq = FakeSequence(cost)
    best, table = autotune.tune(seq, 10, 0, 0, repeats=3, log=lambda *a: None, outputs=fake_outputs)
    assert best["search_by_triplet"][:2] == (64, 1)
    assert table["search_by_triplet"]["64x1x0"] == pytest.approx(0.5)
    # each timing run tunes exactly one step
    assert all(len(t) == 1 for t, _ in seq.runs[1:-1])


def test_tune_rejects_shapes_that_change_the_output():
    want = {a: (s[0][0], s[0][1], 0) for a, s in autotune.SPACE.items()}
    want["fill_candidates"] = (256, 4096, 1024)
    seq = FakeSequence(cost_table(want), breaks=("fill_candidates",))
    with pytest.raises(RuntimeError, match="changed the output"):
        autotune.tune(seq, 10, 0, 0, repeats=1, log=lambda *a: None, outputs=fake_outputs)


def test_tune_batched_starts_from_defaults_and_checks_the_graph_run():
    want = dict(autotune.BATCHED_DEFAULT)
    want["masked_clustering"] = (512, 8192, 4096)
    want["search_by_triplet"] = (64, 1024, 2048)
    want["calculate_phi_and_sort"] = (256, 0, 2048)  # the sort has no tile
    seq = FakeSequence(cost_table(want))
    best, table = autotune.tune_batched(seq, 50, 1, 2, 3, 4, repeats=1, log=lambda *a: None, outputs=fake_outputs)
    assert best == want
    # the search kernel is swept at one wave per workgroup only
    assert all(k.startswith("64x") for k in table["search_by_triplet"])
    # timing runs carry the batched order and the other steps' current best
    tuning, kw = seq.runs[5]
    assert kw["dev_work_order"] == 4 and kw["events_size"] == 3
    assert set(tuning) == set(autotune.STEP_OF)
    # the last run is the batched (graph) run with every tuned shape
    tuning, kw = seq.runs[-1]
    assert tuning == best and kw.get("batched") is True


def test_tune_batched_rejects_a_changed_graph_output():
    seq = FakeSequence(cost_table(dict(autotune.BATCHED_DEFAULT)), breaks=("consolidate",))
    with pytest.raises(RuntimeError, match="changed the output"):
        autotune.tune_batched(seq, 5, 1, 2, 3, 4, repeats=1, log=lambda *a: None, outputs=fake_outputs)


def test_load_round_trip(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"best": {"search_by_triplet": [64, 512, 4096]}}))
    assert autotune.load(p) == {"search_by_triplet": (64, 512, 4096)}


def test_committed_batched_tuning_matches_the_defaults():
    from pathlib import Path
    f = Path(autotune.__file__).parent / "tuning" / "mi355x_uneven_batched.json"
    assert autotune.load(f) == autotune.BATCHED_DEFAULT


def test_batched_reserve_grows_with_the_batch():
    a = autotune.batched_reserve_mb(1000, 1000 * 40000)
    b = autotune.batched_reserve_mb(2000, 2000 * 40000)
    assert 0 < a < b
    # the Kalman fit and the PV steps never need less room
    assert autotune.batched_reserve_mb(1000, 1000 * 40000, velo_kalman_fit=True, pv_beamline=True) >= a


@pytest.mark.gpu
def test_tuned_shapes_reproduce_the_default_output(gpu, geometry, module_zs):
    """A real (small) tuning pass on the GPU: the tuned sequence's device
    output equals the default one (autotune.tune checks it and raises)."""
    import torch
    from cuda_hlt_amd import gen
    from cuda_hlt_amd.stream import VeloSequence
    ev = gen.uneven(300, seed=5)
    raw = torch.from_numpy(ev.events).cuda()
    off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
    seq = VeloSequence(geometry, module_zs, max_events=300, reserve_mb=2048)
    try:
        best, table = autotune.tune(seq, ev.n_events, raw.data_ptr(), off.data_ptr(), repeats=1,
                                    log=lambda *a: None)
        assert set(best) == set(autotune.SPACE)
        for alg, (b, t, g) in best.items():
            assert f"{b}x{t}x{g}" in table[alg]
    finally:
        seq.close()
