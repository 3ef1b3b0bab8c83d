"""Ghost-filter kernels against the numpy oracle, the inline filter in
consolidation against the standalone scorer, and what a trained network
gains over the unfiltered output."""
import numpy as np
import pytest
import torch

from oracle import ghost_oracle as go

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def seq(gpu, geometry, module_zs):
    from cuda_hlt_amd.stream import VeloSequence
    s = VeloSequence(geometry, module_zs, max_events=4000, reserve_mb=4096)
    yield s
    s.close()


@pytest.fixture(scope="module")
def run300(seq, minbias_1000):
    ev = minbias_1000.subset(0, 300)
    seq.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets)
    return ev, seq.host_tracks(), seq.output()


def test_track_features_match_oracle(seq, run300):
    from cuda_hlt_amd import ghost
    ev, host, out = run300
    got = ghost.track_features(out, ev.n_events).cpu().numpy().astype(np.float64)
    want = go.consolidated_features(host.atomics, host.hit_number, host.track_hits, host.states, ev.n_events)
    assert got.shape == want.shape and len(got) > 1000
    np.testing.assert_array_equal(got[:, [0, 1, 2, 8]], want[:, [0, 1, 2, 8]])   # counts and modules: exact
    np.testing.assert_array_equal(got[:, 3:6], want[:, 3:6].astype(np.float32))  # copied from the state
    # residuals (mm): differences of float32 coordinates up to ~40 mm, so a
    # few 1e-6 mm of rounding; phi spread (rad): atan2_poly is within 3 ulp
    np.testing.assert_allclose(got[:, 6:8], want[:, 6:8], rtol=1e-4, atol=2e-5)
    np.testing.assert_allclose(got[:, 9], want[:, 9], rtol=0, atol=5e-6)


def _model(seed=1):
    from cuda_hlt_amd import ghost
    rng = np.random.default_rng(seed)
    m = ghost.GhostModel(ghost.init_params(seed))
    m.params[320:352] = rng.normal(0, 0.1, 32)
    m.params[1376:1408] = rng.normal(0, 0.1, 32)
    m.mean = rng.normal(0, 1, 10).astype(np.float32)
    m.scale = rng.uniform(0.5, 2, 10).astype(np.float32)
    return m


def _data(n, seed=2):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 2, (n, 10)).astype(np.float32)
    y = (rng.random(n) < 0.2).astype(np.float32)
    return x, y


@pytest.mark.parametrize("batch", [1, 63, 64, 700, 4096])
def test_gradient_matches_oracle(gpu, batch):
    from cuda_hlt_amd import ghost
    m = _model()
    x, y = _data(5000)
    tr = ghost.Trainer(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), m, pos_weight=3.0)
    idx = np.random.default_rng(batch).choice(5000, batch, replace=False).astype(np.int32)
    g = tr.gradient(torch.from_numpy(idx).cuda())
    xn = (x[idx].astype(np.float64) - m.mean) * m.scale
    w16 = m.fp16.astype(np.float64)
    want = go.gradient(w16, xn, y[idx].astype(np.float64), 3.0)
    tol = 1e-4 * np.abs(want).max() + 1e-7
    np.testing.assert_allclose(g[: go.N_PARAMS], want, rtol=0, atol=tol)
    assert g[go.N_PARAMS] == pytest.approx(go.loss(w16, xn, y[idx].astype(np.float64), 3.0), rel=1e-4)


def test_adam_step_matches_numpy(gpu):
    from cuda_hlt_amd import ghost
    m = _model()
    x, y = _data(3000)
    tr = ghost.Trainer(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), m)
    idx = torch.arange(2048, dtype=torch.int32, device="cuda")
    master = m.params.astype(np.float64)
    mm = np.zeros_like(master)
    vv = np.zeros_like(master)
    for t in range(1, 4):
        g = tr.gradient(idx)[: go.N_PARAMS].astype(np.float64)
        tr.step(idx, lr=1e-2)
        torch.cuda.synchronize()
        mm = 0.9 * mm + 0.1 * g
        vv = 0.999 * vv + 0.001 * g * g
        master = master - 1e-2 * (mm / (1 - 0.9 ** t)) / (np.sqrt(vv / (1 - 0.999 ** t)) + 1e-8)
        got = tr.master.cpu().numpy()
        np.testing.assert_allclose(got, master, rtol=1e-5, atol=1e-6)
    bits = tr.dev_model.cpu().numpy()
    from cuda_hlt_amd.ghost import ModelStruct
    off = ModelStruct.weights.offset
    w16 = bits[off: off + 2 * go.N_PARAMS].view(np.float16)
    np.testing.assert_array_equal(w16, tr.master.cpu().numpy().astype(np.float16))


def test_inference_matches_oracle(gpu):
    from cuda_hlt_amd import ghost
    m = _model(4)
    x, _ = _data(1000, seed=5)
    score = ghost.infer(m.to_device(), torch.from_numpy(x).cuda()).cpu().numpy()
    z, _ = go.forward(m.fp16.astype(np.float64), (x.astype(np.float64) - m.mean) * m.scale)
    np.testing.assert_allclose(score, go.sigmoid(z), rtol=2e-5, atol=1e-6)


def test_inline_filter_equals_standalone_scores(seq, run300):
    """Consolidation with a model accepts exactly the tracks the standalone
    scorer puts below the threshold, and leaves hits and states unchanged."""
    from cuda_hlt_amd import ghost
    ev, plain, out = run300
    feats = ghost.track_features(out, ev.n_events)
    m = _model(6)
    m.mean, m.scale = ghost.normalisation(feats.cpu().numpy())
    score = ghost.infer(m.to_device(), feats).cpu().numpy()
    m.threshold = float(np.median(score))  # about half the tracks on each side
    dm = m.to_device()
    seq.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets, dev_ghost_model=dm.data_ptr())
    filt = seq.host_tracks()
    for a, b in zip((plain.atomics, plain.hit_number, plain.track_hits, plain.states),
                    (filt.atomics, filt.hit_number, filt.track_hits, filt.states)):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(filt.accepted.astype(bool), score < m.threshold)
    assert 0 < filt.accepted.sum() < len(score)
    assert plain.accepted.all()


def test_trained_filter_cuts_ghosts_within_efficiency_budget(seq, geometry, module_zs):
    """Trained on 3000 events, evaluated on 1000 held-out events: fewer
    ghosts than the unfiltered output at <= 0.5 % efficiency loss."""
    from cuda_hlt_amd import gen, ghost
    train_ev, test_ev = gen.minbias(3000, seed=101), gen.minbias(1000, seed=202)
    report = ghost.train_and_evaluate(seq, train_ev, test_ev, epochs=15, budget=0.004)
    assert report["test"]["efficiency_loss"] <= 0.005
    assert report["test"]["ghost_rate_filtered"] < report["test"]["ghost_rate_plain"]
    assert report["train"]["loss"][-1] < report["train"]["loss"][0]
