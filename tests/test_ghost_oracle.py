"""The ghost-filter oracle pinned by known answers: hand-built tracks for the
features, central finite differences for the gradient."""
import numpy as np
import pytest

from oracle import ghost_oracle as go


def lhcbid(sensor, cx=10, cy=20):
    return 0x80000000 | (sensor << 18) | ((cx % 256) << 8) | cy


def test_features_of_a_straight_track_on_known_modules():
    # hits on modules 30, 28, 26, 22 (pair 13 missing) of a line through the origin
    modules = [30, 28, 26, 22]
    z = np.array([100.0, 80.0, 60.0, 20.0])
    tx, ty = 0.02, -0.01
    x, y = tx * z, ty * z
    ids = [lhcbid(4 * m + 1) for m in modules]
    f = go.track_features(x, y, z, ids, (0.0, 0.0, tx, ty, 0.0, 1.5))
    assert f[0] == 4 and f[1] == 22 and f[2] == 30
    assert f[3] == tx and f[4] == ty and f[5] == 1.5
    assert f[6] == pytest.approx(0, abs=1e-12) and f[7] == pytest.approx(0, abs=1e-12)
    assert f[8] == 1  # pairs 11..15 spanned, pair 12 (modules 24/25) has no hit
    # the seed hits lie on one ray: no phi spread
    assert f[9] == pytest.approx(0, abs=1e-12)


def test_feature_residuals_and_phi_spread():
    z = np.array([0.0, 10.0, 20.0])
    x = np.array([1.0, 0.0, -1.0])   # phi 0, pi/2, pi
    y = np.array([0.0, 1.0, 0.0])
    ids = [lhcbid(4 * m) for m in (10, 11, 12)]
    # line x = y = 0: residuals are the radii (1, 1, 1)
    f = go.track_features(x, y, z, ids, (0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    assert f[6] == pytest.approx(1.0) and f[7] == pytest.approx(1.0)
    assert f[9] == pytest.approx(np.pi)
    assert f[8] == 0  # modules 10, 11, 12: pairs 5 and 6, both hit
    # the wrap-around: phi = 3.0 and -3.0 are 2*pi - 6 apart
    f = go.track_features(np.cos([3.0, -3.0]), np.sin([3.0, -3.0]), [0.0, 1.0], ids[:2], (0, 0, 0, 0, 0, 0))
    assert f[9] == pytest.approx(2 * np.pi - 6.0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = rng.normal(0, 0.3, go.N_PARAMS)
    xn = rng.normal(0, 1, (40, go.N_FEATURES))
    y = (rng.random(40) < 0.3).astype(np.float64)
    g = go.gradient(params, xn, y, pos_weight=2.5)
    h = 1e-6
    for k in list(range(0, go.N_PARAMS, 37)) + [go.N_PARAMS - 1]:
        e = np.zeros(go.N_PARAMS)
        e[k] = h
        num = (go.loss(params + e, xn, y, 2.5) - go.loss(params - e, xn, y, 2.5)) / (2 * h)
        assert g[k] == pytest.approx(num, rel=1e-5, abs=1e-9), k


def test_loss_of_a_zero_network_is_log2():
    xn = np.ones((5, go.N_FEATURES))
    y = np.array([0, 1, 0, 1, 1.0])
    assert go.loss(np.zeros(go.N_PARAMS), xn, y) == pytest.approx(np.log(2))
