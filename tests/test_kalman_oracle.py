"""The velo_kalman_fit restatement pinned by the textbook Kalman filter.

Each step must equal the matrix-form update in float64:
  predict  s' = F s, P' = F P F^T  with F = [[1, dz], [0, 1]]
  update   S = H P' H^T + 1/w, K = P' H^T / S, s'' = s' + K r,
           P'' = (I - K H) P', chi2 = r^2 / S,  H = [1, 0].
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def textbook(z, zhit, xhit, w, x, tx, cxx, cxtx, ctxtx):
    s = np.array([x, tx], np.float64)
    P = np.array([[cxx, cxtx], [cxtx, ctxtx]], np.float64)
    F = np.array([[1.0, zhit - z], [0.0, 1.0]])
    s = F @ s
    P = F @ P @ F.T
    H = np.array([1.0, 0.0])
    S = H @ P @ H + 1.0 / w
    K = P @ H / S
    r = xhit - H @ s
    s = s + K * r
    P = (np.eye(2) - np.outer(K, H)) @ P
    return r * r / S, s[0], s[1], P[0, 0], P[0, 1], P[1, 1]


def test_step_equals_matrix_form():
    """Well-conditioned steps (predicted variance within ~10/w): there the
    reference's float32 update, P'' = (1 - K) P', is exact to rounding.
    (Far outside, 1 - K rounds to 0 in float32 and the reference's
    covariance collapses; the restatement keeps that behaviour.)"""
    rng = np.random.default_rng(1)
    for _ in range(500):
        z = rng.uniform(-300, 700)
        zhit = z + rng.uniform(-50, 50)
        x, tx = rng.normal(0, 20), rng.normal(0, 0.2)
        xhit = x + tx * (zhit - z) + rng.normal(0, 0.05)
        cxx, ctxtx = rng.uniform(1e-5, 1e-3), rng.uniform(1e-9, 1e-6)
        cxtx = rng.uniform(-0.5, 0.5) * np.sqrt(cxx * ctxtx)
        w = 3966.94
        got = po.kalman_step(z, zhit, xhit, w, x, tx, cxx, cxtx, ctxtx)
        want = textbook(np.float32(z), np.float32(zhit), np.float32(xhit), np.float32(w), np.float32(x),
                        np.float32(tx), np.float32(cxx), np.float32(cxtx), np.float32(ctxtx))
        for g, t in zip(got, want):
            assert g == pytest.approx(t, rel=2e-3, abs=1e-6)


def _one_track(z, x, y, beam_z, tx, ty):
    """A consolidated event with one track (reference layouts)."""
    n = len(z)
    atomics = np.array([1, 0, 1], np.uint32)
    hit_number = np.array([0, n], np.uint32)
    hits = np.concatenate([np.asarray(x, np.float32).view(np.uint8), np.asarray(y, np.float32).view(np.uint8),
                           np.asarray(z, np.float32).view(np.uint8), np.zeros(n, np.uint32).view(np.uint8)])
    st = np.zeros(12, np.float32)
    st[2], st[3], st[11] = tx, ty, beam_z
    states = np.concatenate([st.view(np.uint8), np.zeros(1, np.uint8)])
    out = po.velo_kalman_fit(1, atomics, hit_number, hits, states)
    return out[:48].view(np.float32), out[48]


def test_straight_track_is_followed_to_its_upstream_end():
    # forward track, hits in decreasing z (search order), beamline upstream
    z = np.array([600.0, 500.0, 400.0, 300.0, 200.0])
    tx, ty = 0.03, -0.02
    x, y = 0.5 + tx * z, -0.2 + ty * z
    s, bwd = _one_track(z, x, y, beam_z=-10.0, tx=tx, ty=ty)
    # filtered from the downstream end towards the beam: state at z = 200
    assert s[11] == pytest.approx(200.0)
    assert s[0] == pytest.approx(x[-1], abs=1e-4) and s[1] == pytest.approx(y[-1], abs=1e-4)
    assert s[2] == pytest.approx(tx, abs=1e-6) and s[3] == pytest.approx(ty, abs=1e-6)
    assert s[10] == pytest.approx(0.0, abs=1e-6)  # chi2 of exact hits
    assert s[4] > 0 and s[7] > 0 and s[6] > 0 and s[9] > 0  # positive variances
    noise = 1e-8 + 7e-6 * (tx * tx + ty * ty)
    assert s[6] > noise  # the noise of the last layer was added


def test_backward_track_runs_the_other_way():
    z = np.array([-100.0, -150.0, -200.0, -250.0])
    tx, ty = -0.05, 0.01
    x, y = tx * z, ty * z
    s, _ = _one_track(z, x, y, beam_z=10.0, tx=tx, ty=ty)   # beam downstream of the hits: backward
    assert s[11] == pytest.approx(-100.0)


def test_fit_of_the_pipeline_output(minbias_200, geometry, module_zs):
    ev = minbias_200
    r = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    n = ev.n_events
    T = int(r.atomics[2 * n])
    out = po.velo_kalman_fit(n, r.atomics, r.hit_number, r.track_hits, r.states)
    K = out[: 48 * T].view(np.float32).reshape(12, T)
    S = r.states[: 48 * T].view(np.float32).reshape(12, T)
    assert np.isfinite(K).all()
    # slopes agree with the least-squares state (the filter weights the hits
    # differently, so tracks with outlying hits differ more)
    for k in (2, 3):
        d = np.abs(K[k] - S[k])
        assert np.median(d) < 2e-4 and np.percentile(d, 99) < 5e-3
    assert (K[4] > 0).all() and (K[7] > 0).all() and (K[10] >= 0).all()
    np.testing.assert_array_equal(out[48 * T: 49 * T], r.states[48 * T: 49 * T])
