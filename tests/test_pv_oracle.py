"""Beamline PV oracle (oracle_pv_*; cuda/PV/beamlinePV/src/*.cu).

Pinned by:
- exp: libm's exp within 2 ulp over the fit's whole range, and the
  library's own host copy bit for bit;
- the Gaussian integral: its closed-form ends, 0 and 1 at +-sqrt(7) and
  0.5 at 0, its symmetry, monotonicity, and closeness to the normal CDF;
- extrapolation: the closest approach to the beamline is the foot of the
  perpendicular (closed form);
- the histogram: every accepted track adds exactly one unit of mass,
  centred on its z;
- peaks: hand-placed track groups (one, two well separated, two merged, too
  few tracks, an empty event);
- the fit: tracks through a known vertex give back the vertex; the
  5-track and 0.3 mm cuts;
- end to end on generated events: the found vertices lie close to the
  bunch-crossing region.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

SQ7 = math.sqrt(7.0)


def test_exp_against_libm():
    xs = np.concatenate([np.linspace(-103.0, 88.0, 40001), -np.abs(np.random.default_rng(1).normal(0, 5, 20000))])
    worst = 0.0
    for x in xs.astype(np.float32):
        got = po.pv_exp(float(x))
        ref = math.exp(float(x))
        if ref < 1.2e-38:  # subnormal results: absolute error in units of the smallest subnormal
            assert abs(got - ref) <= 2 * 1.4e-45
            continue
        ulp = np.spacing(np.float32(ref))
        worst = max(worst, abs(got - ref) / float(ulp))
    assert worst <= 2.0, worst
    assert po.pv_exp(0.0) == 1.0
    assert po.pv_exp(-200.0) == 0.0 and math.isinf(po.pv_exp(100.0)) and math.isnan(po.pv_exp(float("nan")))


def test_library_exp_equals_oracle():
    from cuda_hlt_amd import native as nat
    L = nat.lib()
    xs = np.random.default_rng(7).uniform(-110, 90, 20000).astype(np.float32)
    for x in xs:
        a, b = np.float32(L.velo_pv_exp(float(x))), np.float32(po.pv_exp(float(x)))
        assert a.tobytes() == b.tobytes(), x


def test_gauss_integral_closed_form():
    gi = po.pv_gauss_integral
    assert gi(0.0) == 0.5
    assert gi(SQ7) == pytest.approx(1.0, abs=1e-6) and gi(-SQ7) == pytest.approx(0.0, abs=1e-6)
    xs = np.linspace(-SQ7, SQ7, 2001)
    v = np.array([gi(float(x)) for x in xs])
    assert np.all(np.diff(v) >= -1e-7)                              # monotone
    np.testing.assert_allclose(v + v[::-1], 1.0, atol=2e-7)         # symmetric about (0, 0.5)
    phi = np.array([0.5 * (1 + math.erf(x / math.sqrt(2))) for x in xs])
    assert np.max(np.abs(v - phi)) < 0.025                          # a Gaussian CDF (compact support)


def _states(rows):
    """Consolidated state SoA (12 floats per track, then the backward byte)."""
    T = len(rows)
    soa = np.zeros((12, T), np.float32)
    for t, r in enumerate(rows):
        for k, key in enumerate(("x", "y", "tx", "ty", "c00", "c20", "c22", "c11", "c31", "c33", "chi2", "z")):
            soa[k, t] = r.get(key, 0.0)
    return np.concatenate([soa.reshape(-1).view(np.uint8), np.zeros(T, np.uint8)])


def _atomics(counts):
    counts = np.asarray(counts, np.uint32)
    return np.concatenate([counts, np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32),
                           [counts.sum()]]).astype(np.uint32)


def test_extrapolate_is_closest_approach():
    rng = np.random.default_rng(3)
    rows = []
    for _ in range(50):
        rows.append({"x": rng.normal(0, 0.5), "y": rng.normal(0, 0.5), "tx": rng.normal(0, 0.2),
                     "ty": rng.normal(0, 0.2), "z": rng.uniform(-100, 100), "c00": 0.01, "c11": 0.02,
                     "c22": 1e-4, "c33": 2e-4, "c20": 0.0, "c31": 0.0})
    tr = po.pv_beamline_extrapolate(1, _atomics([50]), _states(rows))
    for r, p in zip(rows, tr):
        z, x, y, tx, ty, w00, w11 = (float(v) for v in p)
        dz = z - r["z"]
        # the foot of the perpendicular from the beamline: (x, y) orthogonal to (tx, ty)
        assert abs(x * tx + y * ty) < 1e-4 * (1 + abs(dz))
        assert x == pytest.approx(r["x"] + dz * r["tx"], abs=2e-4) and y == pytest.approx(r["y"] + dz * r["ty"], abs=2e-4)
        assert 1 / w00 == pytest.approx(0.01 + dz * dz * 1e-4, rel=1e-4)
        assert 1 / w11 == pytest.approx(0.02 + dz * dz * 2e-4, rel=1e-4)


def _pvtracks(zs, slope=0.1, w=400.0, x0=0.0, y0=0.0):
    """PV tracks through (x0, y0, z) for each z, slopes spread in phi."""
    out = []
    for i, z in enumerate(zs):
        a = 2 * math.pi * i / max(len(zs), 1)
        out.append([z, x0, y0, slope * math.cos(a), slope * math.sin(a), w, w])
    return np.asarray(out, np.float32).reshape(-1, 7)


def test_histogram_mass_and_centre():
    zs = [-120.0, 0.3, 42.0, 250.7]
    tr = _pvtracks(zs)
    h = po.pv_beamline_histo(1, _atomics([len(zs)]), tr)[0]
    assert h.sum() == pytest.approx(len(zs), abs=1e-4)       # one unit per track
    centres = -300 + 0.25 * (np.arange(2400) + 0.5)
    for z in zs:
        sel = np.abs(centres - z) < 10
        assert (h[sel] * centres[sel]).sum() / h[sel].sum() == pytest.approx(z, abs=0.02)
    assert h[0] == 0 and h[-1] == 0
    # outside the z range, or too imprecise in z (zerr >= 1.5 mm): no mass
    bad = np.asarray([[350.0, 0, 0, 0.1, 0.0, 400, 400], [10.0, 0, 0, 0.01, 0.0, 400, 400]], np.float32)
    assert po.pv_beamline_histo(1, _atomics([2]), bad)[0].sum() == 0


@pytest.mark.parametrize("groups,expect", [
    ([(10.0, 20)], [10.0]),
    ([(-50.0, 15), (60.0, 12)], [-50.0, 60.0]),
    ([(0.0, 2)], []),                    # 2 tracks < 2.5 in the seed
    ([], []),
])
def test_peaks_of_track_groups(groups, expect):
    zs = [z for z, n in groups for _ in range(n)]
    tr = _pvtracks(zs) if zs else np.zeros((0, 7), np.float32)
    h = po.pv_beamline_histo(1, _atomics([len(zs)]), tr)
    zp, nz = po.pv_beamline_peak(1, h)
    assert int(nz[0]) == len(expect)
    np.testing.assert_allclose(np.sort(zp[0, :nz[0]]), expect, atol=0.05)


def test_close_groups_merge_into_one_peak():
    zs = [0.0] * 10 + [1.0] * 10        # 1 mm apart at a 0.13 mm z resolution... with a shallow dip
    tr = _pvtracks(zs, w=25.0)          # zerr = 1 / (0.1 * 5) = 2 mm > 1.5: rejected
    assert po.pv_beamline_histo(1, _atomics([20]), tr).sum() == 0
    tr = _pvtracks(zs, w=100.0)         # zerr = 1 mm: the two groups overlap
    zp, nz = po.pv_beamline_peak(1, po.pv_beamline_histo(1, _atomics([20]), tr))
    assert int(nz[0]) == 1 and zp[0, 0] == pytest.approx(0.5, abs=0.3)


def test_fit_finds_the_true_vertices():
    # sigma 0.1 mm at the beamline: the fit starts there, and a track counts
    # only while exp(-chi2/2) over the seeds' sum stays >= 0.3
    truth = [(0.05, -0.03, -30.0), (-0.02, 0.04, 45.0)]
    trs = [_pvtracks([z] * 12, x0=x, y0=y, w=100.0) for x, y, z in truth]
    tr = np.concatenate(trs)
    # tracks through (x0, y0, z): at z their x, y is the vertex's
    at = _atomics([len(tr)])
    h = po.pv_beamline_histo(1, at, tr)
    zp, nz = po.pv_beamline_peak(1, h)
    assert int(nz[0]) == 2
    v, nv = po.pv_beamline_multi_fitter(1, at, tr, zp, nz)
    assert int(nv[0]) == 2
    got = sorted([(float(r["x"]), float(r["y"]), float(r["z"])) for r in v[0, :2]], key=lambda t: t[2])
    # every track passes through its vertex: the weighted least squares
    # minimum is the vertex itself, whatever the weights
    for (x, y, z), (gx, gy, gz) in zip(truth, got):
        assert gz == pytest.approx(z, abs=1e-3) and gx == pytest.approx(x, abs=1e-4) and gy == pytest.approx(y, abs=1e-4)
    assert np.all(v[0, :2]["n_tracks"] == len(tr)) and np.all(v[0, :2]["ndof"] == 0)


def test_fit_cuts():
    # fewer than 5 tracks in the event: no vertex, whatever the seeds
    tr = _pvtracks([10.0] * 4, w=100.0)
    at = _atomics([4])
    zp = np.full((1, 32), 10.0, np.float32)
    v, nv = po.pv_beamline_multi_fitter(1, at, tr, zp, np.array([1], np.uint32))
    assert int(nv[0]) == 0
    # a vertex 0.4 mm from the beamline: rejected; 0.2 mm: kept (sigma 0.5 mm,
    # so the fit reaches it from the beamline)
    for x0, kept in ((0.4, 0), (0.2, 1)):
        tr = _pvtracks([10.0] * 8, x0=x0, w=4.0)
        v, nv = po.pv_beamline_multi_fitter(1, _atomics([8]), tr, zp, np.array([1], np.uint32))
        assert int(nv[0]) == kept


def test_end_to_end_on_generated_events(geometry, module_zs):
    from cuda_hlt_amd import gen
    ev = gen.minbias(200, seed=11)
    r = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    n = ev.n_events
    kal = po.velo_kalman_fit(n, r.atomics, r.hit_number, r.track_hits, r.states)
    pv = po.pv_beamline(n, r.atomics, kal)
    nv = pv["n_vertices"]
    assert 2.0 < nv.mean() < 9.0                       # nu = 7.6 pile-up, 25 tracks per PV
    vs = np.concatenate([pv["vertices"][e, :nv[e]] for e in range(n)])
    assert np.all(vs["x"] ** 2 + vs["y"] ** 2 < 0.09)
    assert np.percentile(np.abs(vs["z"]), 90) < 120    # the generator's PVs: sigma_z = 45 mm
    assert np.all(nv <= 32) and np.all(pv["n_zpeaks"] <= 32)


def test_layouts():
    assert po.PV_VERTEX_DTYPE.itemsize == 52
