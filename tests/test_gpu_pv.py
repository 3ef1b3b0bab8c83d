"""Beamline PV finding on the GPU (velo_pv_beamline_*_launch, include/velo_pv.h)
against the oracle, stage by stage on the oracle's previous-stage output:
PV tracks, histograms, peaks and vertices bit for bit, for one-workgroup-
per-event and persistent launches."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pv_ref(gpu, geometry, module_zs):
    from cuda_hlt_amd import gen
    ev = gen.uneven(300, seed=17)
    r = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    n = ev.n_events
    kal = po.velo_kalman_fit(n, r.atomics, r.hit_number, r.track_hits, r.states)
    return n, np.asarray(r.atomics, np.uint32), kal, po.pv_beamline(n, r.atomics, kal)


def _shapes(n):
    from tests import gpu_stages as gs
    return [("default", (0, 0)), ("persistent", gs.persistent(n, 97, seed=3)[0])]


def _eq(a, b):
    assert np.asarray(a).tobytes() == np.asarray(b).tobytes()


def test_extrapolate(pv_ref):
    from tests import gpu_stages as gs
    n, at, kal, ref = pv_ref
    for name, t in _shapes(n):
        print(f"[pv extrapolate] {name}", flush=True)
        _eq(gs.pv_extrapolate(n, at, kal, t), ref["pvtracks"])


def test_histo(pv_ref):
    from tests import gpu_stages as gs
    n, at, _, ref = pv_ref
    for name, t in _shapes(n):
        print(f"[pv histo] {name}", flush=True)
        _eq(gs.pv_histo(n, at, ref["pvtracks"], t), ref["zhisto"])


def test_peak(pv_ref):
    from tests import gpu_stages as gs
    n, _, _, ref = pv_ref
    for name, t in _shapes(n):
        print(f"[pv peak] {name}", flush=True)
        zp, nz = gs.pv_peak(n, ref["zhisto"], t)
        _eq(nz, ref["n_zpeaks"])
        for e in range(n):
            _eq(zp[e, :nz[e]], ref["zpeaks"][e, :nz[e]])


def test_multi_fitter(pv_ref):
    from tests import gpu_stages as gs
    n, at, _, ref = pv_ref
    assert ref["n_vertices"].sum() > n        # several vertices per event
    for name, t in _shapes(n):
        print(f"[pv multi fitter] {name}", flush=True)
        v, nv = gs.pv_multi_fitter(n, at, ref["pvtracks"], ref["zpeaks"], ref["n_zpeaks"], t)
        _eq(nv, ref["n_vertices"])
        for e in range(n):
            _eq(v[e, :nv[e]], ref["vertices"][e, :nv[e]])


def test_events_without_tracks(gpu):
    """Empty events: an empty histogram, no peaks, no vertices."""
    from tests import gpu_stages as gs
    n = 3
    at = np.zeros(2 * n + 1, np.uint32)
    h = gs.pv_histo(n, at, np.zeros((0, 7), np.float32))
    assert np.all(h == 0)
    zp, nz = gs.pv_peak(n, h)
    assert np.all(nz == 0)
    v, nv = gs.pv_multi_fitter(n, at, np.zeros((0, 7), np.float32), zp, nz)
    assert np.all(nv == 0)


def test_sequence_with_pv_matches_oracle(gpu, geometry, module_zs):
    """The C++ sequence with its four pv_beamline steps (after
    velo_kalman_fit) equals the oracle chain."""
    from cuda_hlt_amd import gen
    from cuda_hlt_amd.stream import VeloSequence
    print("[pv sequence] minbias 400", flush=True)
    ev = gen.minbias(400, seed=23)
    n = ev.n_events
    r = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    kal = po.velo_kalman_fit(n, r.atomics, r.hit_number, r.track_hits, r.states)
    ref = po.pv_beamline(n, r.atomics, kal)
    s = VeloSequence(geometry, module_zs, max_events=n, reserve_mb=2048)
    try:
        s.run(n, host_events=ev.events, host_offsets=ev.offsets, pv_beamline=True)
        out = s.host_tracks()
        T = int(r.atomics[2 * n])
        _eq(out.kalman[:49 * T], kal[:49 * T])
        _eq(out.pv_counts, ref["n_vertices"])
        for e in range(n):
            _eq(out.pv_vertices[e, :out.pv_counts[e]], ref["vertices"][e, :ref["n_vertices"][e]])
        # and without the PV steps nothing of them is reported
        s.run(n, host_events=ev.events, host_offsets=ev.offsets)
        assert s.host_tracks().pv_counts is None
    finally:
        s.close()
