"""The continuous-batching server: every request's result is the same when
it is served alone and when it shares batches with other requests, and
equals the oracle's tracks for its events."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = host(a), host(b)
    np.testing.assert_array_equal(a["tracks_per_event"], b["tracks_per_event"])
    np.testing.assert_array_equal(a["hit_number"], b["hit_number"])
    np.testing.assert_array_equal(a["hits"], b["hits"])
    np.testing.assert_array_equal(a["states"], b["states"])
    if a["accepted"] is not None or b["accepted"] is not None:
        np.testing.assert_array_equal(a["accepted"], b["accepted"])


def host(result):
    """A request's result on the host (device results are copied back)."""
    return result if isinstance(result, dict) else result.host()


@pytest.fixture(scope="module")
def requests(gpu):
    from cuda_hlt_amd import serving
    reqs, _ = serving.make_requests(12, 0, seed=3, min_events=1, max_events=300)
    return reqs


@pytest.mark.parametrize("batched,host_results", [(True, False), (False, False), (True, True)])
def test_results_same_alone_and_in_the_mix(gpu, geometry, module_zs, requests, batched, host_results):
    from cuda_hlt_amd import serving
    srv = serving.Server(geometry, module_zs, hbm_budget_mb=16384, max_slots=2, slot_events=1024,
                         slot_bytes=1024 * 40000, batched=batched, host_results=host_results)
    try:
        reqs = [serving.Request(r.id, r.events, r.offsets, 0.0, pinned=r.pinned) for r in requests]
        srv.serve(reqs)  # all queued at once: requests share batches
        assert any(len(b.requests) > 1 for b in srv.batches)
        assert sorted(i for b in srv.batches for i in b.requests) == [r.id for r in reqs]
        for r in reqs:
            alone = serving.Request(r.id, r.events, r.offsets, 0.0)  # not pinned: staged by the slot
            srv.serve([alone])
            same(alone.result, r.result)
    finally:
        srv.close()


def test_results_match_the_oracle(gpu, geometry, module_zs, requests):
    from cuda_hlt_amd import serving
    srv = serving.Server(geometry, module_zs, hbm_budget_mb=16384, max_slots=2, slot_events=1024,
                         slot_bytes=1024 * 40000)
    try:
        reqs = [serving.Request(r.id, r.events, r.offsets, 0.0, pinned=r.pinned) for r in requests[:4]]
        srv.serve(reqs)
        for r in reqs:
            ref = po.pipeline(r.events, r.offsets, geometry, module_zs)
            n = r.n
            res = host(r.result)
            np.testing.assert_array_equal(res["tracks_per_event"], ref.atomics[:n])
            T = int(ref.atomics[2 * n])
            Ht = int(ref.hit_number[T])
            assert r.result.n_tracks == T
            np.testing.assert_array_equal(res["hit_number"], ref.hit_number[: T + 1])
            np.testing.assert_array_equal(res["hits"], ref.track_hits[: 16 * Ht])
            np.testing.assert_array_equal(res["states"], ref.states[: 49 * T])
            assert res["accepted"].all()
    finally:
        srv.close()


def test_open_loop_run_reports_latency(gpu, geometry, module_zs):
    from cuda_hlt_amd import serving
    srv = serving.Server(geometry, module_zs, hbm_budget_mb=16384, max_slots=2, slot_events=1024,
                         slot_bytes=1024 * 40000)
    try:
        reqs, _ = serving.make_requests(20, 200.0, seed=9, min_events=1, max_events=200)
        wall = srv.serve(reqs)
        s = serving.latency_summary(reqs, srv.batches, wall)
        assert s["events"] == sum(r.n for r in reqs)
        assert 0 < s["request_latency_ms"]["p50"] <= s["request_latency_ms"]["p99"]
        assert all(r.done >= r.arrival for r in reqs)
    finally:
        srv.close()


def test_oversized_request_is_rejected(gpu, geometry, module_zs):
    from cuda_hlt_amd import gen, serving
    srv = serving.Server(geometry, module_zs, hbm_budget_mb=8192, max_slots=1, slot_events=64,
                         slot_bytes=64 * 40000)
    try:
        ev = gen.uneven(100, seed=1)
        with pytest.raises(ValueError, match="exceeds a slot"):
            srv.serve([serving.Request(0, ev.events, ev.offsets, 0.0)])
    finally:
        srv.close()
