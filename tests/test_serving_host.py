"""Host logic of the serving loop (no GPU): the padding event, size
buckets, and cutting a batch's consolidated output into per-request
results, checked on the oracle's output (same layouts)."""
import numpy as np
import pytest

from oracle import pyoracle as po


def test_empty_event_is_a_valid_event_without_hits(geometry, module_zs):
    from cuda_hlt_amd import serving
    e = serving.EMPTY
    assert e.nbytes == 2504
    r = po.pipeline(e, np.array([0, e.nbytes], np.uint32), geometry, module_zs)
    assert int(r.starts[52]) == 0 and int(r.atomics[2]) == 0


def test_buckets():
    from cuda_hlt_amd import serving
    assert serving.bucket_of(1) == 64 and serving.bucket_of(64) == 64 and serving.bucket_of(65) == 128
    assert serving.bucket_of(4096) == 4096 and serving.bucket_of(4097) == 8192
    assert serving.bucket_of(16384) == 16384
    with pytest.raises(ValueError):
        serving.bucket_of(16385)


def test_cut_results_equal_each_request_alone(minbias_200, geometry, module_zs):
    from cuda_hlt_amd import serving
    ev = minbias_200
    parts = [(0, 37), (37, 38), (38, 120), (120, 200)]
    reqs = [serving.Request(k, ev.subset(a, b).events, ev.subset(a, b).offsets) for k, (a, b) in enumerate(parts)]
    # the batch: the requests in order plus padding events
    pad = 5
    events = np.concatenate([ev.events] + [serving.EMPTY] * pad)
    offs = np.concatenate([ev.offsets, ev.offsets[-1] + serving.EMPTY.nbytes * np.arange(1, pad + 1)]).astype(np.uint32)
    batch = po.pipeline(events, offs, geometry, module_zs)
    got = serving.cut_results(batch, ev.n_events + pad, reqs, [a for a, _ in parts])
    for r, res in zip(reqs, got):
        alone = po.pipeline(r.events, r.offsets, geometry, module_zs)
        n = r.n
        T = int(alone.atomics[2 * n])
        Ht = int(alone.hit_number[T])
        np.testing.assert_array_equal(res["tracks_per_event"], alone.atomics[:n])
        np.testing.assert_array_equal(res["hit_number"], alone.hit_number[: T + 1])
        np.testing.assert_array_equal(res["hits"], alone.track_hits[: 16 * Ht])
        np.testing.assert_array_equal(res["states"], alone.states[: 49 * T])
        assert res["accepted"] is None
