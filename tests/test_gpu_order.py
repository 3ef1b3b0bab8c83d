"""velo_event_order_launch: a largest-first order of the selected events
(in 64-byte size buckets) that persistent launches take events in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(order, sizes):
    n = len(sizes)
    assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32))   # a permutation
    buckets = np.minimum(sizes[order] >> 6, 1023)
    assert np.all(np.diff(buckets.astype(np.int64)) <= 0)                  # largest first


@pytest.mark.parametrize("n", [1, 3000, 70000])
def test_order_is_largest_first(gpu, n):
    from tests import gpu_stages as gs
    print(f"[event order] n={n}", flush=True)
    rng = np.random.default_rng(n)
    sizes = rng.integers(2504, 60000, size=n).astype(np.uint32) & ~np.uint32(3)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    _check(gs.event_order(offsets), sizes)


def test_order_of_selected_events(gpu):
    from tests import gpu_stages as gs
    rng = np.random.default_rng(5)
    sizes = rng.integers(2504, 30000, size=5000).astype(np.uint32) & ~np.uint32(3)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    sel = np.flatnonzero(rng.random(5000) < 0.6).astype(np.uint32)
    order = gs.event_order(offsets, sel)
    _check(order, sizes[sel])
