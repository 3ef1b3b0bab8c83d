"""velo_prefix_sum_noalloc_launch (the allocation-free scan) against
velo_prefix_sum_launch on the same inputs, through the C ABI: two calls in a
row, the second smaller, with and without a device-side count."""
import numpy as np
import pytest
import torch

from cuda_hlt_amd import native as nat

pytestmark = pytest.mark.gpu


def _scan_with_aux(values: np.ndarray, count=None) -> np.ndarray:
    """The existing entry point, with its auxiliary array from torch."""
    n = len(values)
    d = torch.from_numpy(np.concatenate([values, [0xDEADBEEF]]).astype(np.uint32).view(np.int32)).cuda()
    aux = torch.zeros(nat.prefix_sum_aux_size(n), dtype=torch.int32, device="cuda")
    t = nat.Tuning(0, 0)
    keep = None
    if count is not None:
        keep = torch.tensor([count], dtype=torch.int32, device="cuda")
        t.dev_count = keep.data_ptr()
    nat.launch("velo_prefix_sum_launch", nat.PrefixSumArgs(d.data_ptr(), aux.data_ptr(), n, t))
    torch.cuda.synchronize()
    return d.cpu().numpy().view(np.uint32)


def _scan_noalloc(values: np.ndarray, count=None) -> np.ndarray:
    n = len(values)
    d = torch.from_numpy(np.concatenate([values, [0xDEADBEEF]]).astype(np.uint32).view(np.int32)).cuda()
    keep = None if count is None else torch.tensor([count], dtype=torch.int32, device="cuda")
    rc = nat.lib().velo_prefix_sum_noalloc_launch(d.data_ptr(), n, None if keep is None else keep.data_ptr(), None)
    nat.check(rc, "velo_prefix_sum_noalloc_launch")
    torch.cuda.synchronize()
    return d.cpu().numpy().view(np.uint32)


def test_noalloc_scan_equals_scan_with_aux_array(gpu):
    rng = np.random.default_rng(17)
    for n in (3_000_017, 40_961, 2048, 1, 0):  # each call smaller than the one before
        v = rng.integers(0, 2000, n, dtype=np.uint32)
        got, want = _scan_noalloc(v), _scan_with_aux(v)
        np.testing.assert_array_equal(got, want, err_msg=f"n={n}")
        expect = np.zeros(n + 1, np.uint64)
        expect[1:] = np.cumsum(v.astype(np.uint64))
        np.testing.assert_array_equal(got.astype(np.uint64), expect, err_msg=f"n={n}")


def test_noalloc_scan_with_device_count(gpu):
    rng = np.random.default_rng(18)
    v = rng.integers(0, 50, 100_000, dtype=np.uint32)
    for count in (77_777, 5):
        got, want = _scan_noalloc(v, count), _scan_with_aux(v, count)
        np.testing.assert_array_equal(got[: count + 1], want[: count + 1])
        assert int(got[count]) == int(v[:count].sum())


def test_noalloc_scan_rejects_too_many_elements(gpu):
    d = torch.zeros(4, dtype=torch.int32, device="cuda")
    rc = nat.lib().velo_prefix_sum_noalloc_launch(d.data_ptr(), 134217728 + 1, None, None)
    assert rc == 1  # hipErrorInvalidValue, nothing launched
    torch.cuda.synchronize()
    assert not d.cpu().numpy().any()
