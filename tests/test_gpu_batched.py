"""Batched runs of the C++ sequence: no host reads between steps, buffers
sized by bounds, the sequence captured in a hipGraph and replayed, and
persistent largest-first launches.  Outputs must equal the plain run's and
the oracle's, bit for bit."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def lpt_order(offsets):
    """Largest event first (ties in event order)."""
    sizes = np.diff(offsets.astype(np.int64))
    return np.argsort(-sizes, kind="stable").astype(np.uint32)


class Resident:
    def __init__(self, ev):
        self.ev = ev
        self.n = ev.n_events
        self.raw = torch.from_numpy(ev.events).cuda()
        self.off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
        self.order = torch.from_numpy(lpt_order(ev.offsets).view(np.int32)).cuda()

    def run(self, seq, **kw):
        seq.run(self.n, dev_events=self.raw.data_ptr(), dev_offsets=self.off.data_ptr(),
                events_size=self.ev.events.nbytes, **kw)
        return seq.host_tracks()


@pytest.fixture(scope="module")
def seq(gpu, geometry, module_zs):
    from cuda_hlt_amd.stream import VeloSequence
    s = VeloSequence(geometry, module_zs, max_events=1000, reserve_mb=4096)
    yield s
    s.close()


@pytest.fixture(scope="module")
def uneven(gpu):
    from cuda_hlt_amd import gen
    return Resident(gen.uneven(600, seed=21))


@pytest.fixture(scope="module")
def uneven_ref(uneven, geometry, module_zs):
    ev = uneven.ev
    return po.pipeline(ev.events, ev.offsets, geometry, module_zs)


def compare(a, b, n):
    np.testing.assert_array_equal(a.atomics[: 2 * n + 1], b.atomics[: 2 * n + 1])
    T = int(b.atomics[2 * n])
    Ht = int(b.hit_number[T])
    np.testing.assert_array_equal(a.hit_number[: T + 1], b.hit_number[: T + 1])
    np.testing.assert_array_equal(a.track_hits[: 16 * Ht], b.track_hits[: 16 * Ht])
    np.testing.assert_array_equal(a.states[: 49 * T], b.states[: 49 * T])


def persistent(grid):
    from cuda_hlt_amd.stream import TUNING_SLOTS
    return {k: (0, 0, grid) for k in TUNING_SLOTS}


def test_batched_matches_plain_and_oracle(seq, uneven, uneven_ref):
    plain = uneven.run(seq)
    batched = uneven.run(seq, batched=True)
    compare(batched, plain, uneven.n)
    compare(batched, uneven_ref, uneven.n)
    o = seq.output()
    assert o.total_tracks == int(uneven_ref.atomics[2 * uneven.n])


def test_graph_replays_are_identical(seq, uneven, uneven_ref):
    uneven.run(seq, batched=True)
    for _ in range(3):
        compare(uneven.run(seq, batched=True), uneven_ref, uneven.n)
    compare(uneven.run(seq, batched=True, repetitions=2), uneven_ref, uneven.n)


@pytest.mark.parametrize("grid", [1, 61, 512])
def test_batched_persistent_largest_first(seq, uneven, uneven_ref, grid):
    out = uneven.run(seq, batched=True, tuning=persistent(grid), dev_work_order=uneven.order.data_ptr())
    compare(out, uneven_ref, uneven.n)
    # replay: the counters were left at zero by the last workgroup
    compare(uneven.run(seq, batched=True, tuning=persistent(grid), dev_work_order=uneven.order.data_ptr()),
            uneven_ref, uneven.n)


def test_plain_run_with_persistent_steps(seq, uneven, uneven_ref):
    compare(uneven.run(seq, tuning=persistent(37), dev_work_order=uneven.order.data_ptr()), uneven_ref, uneven.n)


def test_batched_recaptures_for_new_input(seq, uneven, uneven_ref, minbias_200, geometry, module_zs):
    other = Resident(minbias_200)
    ref = po.pipeline(minbias_200.events, minbias_200.offsets, geometry, module_zs)
    compare(other.run(seq, batched=True), ref, other.n)
    compare(uneven.run(seq, batched=True), uneven_ref, uneven.n)
    compare(uneven.run(seq), uneven_ref, uneven.n)  # plain after batched
    compare(other.run(seq, batched=True), ref, other.n)


def test_batched_needs_resident_input(seq, minbias_200):
    from cuda_hlt_amd import native as nat
    with pytest.raises(nat.VeloError, match="device-resident"):
        seq.run(minbias_200.n_events, host_events=minbias_200.events, host_offsets=minbias_200.offsets, batched=True)


def test_batched_run_fits_the_workspace_it_asks_for(gpu, geometry, module_zs):
    """An arena of exactly velo_workspace_bytes (rounded up to a MiB) holds
    a batched run, whose scheduler reaches the same high-water mark."""
    from cuda_hlt_amd import gen
    from cuda_hlt_amd.stream import VeloSequence, workspace_bytes
    ev = gen.uneven(300, seed=8)
    need = workspace_bytes(ev.n_events, ev.events.nbytes)
    s = VeloSequence(geometry, module_zs, max_events=ev.n_events, reserve_mb=-(-need // 2 ** 20))
    try:
        raw = torch.from_numpy(ev.events).cuda()
        off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
        o = s.run(ev.n_events, dev_events=raw.data_ptr(), dev_offsets=off.data_ptr(), events_size=ev.events.nbytes,
                  batched=True)
        assert o.memory_required == need
        ref = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
        compare(s.host_tracks(), ref, ev.n_events)
    finally:
        s.close()
