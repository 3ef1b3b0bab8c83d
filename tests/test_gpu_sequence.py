"""The C++ sequence (Stream over the configured Velo sequence, through the
velo_sequence C ABI) against the CPU oracle pipeline."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sequence(gpu, geometry, module_zs):
    from cuda_hlt_amd.stream import VeloSequence
    s = VeloSequence(geometry, module_zs, max_events=2000, reserve_mb=2048)
    yield s
    s.close()


def _compare(out, ref, n):
    np.testing.assert_array_equal(out.atomics[: 2 * n + 1], ref.atomics[: 2 * n + 1])
    T = int(ref.atomics[2 * n])
    Ht = int(ref.hit_number[T])
    np.testing.assert_array_equal(out.hit_number[: T + 1], ref.hit_number[: T + 1])
    np.testing.assert_array_equal(out.track_hits[: 16 * Ht], ref.track_hits[: 16 * Ht])
    np.testing.assert_array_equal(out.states[: 49 * T], ref.states[: 49 * T])


def test_sequence_host_input_matches_oracle(sequence, minbias_1000, geometry, module_zs):
    ev = minbias_1000
    sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets)
    ref = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    _compare(sequence.host_tracks(), ref, ev.n_events)


def test_sequence_resident_input_matches_oracle(sequence, geometry, module_zs):
    from cuda_hlt_amd import gen
    ev = gen.uneven(500, seed=5)
    raw = torch.from_numpy(ev.events).cuda()
    off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
    sequence.run(ev.n_events, dev_events=raw.data_ptr(), dev_offsets=off.data_ptr())
    ref = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    _compare(sequence.host_tracks(), ref, ev.n_events)
    o = sequence.output()
    st = torch.empty(ev.n_events, dtype=torch.int32, device="cuda")
    assert o.dev_event_status and o.memory_required > 0


def test_sequence_reruns_are_identical(sequence, minbias_200):
    ev = minbias_200
    sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets, repetitions=3)
    a = sequence.host_tracks()
    sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets)
    b = sequence.host_tracks()
    for x, y in zip((a.atomics, a.hit_number, a.track_hits, a.states), (b.atomics, b.hit_number, b.track_hits, b.states)):
        np.testing.assert_array_equal(x, y)


def test_sequence_arena_exhaustion_raises(gpu, geometry, module_zs, minbias_200):
    from cuda_hlt_amd import native as nat
    from cuda_hlt_amd.stream import VeloSequence
    s = VeloSequence(geometry, module_zs, max_events=200, reserve_mb=1)
    with pytest.raises(nat.VeloError, match="could not be met"):
        s.run(minbias_200.n_events, host_events=minbias_200.events, host_offsets=minbias_200.offsets)
    s.close()


def test_sequence_too_many_events_raises(sequence, minbias_1000):
    from cuda_hlt_amd import native as nat
    big = np.concatenate([minbias_1000.events] * 3)
    offs = np.concatenate([minbias_1000.offsets[:-1] + k * minbias_1000.offsets[-1] for k in range(3)]
                          + [[3 * minbias_1000.offsets[-1]]]).astype(np.uint32)
    with pytest.raises(nat.VeloError):
        sequence.run(3000, host_events=big, host_offsets=offs)


def test_sequence_with_kalman_fit(sequence, minbias_200, geometry, module_zs):
    """The opt-in velo_kalman_fit step: consolidated output unchanged,
    Kalman states equal to the oracle's."""
    ev = minbias_200
    sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets, velo_kalman_fit=True)
    out = sequence.host_tracks()
    ref = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    _compare(out, ref, ev.n_events)
    n = ev.n_events
    T = int(ref.atomics[2 * n])
    want = po.velo_kalman_fit(n, ref.atomics, ref.hit_number, ref.track_hits, ref.states)
    np.testing.assert_array_equal(out.kalman[: 49 * T], want[: 49 * T])
    sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets)
    assert sequence.host_tracks().kalman is None


def test_status_of_the_last_run(sequence, minbias_200, geometry, module_zs):
    """velo_status: the per-event status words equal the oracle's (with
    estimate's bank-order bit) and no error bit is set on minimum bias."""
    ev = minbias_200
    sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets)
    words, combined = sequence.status()
    ref = po.pipeline(ev.events, ev.offsets, geometry, module_zs)
    np.testing.assert_array_equal(words, ref.status)
    assert combined == 0 and np.all(words & po.STATUS_BITS["BANK_ORDER_CHECKED"])


def test_sequence_check_equals_the_host_checker(sequence, minbias_200):
    """SequenceVisitor::check<consolidate_velo_tracks_t> through the C ABI
    (velo_sequence_check) gives the checker's result on the copied-back
    tracks, also for events past the first of the MC truth."""
    from cuda_hlt_amd import checker
    ev = minbias_200
    sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets)
    got = sequence.check(ev.mcps, ev.mc_hits)
    want = checker.check(ev, sequence.host_tracks())
    assert (got.events, got.tracks, got.ghosts) == (want.events, want.tracks, want.ghosts)
    assert got.categories == want.categories
    sub = ev.subset(50, 120)
    sequence.run(sub.n_events, host_events=sub.events, host_offsets=sub.offsets)
    got = sequence.check(ev.mcps, ev.mc_hits, start_event=50)
    want = checker.check(sub, sequence.host_tracks())
    assert got.categories == want.categories and got.ghosts == want.ghosts

