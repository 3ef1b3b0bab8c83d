"""global_event_cut on the GPU (velo_global_event_cut_launch and the
sequence's global_event_cut_t) against the oracle: the kept events are
identical (same set, increasing order), and the sequence run with the cut
equals the sequence run without it on the kept events alone."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _banks(n, seed, kind="minbias"):
    from cuda_hlt_amd import gen
    ev = getattr(gen, kind)(n, seed=seed)
    return ev, gen.gec_banks(ev, seed=seed)


@pytest.mark.parametrize("n,block", [(1, 0), (700, 64), (3000, 256), (5000, 1024)])
def test_gec_kernel_matches_oracle(gpu, n, block):
    from tests import gpu_stages as gs
    print(f"[gec] n={n} block={block}", flush=True)
    _, g = _banks(n, seed=n)
    ref = po.global_event_cut(g.ut_events, g.ut_offsets, g.scifi_events, g.scifi_offsets)
    got, k = gs.global_event_cut(g.ut_events, g.ut_offsets, g.scifi_events, g.scifi_offsets, block=block)
    assert k == len(ref)
    np.testing.assert_array_equal(got, ref)


def test_gec_kernel_boundaries(gpu):
    """The hand-built boundary events of the oracle tests, in one batch."""
    from tests import gpu_stages as gs
    from tests import test_gec_oracle as t
    cases = [(9748, []), (9748, [1]), (9748, [2]), (9750, []), (0, [9749]), (0, [9750]), (0, [5000, 4749]),
             (0, [5000, 4750]), (4000, [3000, 2749]), (4000, [3000, 2750])]
    ut = [t.ut_event(u) for _, u in cases] + [t.ut_event([0]), t.ut_event([9749], high=0xFFFF)]
    sf = [t.scifi_event(s) for s, _ in cases] + [t.blob([]), t.scifi_event(0)]
    ue, uo = t.concat(ut)
    se, so = t.concat(sf)
    ref = po.global_event_cut(ue, uo, se, so)
    assert list(ref) == [0, 1, 4, 6, 8, 11]
    got, _ = gs.global_event_cut(ue, uo, se, so)
    np.testing.assert_array_equal(got, ref)


def test_gec_kernel_no_events(gpu):
    from tests import gpu_stages as gs
    z = np.zeros(8, np.uint8)
    got, k = gs.global_event_cut(z, np.zeros(1, np.uint32), z, np.zeros(1, np.uint32))
    assert k == 0 and len(got) == 0


@pytest.fixture(scope="module")
def sequence(gpu, geometry, module_zs):
    from cuda_hlt_amd.stream import VeloSequence
    s = VeloSequence(geometry, module_zs, max_events=2000, reserve_mb=2048)
    yield s
    s.close()


def _tracks_equal(a, b):
    n = a.n_events
    assert b.n_events == n
    np.testing.assert_array_equal(a.atomics[:2 * n + 1], b.atomics[:2 * n + 1])
    T = int(a.atomics[2 * n])
    Ht = int(a.hit_number[T])
    np.testing.assert_array_equal(a.hit_number[:T + 1], b.hit_number[:T + 1])
    np.testing.assert_array_equal(a.track_hits[:16 * Ht], b.track_hits[:16 * Ht])
    np.testing.assert_array_equal(a.states[:49 * T], b.states[:49 * T])


@pytest.mark.parametrize("resident", [False, True])
def test_sequence_with_gec_equals_sequence_on_kept_events(sequence, resident):
    print(f"[gec sequence] resident={resident}", flush=True)
    ev, g = _banks(1500, seed=21, kind="uneven")
    kept = po.global_event_cut(g.ut_events, g.ut_offsets, g.scifi_events, g.scifi_offsets)
    assert 0 < len(kept) < ev.n_events
    if resident:
        keep = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in
                (ev.events, ev.offsets.view(np.int32), g.ut_events, g.ut_offsets.view(np.int32), g.scifi_events,
                 g.scifi_offsets.view(np.int32))]
        gec = {"ut_events": keep[2].data_ptr(), "ut_offsets": keep[3].data_ptr(),
               "scifi_events": keep[4].data_ptr(), "scifi_offsets": keep[5].data_ptr(),
               "ut_events_size": g.ut_events.nbytes, "scifi_events_size": g.scifi_events.nbytes}
        sequence.run(ev.n_events, dev_events=keep[0].data_ptr(), dev_offsets=keep[1].data_ptr(),
                     events_size=ev.events.nbytes, gec=gec)
    else:
        sequence.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets, gec=g)
    with_cut = sequence.host_tracks()
    np.testing.assert_array_equal(with_cut.event_list, kept)

    parts = [ev.events[ev.offsets[e]:ev.offsets[e + 1]] for e in kept]
    sub_offsets = np.zeros(len(kept) + 1, np.uint32)
    sub_offsets[1:] = np.cumsum([p.nbytes for p in parts])
    sequence.run(len(kept), host_events=np.concatenate(parts), host_offsets=sub_offsets)
    alone = sequence.host_tracks()
    np.testing.assert_array_equal(alone.event_list, np.arange(len(kept), dtype=np.uint32))
    _tracks_equal(with_cut, alone)


def test_sequence_gec_rejected_in_batched_runs(sequence):
    from cuda_hlt_amd import native as nat
    ev, g = _banks(50, seed=3)
    raw = torch.from_numpy(ev.events).cuda()
    off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
    with pytest.raises(nat.VeloError, match="plain runs only"):
        sequence.run(ev.n_events, dev_events=raw.data_ptr(), dev_offsets=off.data_ptr(),
                     events_size=ev.events.nbytes, batched=True, gec=g)
