"""Global event cut oracle (oracle_global_event_cut, GlobalEventCut.cu:3-55).

Pinned by hand-built events whose cluster counts are known by
construction, at both sides of every boundary the cut has: the SciFi
estimate alone (bytes / 2 - 2 against 9750), UT plus SciFi, UT alone, the
4-byte header word left out of each SciFi bank, the high half of the UT
bank word ignored, and the unsigned wrap that cuts events without SciFi
data.  On the reference's own sample (its MDF file's UT and FTCluster
banks, when present) the first event holds 2108 UT hits and an estimate
of 7776 SciFi clusters, 9884 >= 9750, and is the one event cut.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as po
from cuda_hlt_amd import gen

REF_MDF = Path("/root/reference/input/minbias/banks/mdf/upgrade_mc_minbias_scifi_v5.mdf")
MAX = 9750


def blob(banks):
    """Raw event of (source, words) banks."""
    offs, payload = [0], []
    for s, w in banks:
        payload += [s] + list(w)
        offs.append(4 * len(payload))
    return np.array([len(banks)] + offs + payload, np.uint32).view(np.uint8)


def concat(evs):
    offs = np.zeros(len(evs) + 1, np.uint32)
    offs[1:] = np.cumsum([e.nbytes for e in evs])
    return np.concatenate(evs), offs


def scifi_event(n_clusters, n_banks=4):
    """SciFi event whose estimate (data bytes / 2 - 2) is n_clusters (even)."""
    assert n_clusters % 2 == 0
    words = (n_clusters + 2) // 2          # 4 data bytes = 2 clusters
    per = [words // n_banks + (1 if b < words % n_banks else 0) for b in range(n_banks)]
    return blob([(b, [0xABCD0000 + i for i in range(k)]) for b, k in enumerate(per)])


def ut_event(hits_per_bank, high=0x0080):
    """UT event: each bank's first data word holds its hit count in the low
    16 bits (high bits set as in real banks) followed by the hit words."""
    banks = []
    for b, h in enumerate(hits_per_bank):
        banks.append((b, [(high << 16) | h] + [0x12345678] * ((h + 1) // 2)))
    return blob(banks)


def gec(ut_events, scifi_events):
    ut, uo = concat(ut_events)
    sf, so = concat(scifi_events)
    return list(po.global_event_cut(ut, uo, sf, so))


@pytest.mark.parametrize("scifi,ut,keep", [
    (9748, [], True),
    (9748, [1], True),              # 9749
    (9748, [2], False),             # 9750
    (9748, [1, 1], False),
    (9750, [], False),              # the SciFi estimate alone
    (0, [9749], True),
    (0, [9750], False),             # UT alone (per-bank early exit)
    (0, [5000, 4749], True),
    (0, [5000, 4750], False),
    (4000, [3000, 2749], True),
    (4000, [3000, 2750], False),
])
def test_boundaries(scifi, ut, keep):
    assert gec([ut_event(ut)], [scifi_event(scifi)]) == ([0] if keep else [])


def test_scifi_header_word_is_not_counted():
    """Banks holding only the source word add 4 - 4 = 0 bytes: the estimate
    stays that of the data banks."""
    data = scifi_event(9748)
    w = data.view(np.uint32)
    n = int(w[0])
    banks = []
    offs = w[1:2 + n]
    for b in range(n):
        lo, hi = 2 + n + offs[b] // 4, 2 + n + offs[b + 1] // 4
        banks.append((b, list(w[lo + 1:hi])))
    banks += [(100 + k, []) for k in range(30)]
    assert gec([ut_event([1])], [blob(banks)]) == [0]
    assert gec([ut_event([2])], [blob(banks)]) == []


def test_ut_high_bits_ignored():
    assert gec([ut_event([9749], high=0xFFFF)], [scifi_event(0)]) == [0]


@pytest.mark.parametrize("scifi_banks", [[], [(0, [])], [(0, []), (1, [])]])
def test_no_scifi_data_wraps_and_is_cut(scifi_banks):
    """Under 4 data bytes: bytes / 2 - 2 wraps in unsigned arithmetic."""
    assert gec([ut_event([0])], [blob(scifi_banks)]) == []
    # 4 data bytes: estimate 0, kept
    assert gec([ut_event([0])], [blob([(0, [7])])]) == [0]


def test_kept_events_in_increasing_order():
    rng = np.random.default_rng(3)
    uts, sfs, expect = [], [], []
    for e in range(200):
        s = 2 * int(rng.integers(0, 5500))
        u = int(rng.integers(0, 6000))
        split = sorted(rng.integers(0, u + 1, size=3))
        hits = np.diff([0, *split, u]).tolist()
        uts.append(ut_event(hits))
        sfs.append(scifi_event(s, n_banks=int(rng.integers(1, 9))))
        if s + u < MAX:
            expect.append(e)
    assert gec(uts, sfs) == expect


def test_synthetic_gec_banks():
    """gen.gec_banks: layout of real UT/FTCluster events (180 / 240 banks)
    and the stated mix of kept and cut events."""
    ev = gen.minbias(300, seed=5)
    g = gen.gec_banks(ev, seed=5)
    assert len(g.ut_offsets) == len(g.scifi_offsets) == ev.n_events + 1
    kept = po.global_event_cut(g.ut_events, g.ut_offsets, g.scifi_events, g.scifi_offsets)
    frac = 1 - len(kept) / ev.n_events
    assert 0.02 < frac < 0.4
    for e in (0, 17):
        u = g.ut_events[g.ut_offsets[e]:g.ut_offsets[e + 1]].view(np.uint32)
        s = g.scifi_events[g.scifi_offsets[e]:g.scifi_offsets[e + 1]].view(np.uint32)
        assert u[0] == 180 and s[0] == 240
    # the counts it reports are the ones the cut sees
    n = 0
    for e in range(ev.n_events):
        if g.ut_hits[e] + g.scifi_clusters[e] < MAX and g.scifi_clusters[e] < MAX:
            n += 1
    assert n == len(kept)


@pytest.mark.skipif(not REF_MDF.exists(), reason="reference sample absent")
def test_reference_sample():
    from cuda_hlt_amd import mdf
    r = mdf.read_banks(REF_MDF, [mdf.BANK_UT, mdf.BANK_FT_CLUSTER])
    ut, ft = r[mdf.BANK_UT], r[mdf.BANK_FT_CLUSTER]
    kept = po.global_event_cut(ut.events, ut.offsets, ft.events, ft.offsets)
    assert list(kept) == list(range(1, 10))
