"""Pins the CPU oracle (CPU only, no GPU).

Each oracle function is checked against something independent of its own
formula: the reference's own cache_sp_patterns (compiled into oracle/_ref
from the reference sources), brute force over all inputs, a textbook
connected-component clustering, closed forms (scans, least squares, the
affine transform), libm's atan2, and known-answer tracks built from
straight lines through the detector.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import rawevent


# --------------------------------------------------------------------------
# a3: SP pattern tables
# --------------------------------------------------------------------------

def brute_force_sp(sp):
    """8-connected components of the 2x4 SP pattern, ordered by lowest bit."""
    bits = [b for b in range(8) if sp >> b & 1]
    label = {}
    comps = []
    for b in bits:
        if b in label:
            continue
        comp, stack = [], [b]
        label[b] = len(comps)
        while stack:
            p = stack.pop()
            comp.append(p)
            for q in bits:
                if q not in label and abs(q % 4 - p % 4) <= 1 and abs(q // 4 - p // 4) <= 1:
                    label[q] = len(comps)
                    stack.append(q)
        comps.append(comp)
    return comps


def test_sp_lut_brute_force():
    p, s, fx, fy = po.sp_lut()
    for sp in range(256):
        comps = brute_force_sp(sp)
        assert len(comps) <= 2
        want_p = 0
        for k, comp in enumerate(comps):
            x = sum(b // 4 for b in comp)
            y = sum(b % 4 for b in comp)
            n = len(comp)
            cx, cy = x // n, y // n
            want_p |= ((cx << 2) | cy) << (4 * k)
            want_p |= k << 3
            assert s[sp] >> (4 * k) & 0xF == n
            assert fx[2 * sp + k] == np.float32(np.float32(x) / np.float32(n) - np.float32(cx))
            assert fy[2 * sp + k] == np.float32(np.float32(y) / np.float32(n) - np.float32(cy))
        assert p[sp] == want_p, sp


def test_sp_lut_matches_reference_build():
    ref = po.ref_sp_lut()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    mine = po.sp_lut()
    for a, b in zip(mine, ref):
        np.testing.assert_array_equal(a, b)


def test_product_sp_tables_match_oracle():
    from cuda_hlt_amd import native
    for a, b in zip(native.sp_tables(), po.sp_lut()):
        np.testing.assert_array_equal(a, b)


# --------------------------------------------------------------------------
# phi
# --------------------------------------------------------------------------

def _ulp(v):
    return np.spacing(np.abs(np.float32(v))).astype(np.float64)


def test_atan2_accuracy():
    rng = np.random.default_rng(1)
    pts = rng.uniform(-50, 50, size=(20000, 2)).astype(np.float32)
    pts = np.vstack([pts, [[1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, -1], [1e-3, 40], [40, -1e-3],
                           [-5.1, 0.0], [-5.1, -0.0], [3.0, 3.0000002]]]).astype(np.float32)
    L = po.lib()
    worst = 0.0
    for x, y in pts:
        got = L.oracle_atan2f(float(y), float(x))
        want = math.atan2(float(y), float(x))
        worst = max(worst, abs(got - want) / max(_ulp(want), 1e-45))
    assert worst <= 3.0, worst


def test_phi_conventions():
    L = po.lib()
    for x, y in [(10.0, 0.1), (-10.0, 0.1), (-10.0, -0.1), (0.0, -7.0), (3.0, -3.0)]:
        odd = L.oracle_phi_odd(x, y)
        even = L.oracle_phi_even(x, y)
        assert -math.pi <= odd <= math.pi
        assert 0.0 <= even < 2 * math.pi + 1e-6
        assert abs((even - odd) % (2 * math.pi)) < 1e-5 or abs((even - odd) % (2 * math.pi) - 2 * math.pi) < 1e-5


# --------------------------------------------------------------------------
# scans
# --------------------------------------------------------------------------

@pytest.mark.parametrize("size", [0, 1, 511, 512, 513, 2047, 2048, 2049, 1000000])
def test_prefix_sum_closed_form(size):
    a = np.zeros(size + 1, np.uint32)
    a[:size] = np.random.default_rng(size).integers(0, 1 << 10, size)
    got = po.prefix_sum(a)
    want = np.concatenate([[0], np.cumsum(a[:size], dtype=np.uint64)])
    np.testing.assert_array_equal(got.astype(np.uint64), want)


# --------------------------------------------------------------------------
# a4/a6: estimate and clustering
# --------------------------------------------------------------------------

def test_estimate_isolated_sp_all_patterns():
    """One isolated SP per event: the estimate equals its component count."""
    evs = []
    for sp in range(1, 256):
        pix = [(17, 2 * 100 + b // 4, 4 * 30 + b % 4) for b in range(8) if sp >> b & 1]
        evs.append(rawevent.encode(pix))
    events, offsets = rawevent.concat(evs)
    est, num, cnum, cand, st = po.estimate_input_size(events, offsets)
    for i, sp in enumerate(range(1, 256)):
        assert est[i * 52 + 4] == len(brute_force_sp(sp))
        assert est[i * 52: i * 52 + 52].sum() == est[i * 52 + 4]
    assert not cnum.any()


def candidate_pixels_brute(pixels):
    """Pixels whose top, top-right, right and bottom-right neighbours are empty."""
    occ = set(pixels)
    return {(c, r) for (c, r) in pixels
            if not {(c, r + 1), (c + 1, r + 1), (c + 1, r), (c + 1, r - 1)} & occ}


def test_estimate_candidates_brute_force():
    """Candidates of non-isolated SPs on random small sensors."""
    rng = np.random.default_rng(3)
    for trial in range(60):
        pix = set()
        for _ in range(rng.integers(2, 40)):
            pix.add((int(rng.integers(40, 52)), int(rng.integers(20, 36))))
        events, offsets = rawevent.encode([(5, c, r) for c, r in pix])
        est, num, cnum, cand, st = po.estimate_input_size(events, offsets)
        # expected: per non-isolated SP, candidates in its low (rows 0-1) and
        # high (rows 2-3) halves; isolated SPs count their components
        sps = {}
        for c, r in pix:
            sps.setdefault((c // 2, r // 4), []).append((c, r))
        cands = candidate_pixels_brute(pix)
        want_est = 0
        want_n = 0
        for (sc, sr), members in sps.items():
            isolated = not any((sc + dc, sr + dr) in sps for dc in (-1, 0, 1) for dr in (-1, 0, 1) if dc or dr)
            if isolated:
                sp = 0
                for c, r in members:
                    sp |= 1 << ((c % 2) * 4 + r % 4)
                want_est += len(brute_force_sp(sp))
            else:
                mine = [p for p in members if p in cands]
                low = [p for p in mine if p[1] % 4 < 2]
                high = [p for p in mine if p[1] % 4 >= 2]
                assert len(low) <= 1 and len(high) <= 1
                want_est += len(low) + len(high)
                want_n += len(low) + len(high)
        assert est[1] == want_est, trial
        assert cnum[0] == want_n, trial


def ccl_by_module(event):
    ids, mods, sx, sy, n = po.ccl_event(event)
    return ids, mods, sx, sy, n


def test_clustering_equals_connected_components(minbias_200, geometry):
    """Masked clustering equals textbook 8-connected labelling, id for id, on
    every event whose clusters all have <= 3 pixels (they fit the 6x16 window
    and converge), and finds every <= 3-pixel cluster of the other events."""
    ev = minbias_200
    est, *_ = po.estimate_input_size(ev.events, ev.offsets)
    starts = po.prefix_sum(est)
    nums, cont, st = po.masked_clustering(ev.events, ev.offsets, starts, geometry)
    total = int(starts[-1])
    exact_events = 0
    for e in range(ev.n_events):
        lo, hi = int(ev.offsets[e]), int(ev.offsets[e + 1])
        ids, mods, sx, sy, npx = ccl_by_module(ev.events[lo:hi].copy())
        got = []
        for m in range(52):
            s, k = int(starts[e * 52 + m]), int(nums[e * 52 + m])
            got.extend(cont[3 * total + s: 3 * total + s + k].tolist())
        if np.all(npx <= 3):
            assert sorted(got) == sorted(ids.tolist()), e
            exact_events += 1
        else:
            assert set(ids[npx <= 3].tolist()) <= set(got), e
    assert exact_events > ev.n_events // 2


def test_cluster_coordinates_closed_form(minbias_200, geometry):
    """x,y,z = R (local_x[cx] + fx pitch, (cy+0.5+fy) 0.055, 0) + T within 1e-5 mm."""
    ev = minbias_200.subset(0, 40)
    est, *_ = po.estimate_input_size(ev.events, ev.offsets)
    starts = po.prefix_sum(est)
    nums, cont, st = po.masked_clustering(ev.events, ev.offsets, starts, geometry)
    total = int(starts[-1])
    frames, local_x, pitch = rawevent.sensor_frames(geometry)
    checked = 0
    for e in range(ev.n_events):
        lo, hi = int(ev.offsets[e]), int(ev.offsets[e + 1])
        ids, mods, sx, sy, npx = ccl_by_module(ev.events[lo:hi].copy())
        by_id = {int(i): (int(a), int(b), int(c)) for i, a, b, c in zip(ids, sx, sy, npx)}
        for m in range(52):
            s, k = int(starts[e * 52 + m]), int(nums[e * 52 + m])
            for j in range(s, s + k):
                cid = int(cont[3 * total + j])
                if cid not in by_id:
                    continue
                x, y, n = by_id[cid]
                sensor = (cid & 0x0FFFFFFF) >> 18
                cx, cy = x // n, y // n
                lx = local_x[cx] + (x / n - cx) * pitch[cx]
                ly = (cy + 0.5 + (y / n - cy)) * 0.055
                R, T = frames[sensor]
                g = R @ np.array([lx, ly, 0.0]) + T
                got = cont[[j, total + j, 2 * total + j]].view(np.float32).astype(np.float64)
                assert np.max(np.abs(got - g)) < 1e-5
                checked += 1
    assert checked > 1000


def test_sp_index_truncation_flag():
    """A bank with more than 256 SPs is flagged (the reference decodes the SP
    index of a candidate through a uint8_t)."""
    pix = [(9, 2 * c, 4 * r) for c in range(0, 380, 2) for r in range(0, 64, 2)][:600]
    events, offsets = rawevent.encode(pix)
    *_, st = po.estimate_input_size(events, offsets)
    assert st[0] & po.STATUS_BITS["SP_INDEX_TRUNCATED"]


def test_bad_bank_flag():
    events, offsets = rawevent.encode([(3, 10, 10)])
    w = events.view(np.uint32).copy()
    # bank 0 claims to be sensor 1
    w[1 + 209] = 1
    *_, st = po.estimate_input_size(w.view(np.uint8), offsets)
    assert st[0] & po.STATUS_BITS["BAD_BANK"]


def test_huge_bank_is_flagged_and_skipped():
    """A bank of 2^22 SPs or more (16 MB) is flagged BAD_BANK and skipped,
    even when its words lie inside the event: the device decoder sums the SP
    counts of an event's banks in 32 bits."""
    events, offsets = rawevent.encode([(5, 10, 10), (9, 20, 20)])
    w = events.view(np.uint32)
    nb = int(w[0])
    h = nb + 2 + int(w[1 + 5]) // 4
    assert w[h] == 5 and w[h + 1] == 1
    big = np.zeros(len(w) + (1 << 22), np.uint32)
    big[: len(w)] = w
    big[h + 1] = 1 << 22
    est, _, _, _, st = po.estimate_input_size(big.view(np.uint8), np.array([0, 4 * len(big)], np.uint32))
    assert st[0] & po.STATUS_BITS["BAD_BANK"]
    assert est[1] == 0 and est[2] == 1  # sensor 5 (module 1) skipped, sensor 9 (module 2) kept
    big[h + 1] = 1
    est, _, _, _, st = po.estimate_input_size(big.view(np.uint8), np.array([0, 4 * len(big)], np.uint32))
    assert not st[0] & po.STATUS_BITS["BAD_BANK"] and est[1] == 1


# --------------------------------------------------------------------------
# a7/a8: sort and windows
# --------------------------------------------------------------------------

def test_sort_is_stable_phi_order(minbias_200, geometry):
    ev = minbias_200.subset(0, 50)
    r = po.pipeline(ev.events, ev.offsets, geometry, np.zeros(52, np.float32))
    total = int(r.starts[-1])
    phi = r.container[4 * total: 5 * total].view(np.float32)
    x = r.container[5 * total: 6 * total].view(np.float32)
    y = r.container[0: total].view(np.float32)
    for e in range(ev.n_events):
        for m in range(52):
            s, k = int(r.starts[e * 52 + m]), int(r.nums[e * 52 + m])
            p = phi[s:s + k]
            assert np.all(np.diff(p) >= 0)
            want = np.arctan2(y[s:s + k].astype(np.float64), x[s:s + k].astype(np.float64))
            if m % 2 == 0:
                want = np.where(want < 0, want + 2 * np.pi, want)
            assert np.max(np.abs(want - p), initial=0) < 1e-5
        # holes read as NaN
        for m in range(52):
            s, k, nxt = int(r.starts[e * 52 + m]), int(r.nums[e * 52 + m]), int(r.starts[e * 52 + m + 1])
            assert np.all(np.isnan(phi[s + k: nxt]))


def _first_brute(a, v, margin):
    """Reference semantics restated by exhaustive simulation of the probes."""
    found = False
    l, r = 0, len(a) - 1
    while l < r:
        m = (l + r) // 2
        found |= abs(np.float32(v - a[m])) < margin
        if np.float32(v - margin) > a[m]:
            l = m + 1
        else:
            r = m
    found |= abs(np.float32(v - a[l])) < margin
    return l if found else -1


def test_windows_bracket_all_compatible(minbias_200, geometry):
    """[first, last) of every h1 contains every h0 with |dphi| < 0.062 strictly
    inside, and nothing outside the margin except the boundary probe."""
    ev = minbias_200.subset(0, 30)
    est, *_ = po.estimate_input_size(ev.events, ev.offsets)
    starts = po.prefix_sum(est)
    nums, cont, _ = po.masked_clustering(ev.events, ev.offsets, starts, geometry)
    n = ev.n_events
    cont, _ = po.calculate_phi_and_sort(n, starts, nums, cont)
    h0, h2 = po.fill_candidates(n, cont, starts, nums)
    total = int(starts[-1])
    phi = cont[4 * total: 5 * total].view(np.float32)
    margin = np.float32(0.062)
    for e in range(n):
        off = int(starts[e * 52])
        for m in range(2, 50):
            s1, k1 = int(starts[e * 52 + m]), int(nums[e * 52 + m])
            s0, k0 = int(starts[e * 52 + m + 2]), int(nums[e * 52 + m + 2])
            for h1 in range(s1, s1 + k1):
                f, l = int(h0[2 * h1]), int(h0[2 * h1 + 1])  # stored event-relative
                p = phi[h1]
                inside = [j for j in range(s0, s0 + k0) if abs(np.float32(p - phi[j])) < margin]
                if k0 == 0:
                    assert f == -1
                    continue
                fb = _first_brute(phi[s0:s0 + k0], p, margin)
                assert f == (-1 if fb == -1 else fb + s0 - off)
                if inside:
                    assert f != -1
                    assert f + off <= inside[0] and inside[-1] < l + off


# --------------------------------------------------------------------------
# a9-a12, a16: tracking known answers, invariants, fit
# --------------------------------------------------------------------------

def _reco(events, offsets, geometry, module_zs):
    return po.pipeline(events, offsets, geometry, module_zs)


def _track_ids(r, e):
    n = r.n_events
    T = int(r.atomics[2 * n])
    Ht = int(r.hit_number[T])
    ids = r.track_hits[12 * Ht: 16 * Ht].view(np.uint32)
    off = int(r.atomics[n + e])
    return [sorted(ids[r.hit_number[off + t]: r.hit_number[off + t + 1]].tolist())
            for t in range(int(r.atomics[e]))]


def _pixel_id(s, c, r):
    return 0x80000000 + ((s << 18) | ((c // 256) << 16) | ((c % 256) << 8) | r)


def test_single_track_known_answer(geometry, module_zs):
    """A track near phi = 0 crosses only odd (right-half) modules: one track
    holding exactly its clusters.  (A track crossing both halves at one
    station keeps only the better of the two hits: forwarding examines the
    odd and even module of a pair in one step, TrackForwarding.cu:95-124.)"""
    pix = rawevent.line_pixels(geometry, 0.0, 0.0, 0.0, 0.05, 0.002)
    assert len({s // 4 % 2 for s, _, _ in pix}) == 1
    assert len(pix) >= 6
    events, offsets = rawevent.encode(pix)
    r = _reco(events, offsets, geometry, module_zs)
    tr = _track_ids(r, 0)
    assert len(tr) == 1
    assert tr[0] == sorted(_pixel_id(*p) for p in pix)


def test_two_tracks_known_answer(geometry, module_zs):
    a = rawevent.line_pixels(geometry, 0.0, 0.0, 0.0, 0.05, 0.002)
    b = rawevent.line_pixels(geometry, 0.0, 0.0, 10.0, -0.05, -0.003)
    assert {s // 4 % 2 for s, _, _ in a} == {1} and {s // 4 % 2 for s, _, _ in b} == {0}
    events, offsets = rawevent.encode(a + b)
    r = _reco(events, offsets, geometry, module_zs)
    tr = sorted(_track_ids(r, 0))
    assert tr == sorted([sorted(_pixel_id(*p) for p in a), sorted(_pixel_id(*p) for p in b)])


def test_backward_track_known_answer(geometry, module_zs):
    pix = rawevent.line_pixels(geometry, 0.0, 0.0, 0.0, 0.05, -0.002, forward=False)
    assert len(pix) >= 4
    events, offsets = rawevent.encode(pix)
    r = _reco(events, offsets, geometry, module_zs)
    tr = _track_ids(r, 0)
    assert tr == [sorted(_pixel_id(*p) for p in pix)]
    T = int(r.atomics[2])
    backward = r.states[48 * T: 48 * T + T]
    assert backward[0] == 1


def test_three_hit_track_becomes_weak_track(geometry, module_zs):
    """A straight 3-hit track survives as a weak track (chi2 ~ 0 < 20)."""
    pix = rawevent.line_pixels(geometry, 0.0, 0.0, 0.0, 0.05, 0.002)
    mods = sorted({s // 4 for s, _, _ in pix}, reverse=True)
    keep_mods = mods[:3]
    three = [p for p in pix if p[0] // 4 in keep_mods]
    assert len(three) == 3
    events, offsets = rawevent.encode(three)
    r = _reco(events, offsets, geometry, module_zs)
    tr = _track_ids(r, 0)
    assert tr == [sorted(_pixel_id(*p) for p in three)]


def test_track_invariants_and_efficiency(minbias_200, geometry, module_zs):
    ev = minbias_200
    r = _reco(ev.events, ev.offsets, geometry, module_zs)
    n = ev.n_events
    assert not (r.status & po.STATUS_ERROR_MASK).any()
    trk = r.tracks.reshape(n, po.MAX_TRACKS)
    total = int(r.starts[-1])
    ids_slot = r.container[2 * total: 3 * total]
    for e in range(n):
        off = int(r.starts[e * 52])
        for t in range(int(r.atomics[e])):
            k = int(trk[e, t]["hitsNum"])
            assert 3 <= k <= 26
            h = trk[e, t]["hits"][:k].astype(np.int64)
            ids = ids_slot[off + h]
            assert len(set(ids.tolist())) == k
            mods = ((ids & 0x0FFFFFFF) >> 18) // 4
            assert np.all(np.diff(mods.astype(np.int64)) < 0)
    from tests.checker import velo_efficiency
    eff, ghost, _ = velo_efficiency(ev, r)
    assert eff > 0.95 and ghost < 0.03, (eff, ghost)


def test_consolidated_states_least_squares(minbias_200, geometry, module_zs):
    """States equal the closed-form weighted least squares (double precision)."""
    ev = minbias_200.subset(0, 40)
    r = _reco(ev.events, ev.offsets, geometry, module_zs)
    n = ev.n_events
    T = int(r.atomics[2 * n])
    Ht = int(r.hit_number[T])
    h = r.track_hits[: 16 * Ht].view(np.float32)
    xs, ys, zs = h[:Ht], h[Ht:2 * Ht], h[2 * Ht:3 * Ht]
    S = r.states[: 48 * T].view(np.float32).reshape(12, T)
    for t in range(T):
        a, b = int(r.hit_number[t]), int(r.hit_number[t + 1])
        z = zs[a:b].astype(np.float64)
        A = np.vstack([np.ones_like(z), z]).T
        (x0, tx), *_ = np.linalg.lstsq(A, xs[a:b].astype(np.float64), rcond=None)
        (y0, ty), *_ = np.linalg.lstsq(A, ys[a:b].astype(np.float64), rcond=None)
        zb = -(x0 * tx + y0 * ty) / (tx * tx + ty * ty)
        assert abs(S[2, t] - tx) < 1e-4 and abs(S[3, t] - ty) < 1e-4
        assert abs(S[11, t] - zb) < max(0.05, 1e-3 * abs(zb))
        assert abs(S[0, t] - (x0 + tx * zb)) < 5e-3 and abs(S[1, t] - (y0 + ty * zb)) < 5e-3


def test_unsorted_banks_are_flagged(minbias_200):
    """UNSORTED_BANK is set exactly on the events holding a bank whose SP
    addresses do not strictly increase, judged here from the raw words."""
    from tests import rawevent
    off = minbias_200.offsets
    ev = rawevent.scramble_banks(minbias_200.events, off, seed=3)
    st = po.estimate_input_size(ev, off)[-1]
    st0 = po.estimate_input_size(minbias_200.events, off)[-1]
    expect = []
    for e in range(len(off) - 1):
        w = ev[off[e]:off[e + 1]].view(np.uint32)
        nb, bad = int(w[0]), False
        for b in range(nb):
            s = nb + 2 + int(w[1 + b]) // 4
            addr = ((w[s + 2: s + 2 + int(w[s + 1])] >> 8) & 0x7FFF).astype(np.int64)
            bad |= bool(np.any(np.diff(addr) <= 0))
        expect.append(bad)
    assert any(expect)
    np.testing.assert_array_equal((st & po.STATUS_BITS["UNSORTED_BANK"]) != 0, expect)
    assert not np.any(st0 & po.STATUS_BITS["UNSORTED_BANK"])


def test_damaged_banks_are_flagged_and_skipped(minbias_200):
    """A bank whose offset, header or SP words reach past its event is
    flagged (BAD_BANK) and skipped: estimates, candidates and clusters equal
    those of the same event with that bank emptied."""
    from tests import rawevent
    ev, off, cev, coff, damaged = rawevent.malformed(minbias_200.events, minbias_200.offsets)
    est, num, cnum, cand, st = po.estimate_input_size(ev, off)
    cest, cnum0, ccnum, ccand, cst = po.estimate_input_size(cev, coff)
    assert np.all(st & 0x100), st                     # VELO_STATUS_BAD_BANK on every damaged event
    assert not np.any(cst & 0x100)
    np.testing.assert_array_equal(est, cest)
    np.testing.assert_array_equal(cnum, ccnum)
    np.testing.assert_array_equal(cand, ccand)
    starts = po.prefix_sum(est)
    from cuda_hlt_amd import gen
    g = gen.geometry()
    nums, cont, _ = po.masked_clustering(ev, off, starts, g)
    cnums, ccont, _ = po.masked_clustering(cev, coff, starts, g)
    np.testing.assert_array_equal(nums, cnums)
    np.testing.assert_array_equal(cont, ccont)
    assert est[: 2 * 52].sum() > 0 and est[2 * 52: 3 * 52].sum() == 0   # event 2 has no usable bank
