"""Pins the oracle's triplet search at both sides of every cut (CPU only).

Hits are placed by hand as float coordinates (no pixel quantisation): each
event is a list of (module, x, y) hits, z being the module's z.  They go
into the raw cluster container (x, y, z, LHCbID slots) and through the
oracle's own phi sort, candidate windows, search and weak-track stages.
Every expectation is a closed form of the placement, compared with the
reference's cut values (cuda/velo/common/include/VeloDefinitions.cuh:45-77):

* seeding (TrackSeeding.cu:137-151): |x1-x0| and |y1-y0| < 0.4 (z0-z1),
  |dx|, |dy| < 0.6 mm at h2 for the h0-h1 line, scatter < 0.004;
* forwarding (TrackForwarding.cu:12-36, 95-177): the next module's phi
  window of +-0.052 around the extrapolated point, |dx|, |dy| < 0.6 mm,
  a track may miss one module (skip), a tracklet whose forwarding fails
  twice becomes a weak track, hits are marked used once a track has 4;
* weak tracks (WeakTracksAdder.cu:104-111): kept iff no hit is used and
  chi2/ndf < 20 with weights param_w = 3966.94.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

N_MODULES = po.N_MODULES
PARAM_W = 3966.94
MAX_CHI2 = 20.0


def zs():
    from cuda_hlt_amd import gen
    return gen.module_zs().astype(np.float64)


def reco(events, weak_adder=True):
    """events: list of hit lists [(module, x, y), ...].  Returns per event
    a sorted list of tracks (tuples of hit ids, id = position in the
    event's hit list) and the search outputs."""
    z = zs()
    n = len(events)
    counts = np.zeros(n * N_MODULES, np.uint32)
    per_module = []
    for e, hits in enumerate(events):
        mods = [[] for _ in range(N_MODULES)]
        for i, (m, x, y) in enumerate(hits):
            mods[m].append((i, x, y))
        per_module.append(mods)
        for m in range(N_MODULES):
            counts[e * N_MODULES + m] = len(mods[m])
    starts = po.prefix_sum(np.concatenate([counts, [0]]).astype(np.uint32))
    total = int(starts[-1])
    cont = np.zeros(6 * total + 1, np.uint32)
    X = cont[0:total].view(np.float32)
    Y = cont[total:2 * total].view(np.float32)
    Z = cont[2 * total:3 * total].view(np.float32)
    ID = cont[3 * total:4 * total]
    for e in range(n):
        for m in range(N_MODULES):
            s = int(starts[e * N_MODULES + m])
            for k, (i, x, y) in enumerate(per_module[e][m]):
                X[s + k], Y[s + k], Z[s + k], ID[s + k] = x, y, z[m], i
    cont_s, _ = po.calculate_phi_and_sort(n, starts, counts, cont)
    h0, h2 = po.fill_candidates(n, cont_s, starts, counts)
    tracks, weak, used, atomics, st = po.search_by_triplet(n, cont_s, starts, counts, h0, h2, z.astype(np.float32))
    if weak_adder:
        tracks, atomics, _ = po.weak_tracks_adder(n, cont_s, starts, tracks, weak, used, atomics)
    ids = cont_s[2 * total:3 * total]
    out = []
    for e in range(n):
        off = int(starts[e * N_MODULES])
        tr = tracks.reshape(n, po.MAX_TRACKS)[e]
        out.append(sorted(tuple(sorted(int(ids[off + h]) for h in tr[t]["hits"][: tr[t]["hitsNum"]]))
                          for t in range(int(atomics[e]))))
    return out, dict(used=used, weak=weak, atomics=atomics, starts=starts, ids=ids, status=st, cont=cont_s)


def seeds(hits):
    """Number of weak tracks the search leaves for one event (a seeded
    tracklet that is never forwarded ends as one)."""
    _, out = reco([hits], weak_adder=False)
    return int(out["atomics"][1])


def line(modules, tx, ty, x0=0.0, y0=0.0):
    """Hits of the straight line x = x0 + tx z, y = y0 + ty z on the given modules."""
    z = zs()
    return [(m, x0 + tx * z[m], y0 + ty * z[m]) for m in modules]


ODD = [51, 49, 47, 45, 43, 41, 39, 37]   # z 750 ... 387.5 (one half of the detector)


# --------------------------------------------------------------------------
# seeding
# --------------------------------------------------------------------------

@pytest.mark.parametrize("slope,found", [(0.38, True), (0.42, False)])
@pytest.mark.parametrize("axis", [0, 1])
def test_seeding_slope_cut(slope, found, axis):
    """|x1 - x0| < 0.4 (z0 - z1) and the same in y: three hits of a line
    through the origin seed iff its slope on that axis is below 0.4; the
    seed (no residuals) is kept as a 3-hit track."""
    hits = line(ODD[:3], slope, 0.0) if axis == 0 else line(ODD[:3], 0.0, slope)
    assert seeds(hits) == (1 if found else 0)
    tracks, _ = reco([hits])
    assert tracks[0] == ([(0, 1, 2)] if found else [])


@pytest.mark.parametrize("shift,found", [(0.55, True), (0.65, False)])
@pytest.mark.parametrize("axis", [0, 1])
def test_seeding_tolerance_at_h2(shift, found, axis):
    """h2 displaced from the h0-h1 line by `shift` mm on one axis: the
    tolerance |dx|, |dy| < 0.6 mm decides (scatter = shift^2 / 50^2 is far
    below 0.004)."""
    hits = line(ODD[:3], 0.01, 0.02)
    m, x, y = hits[2]
    hits[2] = (m, x + shift, y) if axis == 0 else (m, x, y + shift)
    assert seeds(hits) == (1 if found else 0)


def test_seeding_takes_least_scatter_h2():
    """Two h2 candidates inside the tolerance: the one closer to the line."""
    hits = line(ODD[:3], 0.01, 0.0)
    m, x, y = hits[2]
    hits[2] = (m, x + 0.3, y)
    hits.append((m, x - 0.1, y))   # id 3, closer
    _, out = reco([hits], weak_adder=False)
    assert int(out["atomics"][1]) == 1
    ids = out["ids"]
    assert sorted(int(ids[h]) for h in out["weak"][0]["hits"]) == [0, 1, 3]


# --------------------------------------------------------------------------
# forwarding
# --------------------------------------------------------------------------

@pytest.mark.parametrize("shift,found", [(0.55, True), (0.65, False)])
@pytest.mark.parametrize("axis", [0, 1])
def test_forwarding_tolerance(shift, found, axis):
    """The last hit of a 5-hit line displaced by `shift` mm on one axis
    (inside the phi window): added within 0.6 mm, else the track keeps its
    first four hits."""
    hits = line(ODD[:5], 0.02, -0.01)
    m, x, y = hits[4]
    hits[4] = (m, x + shift, y) if axis == 0 else (m, x, y + shift)
    tracks, _ = reco([hits])
    assert tracks[0] == ([(0, 1, 2, 3, 4)] if found else [(0, 1, 2, 3)])


@pytest.mark.parametrize("dphi,found", [(0.045, True), (0.060, False), (-0.045, True), (-0.060, False)])
def test_forwarding_phi_window(dphi, found):
    """At a radius of about 3 mm the phi window of +-0.052 around the
    extrapolated point decides before the 0.6 mm tolerance: the last hit of
    a 5-hit line rotated by dphi about the beam axis (a displacement of
    0.14-0.19 mm)."""
    hits = line(ODD[:5], 0.0065, 0.0)
    m, x, y = hits[4]
    r, p = math.hypot(x, y), math.atan2(y, x)
    assert 2.5 < r < 4.0 and r * abs(dphi) < 0.3
    hits[4] = (m, r * math.cos(p + dphi), r * math.sin(p + dphi))
    tracks, _ = reco([hits])
    assert tracks[0] == ([(0, 1, 2, 3, 4)] if found else [(0, 1, 2, 3)])


def test_forwarding_kink_loses_the_hits_after_it():
    """A hit accepted 0.55 mm off the line turns the extrapolation: hits on
    the line 150 and 250 mm further are then 1.65 and 2.75 mm off, so the
    track skips once and ends with four hits."""
    hits = line(ODD[:6], 0.02, -0.01)
    m, x, y = hits[3]
    hits[3] = (m, x, y + 0.55)
    tracks, _ = reco([hits])
    assert tracks[0] == [(0, 1, 2, 3)]


def test_forwarding_takes_least_scatter_hit():
    """Two hits inside the tolerance on the last module: the closer one."""
    hits = line(ODD[:5], 0.02, 0.0)
    m, x, y = hits[4]
    hits[4] = (m, x + 0.4, y)
    hits.append((m, x - 0.2, y))  # id 5
    tracks, _ = reco([hits])
    assert tracks[0] == [(0, 1, 2, 3, 5)]


def test_one_missed_module_is_skipped():
    """A track missing one module keeps going (skip once)."""
    hits = [h for k, h in enumerate(line(ODD[:7], 0.015, 0.01)) if k != 4]
    tracks, _ = reco([hits])
    assert tracks[0] == [tuple(range(6))]


def test_two_missed_modules_end_the_track():
    """Missing two modules in a row ends the track: its first three hits
    (never forwarded) become a weak track, kept by the weak-track adder;
    the hits after the gap form a track of their own."""
    full = line(ODD[:8], 0.015, 0.01)
    hits = [full[k] for k in (0, 1, 2, 5, 6, 7)]
    tracks, _ = reco([hits])
    assert tracks[0] == [(0, 1, 2), (3, 4, 5)]
    assert reco([hits], weak_adder=False)[0][0] == []   # both were weak tracks


def test_hits_marked_used_only_from_four_hits():
    """A forwarded track (4+ hits) marks its hits used; a 3-hit tracklet
    that is never forwarded leaves its hits unused."""
    four = line(ODD[:4], 0.01, 0.0)
    three = line([31, 29, 27], -0.03, 0.02)  # z 212.5 ... 162.5, far from the first
    _, out = reco([four + three])
    s = out["starts"]
    used = out["used"][int(s[0]): int(s[N_MODULES])]
    ids = out["ids"][int(s[0]): int(s[N_MODULES])]
    used_ids = sorted(int(i) for i, u in zip(ids, used) if u)
    assert used_ids == [0, 1, 2, 3]


# --------------------------------------------------------------------------
# weak tracks adder (called directly on hand-built weak tracks)
# --------------------------------------------------------------------------

def _weak_setup(used_hit=None):
    """One event, the three hits of a line on modules 51, 49, 47, listed as
    one weak track; `used_hit` (an id) marked used."""
    _, out = reco([line(ODD[:3], 0.01, 0.0)], weak_adder=False)
    total = int(out["starts"][-1])
    ids = out["ids"]
    by_id = {int(ids[h]): h for h in range(total)}
    weak = np.zeros(po.MAX_WEAK_TRACKS, po.TRACKLET_DTYPE)
    weak[0]["hits"] = [by_id[0], by_id[1], by_id[2]]
    atomics = np.zeros(po.NUM_ATOMICS + 1, np.int32)
    atomics[1] = 1  # one weak track
    used = np.zeros(total + 1, np.uint8)
    if used_hit is not None:
        used[by_id[used_hit]] = 1
    return out, np.zeros(po.MAX_TRACKS, po.TRACK_DTYPE), weak, used, atomics


@pytest.mark.parametrize("delta", [0.05, 0.10, 0.118, 0.128, 0.15, 0.25])
def test_weak_track_chi2_cut(delta):
    """Middle hit displaced by delta from the line through three equally
    spaced hits: the least-squares residuals are delta/3 (-1, 2, -1), so
    chi2/ndf = w (2 delta^2 / 3) / 2 = w delta^2 / 3; kept iff < 20
    (delta < 0.1230 mm)."""
    hits = line(ODD[:3], 0.01, 0.0)
    m, x, y = hits[1]
    hits[1] = (m, x + delta, y)
    chi2 = PARAM_W * delta * delta / 3.0
    assert abs(chi2 - MAX_CHI2) > 0.5  # away from the float rounding of the boundary
    tracks, _ = reco([hits])
    assert (tracks[0] == [(0, 1, 2)]) == (chi2 < MAX_CHI2)


@pytest.mark.parametrize("used_hit", [None, 0, 1, 2])
def test_weak_track_with_a_used_hit_is_rejected(used_hit):
    out, tracks, weak, used, atomics = _weak_setup(used_hit)
    t, a, _ = po.weak_tracks_adder(1, out["cont"], out["starts"], tracks, weak, used, atomics)
    assert int(a[0]) == (1 if used_hit is None else 0)
    if used_hit is None:
        assert int(t[0]["hitsNum"]) == 3


# --------------------------------------------------------------------------
# a13: copy_and_prefix_sum
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 7, 1000, 4097])
def test_copy_and_prefix_sum_closed_form(n):
    rng = np.random.default_rng(n)
    atomics = np.zeros(n * po.NUM_ATOMICS + 1, np.int32)
    counts = rng.integers(0, po.MAX_TRACKS + 1, n)
    atomics[:n] = counts
    out = po.copy_and_prefix_sum(n, atomics)
    np.testing.assert_array_equal(out[:n], counts)
    want = np.concatenate([[0], np.cumsum(counts)])
    np.testing.assert_array_equal(out[n: 2 * n + 1].astype(np.int64), want)
