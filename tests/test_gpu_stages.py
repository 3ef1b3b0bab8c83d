"""Stage-by-stage parity of the gfx950 launchers against the CPU oracle.

Every stage gets the oracle's output of the previous stage as input, so a
mismatch names the stage that caused it.  Integer and index outputs must be
identical; float outputs (x, y, z, phi, track states) must be identical bit
for bit, which is stronger than the 1e-5 mm coordinate tolerance of the
specification (asserted separately in test_coordinates_within_tolerance).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

N_MODULES = 52


@pytest.fixture(scope="module")
def stages():
    from tests import gpu_stages
    return gpu_stages


@pytest.fixture(scope="module")
def consts(gpu, geometry, module_zs):
    from cuda_hlt_amd import native as nat
    c = nat.DeviceConstants(geometry, module_zs)
    yield c
    c.free()


def module_mask(n, starts, nums):
    """Boolean mask over [0, total) of slots that hold clusters."""
    total = int(starts[n * N_MODULES])
    m = np.zeros(total, bool)
    s = starts[: n * N_MODULES].astype(np.int64)
    k = nums[: n * N_MODULES].astype(np.int64)
    for a, b in zip(s[k > 0], (s + k)[k > 0]):
        m[a:b] = True
    return m


class Chain:
    """Oracle outputs of every stage for one event sample."""

    def __init__(self, ev, geometry, module_zs):
        self.ev = ev
        self.n = ev.n_events
        n = self.n
        self.est, self.num0, self.cnum, self.cand, self.st_est = po.estimate_input_size(ev.events, ev.offsets)
        self.starts = po.prefix_sum(self.est)
        self.nums, self.cont_raw, self.st_clu = po.masked_clustering(ev.events, ev.offsets, self.starts, geometry)
        self.cont, self.perm = po.calculate_phi_and_sort(n, self.starts, self.nums, self.cont_raw)
        self.h0, self.h2 = po.fill_candidates(n, self.cont, self.starts, self.nums)
        (self.tracks_s, self.weak, self.used, self.atomics_s, self.st_search) = po.search_by_triplet(
            n, self.cont, self.starts, self.nums, self.h0, self.h2, module_zs)
        self.tracks, self.atomics_w, self.st_weak = po.weak_tracks_adder(
            n, self.cont, self.starts, self.tracks_s, self.weak, self.used, self.atomics_s)
        self.atomics = po.copy_and_prefix_sum(n, self.atomics_w)
        self.hn_raw = po.copy_velo_track_hit_number(n, self.tracks, self.atomics)
        self.hn = po.prefix_sum(self.hn_raw)
        self.hits, self.states = po.consolidate(n, self.atomics, self.tracks, self.hn, self.cont, self.starts)
        self.mask = module_mask(n, self.starts, self.nums)


@pytest.fixture(scope="module")
def chain(minbias_1000, geometry, module_zs):
    return Chain(minbias_1000, geometry, module_zs)


@pytest.fixture(scope="module")
def chain_uneven(geometry, module_zs):
    from cuda_hlt_amd import gen
    return Chain(gen.uneven(300, seed=11), geometry, module_zs)


@pytest.fixture(scope="module")
def chain_scrambled(minbias_200, geometry, module_zs):
    """Banks shuffled, reversed or holding a duplicate SP address: the
    decoder's linear-scan path for banks that are not strictly increasing."""
    from cuda_hlt_amd.gen import Events
    from tests.rawevent import scramble_banks
    ev = minbias_200
    scrambled = Events(scramble_banks(ev.events, ev.offsets, seed=5), ev.offsets, ev.mcps, ev.mc_hits)
    return Chain(scrambled, geometry, module_zs)


def tracks_view(raw_bytes, n):
    return np.frombuffer(raw_bytes.tobytes(), po.TRACK_DTYPE)[: n * po.MAX_TRACKS].reshape(n, po.MAX_TRACKS)


def assert_tracks_equal(a, b, counts):
    for e, c in enumerate(counts):
        c = int(c)
        np.testing.assert_array_equal(a[e, :c]["hitsNum"], b[e, :c]["hitsNum"], err_msg=f"event {e}")
        for t in range(c):
            k = int(a[e, t]["hitsNum"])
            assert np.array_equal(a[e, t]["hits"][:k], b[e, t]["hits"][:k]), f"event {e} track {t}"


def _check_estimate(stages, consts, c: Chain, tuning=(0, 0)):
    est, num, cnum, cand, st = stages.estimate(c.ev.events, c.ev.offsets, consts, tuning)
    n = c.n
    np.testing.assert_array_equal(est[: n * N_MODULES], c.est[: n * N_MODULES])
    assert not num.any()
    np.testing.assert_array_equal(cnum, c.cnum)
    for e in range(n):
        k = int(c.cnum[e])
        np.testing.assert_array_equal(cand[e * 3000: e * 3000 + k], c.cand[e * 3000: e * 3000 + k])
    np.testing.assert_array_equal(st, c.st_est)


def test_estimate_input_size(stages, consts, chain):
    _check_estimate(stages, consts, chain)


def test_estimate_input_size_uneven(stages, consts, chain_uneven):
    _check_estimate(stages, consts, chain_uneven)


def test_estimate_input_size_unsorted_banks(stages, consts, chain_scrambled):
    _check_estimate(stages, consts, chain_scrambled)


@pytest.mark.parametrize("block,tile", [(64, 64), (128, 256), (512, 1024)])
def test_estimate_input_size_launch_shapes(stages, consts, chain, block, tile):
    est, *_ = stages.estimate(chain.ev.events, chain.ev.offsets, consts, tuning=(block, tile))
    np.testing.assert_array_equal(est[: chain.n * N_MODULES], chain.est[: chain.n * N_MODULES])


def test_prefix_sum_cluster_offsets(stages, chain):
    np.testing.assert_array_equal(stages.prefix_sum(chain.est), chain.starts)


@pytest.mark.parametrize("size", [0, 1, 2, 511, 512, 513, 2047, 2048, 2049, 4097, 100003, 1000000])
def test_prefix_sum_sizes(stages, size):
    rng = np.random.default_rng(size)
    a = np.zeros(size + 1, np.uint32)
    a[:size] = rng.integers(0, 1000, size)
    got = stages.prefix_sum(a)
    want = np.zeros(size + 1, np.uint64)
    want[1:] = np.cumsum(a[:size].astype(np.uint64))
    np.testing.assert_array_equal(got.astype(np.uint64), want)


def _check_clustering(stages, consts, c: Chain, tuning=(0, 0), cands=None):
    num, cont, st = stages.clustering(c.ev.events, c.ev.offsets, c.starts, consts, tuning, cands, c.st_est)
    n = c.n
    total = int(c.starts[n * N_MODULES])
    np.testing.assert_array_equal(num[: n * N_MODULES], c.nums)
    for s in range(4):
        np.testing.assert_array_equal(cont[s * total: (s + 1) * total][c.mask],
                                      c.cont_raw[s * total: (s + 1) * total][c.mask], err_msg=f"slot {s}")
    np.testing.assert_array_equal(st, c.st_est | c.st_clu)


def test_masked_clustering(stages, consts, chain):
    _check_clustering(stages, consts, chain)


def test_masked_clustering_uneven(stages, consts, chain_uneven):
    _check_clustering(stages, consts, chain_uneven)


def test_masked_clustering_unsorted_banks(stages, consts, chain_scrambled):
    _check_clustering(stages, consts, chain_scrambled)


@pytest.mark.parametrize("block,tile", [(64, 32), (512, 512)])
def test_masked_clustering_launch_shapes(stages, consts, chain, block, tile):
    _check_clustering(stages, consts, chain, (block, tile))


@pytest.mark.parametrize("name", ["chain", "chain_uneven", "chain_scrambled"])
@pytest.mark.parametrize("block,tile", [(512, 0), (128, 2048)])
def test_masked_clustering_from_candidate_list(stages, consts, request, name, block, tile):
    """Candidates read from estimate_input_size's list give what the banks
    give (tile 2048: the larger events do not fit beside their list and take
    the bank path)."""
    c = request.getfixturevalue(name)
    _check_clustering(stages, consts, c, (block, tile), (c.cnum, c.cand))


@pytest.mark.parametrize("status_word", [None, 0x1, 0x200])
def test_masked_clustering_status_not_from_estimate(stages, consts, chain_scrambled, status_word):
    """A status word estimate_input_size did not write (zeroed, or a caller's
    own bits without BANK_ORDER_CHECKED) keeps the bank-order check, so
    unsorted banks still cluster exactly."""
    c = chain_scrambled
    st0 = None if status_word is None else np.full(c.n, status_word, np.uint32)
    num, cont, st = stages.clustering(c.ev.events, c.ev.offsets, c.starts, consts, (0, 0), (c.cnum, c.cand), st0)
    total = int(c.starts[c.n * N_MODULES])
    np.testing.assert_array_equal(num[: c.n * N_MODULES], c.nums)
    for s in range(4):
        np.testing.assert_array_equal(cont[s * total: (s + 1) * total][c.mask],
                                      c.cont_raw[s * total: (s + 1) * total][c.mask], err_msg=f"slot {s}")
    assert st0 is not None or not np.any(st & po.STATUS_BITS["BANK_ORDER_CHECKED"])


def test_masked_clustering_full_candidate_list(stages, consts, chain):
    """An event whose list is at capacity may have lost candidates: it takes
    the bank path, so nothing is lost."""
    cnum = chain.cnum.copy()
    cnum[::7] = 3000
    _check_clustering(stages, consts, chain, (0, 0), (cnum, chain.cand))


def _check_sort(stages, c: Chain, tuning=(0, 0)):
    cont, perm = stages.phi_sort(c.n, c.starts, c.nums, c.cont_raw, tuning)
    total = int(c.starts[c.n * N_MODULES])
    for s in (0, 1, 2, 4, 5):  # sorted slots, holes included (NaN)
        np.testing.assert_array_equal(cont[s * total: (s + 1) * total], c.cont[s * total: (s + 1) * total],
                                      err_msg=f"slot {s}")
    np.testing.assert_array_equal(cont[3 * total: 4 * total][c.mask], c.cont[3 * total: 4 * total][c.mask])
    np.testing.assert_array_equal(perm[:total][c.mask], c.perm[:total][c.mask])


def test_calculate_phi_and_sort(stages, chain):
    _check_sort(stages, chain)


@pytest.mark.parametrize("block,tile", [(64, 1), (256, 1024), (512, 4096), (1024, 0)])
def test_calculate_phi_and_sort_launch_shapes(stages, chain_uneven, block, tile):
    """Every workgroup size gives the same result: events taken in one or
    several module groups (two slots per thread each), and at 64 threads
    modules over 128 slots sorted through global memory.  The tile is not
    used by the sort."""
    _check_sort(stages, chain_uneven, (block, tile))


def test_calculate_phi_and_sort_uneven(stages, chain_uneven):
    _check_sort(stages, chain_uneven, (512, 0))


@pytest.mark.parametrize("tile", [0, 1])
def test_calculate_phi_and_sort_equal_phis(stages, chain, tile):
    """Hits sharing a phi keep their slot order, and a phi of -0 sorts level
    with +0 and keeps its sign: in every module of the first events, hits 1-3
    copy hit 0's x and y, and two hits sit at y = -0 and y = +0; 65 hits of
    one module share one phi, so the tie spans two waves of ranks."""
    n = 20
    starts = chain.starts[: n * N_MODULES + 1].copy()
    nums = chain.nums[: n * N_MODULES].copy()
    total_all = int(chain.starts[chain.n * N_MODULES])
    total = int(starts[n * N_MODULES])
    cont = np.concatenate([chain.cont_raw[k * total_all: k * total_all + total] for k in range(6)]).copy()
    x = cont[0:total].view(np.float32)
    y = cont[total: 2 * total].view(np.float32)
    changed, big = 0, False
    for e in range(n):
        for m in range(N_MODULES):
            s, k = int(starts[e * N_MODULES + m]), int(nums[e * N_MODULES + m])
            if k >= 6:
                x[s + 1: s + 4] = x[s]
                y[s + 1: s + 4] = y[s]
                x[s + 4] = x[s + 5] = 5.0
                y[s + 4] = np.float32(-0.0)
                y[s + 5] = np.float32(0.0)
                changed += 1
            if k >= 70 and not big:
                x[s: s + 65] = x[s]
                y[s: s + 65] = y[s]
                big = True
    assert changed > 100 and big
    want, want_perm = po.calculate_phi_and_sort(n, starts, nums, cont)
    got, perm = stages.phi_sort(n, starts, nums, cont, (256, tile))
    mask = module_mask(n, starts, nums)
    for s in (0, 1, 2, 4, 5):
        np.testing.assert_array_equal(got[s * total: (s + 1) * total], want[s * total: (s + 1) * total],
                                      err_msg=f"slot {s}")
    np.testing.assert_array_equal(perm[:total][mask], want_perm[:total][mask])
    phi = want[4 * total: 5 * total][mask]
    assert np.any(phi == 0x80000000) and np.any(phi == 0)   # both zeros present


def _check_fill(stages, c: Chain, tuning=(0, 0)):
    h0, h2 = stages.fill_candidates(c.n, c.cont, c.starts, c.nums, tuning)
    total = int(c.starts[c.n * N_MODULES])
    sel = np.zeros(total, bool)
    for e in range(c.n):
        for m in range(2, 50):
            s, k = int(c.starts[e * 52 + m]), int(c.nums[e * 52 + m])
            sel[s:s + k] = True
    idx = np.nonzero(sel)[0]
    both = np.concatenate([2 * idx, 2 * idx + 1])
    np.testing.assert_array_equal(h0[both], c.h0[both])
    np.testing.assert_array_equal(h2[both], c.h2[both])


def test_fill_candidates(stages, chain):
    _check_fill(stages, chain)


def test_fill_candidates_uneven(stages, chain_uneven):
    _check_fill(stages, chain_uneven)


def _check_search(stages, consts, c: Chain, tuning=(0, 0)):
    n = c.n
    tracks, weak, used, atomics, st = stages.search_by_triplet(n, c.cont, c.starts, c.nums, c.h0, c.h2, consts,
                                                               tuning)
    atomics = atomics.view(np.int32)
    np.testing.assert_array_equal(atomics[:n], c.atomics_s[:n])
    assert_tracks_equal(tracks_view(tracks, n), c.tracks_s.reshape(n, po.MAX_TRACKS), c.atomics_s[:n])
    nw = c.atomics_s[n: 5 * n: 4]
    np.testing.assert_array_equal(atomics[n: 5 * n: 4], nw)
    wv = np.frombuffer(weak.tobytes(), po.TRACKLET_DTYPE).reshape(n, po.MAX_WEAK_TRACKS)
    ow = c.weak.reshape(n, po.MAX_WEAK_TRACKS)
    for e in range(n):
        np.testing.assert_array_equal(wv[e, :nw[e]]["hits"], ow[e, :nw[e]]["hits"], err_msg=f"event {e}")
    total = int(c.starts[n * N_MODULES])
    np.testing.assert_array_equal(used[:total], c.used[:total])
    np.testing.assert_array_equal(st, c.st_search)
    return tracks


def test_search_by_triplet(stages, consts, chain):
    _check_search(stages, consts, chain)


def test_search_by_triplet_uneven(stages, consts, chain_uneven):
    _check_search(stages, consts, chain_uneven)


def window_modes(starts, n, cap):
    """Per step of every event: 1 when its six modules fit a window of cap
    hits, 2 when only the four h1 / h2 modules do, 0 when neither (the
    kernel's three step modes)."""
    st = np.asarray(starts[: n * N_MODULES + 1], np.int64)
    modes = []
    for e in range(n):
        s = np.append(st[e * N_MODULES: (e + 1) * N_MODULES], st[(e + 1) * N_MODULES])
        for first in range(N_MODULES - 1, 4, -2):
            lo = s[first - 5]
            modes.append(1 if s[first + 1] - lo <= cap else (2 if s[first - 1] - lo <= cap else 0))
    return np.array(modes)


@pytest.mark.parametrize("tile", [64, 200, 256])
def test_search_by_triplet_small_windows(stages, consts, chain_uneven, tile):
    """Windows too small for most steps (hits read from global memory), a
    tile that rounds down to 128, and one that fits part of the steps: each
    of the three step modes occurs."""
    cap = 1 << (tile.bit_length() - 1)
    modes = window_modes(chain_uneven.starts, chain_uneven.n, cap)
    if tile == 200:
        assert {0, 1, 2} <= set(modes.tolist()), "every step mode occurs"
    _check_search(stages, consts, chain_uneven, (64, tile))


def test_weak_tracks_adder(stages, chain):
    n = chain.n
    tracks_in = chain.tracks_s.view(np.uint8)
    weak_in = chain.weak.view(np.uint8)
    tracks, atomics, st = stages.weak_tracks_adder(n, chain.cont, chain.starts, tracks_in, weak_in, chain.used,
                                                   chain.atomics_s)
    atomics = atomics.view(np.int32)
    np.testing.assert_array_equal(atomics[:n], chain.atomics_w[:n])
    assert_tracks_equal(tracks_view(tracks, n), chain.tracks.reshape(n, po.MAX_TRACKS), chain.atomics_w[:n])
    np.testing.assert_array_equal(st, chain.st_weak)


def test_copy_and_prefix_sum(stages, chain):
    n = chain.n
    got = stages.copy_and_prefix_sum(n, chain.atomics_w.astype(np.uint32))
    np.testing.assert_array_equal(got[: 2 * n + 1], chain.atomics[: 2 * n + 1])


def test_copy_track_hit_number_and_scan(stages, chain):
    n = chain.n
    hn = stages.copy_track_hit_number(n, chain.tracks.view(np.uint8), chain.atomics)
    total = int(chain.atomics[2 * n])
    np.testing.assert_array_equal(hn[:total], chain.hn_raw[:total])
    np.testing.assert_array_equal(stages.prefix_sum(chain.hn_raw), chain.hn)


def test_consolidate(stages, chain):
    n = chain.n
    hits, states, acc = stages.consolidate(n, chain.atomics, chain.tracks.view(np.uint8), chain.hn, chain.cont,
                                           chain.starts, chain.nums)
    T = int(chain.atomics[2 * n])
    Ht = int(chain.hn[T])
    np.testing.assert_array_equal(hits[: 16 * Ht], chain.hits[: 16 * Ht])
    np.testing.assert_array_equal(states[: 49 * T], chain.states[: 49 * T])
    assert acc[:T].all()


def _check_kalman(stages, c: Chain, tuning=(0, 0)):
    n = c.n
    T = int(c.atomics[2 * n])
    got = stages.velo_kalman_fit(n, c.atomics, c.hn, c.hits, c.states, tuning)
    want = po.velo_kalman_fit(n, c.atomics, c.hn, c.hits, c.states)
    np.testing.assert_array_equal(got[: 49 * T], want[: 49 * T])


def test_velo_kalman_fit(stages, chain):
    _check_kalman(stages, chain)


def test_velo_kalman_fit_uneven_persistent(stages, chain_uneven):
    _check_kalman(stages, chain_uneven)
    t, counters = stages.persistent(chain_uneven.n, 37, seed=4)
    _check_kalman(stages, chain_uneven, t)
    assert not stages.host(counters).any()


def test_coordinates_within_tolerance(stages, consts, chain):
    """Cluster x, y, z agree with the oracle within 1e-5 mm."""
    n = chain.n
    _, cont, _ = stages.clustering(chain.ev.events, chain.ev.offsets, chain.starts, consts)
    total = int(chain.starts[n * N_MODULES])
    for s in range(3):
        g = cont[s * total: (s + 1) * total][chain.mask].view(np.float32)
        o = chain.cont_raw[s * total: (s + 1) * total][chain.mask].view(np.float32)
        assert np.max(np.abs(g.astype(np.float64) - o)) <= 1e-5


@pytest.mark.parametrize("count,capacity", [(0, 5), (1, 5), (2048, 5000), (2049, 5000), (4096, 4097), (9000, 9000)])
def test_prefix_sum_device_count(stages, count, capacity):
    """The scan reads its element count on the device; elements past it are
    neither read nor written, and the total lands at [count]."""
    rng = np.random.default_rng(count)
    a = rng.integers(0, 1000, capacity + 1).astype(np.uint32)
    got = stages.prefix_sum(a, device_count=count)
    want = a.copy()
    want[0] = 0
    want[1: count + 1] = np.cumsum(a[:count].astype(np.uint64)).astype(np.uint32)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("grid", [1, 7, 256])
def test_stages_persistent_shuffled_order(stages, consts, chain_uneven, grid):
    """Persistent launches (grid workgroups taking events from a counter in
    a shuffled order) give the one-workgroup-per-event outputs, and every
    launch leaves the shared counter pair at zero for the next."""
    c = chain_uneven
    n = c.n
    t, counters = stages.persistent(n, grid, seed=grid)
    _check_estimate(stages, consts, c, t)
    _check_clustering(stages, consts, c, t)
    _check_sort(stages, c, t)
    _check_fill(stages, c, t)
    _check_search(stages, consts, c, t)
    tracks, atomics, st = stages.weak_tracks_adder(n, c.cont, c.starts, c.tracks_s.view(np.uint8), c.weak.view(np.uint8),
                                                   c.used, c.atomics_s, t)
    np.testing.assert_array_equal(atomics.view(np.int32)[:n], c.atomics_w[:n])
    T = int(c.atomics[2 * n])
    hn = stages.copy_track_hit_number(n, c.tracks.view(np.uint8), c.atomics, t)
    np.testing.assert_array_equal(hn[:T], c.hn_raw[:T])
    hits, states, acc = stages.consolidate(n, c.atomics, c.tracks.view(np.uint8), c.hn, c.cont, c.starts, c.nums, t)
    Ht = int(c.hn[T])
    np.testing.assert_array_equal(hits[: 16 * Ht], c.hits[: 16 * Ht])
    np.testing.assert_array_equal(states[: 49 * T], c.states[: 49 * T])
    assert not stages.host(counters).any()


def test_persistent_launch_without_counters_is_rejected(stages, consts, chain):
    from cuda_hlt_amd import native as nat
    with pytest.raises(nat.VeloError):
        stages.estimate(chain.ev.events, chain.ev.offsets, consts, tuning=(0, 0, 64))


def test_empty_batch_is_a_no_op(stages, consts):
    from cuda_hlt_amd import native as nat
    a = nat.EstimateInputSizeArgs()
    a.n_events = 0
    a.dev_raw_input = a.dev_raw_input_offsets = a.dev_estimated_input_size = 1
    a.dev_module_cluster_num = a.dev_event_candidate_num = a.dev_cluster_candidates = 1
    nat.launch("velo_estimate_input_size_launch", a)


def test_null_argument_is_rejected(stages):
    from cuda_hlt_amd import native as nat
    with pytest.raises(nat.VeloError):
        nat.launch("velo_search_by_triplet_launch", nat.SearchByTripletArgs())


def _flat_edge_events():
    """Events for the flat bank table: one SP in every bank (a 64-SP chunk
    spans 64 banks), one bank of 300 SPs (a bank over several chunks and
    waves), an empty event, SP pairs in every third bank, and a mix."""
    from tests import rawevent
    one_each = [(s, 10 + 2 * (s % 7), 20 + 4 * (s % 5)) for s in range(208)]
    big = [(17, 2 * c, 4 * r) for c in range(30) for r in range(10)] + [(40, 5, 5), (41, 100, 30)]
    pairs = [(s, 2 * (s % 50) + 1, 8) for s in range(0, 208, 3)] + [(s, 2 * (s % 50) + 2, 8) for s in range(0, 208, 3)]
    mix = one_each[::2] + big + pairs[::5]
    return rawevent.concat([rawevent.encode(p) for p in (one_each, big, [], pairs, mix)])


@pytest.mark.parametrize("block,tile", [(64, 0), (256, 0), (1024, 0), (256, 32)])
def test_flat_bank_table_edges(stages, consts, geometry, block, tile):
    """estimate_input_size and masked clustering on the flat bank table's
    edge cases equal the oracle, from LDS (tile 0) and global memory (32)."""
    ev, off = _flat_edge_events()
    n = len(off) - 1
    oest, _, ocnum, ocand, ost = po.estimate_input_size(ev, off)
    est, num, cnum, cand, st = stages.estimate(ev, off, consts, tuning=(block, tile))
    np.testing.assert_array_equal(est[: n * N_MODULES], oest[: n * N_MODULES])
    assert not np.any(num[: n * N_MODULES])
    np.testing.assert_array_equal(cnum, ocnum)
    for e in range(n):
        k = int(ocnum[e])
        np.testing.assert_array_equal(cand[e * 3000: e * 3000 + k], ocand[e * 3000: e * 3000 + k])
    np.testing.assert_array_equal(st, ost)
    starts = po.prefix_sum(oest)
    onums, ocont, _ = po.masked_clustering(ev, off, starts, geometry)
    total = int(starts[n * N_MODULES])
    mask = module_mask(n, starts, onums)
    for cands in (None, (ocnum, ocand)):
        nums, cont, _ = stages.clustering(ev, off, starts, consts, (block, tile), cands, status0=ost)
        np.testing.assert_array_equal(nums[: n * N_MODULES], onums)
        for k in range(4):
            np.testing.assert_array_equal(cont[k * total: (k + 1) * total][mask], ocont[k * total: (k + 1) * total][mask])


def test_huge_bank_is_skipped(stages, consts, geometry):
    """A bank of 2^22 SPs (inside its 16 MB event) is flagged BAD_BANK and
    skipped as the oracle does; the other banks decode as usual."""
    from tests import rawevent
    events, offsets = rawevent.encode([(5, 10, 10), (9, 20, 20), (9, 21, 21)])
    w = events.view(np.uint32)
    h = int(w[0]) + 2 + int(w[1 + 5]) // 4
    big = np.zeros(len(w) + (1 << 22), np.uint32)
    big[: len(w)] = w
    big[h + 1] = 1 << 22
    ev, off = big.view(np.uint8), np.array([0, 4 * len(big)], np.uint32)
    oest, _, ocnum, ocand, ost = po.estimate_input_size(ev, off)
    est, _, cnum, cand, st = stages.estimate(ev, off, consts)
    assert ost[0] & 0x100 and st[0] == ost[0]
    np.testing.assert_array_equal(est[:N_MODULES], oest[:N_MODULES])
    assert cnum[0] == ocnum[0]
    np.testing.assert_array_equal(cand[: ocnum[0]], ocand[: ocnum[0]])
    starts = po.prefix_sum(oest)
    onums, ocont, _ = po.masked_clustering(ev, off, starts, geometry)
    nums, cont, _ = stages.clustering(ev, off, starts, consts, (256, 0), None)
    np.testing.assert_array_equal(nums[:N_MODULES], onums)
    total = int(starts[N_MODULES])
    np.testing.assert_array_equal(cont[: 4 * total], ocont[: 4 * total])


@pytest.mark.parametrize("tile", [0, 32])
def test_damaged_banks(stages, consts, minbias_200, geometry, tile):
    """Banks whose offset, header or SP words reach past their event are
    flagged and skipped as the oracle does, the event read from LDS (tile 0)
    or from global memory (tile 32); every read stays inside the event."""
    from tests import rawevent
    ev, off, *_ = rawevent.malformed(minbias_200.events, minbias_200.offsets)
    n = len(off) - 1
    oest, onum, ocnum, ocand, ost = po.estimate_input_size(ev, off)
    est, num, cnum, cand, st = stages.estimate(ev, off, consts, tuning=(256, tile))
    np.testing.assert_array_equal(est[: n * N_MODULES], oest[: n * N_MODULES])
    np.testing.assert_array_equal(cnum, ocnum)
    for e in range(n):
        k = int(ocnum[e])
        np.testing.assert_array_equal(cand[e * 3000: e * 3000 + k], ocand[e * 3000: e * 3000 + k])
    np.testing.assert_array_equal(st & 0x100, ost & 0x100)
    assert np.all(ost & 0x100)
    starts = po.prefix_sum(oest)
    onums, ocont, _ = po.masked_clustering(ev, off, starts, geometry)
    total = int(starts[n * N_MODULES])
    mask = module_mask(n, starts, onums)
    for cands in (None, (ocnum, ocand)):
        nums, cont, _ = stages.clustering(ev, off, starts, consts, (512, tile), cands)
        np.testing.assert_array_equal(nums[: n * N_MODULES], onums)
        for s in range(4):
            np.testing.assert_array_equal(cont[s * total: (s + 1) * total][mask], ocont[s * total: (s + 1) * total][mask])
