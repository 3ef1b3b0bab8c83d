"""Runs single launchers of libvelo_hlt.so on numpy inputs (GPU tests only).

Each function uploads its inputs, calls exactly one C-ABI launcher (through
cuda_hlt_amd.native) and downloads the outputs, so every stage can be fed
the oracle's output of the previous stage and compared on its own.
"""
from __future__ import annotations

import numpy as np
import torch

from cuda_hlt_amd import native as nat

N_MODULES = nat.N_MODULES


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy()


def u32(n):
    return torch.zeros(max(int(n), 1), dtype=torch.int32, device="cuda")


def garbage_u32(n, seed=1):
    """Buffers the kernels must fully define are pre-filled with noise."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randint(-2**31, 2**31 - 1, (max(int(n), 1),), generator=g, dtype=torch.int32).cuda()


def sync():
    torch.cuda.synchronize()


_keep = []


def tuning_of(t):
    return t if isinstance(t, nat.Tuning) else nat.Tuning(*t)


def persistent(n, grid, block=0, tile=0, seed=0):
    """Persistent launch shape: `grid` workgroups take the n events in a
    seeded shuffled order.  Returns (tuning, counter tensor)."""
    counters = u32(2)
    order = dev(np.random.default_rng(seed).permutation(n).astype(np.uint32).view(np.int32))
    _keep.append((counters, order))
    return nat.Tuning(block, tile, grid, 0, counters.data_ptr(), order.data_ptr(), None), counters


def estimate(events, offsets, consts: nat.DeviceConstants, tuning=(0, 0)):
    n = len(offsets) - 1
    raw, off = dev(events), dev(offsets.view(np.int32))
    est, num, cnum = garbage_u32(n * N_MODULES + 1), garbage_u32(n * N_MODULES), garbage_u32(n)
    cand, status = garbage_u32(n * nat.MAX_CANDIDATES_EVENT), u32(n)
    a = nat.EstimateInputSizeArgs(raw.data_ptr(), off.data_ptr(), est.data_ptr(), num.data_ptr(), cnum.data_ptr(),
                                  cand.data_ptr(), None, None, consts.c.dev_velo_candidate_ks, status.data_ptr(),
                                  n, tuning_of(tuning))
    nat.launch("velo_estimate_input_size_launch", a)
    sync()
    return (host(est).view(np.uint32), host(num).view(np.uint32), host(cnum).view(np.uint32),
            host(cand).view(np.uint32), host(status).view(np.uint32))


def prefix_sum(arr: np.ndarray, device_count=None) -> np.ndarray:
    """Exclusive scan of arr[:-1]; total written at [-1].  With device_count
    the launch gets capacity len(arr) - 1 and reads the count on the device."""
    n = len(arr) - 1
    d = dev(arr.view(np.int32))
    aux = garbage_u32(nat.prefix_sum_aux_size(n))
    t = nat.Tuning(0, 0)
    if device_count is not None:
        cnt = dev(np.array([device_count], np.uint32).view(np.int32))
        t.dev_count = cnt.data_ptr()
    nat.launch("velo_prefix_sum_launch", nat.PrefixSumArgs(d.data_ptr(), aux.data_ptr(), n, t))
    sync()
    return host(d).view(np.uint32)


def clustering(events, offsets, starts, consts: nat.DeviceConstants, tuning=(0, 0), cands=None, status0=None):
    """cands: (per-event candidate counts, candidate list) from
    estimate_input_size, or None to derive the candidates from the banks.
    status0: the event status words estimate_input_size wrote (their
    UNSORTED_BANK bit tells clustering which events to check for bank
    order); None for zeros, which is right only when every bank is sorted."""
    n = len(offsets) - 1
    total = int(starts[n * N_MODULES])
    raw, off, st = dev(events), dev(offsets.view(np.int32)), dev(starts.view(np.int32))
    num, status = garbage_u32(n * N_MODULES), u32(n)
    if status0 is not None:
        status = dev(np.ascontiguousarray(status0, np.uint32).view(np.int32))
    cont = garbage_u32(6 * total + 1)
    c = consts.c
    cnum_p = cand_p = None
    if cands is not None:
        cnum_d, cand_d = dev(cands[0].view(np.int32)), dev(cands[1].view(np.int32))
        cnum_p, cand_p = cnum_d.data_ptr(), cand_d.data_ptr()
    a = nat.MaskedClusteringArgs(raw.data_ptr(), off.data_ptr(), st.data_ptr(), num.data_ptr(), cnum_p, cand_p,
                                 cont.data_ptr(), None, None, c.dev_velo_geometry, c.dev_velo_sp_patterns,
                                 c.dev_velo_sp_fx, c.dev_velo_sp_fy, status.data_ptr(), n, tuning_of(tuning))
    nat.launch("velo_masked_clustering_launch", a)
    sync()
    return host(num).view(np.uint32), host(cont).view(np.uint32), host(status).view(np.uint32)


def phi_sort(n, starts, nums, container, tuning=(0, 0)):
    total = int(starts[n * N_MODULES])
    st, nm, cont = dev(starts.view(np.int32)), dev(nums.view(np.int32)), dev(container.view(np.int32))
    perm = garbage_u32(total + 1)
    a = nat.CalculatePhiAndSortArgs(st.data_ptr(), nm.data_ptr(), cont.data_ptr(), perm.data_ptr(), n,
                                    tuning_of(tuning))
    nat.launch("velo_calculate_phi_and_sort_launch", a)
    sync()
    return host(cont).view(np.uint32), host(perm).view(np.uint32)


def fill_candidates(n, container, starts, nums, tuning=(0, 0)):
    total = int(starts[n * N_MODULES])
    st, nm, cont = dev(starts.view(np.int32)), dev(nums.view(np.int32)), dev(container.view(np.int32))
    h0 = torch.full((2 * total + 2,), -7, dtype=torch.int16, device="cuda")
    h2 = torch.full((2 * total + 2,), -7, dtype=torch.int16, device="cuda")
    a = nat.FillCandidatesArgs(cont.data_ptr(), st.data_ptr(), nm.data_ptr(), h0.data_ptr(), h2.data_ptr(), n,
                               tuning_of(tuning))
    nat.launch("velo_fill_candidates_launch", a)
    sync()
    return host(h0), host(h2)


def search_by_triplet(n, container, starts, nums, h0, h2, consts: nat.DeviceConstants, tuning=(0, 0)):
    total = int(starts[n * N_MODULES])
    st, nm, cont = dev(starts.view(np.int32)), dev(nums.view(np.int32)), dev(container.view(np.int32))
    d0, d2 = dev(h0), dev(h2)
    tracks = torch.zeros(n * nat.MAX_TRACKS * nat.TRACK_HITS_BYTES, dtype=torch.uint8, device="cuda")
    weak = torch.zeros(n * nat.MAX_WEAK_TRACKS * nat.TRACKLET_BYTES, dtype=torch.uint8, device="cuda")
    ttf = garbage_u32(n * nat.TTF_MODULO * nat.TTF_ENTRY_WORDS)
    used = torch.full((total + 1,), 7, dtype=torch.uint8, device="cuda")
    atomics, status = garbage_u32(n * nat.NUM_ATOMICS + 1), u32(n)
    a = nat.SearchByTripletArgs(cont.data_ptr(), st.data_ptr(), nm.data_ptr(), tracks.data_ptr(), None,
                                ttf.data_ptr(), weak.data_ptr(), used.data_ptr(), atomics.data_ptr(), d0.data_ptr(),
                                d2.data_ptr(), None, consts.c.dev_velo_module_zs, status.data_ptr(), n,
                                tuning_of(tuning))
    nat.launch("velo_search_by_triplet_launch", a)
    sync()
    return (host(tracks), host(weak), host(used), host(atomics), host(status).view(np.uint32))


def weak_tracks_adder(n, container, starts, tracks, weak, used, atomics, tuning=(0, 0)):
    st, cont = dev(starts.view(np.int32)), dev(container.view(np.int32))
    dt, dw, du, da = dev(tracks), dev(weak), dev(used), dev(atomics.view(np.int32))
    status = u32(n)
    a = nat.WeakTracksAdderArgs(cont.data_ptr(), st.data_ptr(), dt.data_ptr(), dw.data_ptr(), du.data_ptr(),
                                da.data_ptr(), status.data_ptr(), n, tuning_of(tuning))
    nat.launch("velo_weak_tracks_adder_launch", a)
    sync()
    return host(dt), host(da), host(status).view(np.uint32)


def copy_and_prefix_sum(n, atomics):
    da = dev(atomics.view(np.int32))
    base = da.data_ptr()
    a = nat.CopyAndPrefixSumArgs(base + 8 * n, base, base + 4 * n, n, nat.Tuning(0, 0))
    nat.launch("velo_copy_and_prefix_sum_single_block_launch", a)
    sync()
    return host(da).view(np.uint32)


def copy_track_hit_number(n, tracks, atomics, tuning=(0, 0)):
    total = int(atomics[2 * n])
    dt, da = dev(tracks), dev(atomics.view(np.int32))
    hn = garbage_u32(total + 1)
    a = nat.CopyTrackHitNumberArgs(dt.data_ptr(), da.data_ptr(), hn.data_ptr(), n, tuning_of(tuning))
    nat.launch("velo_copy_velo_track_hit_number_launch", a)
    sync()
    return host(hn).view(np.uint32)[:total + 1]


def consolidate(n, atomics, tracks, hit_number, container, starts, nums, tuning=(0, 0)):
    total_tracks = int(atomics[2 * n])
    total_hits = int(hit_number[total_tracks])
    da, dt, dh = dev(atomics.view(np.int32)), dev(tracks), dev(hit_number.view(np.int32))
    cont, st, nm = dev(container.view(np.int32)), dev(starts.view(np.int32)), dev(nums.view(np.int32))
    hits = torch.zeros(16 * total_hits + 16, dtype=torch.uint8, device="cuda")
    states = torch.zeros(49 * total_tracks + 49, dtype=torch.uint8, device="cuda")
    acc = torch.zeros(total_tracks + 1, dtype=torch.uint8, device="cuda")
    a = nat.ConsolidateArgs(da.data_ptr(), dt.data_ptr(), dh.data_ptr(), cont.data_ptr(), st.data_ptr(),
                            nm.data_ptr(), hits.data_ptr(), states.data_ptr(), acc.data_ptr(), None, None, n,
                            tuning_of(tuning))
    nat.launch("velo_consolidate_velo_tracks_launch", a)
    sync()
    return host(hits), host(states), host(acc)


def velo_kalman_fit(n, atomics, hit_number, track_hits, states, tuning=(0, 0)):
    T = int(atomics[2 * n])
    da, dh = dev(atomics.view(np.int32)), dev(hit_number.view(np.int32))
    dhits, dst = dev(track_hits), dev(states)
    out = torch.full((49 * T + 49,), 0x5A, dtype=torch.uint8, device="cuda")
    a = nat.KalmanFitArgs(da.data_ptr(), dh.data_ptr(), dhits.data_ptr(), dst.data_ptr(), out.data_ptr(), n,
                          tuning_of(tuning))
    nat.launch("velo_kalman_fit_launch", a)
    sync()
    return host(out)


def global_event_cut(ut_events, ut_offsets, scifi_events, scifi_offsets, block=0):
    """(kept events, count) of velo_global_event_cut_launch; the list buffer
    starts as noise."""
    n = len(ut_offsets) - 1
    du, duo = dev(ut_events), dev(np.asarray(ut_offsets, np.uint32).view(np.int32))
    ds, dso = dev(scifi_events), dev(np.asarray(scifi_offsets, np.uint32).view(np.int32))
    lst, cnt = garbage_u32(n), garbage_u32(1, seed=2)
    a = nat.GlobalEventCutArgs(du.data_ptr(), duo.data_ptr(), ds.data_ptr(), dso.data_ptr(), cnt.data_ptr(),
                               lst.data_ptr(), int(np.asarray(ut_events).nbytes),
                               int(np.asarray(scifi_events).nbytes), n, nat.Tuning(block, 0, 0, 0, None, None, None))
    nat.launch("velo_global_event_cut_launch", a)
    sync()
    k = int(host(cnt)[0])
    return host(lst).view(np.uint32)[:k].copy(), k


def pv_extrapolate(n, atomics, kalman_states, tuning=(0, 0)):
    T = int(atomics[2 * n])
    da, ds = dev(np.asarray(atomics, np.uint32).view(np.int32)), dev(kalman_states)
    out = torch.full((max(T, 1) * 7,), float("nan"), dtype=torch.float32, device="cuda")
    a = nat.PvExtrapolateArgs(ds.data_ptr(), da.data_ptr(), None, out.data_ptr(), n, tuning_of(tuning))
    nat.launch("velo_pv_beamline_extrapolate_launch", a)
    sync()
    return host(out)[:7 * T].reshape(-1, 7)


def pv_histo(n, atomics, pvtracks, tuning=(0, 0)):
    da = dev(np.asarray(atomics, np.uint32).view(np.int32))
    dt = dev(np.ascontiguousarray(pvtracks, np.float32).reshape(-1) if len(pvtracks) else np.zeros(7, np.float32))
    out = torch.full((max(n, 1) * 2400,), float("nan"), dtype=torch.float32, device="cuda")
    a = nat.PvHistoArgs(da.data_ptr(), None, dt.data_ptr(), out.data_ptr(), n, tuning_of(tuning))
    nat.launch("velo_pv_beamline_histo_launch", a)
    sync()
    return host(out)[:n * 2400].reshape(n, 2400)


def pv_peak(n, zhisto, tuning=(0, 0)):
    dh = dev(np.ascontiguousarray(zhisto, np.float32).reshape(-1))
    zp = torch.full((max(n, 1) * 32,), float("nan"), dtype=torch.float32, device="cuda")
    cnt = garbage_u32(n)
    a = nat.PvPeakArgs(dh.data_ptr(), zp.data_ptr(), cnt.data_ptr(), n, tuning_of(tuning))
    nat.launch("velo_pv_beamline_peak_launch", a)
    sync()
    return host(zp)[:n * 32].reshape(n, 32), host(cnt).view(np.uint32)[:n]


def pv_multi_fitter(n, atomics, pvtracks, zpeaks, n_zpeaks, tuning=(0, 0)):
    from oracle.pyoracle import PV_VERTEX_DTYPE
    da = dev(np.asarray(atomics, np.uint32).view(np.int32))
    dt = dev(np.ascontiguousarray(pvtracks, np.float32).reshape(-1) if len(pvtracks) else np.zeros(7, np.float32))
    dz = dev(np.ascontiguousarray(zpeaks, np.float32).reshape(-1))
    dn = dev(np.asarray(n_zpeaks, np.uint32).view(np.int32))
    v = torch.zeros((max(n, 1) * 32 * 13,), dtype=torch.int32, device="cuda")
    cnt = garbage_u32(n)
    a = nat.PvMultiFitterArgs(da.data_ptr(), None, dt.data_ptr(), dz.data_ptr(), dn.data_ptr(), v.data_ptr(),
                              cnt.data_ptr(), n, tuning_of(tuning))
    nat.launch("velo_pv_beamline_multi_fitter_launch", a)
    sync()
    return host(v).view(PV_VERTEX_DTYPE)[:n * 32].reshape(n, 32), host(cnt).view(np.uint32)[:n]


def event_order(offsets, event_list=None):
    n = len(offsets) - 1 if event_list is None else len(event_list)
    do = dev(np.asarray(offsets, np.uint32).view(np.int32))
    dl = None if event_list is None else dev(np.asarray(event_list, np.uint32).view(np.int32))
    out = garbage_u32(n)
    a = nat.EventOrderArgs(do.data_ptr(), None if dl is None else dl.data_ptr(), out.data_ptr(), n,
                           nat.Tuning(0, 0, 0, 0, None, None, None))
    nat.launch("velo_event_order_launch", a)
    sync()
    return host(out).view(np.uint32)[:n]
