"""The plain VELO sequence driven from Python over the C ABI.

Runs the reference's Velo sequence (configuration/sequences/Velo.h:5-18,
without the global event cut) on device buffers owned by torch, one
launcher per algorithm, sizing each buffer from the host counters the way
the reference's visitors do (stream/visitors/velo/src/*.cu): three
device-to-host reads of totals (cluster estimate, track count, track hits).

This is the PyTorch-facing entry point; the C++ sequence (libvelo_hlt.so,
velo_sequence_*) is the drop-in for a C++ host.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import native as nat

N_MODULES = nat.N_MODULES


@dataclass
class VeloOutput:
    """Consolidated VELO tracks (VeloConsolidated.cuh:12-144 layouts)."""
    n_events: int
    atomics: torch.Tensor             # int32[2N+1]: counts, offsets, total
    track_hit_number: torch.Tensor    # uint32 as int32[T+1] exclusive offsets
    track_hits: torch.Tensor          # uint8[16 Ht]: SoA x, y, z, LHCbID
    states: torch.Tensor              # uint8[49 T]: SoA 12 floats + bool
    accepted: torch.Tensor            # uint8[T]
    status: torch.Tensor              # uint32 as int32[N]
    module_cluster_start: torch.Tensor
    module_cluster_num: torch.Tensor
    cluster_container: torch.Tensor
    total_tracks: int
    total_track_hits: int

    def to_host(self):
        return {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in self.__dict__.items()}


class PlainSequence:
    """init_event_list .. consolidate_velo_tracks on one stream."""

    def __init__(self, geometry: bytes, module_zs=None, device="cuda:0", tuning=None):
        self.device = torch.device(device)
        self.consts = nat.DeviceConstants(geometry, module_zs)
        self.tuning = tuning or {}

    def _t(self, name):
        b, t = self.tuning.get(name, (0, 0))
        return nat.Tuning(b, t)

    def _u32(self, n, zero=False):
        f = torch.zeros if zero else torch.empty
        return f(max(int(n), 1), dtype=torch.int32, device=self.device)

    def run(self, raw: torch.Tensor, offsets: torch.Tensor, n_events: int, stream=None) -> VeloOutput:
        """raw: uint8 device tensor of events; offsets: int32[N+1] device tensor."""
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream(self.device).cuda_stream
        N = int(n_events)
        c = self.consts.c
        status = self._u32(N, zero=True)
        event_list = self._u32(N)
        nat.launch("velo_init_event_list_launch", nat.InitEventListArgs(event_list.data_ptr(), N, nat.Tuning()), s)
        # velo_estimate_input_size_t
        est = self._u32(N * N_MODULES + 1)
        num = self._u32(N * N_MODULES)
        cnum = self._u32(N)
        cand = self._u32(N * nat.MAX_CANDIDATES_EVENT)
        nat.launch("velo_estimate_input_size_launch", nat.EstimateInputSizeArgs(
            raw.data_ptr(), offsets.data_ptr(), est.data_ptr(), num.data_ptr(), cnum.data_ptr(), cand.data_ptr(),
            event_list.data_ptr(), None, c.dev_velo_candidate_ks, status.data_ptr(), N,
            self._t("estimate_input_size")), s)
        # prefix_sum_velo_clusters_t (+ host read of the total, sync #1)
        aux = self._u32(nat.prefix_sum_aux_size(N * N_MODULES))
        nat.launch("velo_prefix_sum_launch", nat.PrefixSumArgs(est.data_ptr(), aux.data_ptr(), N * N_MODULES,
                                                               nat.Tuning()), s)
        total = int(est[N * N_MODULES].item())
        # velo_masked_clustering_t
        cont = self._u32(6 * total)
        nat.launch("velo_masked_clustering_launch", nat.MaskedClusteringArgs(
            raw.data_ptr(), offsets.data_ptr(), est.data_ptr(), num.data_ptr(), cnum.data_ptr(), cand.data_ptr(),
            cont.data_ptr(), event_list.data_ptr(), None, c.dev_velo_geometry, c.dev_velo_sp_patterns,
            c.dev_velo_sp_fx, c.dev_velo_sp_fy, status.data_ptr(), N, self._t("masked_clustering")), s)
        # velo_calculate_phi_and_sort_t
        perm = self._u32(total)
        nat.launch("velo_calculate_phi_and_sort_launch", nat.CalculatePhiAndSortArgs(
            est.data_ptr(), num.data_ptr(), cont.data_ptr(), perm.data_ptr(), N, self._t("calculate_phi_and_sort")), s)
        # velo_fill_candidates_t
        h0 = torch.empty(max(2 * total, 1), dtype=torch.int16, device=self.device)
        h2 = torch.empty(max(2 * total, 1), dtype=torch.int16, device=self.device)
        nat.launch("velo_fill_candidates_launch", nat.FillCandidatesArgs(
            cont.data_ptr(), est.data_ptr(), num.data_ptr(), h0.data_ptr(), h2.data_ptr(), N,
            self._t("fill_candidates")), s)
        # velo_search_by_triplet_t
        tracks = torch.empty(N * nat.MAX_TRACKS * nat.TRACK_HITS_BYTES, dtype=torch.uint8, device=self.device)
        weak = torch.empty(N * nat.MAX_WEAK_TRACKS * nat.TRACKLET_BYTES, dtype=torch.uint8, device=self.device)
        ttf = self._u32(N * nat.TTF_MODULO * nat.TTF_ENTRY_WORDS)
        used = torch.empty(max(total, 1), dtype=torch.uint8, device=self.device)
        atomics = self._u32(N * nat.NUM_ATOMICS + 1)
        nat.launch("velo_search_by_triplet_launch", nat.SearchByTripletArgs(
            cont.data_ptr(), est.data_ptr(), num.data_ptr(), tracks.data_ptr(), None, ttf.data_ptr(),
            weak.data_ptr(), used.data_ptr(), atomics.data_ptr(), h0.data_ptr(), h2.data_ptr(), None,
            c.dev_velo_module_zs, status.data_ptr(), N, self._t("search_by_triplet")), s)
        # velo_weak_tracks_adder_t
        nat.launch("velo_weak_tracks_adder_launch", nat.WeakTracksAdderArgs(
            cont.data_ptr(), est.data_ptr(), tracks.data_ptr(), weak.data_ptr(), used.data_ptr(),
            atomics.data_ptr(), status.data_ptr(), N, nat.Tuning()), s)
        # copy_and_prefix_sum_single_block_velo_t (+ host read, sync #2)
        base = atomics.data_ptr()
        nat.launch("velo_copy_and_prefix_sum_single_block_launch", nat.CopyAndPrefixSumArgs(
            base + 8 * N, base, base + 4 * N, N, nat.Tuning()), s)
        n_tracks = int(atomics[2 * N].item())
        # copy_velo_track_hit_number_t
        hn = self._u32(n_tracks + 1)
        nat.launch("velo_copy_velo_track_hit_number_launch", nat.CopyTrackHitNumberArgs(
            tracks.data_ptr(), atomics.data_ptr(), hn.data_ptr(), N, nat.Tuning()), s)
        # prefix_sum_velo_track_hit_number_t (+ host read, sync #3)
        aux2 = self._u32(nat.prefix_sum_aux_size(n_tracks))
        nat.launch("velo_prefix_sum_launch", nat.PrefixSumArgs(hn.data_ptr(), aux2.data_ptr(), n_tracks,
                                                               nat.Tuning()), s)
        n_hits = int(hn[n_tracks].item())
        # consolidate_velo_tracks_t
        track_hits = torch.empty(max(16 * n_hits, 1), dtype=torch.uint8, device=self.device)
        states = torch.empty(max(49 * n_tracks, 1), dtype=torch.uint8, device=self.device)
        accepted = torch.empty(max(n_tracks, 1), dtype=torch.uint8, device=self.device)
        nat.launch("velo_consolidate_velo_tracks_launch", nat.ConsolidateArgs(
            atomics.data_ptr(), tracks.data_ptr(), hn.data_ptr(), cont.data_ptr(), est.data_ptr(), num.data_ptr(),
            track_hits.data_ptr(), states.data_ptr(), accepted.data_ptr(), None, None, N, self._t("consolidate")), s)
        return VeloOutput(N, atomics[: 2 * N + 1], hn, track_hits, states, accepted, status, est, num, cont,
                          n_tracks, n_hits)


def upload_events(events: np.ndarray, offsets: np.ndarray, device="cuda:0"):
    raw = torch.from_numpy(np.ascontiguousarray(events)).to(device)
    off = torch.from_numpy(np.ascontiguousarray(offsets).view(np.int32)).to(device)
    return raw, off
