"""Autotuner for the VELO sequence launch shapes.

For each tunable algorithm it sweeps the workgroup size and the LDS tile
(the staging capacity each kernel uses: raw event words for estimate and
clustering, phi floats for fill_candidates, window hits for
search_by_triplet, staged hits for consolidation), times the algorithm's
step with HIP events on the sequence's stream (median of repeats), keeps the
fastest, and checks that the tuned sequence reproduces the default output
bit for bit.  The result is a JSON file usable by bench.py and the serving
loop (`--tuning`).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

SPACE = {
    "estimate_input_size": [(b, t) for b in (128, 256, 512, 1024) for t in (2048, 3072, 4096, 5632, 6144, 8192)],
    "masked_clustering": [(b, t) for b in (128, 256, 512, 1024) for t in (2048, 3072, 4096, 6144, 8192)],
    "calculate_phi_and_sort": [(b, 0) for b in (128, 256, 512, 1024)],
    "fill_candidates": [(b, t) for b in (64, 128, 256, 512) for t in (1024, 2048, 4096)],
    "search_by_triplet": [(64, t) for t in (1, 256, 512, 768, 1024, 2048)],
    "consolidate": [(b, t) for b in (64, 128, 256, 512) for t in (1, 1024, 2048, 3072)],
}
# Uneven batches: workgroup size x LDS tile first, then the persistent grid
# (0 = one workgroup per event) at the best shape; events in largest-first
# order.  The search kernel is one wave per event.
BATCHED_BLOCKS = (64, 128, 256, 512)
BATCHED_TILES = {
    "estimate_input_size": (4096, 5632, 6144, 8192, 12288),
    "masked_clustering": (4096, 6144, 8192, 12288),
    "calculate_phi_and_sort": (0,),
    "fill_candidates": (1024, 2048, 4096, 8192),
    "search_by_triplet": (256, 512, 1024, 2048, 4096),
    "consolidate": (256, 1024, 2048, 3072),
}
BATCHED_GRIDS = (0, 1024, 2048, 4096, 8192)

# Launch shapes of the batched path when no tuning file is given: the
# autotuner's result on MI355X for 20 000 uneven events
# (cuda_hlt_amd/tuning/mi355x_uneven_batched.json)
BATCHED_DEFAULT = {
    "estimate_input_size": (256, 5632, 0),
    "masked_clustering": (256, 4096, 0),
    "calculate_phi_and_sort": (256, 0, 0),
    "fill_candidates": (256, 4096, 0),
    "search_by_triplet": (64, 256, 4096),
    "consolidate": (512, 3072, 0),
}

STEP_OF = {
    "estimate_input_size": "velo_estimate_input_size_t",
    "masked_clustering": "velo_masked_clustering_t",
    "calculate_phi_and_sort": "velo_calculate_phi_and_sort_t",
    "fill_candidates": "velo_fill_candidates_t",
    "search_by_triplet": "velo_search_by_triplet_t",
    "consolidate": "consolidate_velo_tracks_t",
}


def _outputs(seq, n):
    o = seq.output()
    from . import native as nat
    T, Ht = o.total_tracks, o.total_track_hits
    return (nat.d2h(o.dev_atomics_velo, 4 * (2 * n + 1)).tobytes(),
            nat.d2h(o.dev_track_hit_number, 4 * (T + 1)).tobytes(),
            nat.d2h(o.dev_track_hits, 16 * Ht).tobytes(),
            nat.d2h(o.dev_states, 49 * T).tobytes())


def tune(seq, n_events, dev_events, dev_offsets, repeats=3, log=print, outputs=_outputs):
    """Per step: workgroup size x LDS tile (SPACE), then the persistent grid
    at the best shape (BATCHED_GRIDS, 0 = one workgroup per event).
    Returns {algorithm: (block_dim, lds_tile, grid_dim)} and the times.
    outputs(seq, n) returns the run's consolidated output as comparable
    bytes (the device buffers by default)."""
    seq.enable_step_timing(True)
    seq.run(n_events, dev_events=dev_events, dev_offsets=dev_offsets, copy_outputs=False)
    reference = outputs(seq, n_events)
    best, table = {}, {}

    def time_cfg(alg, cfg):
        ms = []
        for _ in range(repeats):
            seq.run(n_events, dev_events=dev_events, dev_offsets=dev_offsets, copy_outputs=False, tuning={alg: cfg})
            ms.append(dict(seq.step_times())[STEP_OF[alg]])
        return float(np.median(ms))

    for alg, space in SPACE.items():
        times = {(b, t, 0): time_cfg(alg, (b, t, 0)) for b, t in space}
        b, t, _ = min(times, key=times.get)
        for g in BATCHED_GRIDS[1:]:
            times[(b, t, g)] = time_cfg(alg, (b, t, g))
        cfg = min(times, key=times.get)
        best[alg] = cfg
        table[alg] = {f"{b}x{t}x{g}": round(v, 4) for (b, t, g), v in times.items()}
        first = (*space[0], 0)
        log(f"[autotune] {alg}: best block={cfg[0]} tile={cfg[1]} grid={cfg[2]} {times[cfg]:.3f} ms "
            f"(first shape {times.get(first, float('nan')):.3f})")
    seq.run(n_events, dev_events=dev_events, dev_offsets=dev_offsets, copy_outputs=False, tuning=best)
    if outputs(seq, n_events) != reference:
        raise RuntimeError("tuned launch shapes changed the output")
    return best, table


def batched_reserve_mb(n_events, raw_bytes, velo_kalman_fit=False, pv_beamline=False):
    """Device arena (MiB) a batched run needs: the library's scheduler lays
    the sequence out on the batched bounds (velo_workspace_bytes: clusters
    <= raw bytes / 2, 1200 tracks and 1200 * 26 track hits per event) and
    reports the high-water mark; rounded up to whole MiB."""
    from .stream import workspace_bytes
    b = workspace_bytes(n_events, raw_bytes, velo_kalman_fit, pv_beamline)
    if b == 0:
        raise ValueError(f"no batched layout for {n_events} events of {raw_bytes} bytes")
    return -(-b // 2 ** 20)


def tune_batched(seq, n_events, dev_events, dev_offsets, events_size, dev_order, repeats=3, log=print,
                 outputs=_outputs):
    """Launch shapes for uneven batches: per step, workgroup size x LDS tile,
    then the persistent grid, timed with HIP events in plain runs that use
    the batched path's largest-first order; checks the batched (graph) run
    with the result against the default output.  Returns (best, table)."""
    seq.enable_step_timing(True)
    kw = dict(dev_events=dev_events, dev_offsets=dev_offsets, events_size=events_size, copy_outputs=False,
              dev_work_order=dev_order)
    seq.run(n_events, **kw)
    reference = outputs(seq, n_events)
    best = dict(BATCHED_DEFAULT)
    table = {}

    def time_cfg(alg, cfg):
        ms = []
        for _ in range(repeats):
            seq.run(n_events, tuning={**best, alg: cfg}, **kw)
            ms.append(dict(seq.step_times())[STEP_OF[alg]])
        return float(np.median(ms))

    for alg in STEP_OF:
        blocks = (64,) if alg == "search_by_triplet" else BATCHED_BLOCKS
        times = {(b, t, 0): time_cfg(alg, (b, t, 0)) for b in blocks for t in BATCHED_TILES[alg]}
        b, t, _ = min(times, key=times.get)
        for g in BATCHED_GRIDS[1:]:
            times[(b, t, g)] = time_cfg(alg, (b, t, g))
        cfg = min(times, key=times.get)
        best[alg] = cfg
        table[alg] = {f"{b}x{t}x{g}": round(v, 4) for (b, t, g), v in times.items()}
        log(f"[autotune] {alg}: best block={cfg[0]} tile={cfg[1]} grid={cfg[2]} {times[cfg]:.3f} ms")
    seq.run(n_events, tuning=best, batched=True, **kw)
    if outputs(seq, n_events) != reference:
        raise RuntimeError("tuned batched run changed the output")
    return best, table


def load(path):
    with open(path) as f:
        d = json.load(f)
    return {k: tuple(v) for k, v in d["best"].items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--events", type=int, default=10000)
    ap.add_argument("--uneven", action="store_true", help="per-event nu ~ U(2, 11)")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out", default="tuning.json")
    ap.add_argument("--batched", action="store_true",
                    help="tune the batched path (persistent grids, largest-first order)")
    args = ap.parse_args()
    import torch
    from . import gen
    from .stream import VeloSequence
    ev = gen.uneven(args.events, seed=args.seed) if args.uneven else gen.minbias(args.events, seed=args.seed)
    raw = torch.from_numpy(ev.events).cuda()
    off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
    t0 = time.time()
    say = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.batched:
        sizes = np.diff(ev.offsets.astype(np.int64))
        order = torch.from_numpy(np.argsort(-sizes, kind="stable").astype(np.uint32).view(np.int32)).cuda()
        seq = VeloSequence(gen.geometry(), gen.module_zs(), max_events=args.events,
                           reserve_mb=batched_reserve_mb(args.events, ev.events.nbytes))
        best, table = tune_batched(seq, ev.n_events, raw.data_ptr(), off.data_ptr(), ev.events.nbytes,
                                   order.data_ptr(), log=say)
    else:
        seq = VeloSequence(gen.geometry(), gen.module_zs(), max_events=args.events, reserve_mb=8192)
        best, table = tune(seq, ev.n_events, raw.data_ptr(), off.data_ptr(), log=say)
    out = {"best": {k: list(v) for k, v in best.items()}, "ms": table,
           "workload": f"{'uneven nu~U(2,11)' if args.uneven else 'minbias nu=7.6'} {args.events} events seed {args.seed}",
           "seconds": round(time.time() - t0, 1)}
    Path(args.out).write_text(json.dumps(out, indent=1))
    print(json.dumps(out["best"]))


if __name__ == "__main__":
    main()
