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
    "estimate_input_size": [(b, t) for b in (128, 256, 512, 1024) for t in (2048, 3072, 4096, 6144, 8192)],
    "masked_clustering": [(b, t) for b in (128, 256, 512, 1024) for t in (2048, 3072, 4096, 6144, 8192)],
    "calculate_phi_and_sort": [(b, 0) for b in (64, 128, 256, 512, 1024)],
    "fill_candidates": [(b, t) for b in (64, 128, 256, 512) for t in (1024, 2048, 4096)],
    "search_by_triplet": [(64, t) for t in (1, 256, 512, 768, 1024, 2048)],
    "consolidate": [(b, t) for b in (64, 128, 256, 512) for t in (1, 1024, 2048, 3072)],
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


def tune(seq, n_events, dev_events, dev_offsets, repeats=3, log=print):
    """Returns {algorithm: (block_dim, lds_tile)} and the per-config times."""
    seq.enable_step_timing(True)
    seq.run(n_events, dev_events=dev_events, dev_offsets=dev_offsets, copy_outputs=False)
    reference = _outputs(seq, n_events)
    best, table = {}, {}
    for alg, space in SPACE.items():
        times = {}
        for cfg in space:
            ms = []
            for _ in range(repeats):
                seq.run(n_events, dev_events=dev_events, dev_offsets=dev_offsets, copy_outputs=False,
                        tuning={alg: cfg})
                ms.append(dict(seq.step_times())[STEP_OF[alg]])
            times[cfg] = float(np.median(ms))
        cfg = min(times, key=times.get)
        best[alg] = cfg
        table[alg] = {f"{b}x{t}": round(v, 4) for (b, t), v in times.items()}
        log(f"[autotune] {alg}: best block={cfg[0]} tile={cfg[1]} {times[cfg]:.3f} ms "
            f"(default {times.get(space[0], float('nan')):.3f})")
    seq.run(n_events, dev_events=dev_events, dev_offsets=dev_offsets, copy_outputs=False, tuning=best)
    if _outputs(seq, n_events) != reference:
        raise RuntimeError("tuned launch shapes changed the output")
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
    args = ap.parse_args()
    import torch
    from . import gen
    from .stream import VeloSequence
    ev = gen.uneven(args.events, seed=args.seed) if args.uneven else gen.minbias(args.events, seed=args.seed)
    raw = torch.from_numpy(ev.events).cuda()
    off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
    seq = VeloSequence(gen.geometry(), gen.module_zs(), max_events=args.events, reserve_mb=8192)
    t0 = time.time()
    best, table = tune(seq, ev.n_events, raw.data_ptr(), off.data_ptr(), log=lambda *a: print(*a, file=sys.stderr))
    out = {"best": {k: list(v) for k, v in best.items()}, "ms": table,
           "workload": f"{'uneven nu~U(2,11)' if args.uneven else 'minbias nu=7.6'} {args.events} events seed {args.seed}",
           "seconds": round(time.time() - t0, 1)}
    Path(args.out).write_text(json.dumps(out, indent=1))
    print(json.dumps(out["best"]))


if __name__ == "__main__":
    main()
