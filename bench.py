#!/usr/bin/env python3
"""VELO reconstruction throughput on one MI355X.

One step = one launch of the plain VELO sequence (raw SuperPixel banks to
consolidated tracks, the C++ Stream over the configured Velo sequence) on a
batch of --events-per-launch synthetic minimum-bias events (nu = 7.6) that
are resident in HBM before the timed region.  The timed region runs
--steps launches on distinct batches; value = events / second.

Also reported: per-launch latency p50/p99, per-step kernel times from HIP
events recorded on the sequence's stream, the HBM roofline of the dominant
kernel from its algorithmic bytes, the tracking efficiency of the output,
and the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "VELO events/s on 1 MI355X, plain vs batched+NN at equal eff.; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(S, H, C, T, Ht, N, W=0, A=0):
    """HBM bytes each step must move at least once (per batch).
    S SP words, H clusters, C candidates, T tracks, Ht hits on tracks, N
    events, W weak tracks left by the search, A of them accepted."""
    raw = 4 * S + 2504 * N
    return {
        # raw words in; estimates, cluster counts, candidate list and its count, status out
        "velo_estimate_input_size_t": raw + 4 * C + N * (2 * 52 * 4 + 4 + 4),
        "prefix_sum_velo_clusters_t": 8 * (52 * N),
        # raw words and the candidate list in; clusters (x, y, z, id) and counts out
        "velo_masked_clustering_t": raw + 4 * C + 4 * N + 16 * H + N * (2 * 52 * 4),
        "velo_calculate_phi_and_sort_t": 16 * H + 4 * H + 16 * H + 4 * H,
        "velo_fill_candidates_t": 4 * H + 8 * H,
        "velo_search_by_triplet_t": 16 * H + 12 * H + 54 * T,
        # per weak track its 3 hit numbers, their used flags and x, y, z; an
        # accepted one appended as a 54-byte track; per event two counters
        "velo_weak_tracks_adder_t": W * (6 + 3 + 36) + 54 * A + 8 * N,
        "copy_and_prefix_sum_single_block_velo_t": 8 * N,
        "copy_velo_track_hit_number_t": 2 * T + 4 * T,
        "prefix_sum_velo_track_hit_number_t": 8 * T,
        "consolidate_velo_tracks_t": 54 * T + 16 * Ht + 16 * Ht + 49 * T,
    }


# kernels each step launches (names as rocprofv3 reports them)
STEP_KERNELS = {
    "velo_estimate_input_size_t": ["estimate_input_size_kernel"],
    "velo_masked_clustering_t": ["masked_clustering_kernel"],
    "velo_calculate_phi_and_sort_t": ["calculate_phi_and_sort_kernel"],
    "velo_fill_candidates_t": ["fill_candidates_kernel"],
    "velo_search_by_triplet_t": ["search_by_triplet_kernel"],
    "velo_weak_tracks_adder_t": ["weak_tracks_adder_kernel"],
    "consolidate_velo_tracks_t": ["consolidate_kernel<false>"],
}


def measured_traffic(path, step, events_per_launch):
    """HBM bytes per launch of the step's kernels from a rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None
    when no summary for this launch size exists."""
    try:
        summary = json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return None
    if summary.get("events_per_launch") != events_per_launch:
        return None
    total = 0.0
    for k in STEP_KERNELS.get(step, []):
        rec = summary.get("kernels", {}).get(k)
        if rec is None:
            return None
        total += rec["fetch_bytes"] + rec["write_bytes"]
    return round(total) if total else None


def host_cpu_info():
    """Model name, physical cores and hardware threads of the host (lscpu's
    fields, read from /proc/cpuinfo) and the threads this process may use."""
    model, cores, threads = "", set(), 0
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and not model:
                model = v
            elif k == "processor":
                threads += 1
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
                cores.add((phys, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "physical_cores": len(cores) or None, "hardware_threads": threads or None,
            "usable_threads": usable}


def cpu_baseline(ev, geometry, zs, budget_s=18.0, threads=None, reps=3):
    """The oracle (C) on the first events of the workload, as BASELINE.md
    plans it: the median of `reps` timed passes single-threaded, then the
    median of `reps` passes event-parallel over the host threads this
    process may use (at most 16; tasks of 32 events; the C calls release
    the GIL).  Input is in memory; events/s over each pass."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle as po
    # size the single-threaded sample to about budget / (2 reps) seconds in all
    n = 100
    while True:
        sub = ev.subset(0, min(n, ev.n_events))
        t0 = time.perf_counter()
        po.pipeline_c(sub.events, sub.offsets, geometry, zs)
        dt = time.perf_counter() - t0
        target = budget_s / (2 * reps)
        if dt > target / 2 or n >= ev.n_events:
            break
        n = min(ev.n_events, int(n * max(2.0, target / max(dt, 1e-3))))
    one = []
    for _ in range(reps):
        t0 = time.perf_counter()
        po.pipeline_c(sub.events, sub.offsets, geometry, zs)
        one.append(sub.n_events / (time.perf_counter() - t0))
    one_med = float(np.median(one))
    info = host_cpu_info()
    threads = threads or min(16, info["usable_threads"])
    m = min(ev.n_events, max(sub.n_events, int(one_med * threads * budget_s / (2 * reps))))
    big = ev.subset(0, m)
    chunks = [big.subset(a, min(a + 32, m)) for a in range(0, m, 32)]
    multi = []
    with ThreadPoolExecutor(threads) as pool:
        for _ in range(reps):
            t0 = time.perf_counter()
            list(pool.map(lambda c: po.pipeline_c(c.events, c.offsets, geometry, zs), chunks))
            multi.append(m / (time.perf_counter() - t0))
    return {"value": round(float(np.median(multi)), 1), "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"first {m} events of the timed workload, C oracle (single-threaded restatement of the "
                      f"reference's clustering, triplet search, scans and consolidation) in {threads} threads "
                      f"(32-event tasks); median of {reps} passes",
            "passes": [round(v, 1) for v in multi],
            "single_thread": {"value": round(one_med, 1), "cores": 1, "passes": [round(v, 1) for v in one],
                              "sample": f"first {sub.n_events} events, median of {reps} passes"},
            "host_cpu": info}


GHOST_MODEL = ROOT / "cuda_hlt_amd" / "filters" / "velo_ghost_mi355x.npz"


def batched_run(args, geometry, zs):
    """Uneven events (nu ~ U(2,11) per event) in one batch: the sequence
    captured once in a hipGraph, persistent steps in largest-first order,
    one graph launch per step.  Timed without and with the ghost filter
    (inline in consolidation); the checker compares the filtered output's
    efficiency and ghost rate with the unfiltered one."""
    import torch
    from cuda_hlt_amd import autotune, checker, gen
    from cuda_hlt_amd.ghost import GhostModel
    from cuda_hlt_amd.stream import VeloSequence
    n = args.batched_events
    t0 = time.perf_counter()
    ev = gen.uneven(n, seed=args.seed + 1)
    log(f"[bench] batched: generated {n} uneven events in {time.perf_counter() - t0:.1f} s")
    raw = torch.from_numpy(ev.events).cuda()
    off = torch.from_numpy(ev.offsets.view(np.int32)).cuda()
    sizes = np.diff(ev.offsets.astype(np.int64))
    order = torch.from_numpy(np.argsort(-sizes, kind="stable").astype(np.uint32).view(np.int32)).cuda()
    tuning = autotune.load(args.batched_tuning) if args.batched_tuning else autotune.BATCHED_DEFAULT
    need_mb = autotune.batched_reserve_mb(n, ev.events.nbytes)
    seq = VeloSequence(geometry, zs, max_events=n, reserve_mb=need_mb)
    seq.enable_step_timing(True)
    model_path = Path(args.ghost_model) if args.ghost_model else GHOST_MODEL
    model = GhostModel.load(model_path) if model_path.exists() else None
    model_dev = model.to_device() if model is not None else None

    def run(filtered, copy=False):
        return seq.run(n, dev_events=raw.data_ptr(), dev_offsets=off.data_ptr(), events_size=ev.events.nbytes,
                       copy_outputs=copy, batched=True, tuning=tuning, dev_work_order=order.data_ptr(),
                       dev_ghost_model=model_dev.data_ptr() if filtered else None)

    def timed(filtered):
        t0 = time.perf_counter()
        run(filtered)  # capture + first replay
        capture_s = time.perf_counter() - t0
        run(filtered)
        torch.cuda.synchronize()
        lat, graph_ms = [], []
        t_start = time.perf_counter()
        for _ in range(args.batched_steps):
            t1 = time.perf_counter()
            o = run(filtered)
            lat.append(time.perf_counter() - t1)
            graph_ms.append(seq.step_times()[0][1])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t_start
        return o, {"value": round(args.batched_steps * n / elapsed, 1), "unit": "events/s",
                   "ms_per_step": round(1e3 * elapsed / args.batched_steps, 3),
                   "graph_ms_p50": round(float(np.percentile(graph_ms, 50)), 3),
                   "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                                  "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
                   "capture_and_first_run_s": round(capture_s, 3)}

    o, plain = timed(False)
    out = {**plain, "events_per_launch": n, "steps": args.batched_steps,
           "workload": f"uneven nu~U(2,11), {n} events (seed {args.seed + 1}), sequence captured in one hipGraph, "
                       f"persistent steps in largest-first order",
           "tuning": tuning, "reserve_mb": need_mb, "tracks": int(o.total_tracks),
           "raw_bytes": int(ev.events.nbytes), "estimated_clusters": int(o.total_estimated_hits)}
    run(False, copy=True)
    unfiltered = checker.check(ev, seq.host_tracks())
    out["efficiency"] = {"category": checker.HEADLINE, "value": round(unfiltered.efficiency(), 5),
                         "ghost_rate": round(unfiltered.ghost_rate, 5)}
    if model is not None:
        _, filt = timed(True)
        run(True, copy=True)
        host = seq.host_tracks()
        f = checker.check(ev, host, accepted=host.accepted)
        filt["efficiency"] = {"category": checker.HEADLINE, "value": round(f.efficiency(), 5),
                              "ghost_rate": round(f.ghost_rate, 5), "tracks_kept": int(host.accepted.sum()),
                              "efficiency_loss": round(unfiltered.efficiency() - f.efficiency(), 5)}
        filt["model"] = str(model_path.relative_to(ROOT)) if model_path.is_relative_to(ROOT) else str(model_path)
        filt["threshold"] = model.threshold
        out["filtered"] = filt
    seq.close()
    return out


def serving_run(args):
    """The continuous-batching server (batched graphs with the ghost filter,
    results kept in device memory): peak events/s with every request queued
    at once, then Poisson arrivals at 20/50/90 % of the peak, with request
    and batch latencies."""
    from cuda_hlt_amd import serving
    from cuda_hlt_amd.ghost import GhostModel
    model = GhostModel.load(args.ghost_model or str(GHOST_MODEL))
    t0 = time.perf_counter()
    out = serving.run_loads((0.2, 0.5, 0.9), n_requests=args.serving_requests, seed=5, batched=True,
                            ghost_model=model, log=log)
    out["requests_per_run"] = args.serving_requests
    out["wall_s"] = round(time.perf_counter() - t0, 1)
    return out


def extended_run(args, ev, seq, raw, batch_offsets, B, dev):
    """The sequence's opt-in steps on the first batch of the same workload:
    the global event cut (UT and SciFi banks from gen.gec_banks),
    velo_kalman_fit and the four pv_beamline steps.  Per-step times from HIP
    events over --steps runs, algorithmic bytes and the fraction of HBM
    peak they reach."""
    import torch
    from cuda_hlt_amd import gen
    g = gen.gec_banks(ev.subset(0, B), seed=args.seed)
    keep = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in (
        ("ut_events", g.ut_events), ("ut_offsets", g.ut_offsets.view(np.int32)),
        ("scifi_events", g.scifi_events), ("scifi_offsets", g.scifi_offsets.view(np.int32)))}
    gec = {k: v.data_ptr() for k, v in keep.items()}
    gec.update(ut_events_size=g.ut_events.nbytes, scifi_events_size=g.scifi_events.nbytes)
    torch.cuda.synchronize()
    base, off = batch_offsets[0]

    def run(copy_outputs=False):
        return seq.run(B, dev_events=raw.data_ptr() + base, dev_offsets=off.data_ptr(), gec=gec, pv_beamline=True,
                       copy_outputs=copy_outputs)

    for _ in range(args.warmup):
        run()
    steps = {}
    for _ in range(args.steps):
        run()
        for name, ms in seq.step_times():
            steps[name] = steps.get(name, 0.0) + ms / args.steps
    run(copy_outputs=True)
    ht = seq.host_tracks()
    n = ht.n_events
    T = int(ht.atomics[2 * n])
    Ht = int(ht.hit_number[T])
    V = int(ht.pv_counts.sum())
    ut_banks, scifi_banks = gen.UT_BANKS, gen.SCIFI_BANKS
    ab = {
        # offset tables and the first word of every UT bank; one list word per kept event
        "global_event_cut_t": B * (4 * (3 + ut_banks) + 4 * ut_banks + 4 * (3 + scifi_banks)) + 4 * n + 4,
        # hits (x, y, z) and tx, ty, z, backward of each track read; 49 B state written
        "velo_kalman_fit_t": 12 * Ht + 62 * T,
        "pv_beamline_extrapolate_t": 72 * T,
        "pv_beamline_histo_t": 28 * T + 4 * 2400 * n,
        "pv_beamline_peak_t": 4 * 2400 * n + 132 * n,
        "pv_beamline_multi_fitter_t": 28 * T + 136 * n + 52 * V,
    }
    kernels = {}
    for k, ms in steps.items():
        if k == "end":
            continue
        rec = {"ms": round(ms, 4)}
        if k in ab and ms > 0:
            gbps = ab[k] / (ms * 1e-3) / 1e9
            rec.update(algorithmic_bytes=ab[k], GBps=round(gbps, 1), frac=round(gbps / HBM_PEAK_GBS, 5))
        kernels[k] = rec
    return {"workload": f"first {B} events with the global event cut, velo_kalman_fit and beamline PV finding",
            "events_in": B, "events_kept": n, "tracks": T, "track_hits": Ht, "vertices": V,
            "vertices_per_event": round(V / max(n, 1), 3), "kernels": kernels}


def oracle_sizes(ev, geometry, zs, sample=500):
    """Candidate-list entries of the batch (the oracle's estimate, exact)
    and its weak tracks before and after the weak-track adder (the oracle's
    search on the first `sample` events, scaled to the batch)."""
    from oracle import pyoracle as po
    _, _, cnum, _, _ = po.estimate_input_size(ev.events, ev.offsets)
    sub = ev.subset(0, min(sample, ev.n_events))
    n = sub.n_events
    est, _, _, _, _ = po.estimate_input_size(sub.events, sub.offsets)
    starts = po.prefix_sum(est)
    nums, cont, _ = po.masked_clustering(sub.events, sub.offsets, starts, geometry)
    cont, _ = po.calculate_phi_and_sort(n, starts, nums, cont)
    h0, h2 = po.fill_candidates(n, cont, starts, nums)
    tracks, weak, used, atomics, _ = po.search_by_triplet(n, cont, starts, nums, h0, h2, zs)
    _, atomics_w, _ = po.weak_tracks_adder(n, cont, starts, tracks, weak, used, atomics)
    w = int(sum(atomics[n + 4 * e] for e in range(n)))
    a = int(atomics_w[:n].sum() - atomics[:n].sum())
    scale = ev.n_events / n
    return {"candidates": int(cnum.sum()), "weak_tracks": round(w * scale), "weak_accepted": round(a * scale)}


def verify_slice(ev, host, geometry, zs, n_check):
    """The first n_check events of a timed batch's output against the
    oracle, bit for bit (per-event track counts, hit offsets, hits,
    states): outputs of an event do not depend on the other events."""
    from oracle import pyoracle as po
    if n_check <= 0:
        return "not checked"
    sub = ev.subset(0, n_check)
    ref = po.pipeline(sub.events, sub.offsets, geometry, zs)
    n, N = n_check, host.n_events
    T_all, Ht_all = int(host.atomics[2 * N]), int(host.hit_number[int(host.atomics[2 * N])])
    T = int(host.atomics[N + n])
    Ht = int(host.hit_number[T])
    ok = np.array_equal(host.atomics[:n], ref.atomics[:n]) and T == int(ref.atomics[2 * n])
    ok = ok and np.array_equal(host.hit_number[: T + 1], ref.hit_number[: T + 1])
    hits = host.track_hits[: 16 * Ht_all].view(np.uint32).reshape(4, Ht_all)[:, :Ht]
    ok = ok and np.array_equal(hits, ref.track_hits[: 16 * Ht].view(np.uint32).reshape(4, Ht))
    states = [host.states[4 * T_all * k: 4 * T_all * k + 4 * T] for k in range(12)]
    states.append(host.states[48 * T_all: 48 * T_all + T])
    ref_states = [ref.states[4 * T * k: 4 * T * (k + 1)] for k in range(12)] + [ref.states[48 * T: 49 * T]]
    ok = ok and all(np.array_equal(a, b) for a, b in zip(states, ref_states))
    if not ok:
        raise SystemExit(f"bench: the first {n_check} events of a timed batch differ from the oracle")
    return f"first {n_check} events bit-exact with the oracle ({T} tracks)"


def pcie_run(args, ev, geometry, zs, B, tuning):
    """The reference's data path: raw events in (pinned) host memory copied
    in by init_event_list, the consolidated tracks copied back
    (InitEventListVisitor.cu:43-78, ConsolidateVeloVisitor.cu:47-66), one
    launch per batch.  Not the headline value."""
    import torch
    from cuda_hlt_amd.stream import VeloSequence
    steps = args.pcie_steps
    if steps <= 0:
        return None
    host = []
    for b in range(steps + 1):
        sub = ev.subset(b * B, (b + 1) * B)
        e = torch.from_numpy(sub.events).pin_memory()
        o = torch.from_numpy(sub.offsets.view(np.int32)).pin_memory()
        host.append((e, o))
    seq = VeloSequence(geometry, zs, max_events=B, reserve_mb=int(os.environ.get("VELO_RESERVE_MB", "8192")))

    def run(b):
        e, o = host[b]
        return seq.run(B, host_events=e.numpy(), host_offsets=o.numpy().view(np.uint32), copy_outputs=True,
                       tuning=tuning)

    run(steps)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    h2d = d2h = 0
    for b in range(steps):
        t1 = time.perf_counter()
        o = run(b)
        lat.append(time.perf_counter() - t1)
        h2d += host[b][0].numel() + 4 * (B + 1)
        d2h += 4 * (2 * B + 1) + 4 * (o.total_tracks + 1) + 16 * o.total_track_hits + 49 * o.total_tracks + \
            o.total_tracks
    elapsed = time.perf_counter() - t0
    seq.close()
    return {"value": round(steps * B / elapsed, 1), "unit": "events/s", "steps": steps,
            "ms_per_step": round(1e3 * elapsed / steps, 3),
            "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                           "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
            "h2d_bytes_per_step": h2d // steps, "d2h_bytes_per_step": d2h // steps,
            "workload": f"plain sequence on {B}-event batches in pinned host memory, outputs copied back"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--events-per-launch", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tuning", default=None, help="launch shapes from cuda_hlt_amd.autotune (JSON)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="first run one int32 add-one kernel over this many MiB (reads and writes exactly that "
                         "much): the known-size reference for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--batched-events", type=int, default=50000,
                    help="events of the batched run (uneven nu ~ U(2,11)); 0 skips it")
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--batched-tuning", default=None,
                    help="launch shapes of the batched run (cuda_hlt_amd.autotune JSON)")
    ap.add_argument("--ghost-model", default=None,
                    help="ghost filter for the batched run (.npz from python -m cuda_hlt_amd.ghost)")
    ap.add_argument("--verify-events", type=int, default=1000,
                    help="events of timed batch 0 compared bit for bit with the oracle")
    ap.add_argument("--pcie-steps", type=int, default=5,
                    help="launches of the host-input run (input copied in, outputs copied back); 0 skips it")
    ap.add_argument("--serving-requests", type=int, default=150,
                    help="requests per serving run (peak, then 20/50/90 %% of it; batched + filtered, "
                         "results in device memory); 0 skips serving")
    ap.add_argument("--no-extended", action="store_true",
                    help="skip the run with the global event cut, velo_kalman_fit and beamline PV finding")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "traffic_minbias_10k.json"),
                    help="per-kernel HBM traffic measured by rocprofv3 --pmc (cuda_hlt_amd.profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(events would shard across GPUs with no exchange: each GPU runs this same sequence "
              "on its own share of the events).", file=sys.stderr)
        sys.exit(2)

    import torch
    from cuda_hlt_amd import gen
    from cuda_hlt_amd.stream import VeloSequence

    B = args.events_per_launch
    n_batches = args.steps + args.warmup
    t0 = time.perf_counter()
    ev = gen.minbias(B * n_batches, seed=args.seed)
    log(f"[bench] generated {ev.n_events} events in {time.perf_counter() - t0:.1f} s "
        f"({ev.events.nbytes / 1e9:.2f} GB raw)")
    geometry, zs = gen.geometry(), gen.module_zs()

    dev = torch.device("cuda:0")
    if args.calibrate_copy_mb:
        words = args.calibrate_copy_mb << 18
        src = torch.ones(words, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.add(src, 1, out=dst)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {4 * words} bytes read and written")
        del src, dst
    raw = torch.from_numpy(ev.events).to(dev)
    batch_offsets = []
    for b in range(n_batches):
        o = ev.offsets[b * B: (b + 1) * B + 1].astype(np.int64)
        batch_offsets.append((int(o[0]), torch.from_numpy((o - o[0]).astype(np.uint32).view(np.int32)).to(dev)))
    torch.cuda.synchronize()

    seq = VeloSequence(geometry, zs, max_events=B, reserve_mb=int(os.environ.get("VELO_RESERVE_MB", "8192")))
    seq.enable_step_timing(True)
    tuning = None
    if args.tuning:
        from cuda_hlt_amd import autotune
        tuning = autotune.load(args.tuning)

    def run(b, copy_outputs=False):
        base, off = batch_offsets[b]
        return seq.run(B, dev_events=raw.data_ptr() + base, dev_offsets=off.data_ptr(), copy_outputs=copy_outputs,
                       tuning=tuning)

    for w in range(args.warmup):
        run(args.steps + w)
    torch.cuda.synchronize()
    lat, steps_ms = [], {}
    totals = {"T": 0, "Ht": 0, "H_est": 0}
    t_start = time.perf_counter()
    for k in range(args.steps):
        t1 = time.perf_counter()
        o = run(k)
        lat.append(time.perf_counter() - t1)
        for name, ms in seq.step_times():
            steps_ms[name] = steps_ms.get(name, 0.0) + ms
        totals["T"] += o.total_tracks
        totals["Ht"] += o.total_track_hits
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    value = args.steps * B / elapsed

    # per-batch sizes for the algorithmic bytes (batch 0 stands for all)
    from cuda_hlt_amd import checker
    from cuda_hlt_amd import native as nat
    o0 = run(0, copy_outputs=True)
    host0 = seq.host_tracks()
    chk = checker.check(ev.subset(0, B), host0)
    host_nums = nat.d2h(o0.dev_module_cluster_num, 4 * B * 52).view(np.uint32)
    H = int(host_nums.sum())
    S = int(ev.sp_counts()[:B].sum())
    sizes = oracle_sizes(ev.subset(0, B), geometry, zs)
    verified = verify_slice(ev, host0, geometry, zs, args.verify_events)
    log(f"[bench] batch 0: {verified}")
    ab = algorithmic_bytes(S, H, sizes["candidates"], o0.total_tracks, o0.total_track_hits, B,
                           sizes["weak_tracks"], sizes["weak_accepted"])
    per_step = {k: v / args.steps for k, v in steps_ms.items()}
    dominant = max((k for k in per_step if k != "end"), key=lambda k: per_step[k])
    kernels = {}
    for k, ms in per_step.items():
        b = ab.get(k, 0)
        gbps = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        kernels[k] = {"ms": round(ms, 4), "algorithmic_bytes": b,
                      "GBps": round(gbps, 1) if gbps is not None else None,
                      "frac": round(gbps / HBM_PEAK_GBS, 5) if gbps is not None else None}
        t = measured_traffic(args.traffic, k, B)
        if t is not None and ms > 0:
            kernels[k]["traffic_bytes"] = t
            kernels[k]["traffic_GBps"] = round(t / (ms * 1e-3) / 1e9, 1)
    dom_ms = per_step[dominant]
    achieved = ab[dominant] / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(args.traffic, dominant, B)}
    # decode + clustering together: R + 16 H + 416 bytes per event over the two steps' time
    dc_ms = per_step.get("velo_estimate_input_size_t", 0.0) + per_step.get("velo_masked_clustering_t", 0.0)
    dc_bytes = 4 * S + 2504 * B + 16 * H + 416 * B
    dc = {"bytes": dc_bytes, "ms": round(dc_ms, 4),
          "GBps": round(dc_bytes / (dc_ms * 1e-3) / 1e9, 1) if dc_ms > 0 else None,
          "frac": round(dc_bytes / (dc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if dc_ms > 0 else None}

    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain VELO sequence, minbias nu=7.6, {B}-event launches, "
                               f"{args.steps * B} distinct events timed (seed {args.seed})",
                   "events_per_launch": B, "path": "plain", "tuning": tuning},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "roofline": roofline,
        "decode_clustering_roofline": dc,
        "efficiency": {"category": checker.HEADLINE, "value": round(chk.efficiency(), 5),
                       "ghost_rate": round(chk.ghost_rate, 5), "events": B, "tracks": chk.tracks},
        "kernels": kernels,
        "batch": {"sp_words": S, "clusters": H, "tracks": o0.total_tracks, "track_hits": o0.total_track_hits,
                  "candidates": sizes["candidates"], "weak_tracks_est": sizes["weak_tracks"],
                  "weak_accepted_est": sizes["weak_accepted"]},
        "verified": verified,
    }
    if not args.no_extended:
        try:
            result["extended"] = extended_run(args, ev, seq, raw, batch_offsets, B, dev)
        except Exception as e:  # the headline line is printed regardless
            result["extended"] = {"error": repr(e)}
    seq.close()
    del raw, batch_offsets
    result["pcie_inclusive"] = pcie_run(args, ev, geometry, zs, B, tuning)
    if args.batched_events:
        result["batched"] = batched_run(args, geometry, zs)
    if args.serving_requests:
        try:
            result["serving"] = serving_run(args)
        except Exception as e:  # the headline line is printed regardless
            result["serving"] = {"error": repr(e)}
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(ev.subset(0, B), geometry, zs)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
