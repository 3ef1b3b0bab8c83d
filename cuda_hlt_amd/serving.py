"""Trigger-style serving loop with continuous batching.

Requests (1-2000 raw events each) arrive over time.  The server keeps K
slots, each with resident device buffers for one batch (raw events,
offsets, largest-first order) and its own sequence Stream (HIP stream,
device arena, captured graphs).  Whenever a slot is free, the dispatcher
packs every queued request that fits the slot (events and bytes) into one
batch, in arrival order, and hands it to the slot's worker thread, which
copies it to the device, replays the captured sequence and cuts the
consolidated output back into per-request results.  Requests that arrive
while batches run wait only until the next slot frees up, and the slots'
work overlaps on the GPU.

Admission is by HBM footprint: a slot's arena is sized once, at start-up,
for the largest batch it admits (the batched path's bounds for
slot_events events and slot_bytes raw bytes); the number of slots is the
HBM budget divided by that footprint (at most max_slots).  Batches are
padded with empty events to a size bucket so a slot replays one of a few
captured graphs; empty events produce no hits and no tracks, and every
event's output is independent of the other events in its batch.
"""
from __future__ import annotations

import argparse
import json
import queue
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import autotune, gen
from .stream import VeloSequence

BUCKETS = (64, 128, 256, 512, 1024, 2048, 4096)


def empty_event() -> np.ndarray:
    """A raw VP event with 208 empty banks (2504 bytes)."""
    w = np.zeros(1 + 209 + 2 * 208, np.uint32)
    w[0] = 208
    w[1:210] = 8 * np.arange(209)
    w[210::2] = np.arange(208)
    return w.view(np.uint8)


EMPTY = empty_event()


def bucket_of(n: int) -> int:
    for b in BUCKETS:
        if n <= b:
            return b
    raise ValueError(f"batch of {n} events exceeds the largest bucket")


@dataclass
class Request:
    id: int
    events: np.ndarray       # uint8 raw events, concatenated
    offsets: np.ndarray      # uint32[n+1]
    arrival: float = 0.0     # seconds since the start of the run
    done: float = 0.0
    result: dict | None = None

    @property
    def n(self) -> int:
        return len(self.offsets) - 1


@dataclass
class BatchRecord:
    slot: int
    requests: list
    events: int
    bucket: int
    start: float
    end: float


def cut_results(host, n_events: int, requests, first_event: list):
    """Per-request slices of a batch's consolidated output: per event the
    track count, and per track its hits (x, y, z, LHCbID), state and
    accepted flag, in the reference layouts."""
    N = n_events
    T = int(host.atomics[2 * N])
    Ht = int(host.hit_number[T])
    soa = host.track_hits[: 16 * Ht]
    S = host.states[: 49 * T]
    out = []
    for r, e0 in zip(requests, first_event):
        t0, t1 = int(host.atomics[N + e0]), int(host.atomics[N + e0 + r.n])
        h0, h1 = int(host.hit_number[t0]), int(host.hit_number[t1])
        hits = np.concatenate([soa[k * 4 * Ht + 4 * h0: k * 4 * Ht + 4 * h1] for k in range(4)])
        states = np.concatenate([S[k * 4 * T + 4 * t0: k * 4 * T + 4 * t1] for k in range(12)] + [S[48 * T + t0: 48 * T + t1]])
        out.append({"tracks_per_event": host.atomics[e0: e0 + r.n].copy(),
                    "hit_number": (host.hit_number[t0: t1 + 1] - h0).astype(np.uint32),
                    "hits": hits, "states": states,
                    "accepted": None if getattr(host, "accepted", None) is None else host.accepted[t0:t1].copy()})
    return out


class Slot:
    """Resident buffers + one sequence Stream for batches of up to
    max_events events / max_bytes raw bytes."""

    def __init__(self, index, geometry, zs, max_events, max_bytes, batched, tuning, ghost_model_dev):
        import torch
        self.index = index
        self.max_events = max_events
        self.max_bytes = max_bytes
        self.batched = batched
        self.tuning = tuning
        self.ghost = ghost_model_dev
        self.reserve_mb = autotune.batched_reserve_mb(max_events, max_bytes)
        self.seq = VeloSequence(geometry, zs, max_events=max_events, reserve_mb=self.reserve_mb)
        self.stream = torch.cuda.ExternalStream(self.seq.stream)
        self.dev_events = torch.zeros(max_bytes, dtype=torch.uint8, device="cuda")
        self.dev_offsets = torch.zeros(max_events + 1, dtype=torch.int32, device="cuda")
        self.dev_order = torch.zeros(max_events, dtype=torch.int32, device="cuda")
        self.host_events = torch.zeros(max_bytes, dtype=torch.uint8).pin_memory()
        self.host_offsets = torch.zeros(max_events + 1, dtype=torch.int32).pin_memory()
        self.host_order = torch.zeros(max_events, dtype=torch.int32).pin_memory()

    def run(self, requests):
        """Packs the requests (plus empty events up to the bucket), runs
        the sequence and returns (per-request results, bucket)."""
        import torch
        n_real = sum(r.n for r in requests)
        bucket = bucket_of(n_real) if self.batched else n_real
        first, offs, pos, e = [], [0], 0, 0
        he = self.host_events.numpy()
        for r in requests:
            first.append(e)
            nb = int(r.offsets[-1])
            he[pos: pos + nb] = r.events[: nb]
            offs.extend((pos + r.offsets[1:].astype(np.int64)).tolist())
            pos += nb
            e += r.n
        for _ in range(bucket - n_real):
            he[pos: pos + EMPTY.nbytes] = EMPTY
            pos += EMPTY.nbytes
            offs.append(pos)
        if pos > self.max_bytes:
            raise ValueError("batch exceeds the slot's raw-byte capacity")
        off = np.asarray(offs, np.int64)
        self.host_offsets.numpy()[: bucket + 1] = off.astype(np.uint32).view(np.int32)
        sizes = np.diff(off)
        self.host_order.numpy()[:bucket] = np.argsort(-sizes, kind="stable").astype(np.int32)
        with torch.cuda.stream(self.stream):
            self.dev_events[:pos].copy_(self.host_events[:pos], non_blocking=True)
            self.dev_offsets[: bucket + 1].copy_(self.host_offsets[: bucket + 1], non_blocking=True)
            self.dev_order[:bucket].copy_(self.host_order[:bucket], non_blocking=True)
        kw = dict(dev_events=self.dev_events.data_ptr(), dev_offsets=self.dev_offsets.data_ptr(), copy_outputs=True,
                  tuning=self.tuning, dev_ghost_model=None if self.ghost is None else self.ghost.data_ptr())
        if self.batched:
            kw.update(batched=True, events_size=self.max_bytes, dev_work_order=self.dev_order.data_ptr())
        else:
            kw.update(events_size=pos)
        self.seq.run(bucket, **kw)
        return cut_results(self.seq.host_tracks(), bucket, requests, first), bucket

    def close(self):
        """Frees the buffers before the sequence's stream: the pinned staging
        buffers were copied from on that stream, and freeing them records an
        event on it."""
        import torch
        torch.cuda.synchronize()
        self.host_events = self.host_offsets = self.host_order = None
        self.dev_events = self.dev_offsets = self.dev_order = None
        torch.cuda.synchronize()
        self.stream = None
        self.seq.close()


class Server:
    """Continuous-batching server over K slots (see the module docstring)."""

    def __init__(self, geometry, zs, hbm_budget_mb=65536, max_slots=4, slot_events=4096, slot_bytes=None,
                 batched=True, tuning=None, ghost_model=None):
        slot_bytes = slot_bytes or slot_events * 40000
        if tuning is None and batched:
            tuning = dict(autotune.BATCHED_DEFAULT)
        per_slot = autotune.batched_reserve_mb(slot_events, slot_bytes)
        self.n_slots = max(1, min(max_slots, hbm_budget_mb // per_slot))
        self.slot_events, self.slot_bytes = slot_events, slot_bytes
        self.footprint_mb = per_slot
        ghost_dev = ghost_model.to_device() if ghost_model is not None else None
        self._ghost = ghost_dev
        self.slots = [Slot(i, geometry, zs, slot_events, slot_bytes, batched, tuning, ghost_dev)
                      for i in range(self.n_slots)]
        self.free = queue.Queue()
        for s in self.slots:
            self.free.put(s)
        self.pending = []
        self.cv = threading.Condition()
        self.batches: list[BatchRecord] = []
        self.errors: list[BaseException] = []

    def fits(self, reqs, r) -> bool:
        n = sum(x.n for x in reqs) + r.n
        b = sum(int(x.offsets[-1]) for x in reqs) + int(r.offsets[-1])
        pad = (bucket_of(n) - n) * EMPTY.nbytes if n <= BUCKETS[-1] else 0
        return n <= self.slot_events and b + pad <= self.slot_bytes

    def _work(self, slot, reqs, t0):
        start = time.perf_counter() - t0
        try:
            results, bucket = slot.run(reqs)
            end = time.perf_counter() - t0
            for r, res in zip(reqs, results):
                r.result, r.done = res, end
            with self.cv:
                self.batches.append(BatchRecord(slot.index, [r.id for r in reqs], sum(r.n for r in reqs), bucket,
                                                start, end))
        except BaseException as e:  # reported by serve()
            with self.cv:
                self.errors.append(e)
        finally:
            self.free.put(slot)
            with self.cv:
                self.cv.notify_all()

    def serve(self, requests):
        """Open-loop run: each request is admitted at its arrival time
        (seconds from now); returns when all are done."""
        for r in requests:
            if r.n > self.slot_events or int(r.offsets[-1]) > self.slot_bytes:
                raise ValueError(f"request {r.id} exceeds a slot")
        todo = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        threads = []
        i = 0
        while i < len(todo) or self.pending:
            now = time.perf_counter() - t0
            while i < len(todo) and todo[i].arrival <= now:
                self.pending.append(todo[i])
                i += 1
            if self.errors:
                break
            if self.pending:
                try:
                    slot = self.free.get_nowait()
                except queue.Empty:
                    slot = None
                if slot is not None:
                    batch = []
                    while self.pending and self.fits(batch, self.pending[0]):
                        batch.append(self.pending.pop(0))
                    th = threading.Thread(target=self._work, args=(slot, batch, t0), daemon=True)
                    th.start()
                    threads.append(th)
                    continue
            # wait for an arrival or a free slot
            nxt = todo[i].arrival - (time.perf_counter() - t0) if i < len(todo) else 0.05
            with self.cv:
                self.cv.wait(timeout=max(0.0, min(nxt, 0.05)))
        for th in threads:
            th.join()
        if self.errors:
            raise self.errors[0]
        return time.perf_counter() - t0

    def close(self):
        for s in self.slots:
            s.close()


def make_requests(n_requests, rate_hz, seed=0, min_events=1, max_events=2000):
    """Poisson arrivals at rate_hz of requests of U(min, max) uneven events
    (per-event nu ~ U(2, 11))."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(min_events, max_events + 1, n_requests)
    ev = gen.uneven(int(sizes.sum()), seed=seed + 1)
    gaps = rng.exponential(1.0 / rate_hz, n_requests) if rate_hz > 0 else np.zeros(n_requests)
    t = np.cumsum(gaps) - gaps[0]
    reqs, e = [], 0
    for k, (s, a) in enumerate(zip(sizes, t)):
        sub = ev.subset(e, e + int(s))
        reqs.append(Request(k, sub.events, sub.offsets, float(a)))
        e += int(s)
    return reqs, ev


def latency_summary(requests, batches, wall):
    lat = np.array([r.done - r.arrival for r in requests])
    blat = np.array([b.end - b.start for b in batches])
    events = sum(r.n for r in requests)
    return {"events": events, "requests": len(requests), "batches": len(batches), "wall_s": round(wall, 4),
            "events_per_s": round(events / wall, 1),
            "request_latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                                   "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
            "batch_latency_ms": {"p50": round(1e3 * float(np.percentile(blat, 50)), 3),
                                 "p99": round(1e3 * float(np.percentile(blat, 99)), 3)},
            "mean_batch_events": round(float(np.mean([b.events for b in batches])), 1)}


def run_loads(loads=(0.2, 0.5, 0.9), n_requests=300, seed=0, batched=True, ghost_model=None, tuning=None,
              hbm_budget_mb=65536, max_slots=4, log=None):
    """Peak throughput (every request queued at t = 0), then Poisson
    arrivals at the given fractions of it."""
    geometry, zs = gen.geometry(), gen.module_zs()
    srv = Server(geometry, zs, hbm_budget_mb=hbm_budget_mb, max_slots=max_slots, batched=batched, tuning=tuning,
                 ghost_model=ghost_model)
    try:
        warm, _ = make_requests(max(8, srv.n_slots * 2), 0, seed=seed + 7)
        srv.serve(warm)
        srv.batches.clear()
        reqs, _ = make_requests(n_requests, 0, seed=seed)
        wall = srv.serve(reqs)
        peak = latency_summary(reqs, srv.batches, wall)
        if log:
            log(f"[serving] peak {peak['events_per_s']} events/s over {srv.n_slots} slots")
        mean_events = np.mean([r.n for r in reqs])
        out = {"slots": srv.n_slots, "slot_footprint_mb": srv.footprint_mb, "batched": batched,
               "filtered": ghost_model is not None, "peak": peak, "loads": {}}
        for f in loads:
            srv.batches.clear()
            rate = f * peak["events_per_s"] / mean_events
            reqs, _ = make_requests(n_requests, rate, seed=seed + int(100 * f))
            wall = srv.serve(reqs)
            out["loads"][str(f)] = latency_summary(reqs, srv.batches, wall) | {"offered_requests_per_s": round(rate, 2)}
            if log:
                log(f"[serving] load {f}: {out['loads'][str(f)]}")
        return out
    finally:
        srv.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--loads", type=float, nargs="*", default=[0.2, 0.5, 0.9])
    ap.add_argument("--plain", action="store_true", help="plain path (host reads between steps, no filter)")
    ap.add_argument("--ghost-model", default=None, help="trained filter (.npz) for the batched path")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    model = None
    if a.ghost_model and not a.plain:
        from .ghost import GhostModel
        model = GhostModel.load(a.ghost_model)
    say = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    res = run_loads(a.loads, a.requests, batched=not a.plain, ghost_model=model, log=say)
    text = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
