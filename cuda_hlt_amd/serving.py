"""Trigger-style serving loop with continuous batching.

Requests (1-2000 raw events each) arrive over time, their events in pinned
host memory (as a DMA-capable receive buffer would hold them).  The server
keeps K slots, each with resident device buffers for one batch (raw events,
offsets, largest-first order) and its own sequence Stream (HIP stream,
device arena, captured graphs).  Whenever a slot is free, the dispatcher
takes every queued request that fits the slot (events and bytes), in
arrival order, as one batch and hands it to the slot's worker thread, which
copies each request's events straight from its pinned buffer to its place
in the slot's device buffer, replays the captured sequence and copies the
consolidated output into a device buffer of the batch that the requests'
results view.  Requests that arrive while batches run wait only until the
next slot frees up, and the slots' work overlaps on the GPU.

Results stay in device memory (DeviceResult: views of the batch's output
for the request's events, in the reference layouts, ready for a consumer on
the GPU); DeviceResult.host() copies one to the host.  Server(...,
host_results=True) copies every batch back and cuts it on the host instead.

Admission is by HBM footprint: a slot's arena is sized once, at start-up,
for the largest batch it admits (the batched path's bounds for
slot_events events and slot_bytes raw bytes); the number of slots is the
HBM budget divided by that footprint (at most max_slots).  Batches are
padded with empty events to a size bucket (64 to 16384 events) so a slot
replays one of a few graphs, captured for every bucket before the first
request (Server.warm); empty events produce no hits and no tracks, and
every event's output is independent of the other events in its batch.
Large slots keep the host's per-batch work (placing requests, cutting
results) small against the device time of the batch.
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

BUCKETS = (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)


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
    result: object = None    # DeviceResult, or the host dict of cut_results
    pinned: object = None    # the events in pinned host memory (torch uint8), if any

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


class _DeviceBytes:
    """A device address range seen by torch (no copy, no ownership)."""

    def __init__(self, ptr: int, nbytes: int):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (int(ptr), False),
                                         "version": 2}


def device_view(ptr: int, nbytes: int):
    import torch
    if nbytes == 0:
        return torch.empty(0, dtype=torch.uint8, device="cuda")
    return torch.as_tensor(_DeviceBytes(ptr, nbytes), device="cuda")


class BatchOutput:
    """A batch's consolidated output copied out of the slot's arena into a
    device buffer of its own: atomics u32[2N+1], hit numbers u32[T+1],
    hits x, y, z, id (16 B per hit, SoA), states (12 floats SoA + backward
    byte per track), accepted byte per track."""

    def __init__(self, seq, n_events: int, stream):
        import torch
        o = seq.output()
        self.n, self.T, self.Ht = n_events, int(o.total_tracks), int(o.total_track_hits)
        n, T, Ht = self.n, self.T, self.Ht
        parts = [(o.dev_atomics_velo, 4 * (2 * n + 1)), (o.dev_track_hit_number, 4 * (T + 1)),
                 (o.dev_track_hits, 16 * Ht), (o.dev_states, 49 * T), (o.dev_accepted, T)]
        sizes = [b for _, b in parts]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        with torch.cuda.stream(stream):
            self.buf = torch.empty(int(self.offsets[-1]), dtype=torch.uint8, device="cuda")
            for (ptr, nb), off in zip(parts, self.offsets[:-1]):
                if nb:
                    self.buf[int(off): int(off) + nb].copy_(device_view(ptr, nb), non_blocking=True)
            # the track and hit offsets of every event, for cutting
            self.atomics = self.part(0).view(torch.int32).cpu().numpy().view(np.uint32)
        self.hit_number = self.part(1).view(torch.int32)

    def part(self, k: int):
        return self.buf[int(self.offsets[k]): int(self.offsets[k + 1])]


class DeviceResult:
    """One request's output in device memory: views of its batch's
    BatchOutput for events [e0, e0 + n) (tracks [t0, t1), hits [h0, h1))."""

    def __init__(self, batch: BatchOutput, e0: int, n: int, h0: int, h1: int):
        N = batch.n
        self.batch, self.e0, self.n = batch, e0, n
        self.t0, self.t1 = int(batch.atomics[N + e0]), int(batch.atomics[N + e0 + n])
        self.h0, self.h1 = h0, h1

    @property
    def n_tracks(self) -> int:
        return self.t1 - self.t0

    def hits(self):
        """x, y, z (float32) and LHCbID (int32) of the request's track hits."""
        import torch
        b, Ht = self.batch, self.batch.Ht
        soa = b.part(2)
        return [soa[4 * (k * Ht + self.h0): 4 * (k * Ht + self.h1)].view(torch.float32 if k < 3 else torch.int32)
                for k in range(4)]

    def host(self) -> dict:
        """The request's result on the host, as cut_results gives it."""
        import torch
        b, T = self.batch, self.batch.T
        t0, t1, h0, h1 = self.t0, self.t1, self.h0, self.h1
        hits = torch.cat([x.view(torch.uint8) for x in self.hits()]).cpu().numpy()
        S = b.part(3)
        states = torch.cat([S[4 * (k * T + t0): 4 * (k * T + t1)] for k in range(12)] + [S[48 * T + t0: 48 * T + t1]])
        hn = b.hit_number[t0: t1 + 1].cpu().numpy().view(np.uint32)
        return {"tracks_per_event": b.atomics[self.e0: self.e0 + self.n].copy(),
                "hit_number": (hn - np.uint32(h0)).astype(np.uint32),
                "hits": hits, "states": states.cpu().numpy(),
                "accepted": b.part(4)[t0:t1].cpu().numpy()}


def cut_device(batch: BatchOutput, requests, first_event: list):
    """Per-request DeviceResults of a batch (one small gather of the hit
    offsets at the requests' track boundaries)."""
    import torch
    N = batch.n
    idx = []
    for r, e0 in zip(requests, first_event):
        idx += [int(batch.atomics[N + e0]), int(batch.atomics[N + e0 + r.n])]
    hn = batch.hit_number[torch.as_tensor(idx, dtype=torch.int64, device="cuda")].cpu().numpy().view(np.uint32)
    return [DeviceResult(batch, e0, r.n, int(hn[2 * k]), int(hn[2 * k + 1]))
            for k, (r, e0) in enumerate(zip(requests, first_event))]


class Slot:
    """Resident buffers + one sequence Stream for batches of up to
    max_events events / max_bytes raw bytes."""

    def __init__(self, index, geometry, zs, max_events, max_bytes, batched, tuning, ghost_model_dev,
                 host_results=False):
        import torch
        self.index = index
        self.max_events = max_events
        self.max_bytes = max_bytes
        self.batched = batched
        self.tuning = tuning
        self.ghost = ghost_model_dev
        self.host_results = host_results
        self.reserve_mb = autotune.batched_reserve_mb(max_events, max_bytes)
        self.seq = VeloSequence(geometry, zs, max_events=max_events, reserve_mb=self.reserve_mb)
        self.stream = torch.cuda.ExternalStream(self.seq.stream)
        # every torch copy and allocation of the slot runs on a stream torch
        # owns (pinned buffers record their copies on it, and it outlives
        # the sequence's stream); the two streams wait for each other
        self.copy_stream = torch.cuda.Stream()
        self.dev_events = torch.zeros(max_bytes, dtype=torch.uint8, device="cuda")
        self.dev_offsets = torch.zeros(max_events + 1, dtype=torch.int32, device="cuda")
        self.dev_order = torch.zeros(max_events, dtype=torch.int32, device="cuda")
        # padding events, copied on the device
        n_pad = min(max_events, max_bytes // EMPTY.nbytes)
        self.dev_empty = torch.from_numpy(np.tile(EMPTY, max(n_pad, 1))).cuda()
        self.host_events = torch.zeros(max_bytes, dtype=torch.uint8).pin_memory()
        self.host_offsets = torch.zeros(max_events + 1, dtype=torch.int32).pin_memory()
        self.host_order = torch.zeros(max_events, dtype=torch.int32).pin_memory()

    def run(self, requests, bucket=None):
        """Places the requests (plus empty events up to the bucket, at least
        `bucket` when given) in the slot's device buffers, runs the sequence
        and returns (per-request results, bucket)."""
        import torch
        n_real = sum(r.n for r in requests)
        bucket = max(bucket or 0, bucket_of(n_real)) if self.batched else n_real
        first, offs, pos, e = [], [np.zeros(1, np.int64)], 0, 0
        he = None
        with torch.cuda.stream(self.copy_stream):
            for r in requests:
                first.append(e)
                nb = int(r.offsets[-1])
                if pos + nb > self.max_bytes:
                    raise ValueError("batch exceeds the slot's raw-byte capacity")
                if r.pinned is not None:
                    self.dev_events[pos: pos + nb].copy_(r.pinned[:nb], non_blocking=True)
                else:
                    # not pinned: staged through the slot's pinned buffer
                    he = self.host_events.numpy() if he is None else he
                    he[pos: pos + nb] = r.events[:nb]
                    self.dev_events[pos: pos + nb].copy_(self.host_events[pos: pos + nb], non_blocking=True)
                offs.append(pos + r.offsets[1:].astype(np.int64))
                pos += nb
                e += r.n
            pad = bucket - n_real
            if pad:
                nbp = pad * EMPTY.nbytes
                if pos + nbp > self.max_bytes or nbp > self.dev_empty.numel():
                    raise ValueError("batch exceeds the slot's raw-byte capacity")
                self.dev_events[pos: pos + nbp].copy_(self.dev_empty[:nbp], non_blocking=True)
                offs.append(pos + EMPTY.nbytes * np.arange(1, pad + 1, dtype=np.int64))
                pos += nbp
            off = np.concatenate(offs)
            self.host_offsets.numpy()[: bucket + 1] = off.astype(np.uint32).view(np.int32)
            sizes = np.diff(off)
            self.host_order.numpy()[:bucket] = np.argsort(-sizes, kind="stable").astype(np.int32)
            self.dev_offsets[: bucket + 1].copy_(self.host_offsets[: bucket + 1], non_blocking=True)
            self.dev_order[:bucket].copy_(self.host_order[:bucket], non_blocking=True)
        self.stream.wait_stream(self.copy_stream)
        kw = dict(dev_events=self.dev_events.data_ptr(), dev_offsets=self.dev_offsets.data_ptr(),
                  copy_outputs=self.host_results, tuning=self.tuning,
                  dev_ghost_model=None if self.ghost is None else self.ghost.data_ptr())
        if self.batched:
            kw.update(batched=True, events_size=self.max_bytes, dev_work_order=self.dev_order.data_ptr())
        else:
            kw.update(events_size=pos)
        self.seq.run(bucket, **kw)
        if self.host_results:
            return cut_results(self.seq.host_tracks(), bucket, requests, first), bucket
        self.copy_stream.wait_stream(self.stream)
        batch = BatchOutput(self.seq, bucket, self.copy_stream)
        with torch.cuda.stream(self.copy_stream):
            results = cut_device(batch, requests, first)
        self.copy_stream.synchronize()
        return results, bucket

    def close(self):
        """Frees the slot's buffers, then the sequence (and its stream).
        Torch never used the sequence's stream for a copy or an
        allocation, so nothing records an event on it afterwards."""
        import torch
        torch.cuda.synchronize()
        self.host_events = self.host_offsets = self.host_order = None
        self.dev_events = self.dev_offsets = self.dev_order = self.dev_empty = None
        torch.cuda.synchronize()
        self.stream = None
        self.seq.close()


class Server:
    """Continuous-batching server over K slots (see the module docstring)."""

    def __init__(self, geometry, zs, hbm_budget_mb=131072, max_slots=6, slot_events=16384, slot_bytes=None,
                 batched=True, tuning=None, ghost_model=None, host_results=False):
        slot_bytes = slot_bytes or slot_events * 40000
        if tuning is None and batched:
            tuning = dict(autotune.BATCHED_DEFAULT)
        per_slot = autotune.batched_reserve_mb(slot_events, slot_bytes)
        self.n_slots = max(1, min(max_slots, hbm_budget_mb // per_slot))
        self.slot_events, self.slot_bytes = slot_events, slot_bytes
        self.footprint_mb = per_slot
        ghost_dev = ghost_model.to_device() if ghost_model is not None else None
        self._ghost = ghost_dev
        self.host_results = host_results
        self.slots = [Slot(i, geometry, zs, slot_events, slot_bytes, batched, tuning, ghost_dev, host_results)
                      for i in range(self.n_slots)]
        self.free = queue.Queue()
        for s in self.slots:
            self.free.put(s)
        self.pending = []
        self.cv = threading.Condition()
        self.batches: list[BatchRecord] = []
        self.errors: list[BaseException] = []

    def warm(self):
        """Captures every slot's graph for every bucket it can take (a batch
        of empty events each), so no request waits for a capture."""
        if not self.slots[0].batched:
            return
        for slot in self.slots:
            for b in BUCKETS:
                if b <= self.slot_events and b * EMPTY.nbytes <= self.slot_bytes:
                    slot.run([], bucket=b)

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


def make_requests(n_requests, rate_hz, seed=0, min_events=1, max_events=2000, pinned=True, pool=None):
    """Poisson arrivals at rate_hz of requests of U(min, max) uneven events
    (per-event nu ~ U(2, 11)), their events in pinned host memory unless
    pinned=False.  pool: (events, pinned copy or None) from make_pool to
    draw the requests' events from instead of generating them."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(min_events, max_events + 1, n_requests)
    if pool is None:
        ev = gen.uneven(int(sizes.sum()), seed=seed + 1)
        pinned_all = pin(ev.events) if pinned else None
    else:
        ev, pinned_all = pool
        if int(sizes.sum()) > ev.n_events:
            raise ValueError("the pool holds fewer events than the requests need")
    gaps = rng.exponential(1.0 / rate_hz, n_requests) if rate_hz > 0 else np.zeros(n_requests)
    t = np.cumsum(gaps) - gaps[0]
    reqs, e = [], 0
    for k, (s, a) in enumerate(zip(sizes, t)):
        sub = ev.subset(e, e + int(s))
        r = Request(k, sub.events, sub.offsets, float(a))
        if pinned_all is not None:
            r.pinned = pinned_all[int(ev.offsets[e]): int(ev.offsets[e + int(s)])]
        reqs.append(r)
        e += int(s)
    return reqs, ev


def make_pool(n_events, seed=0, pinned=True):
    """Uneven events (and their pinned copy) for make_requests(pool=...)."""
    ev = gen.uneven(n_events, seed=seed)
    return ev, (pin(ev.events) if pinned else None)


def pin(events: np.ndarray):
    """A pinned host copy of raw events (torch uint8)."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(events)).pin_memory()


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
              hbm_budget_mb=131072, max_slots=6, log=None, host_results=False, slot_events=16384):
    """Peak throughput (every request queued at t = 0), then Poisson
    arrivals at the given fractions of it."""
    geometry, zs = gen.geometry(), gen.module_zs()
    srv = Server(geometry, zs, hbm_budget_mb=hbm_budget_mb, max_slots=max_slots, batched=batched, tuning=tuning,
                 host_results=host_results, ghost_model=ghost_model, slot_events=slot_events)
    try:
        # one pool of events for every run: requests of U(1, 2000) events
        # need 1000 per request on average (the margin covers > 4 sigma)
        pool = make_pool(int(n_requests * 1000 * 1.2) + 2000, seed=seed + 1)
        srv.warm()
        warm, _ = make_requests(max(8, srv.n_slots * 2), 0, seed=seed + 7, pool=pool)
        srv.serve(warm)
        srv.batches.clear()
        reqs, _ = make_requests(n_requests, 0, seed=seed, pool=pool)
        wall = srv.serve(reqs)
        peak = latency_summary(reqs, srv.batches, wall)
        if log:
            log(f"[serving] peak {peak['events_per_s']} events/s over {srv.n_slots} slots")
        mean_events = np.mean([r.n for r in reqs])
        out = {"slots": srv.n_slots, "slot_footprint_mb": srv.footprint_mb, "batched": batched,
               "filtered": ghost_model is not None, "results": "host" if host_results else "device",
               "peak": peak, "loads": {}}
        for f in loads:
            srv.batches.clear()
            rate = f * peak["events_per_s"] / mean_events
            reqs, _ = make_requests(n_requests, rate, seed=seed + int(100 * f), pool=pool)
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
    ap.add_argument("--max-slots", type=int, default=6)
    ap.add_argument("--slot-events", type=int, default=16384, help="largest batch a slot takes")
    ap.add_argument("--host-results", action="store_true",
                    help="copy every batch's output to the host and cut it there")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    model = None
    if a.ghost_model and not a.plain:
        from .ghost import GhostModel
        model = GhostModel.load(a.ghost_model)
    say = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    res = run_loads(a.loads, a.requests, batched=not a.plain, ghost_model=model, log=say,
                    host_results=a.host_results, max_slots=a.max_slots, slot_events=a.slot_events)
    text = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
