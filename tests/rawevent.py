"""Builds raw VP events from pixel lists (test helper).

Encodes the reference's raw layout (ClusteringDefinitions.cu:3-21): 208 banks
in sensor order, SP words sorted by address, bit 31 set exactly when none of
the 8 neighbouring SPs of the same sensor is occupied.
"""
from __future__ import annotations

import numpy as np


def encode(pixels, n_sensors=208, sort=True):
    """pixels: iterable of (sensor, col, row). Returns (events uint8, offsets uint32)."""
    sps = {}
    for s, c, r in pixels:
        key = (int(s), (int(c) // 2 << 6) | (int(r) // 4))
        sps[key] = sps.get(key, 0) | (1 << ((int(c) % 2) * 4 + int(r) % 4))
    words = []
    payload = []
    offs = []
    pos = 0
    for s in range(n_sensors):
        keys = [k for k in sps if k[0] == s]
        if sort:
            keys.sort(key=lambda k: k[1])
        occ = {k[1] for k in keys}
        bank = [s, len(keys)]
        for k in keys:
            addr = k[1]
            col, row = addr >> 6, addr & 0x3F
            neigh = any(((col + dc) << 6 | (row + dr)) in occ
                        for dc in (-1, 0, 1) for dr in (-1, 0, 1)
                        if (dc or dr) and 0 <= col + dc < 384 and 0 <= row + dr < 64)
            bank.append((0 if neigh else 0x80000000) | (addr << 8) | sps[k])
        offs.append(pos * 4)
        payload.extend(bank)
        pos += len(bank)
    offs.append(pos * 4)
    ev = np.array([n_sensors] + offs + payload, dtype=np.uint32)
    return ev.view(np.uint8), np.array([0, ev.nbytes], np.uint32)


def concat(events):
    """Concatenate (events, offsets) pairs of single events."""
    parts = [e for e, _ in events]
    offs = np.zeros(len(parts) + 1, np.uint32)
    offs[1:] = np.cumsum([p.nbytes for p in parts])
    return np.concatenate(parts) if parts else np.zeros(0, np.uint8), offs


def sensor_frames(geometry: bytes):
    """(R 3x3, T 3) per sensor from the geometry blob's ltg (row-major R, T at 9..11)."""
    g = np.frombuffer(geometry, np.float32)
    ltg = g[1536:].reshape(208, 16).astype(np.float64)
    return [(l[:9].reshape(3, 3), l[9:12]) for l in ltg], g[:768].astype(np.float64), g[768:1536].astype(np.float64)


def line_pixels(geometry: bytes, x0, y0, z0, tx, ty, forward=True, rmin=5.1, rmax=42.0):
    """Pixels (sensor, col, row) where the line through (x0,y0,z0) with slopes
    (tx, ty) crosses an active sensor; one pixel per crossing."""
    frames, local_x, pitch = sensor_frames(geometry)
    edges = np.concatenate([[0.0], np.cumsum(pitch)])
    out = []
    for s, (R, T) in enumerate(frames):
        z = T[2]
        if (z - z0 <= 0) == forward:
            continue
        x = x0 + tx * (z - z0)
        y = y0 + ty * (z - z0)
        r = np.hypot(x, y)
        if r < rmin or r > rmax:
            continue
        d = np.array([x, y, z]) - T
        lx = R[0, 0] * d[0] + R[1, 0] * d[1]
        ly = R[0, 1] * d[0] + R[1, 1] * d[1]
        if lx < 0 or lx >= edges[-1] or ly < 0 or ly >= 256 * 0.055:
            continue
        col = int(np.searchsorted(edges, lx, side="right") - 1)
        row = int(ly / 0.055)
        out.append((s, col, row))
    return out


def scramble_banks(events: np.ndarray, offsets: np.ndarray, seed: int = 0) -> np.ndarray:
    """Copy of the raw events with the SP words of some banks out of order.

    Per bank, at random: unchanged, shuffled, reversed, or one SP word copied
    over its successor (a duplicate address).  Bank sizes and offsets stay
    the same, so the result is a valid raw event of the same layout whose
    banks are no longer strictly increasing in SP address.
    """
    rng = np.random.default_rng(seed)
    out = events.copy()
    for e in range(len(offsets) - 1):
        lo, hi = int(offsets[e]), int(offsets[e + 1])
        w = out[lo:hi].view(np.uint32)
        n_banks = int(w[0])
        base = n_banks + 2
        for b in range(n_banks):
            s = base + int(w[1 + b]) // 4
            n = int(w[s + 1])
            if n < 2:
                continue
            sps = w[s + 2: s + 2 + n]
            mode = rng.integers(0, 4)
            if mode == 1:
                sps[:] = sps[rng.permutation(n)]
            elif mode == 2:
                sps[:] = sps[::-1].copy()
            elif mode == 3:
                j = int(rng.integers(0, n - 1))
                sps[j + 1] = sps[j]
    return out


def malformed(events: np.ndarray, offsets: np.ndarray):
    """Damaged copies of the first four events, in four ways, plus the
    same events with every damaged bank emptied (its SP count set to 0),
    which is what a decoder that skips damaged banks must see:
    0: the event cut inside the bank stored last;
    1: bank 7's offset points past the event;
    2: a bank count the event's offset table cannot hold (every bank damaged);
    3: bank 100's SP count larger than the event.
    Returns (events, offsets, clean_events, clean_offsets, damaged) with
    damaged[e] the set of damaged bank numbers."""
    out, clean, damaged = [], [], []
    for e in range(4):
        w = np.frombuffer(events[offsets[e]:offsets[e + 1]].tobytes(), np.uint32).copy()
        nb = int(w[0])
        header = lambda b: nb + 2 + int(w[1 + b]) // 4
        c = w.copy()
        if e == 0:
            last = max(range(nb), key=header)  # the bank stored last
            end = header(last) + 2 + int(w[header(last) + 1])
            c[header(last) + 1] = 0
            w = w[: end - 1] if end - 1 > header(last) + 1 else w[: header(last) + 1]
            bad = {last}
        elif e == 1:
            c[header(7) + 1] = 0
            w[1 + 7] = 4 * len(w) + 400
            bad = {7}
        elif e == 2:
            w[0] = len(w)
            c = np.array([0], np.uint32)  # no banks at all
            bad = set(range(min(nb, 208)))
        else:
            w[header(100) + 1] = len(w)
            c[header(100) + 1] = 0
            bad = {100}
        out.append((w.view(np.uint8), None))
        clean.append((c.view(np.uint8), None))
        damaged.append(bad)
    ev, off = concat(out)
    cev, coff = concat(clean)
    return ev, off, cev, coff, damaged
