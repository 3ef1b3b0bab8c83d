"""VELO track checker semantics (test-side restatement).

MCAssociator.h:166-182 / MCAssociator.cpp:13-60: each track LHCbID is
looked up in the sorted (LHCbID -> MC particle) map (first entry wins);
weight(p) = hits matched to p / hits matched to any particle.
TrackChecker.cpp:349-405: a track is a ghost when nothing matches or the
best weight is below 0.7 (TrackChecker.h:137); otherwise its particle is
found (repeats are clones).  Efficiency for "Not electron Velo"
(CategoriesTrackChecker.cpp:215-216) = found / (hasVelo and not electron).
"""
from __future__ import annotations

import numpy as np

HAS_VELO, IS_ELECTRON = 0x1, 0x4


def track_lists(r):
    """Per event, list of uint32 LHCbID arrays of the consolidated tracks."""
    n = r.n_events
    T = int(r.atomics[2 * n])
    Ht = int(r.hit_number[T])
    ids = r.track_hits[12 * Ht: 16 * Ht].view(np.uint32)
    out = []
    for e in range(n):
        off = int(r.atomics[n + e])
        out.append([ids[r.hit_number[off + t]: r.hit_number[off + t + 1]] for t in range(int(r.atomics[e]))])
    return out


def velo_efficiency(ev, r, accepted=None):
    """(efficiency, ghost rate, per-track ghost flags) over all events."""
    per_event = track_lists(r)
    mcp_event = ev.mcps["event"]
    found = accepted_n = ghosts = ntracks = 0
    flags = []
    k = 0
    for e, tracks in enumerate(per_event):
        sel = np.nonzero(mcp_event == e)[0]
        id2p = {}
        for p in sel:
            m = ev.mcps[p]
            for h in ev.mc_hits[m["hit_offset"]: m["hit_offset"] + m["n_hits"]]:
                id2p.setdefault(int(h), int(p))
        seen = set()
        for t in tracks:
            if accepted is not None and not accepted[k]:
                k += 1
                continue
            k += 1
            ntracks += 1
            cnt = {}
            matched = 0
            for i in t:
                p = id2p.get(int(i))
                if p is None:
                    continue
                matched += 1
                cnt[p] = cnt.get(p, 0) + 1
            ghost = True
            if cnt:
                p, c = max(cnt.items(), key=lambda kv: (kv[1], -kv[0]))
                if c / matched >= 0.7:
                    ghost = False
                    m = ev.mcps[p]
                    if (m["flags"] & HAS_VELO) and not (m["flags"] & IS_ELECTRON):
                        seen.add(p)
            ghosts += ghost
            flags.append(ghost)
        acc = [p for p in sel if (ev.mcps[p]["flags"] & HAS_VELO) and not (ev.mcps[p]["flags"] & IS_ELECTRON)]
        accepted_n += len(acc)
        found += len(seen)
    eff = found / accepted_n if accepted_n else 0.0
    ghost_rate = ghosts / ntracks if ntracks else 0.0
    return eff, ghost_rate, np.array(flags, bool)
