"""Summaries of rocprofv3 CSV output for the VELO kernels.

rocprofv3 writes one row per (dispatch, counter) with ``--pmc`` and one row
per kernel with ``--stats``.  This module turns those files into the short
tables kept beside the benchmark:

* ``stats``    kernel time per kernel (calls, total, average, share);
* ``counters`` the mean of each counter per dispatch, per kernel;
* ``trace``    mean duration per kernel and launch shape (--kernel-trace);
* ``traffic``  HBM bytes per dispatch from two ``--pmc`` runs (FETCH_SIZE,
  then WRITE_SIZE), each scaled by what the same counter reads for a copy
  kernel of known size run in the same process.

Usage::

    python -m cuda_hlt_amd.profiling stats --csv DIR --command "..." --title "..."
    python -m cuda_hlt_amd.profiling counters --csv DIR --command "..." --title "..."
    python -m cuda_hlt_amd.profiling trace --csv DIR --command "..." --title "..."
    python -m cuda_hlt_amd.profiling traffic --fetch DIR --write DIR --copy-bytes N \\
        --events-per-launch 10000 --command "..." --out-json F --out-md F
"""
from __future__ import annotations

import argparse
import csv
import json
import re
import sys
from collections import defaultdict
from pathlib import Path


def short_name(name: str) -> str:
    """'velo::search_by_triplet_kernel(args)' -> 'search_by_triplet_kernel',
    'void velo::consolidate_kernel<false>(args)' -> 'consolidate_kernel<false>'."""
    head = name.split("(")[0].strip()
    if head.startswith("void "):
        head = head[5:]
    if head.startswith("velo::"):
        return head[len("velo::"):]
    if "<" in head:  # templated library kernels keep a readable prefix
        return re.sub(r"\s+", " ", head)[:120]
    return head.split("::")[-1]


def _is_velo_kernel(name: str) -> bool:
    return "_kernel" in name and not name.startswith(("at::", "__amd"))


def _find(path: str | Path, suffix: str) -> Path:
    p = Path(path)
    if p.is_file():
        return p
    hits = sorted(p.rglob(f"*{suffix}"))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {p}")
    return hits[0]


def read_counters(path: str | Path) -> dict[str, dict[str, float]]:
    """Mean value per dispatch of every counter, per kernel (short name)."""
    f = _find(path, "counter_collection.csv")
    per_dispatch: dict[tuple[str, str, str], float] = defaultdict(float)
    with open(f, newline="") as fh:
        for row in csv.DictReader(fh):
            key = (short_name(row["Kernel_Name"]), row.get("Dispatch_Id", row.get("Correlation_Id", "")),
                   row["Counter_Name"])
            per_dispatch[key] += float(row["Counter_Value"])
    sums: dict[str, dict[str, list[float]]] = defaultdict(lambda: defaultdict(list))
    for (k, _, c), v in per_dispatch.items():
        sums[k][c].append(v)
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} | {"dispatches": max(len(v) for v in cs.values())}
            for k, cs in sums.items()}


def read_stats(path: str | Path) -> list[dict]:
    f = _find(path, "kernel_stats.csv")
    out = []
    with open(f, newline="") as fh:
        for row in csv.DictReader(fh):
            out.append({"kernel": short_name(row["Name"]), "calls": int(row["Calls"]),
                        "total_us": float(row["TotalDurationNs"]) / 1e3,
                        "average_us": float(row["AverageNs"]) / 1e3, "percent": float(row["Percentage"])})
    return out


def read_trace(path: str | Path) -> list[dict]:
    """Per (kernel, workgroups, workgroup size): dispatches and mean duration
    from a --kernel-trace CSV, largest total first."""
    f = _find(path, "kernel_trace.csv")
    groups: dict[tuple, list[float]] = defaultdict(list)
    with open(f, newline="") as fh:
        for row in csv.DictReader(fh):
            wg = int(row.get("Workgroup_Size_X") or row.get("Workgroup_Size") or 1)
            grid = int(row.get("Grid_Size_X") or row.get("Grid_Size") or 0)
            dur = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3
            groups[(short_name(row["Kernel_Name"]), grid // max(wg, 1), wg)].append(dur)
    out = [{"kernel": k, "workgroups": g, "workgroup_size": w, "dispatches": len(v), "average_us": sum(v) / len(v),
            "total_us": sum(v)} for (k, g, w), v in groups.items()]
    return sorted(out, key=lambda r: -r["total_us"])


def _header(title: str, command: str) -> list[str]:
    return [f"# {title}", "", "Command (run from the repository root):", "", "```", command, "```", ""]


def stats_markdown(path, title, command) -> str:
    lines = _header(title, command) + ["| kernel | calls | total us | average us | % |", "|---|---|---|---|---|"]
    for r in read_stats(path):
        lines.append(f"| `{r['kernel']}` | {r['calls']} | {r['total_us']:.1f} | {r['average_us']:.1f} | "
                     f"{r['percent']:.2f} |")
    return "\n".join(lines) + "\n"


def trace_markdown(path, title, command) -> str:
    lines = _header(title, command) + ["| kernel | workgroups | workgroup size | dispatches | average us | total us |",
                                       "|---|---|---|---|---|---|"]
    for r in read_trace(path):
        lines.append(f"| `{r['kernel']}` | {r['workgroups']} | {r['workgroup_size']} | {r['dispatches']} | "
                     f"{r['average_us']:.1f} | {r['total_us']:.1f} |")
    return "\n".join(lines) + "\n"


def counters_markdown(path, title, command) -> str:
    data = read_counters(path)
    names = sorted({c for v in data.values() for c in v if c != "dispatches"})
    lines = _header(title, command) + ["Mean per dispatch.", "",
                                       "| kernel | dispatches | " + " | ".join(names) + " |",
                                       "|---|---|" + "---|" * len(names)]
    for k, v in sorted(data.items()):
        lines.append(f"| `{k}` | {v['dispatches']} | " + " | ".join(f"{v.get(c, float('nan')):.4g}" for c in names)
                     + " |")
    return "\n".join(lines) + "\n"


def traffic(fetch_dir, write_dir, copy_bytes: int, copy_match: str = "add") -> dict:
    """Calibrated HBM bytes per dispatch, per kernel."""
    fetch = read_counters(fetch_dir)
    write = read_counters(write_dir)

    def calibration(data, counter):
        cands = [(v[counter], k) for k, v in data.items() if counter in v and copy_match in k.lower()
                 and not _is_velo_kernel(k)]
        if not cands:
            raise ValueError(f"no copy kernel matching '{copy_match}' with {counter}")
        kib, name = max(cands)
        return copy_bytes / (kib * 1024.0), name, kib

    ff, fname, fkib = calibration(fetch, "FETCH_SIZE")
    wf, wname, wkib = calibration(write, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        if not _is_velo_kernel(k):
            continue
        fb = fetch.get(k, {}).get("FETCH_SIZE")
        wb = write.get(k, {}).get("WRITE_SIZE")
        if fb is None or wb is None:
            continue
        kernels[k] = {"fetch_kib_raw": fb, "write_kib_raw": wb,
                      "fetch_bytes": round(fb * 1024.0 * ff), "write_bytes": round(wb * 1024.0 * wf)}
    return {"copy_kernel": {"fetch": fname, "write": wname, "bytes": copy_bytes, "fetch_kib": fkib,
                            "write_kib": wkib, "fetch_factor": ff, "write_factor": wf},
            "kernels": kernels}


def traffic_markdown(t: dict, title: str, command: str) -> str:
    c = t["copy_kernel"]
    lines = _header(title, command) + [
        f"Calibration: a copy kernel that reads and writes {c['bytes']} bytes reads FETCH_SIZE = "
        f"{c['fetch_kib']:.0f} KiB (factor {c['fetch_factor']:.4f}) and WRITE_SIZE = {c['write_kib']:.0f} KiB "
        f"(factor {c['write_factor']:.4f}).", "",
        "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | fetch bytes | write bytes |", "|---|---|---|---|---|"]
    for k, v in t["kernels"].items():
        lines.append(f"| `{k}` | {v['fetch_kib_raw']:.0f} | {v['write_kib_raw']:.0f} | {v['fetch_bytes']} | "
                     f"{v['write_bytes']} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("stats", "counters", "trace"):
        p = sub.add_parser(name)
        p.add_argument("--csv", required=True)
        p.add_argument("--command", required=True)
        p.add_argument("--title", required=True)
    p = sub.add_parser("traffic")
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--copy-bytes", type=int, required=True)
    p.add_argument("--copy-match", default="add")
    p.add_argument("--events-per-launch", type=int, required=True)
    p.add_argument("--command", required=True)
    p.add_argument("--title", default="HBM traffic per dispatch")
    p.add_argument("--out-json", required=True)
    p.add_argument("--out-md", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.title, a.command))
    elif a.cmd == "counters":
        sys.stdout.write(counters_markdown(a.csv, a.title, a.command))
    elif a.cmd == "trace":
        sys.stdout.write(trace_markdown(a.csv, a.title, a.command))
    else:
        t = traffic(a.fetch, a.write, a.copy_bytes, a.copy_match)
        t["events_per_launch"] = a.events_per_launch
        t["command"] = a.command
        Path(a.out_json).write_text(json.dumps(t, indent=1) + "\n")
        Path(a.out_md).write_text(traffic_markdown(t, a.title, a.command))


if __name__ == "__main__":
    main()
