"""Summaries of rocprofv3 CSV output for the VELO kernels.

rocprofv3 writes one row per (dispatch, counter) with ``--pmc`` and one row
per kernel with ``--stats``.  This module turns those files into the short
tables kept beside the benchmark:

* ``stats``    kernel time per kernel (calls, total, average, share);
* ``counters`` the mean of each counter per dispatch, per kernel;
* ``trace``    mean duration per kernel and launch shape (--kernel-trace);
* ``traffic``  HBM bytes per dispatch from two ``--pmc`` runs (FETCH_SIZE,
  then WRITE_SIZE), each scaled by what the same counter reads for a copy
  kernel of known size run in the same process;
* ``lds``      LDS throughput, bank conflicts and occupancy per kernel from
  one ``--pmc`` run of SQ_INSTS_LDS_LOAD_BANDWIDTH,
  SQ_INSTS_LDS_STORE_BANDWIDTH, SQ_LDS_BANK_CONFLICT, SQ_WAVE_CYCLES and
  GRBM_GUI_ACTIVE, normalised as rocprofv3's derived LDSBankConflict and
  OccupancyPercent are, beside the occupancy the kernel's registers and
  LDS allow.

Usage::

    python -m cuda_hlt_amd.profiling stats --csv DIR --command "..." --title "..."
    python -m cuda_hlt_amd.profiling counters --csv DIR --command "..." --title "..."
    python -m cuda_hlt_amd.profiling trace --csv DIR --command "..." --title "..."
    python -m cuda_hlt_amd.profiling traffic --fetch DIR --write DIR --copy-bytes N \\
        --events-per-launch 10000 --command "..." --out-json F --out-md F
    python -m cuda_hlt_amd.profiling lds --csv DIR [--trace DIR] --command "..." --title "..." \\
        [--out-json F]
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


def read_dispatches(path: str | Path, kernel: str) -> list[float]:
    """Durations (us) of every dispatch of `kernel`, in dispatch order."""
    f = _find(path, "kernel_trace.csv")
    with open(f, newline="") as fh:
        rows = [r for r in csv.DictReader(fh) if short_name(r["Kernel_Name"]) == kernel]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows]


def _header(title: str, command: str) -> list[str]:
    return [f"# {title}", "", "Command (run from the repository root):", "", "```", command, "```", ""]


def stats_markdown(path, title, command) -> str:
    lines = _header(title, command) + ["| kernel | calls | total us | average us | % |", "|---|---|---|---|---|"]
    for r in read_stats(path):
        lines.append(f"| `{r['kernel']}` | {r['calls']} | {r['total_us']:.1f} | {r['average_us']:.1f} | "
                     f"{r['percent']:.2f} |")
    return "\n".join(lines) + "\n"


def trace_markdown(path, title, command, per_dispatch=None, first=0) -> str:
    """per_dispatch: a kernel whose dispatches are also listed one by one,
    the first `first` of them (the plain sequence's) averaged apart."""
    lines = _header(title, command) + ["| kernel | workgroups | workgroup size | dispatches | average us | total us |",
                                       "|---|---|---|---|---|---|"]
    for r in read_trace(path):
        lines.append(f"| `{r['kernel']}` | {r['workgroups']} | {r['workgroup_size']} | {r['dispatches']} | "
                     f"{r['average_us']:.1f} | {r['total_us']:.1f} |")
    if per_dispatch:
        d = read_dispatches(path, per_dispatch)
        lines += ["", f"`{per_dispatch}` per dispatch, in order (us): " + ", ".join(f"{x:.1f}" for x in d) + "."]
        if 0 < first <= len(d):
            lines.append(f"Mean of the first {first} (the plain sequence): {sum(d[:first]) / first:.1f} us; "
                         f"of the rest: {sum(d[first:]) / max(1, len(d) - first):.1f} us.")
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


def read_agent(path: str | Path) -> dict:
    """The GPU agent's row of rocprofv3's agent_info.csv."""
    f = _find(path, "agent_info.csv")
    with open(f, newline="") as fh:
        for row in csv.DictReader(fh):
            if row.get("Agent_Type", "").upper() == "GPU" or int(row.get("Simd_Count") or 0) > 0:
                return row
    raise ValueError(f"no GPU agent in {f}")


def theoretical_waves_per_cu(wg_size: int, vgprs: int, agprs: int, lds_bytes: int, agent: dict) -> dict:
    """Resident waves per CU the kernel's resources allow: the register file
    (512 registers per lane per SIMD, VGPR and AGPR together, granule 8),
    the LDS per CU and the wave slots."""
    simds = int(agent.get("Simd_Per_Cu") or 4)
    slots = int(agent.get("Max_Waves_Per_Simd") or 8)
    lds_cu = int(float(agent.get("Lds_Size_In_Kb") or 160) * 1024)
    waves_per_wg = max(1, -(-wg_size // 64))
    regs = -(-(vgprs + agprs) // 8) * 8
    by_regs = min(slots, 512 // regs) if regs else slots
    wg_by_waves = (simds * by_regs) // waves_per_wg
    wg_by_lds = lds_cu // lds_bytes if lds_bytes else wg_by_waves
    wgs = min(wg_by_waves, wg_by_lds)
    limit = "lds" if wg_by_lds < wg_by_waves else ("registers" if by_regs < slots else "wave slots")
    return {"waves": wgs * waves_per_wg, "max_waves": simds * slots, "limit": limit}


def lds_occupancy(path: str | Path, trace: str | Path | None = None) -> dict:
    """Per kernel (mean per dispatch): LDS bytes moved, LDS throughput,
    bank-conflict share of active cycles, mean resident waves per CU and
    their share of the wave slots.  (rocprofv3's per-dispatch LDS field
    holds the static allocation only, so the occupancy the resources allow
    is left to theoretical_waves_per_cu with the launch's dynamic LDS.)"""
    f = _find(path, "counter_collection.csv")
    agent = read_agent(path)
    cus = int(agent["Cu_Count"])
    xcc = int(agent.get("Num_Xcc") or 1)
    clk_mhz = float(agent.get("Max_Engine_Clk_Fcompute") or 0)
    slots_cu = int(agent.get("Simd_Per_Cu") or 4) * int(agent.get("Max_Waves_Per_Simd") or 8)
    disp: dict[tuple, dict] = {}
    with open(f, newline="") as fh:
        for row in csv.DictReader(fh):
            key = (short_name(row["Kernel_Name"]), row["Dispatch_Id"])
            d = disp.setdefault(key, {"ns": int(row["End_Timestamp"]) - int(row["Start_Timestamp"]),
                                      "wg": int(row.get("Workgroup_Size") or 64),
                                      "vgpr": int(row.get("VGPR_Count") or 0),
                                      "agpr": int(row.get("Accum_VGPR_Count") or 0),
                                      "lds": int(row.get("LDS_Block_Size") or 0)})
            d[row["Counter_Name"]] = d.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    durations = {}
    if trace is not None:
        for r in read_trace(trace):
            durations.setdefault(r["kernel"], r["average_us"])
    per_kernel: dict[str, list[dict]] = defaultdict(list)
    for (k, _), d in disp.items():
        if not _is_velo_kernel(k) or "GRBM_GUI_ACTIVE" not in d:
            continue
        gui = d["GRBM_GUI_ACTIVE"]
        # one GRBM_GUI_ACTIVE per XCD summed into the row: keep one of them
        if clk_mhz and d["ns"] and gui > 1.5 * d["ns"] * clk_mhz * 1e-3:
            gui /= xcc
        lds_bytes = 64.0 * (d.get("SQ_INSTS_LDS_LOAD_BANDWIDTH", 0.0) + d.get("SQ_INSTS_LDS_STORE_BANDWIDTH", 0.0))
        waves = 4.0 * d.get("SQ_WAVE_CYCLES", 0.0) / gui / cus if gui else 0.0
        per_kernel[k].append({"ns": d["ns"], "gui": gui, "lds_bytes": lds_bytes,
                              "conflict": 100.0 * d.get("SQ_LDS_BANK_CONFLICT", 0.0) / gui / cus if gui else 0.0,
                              "waves": waves, "wg": d["wg"], "vgpr": d["vgpr"], "agpr": d["agpr"], "lds": d["lds"]})
    out = {}
    for k, rs in per_kernel.items():
        n = len(rs)
        mean = {f: sum(r[f] for r in rs) / n for f in ("ns", "gui", "lds_bytes", "conflict", "waves")}
        us = durations.get(k, mean["ns"] / 1e3)
        out[k] = {"dispatches": n, "duration_us": round(us, 2), "lds_bytes": round(mean["lds_bytes"]),
                  "lds_TBps": round(mean["lds_bytes"] / (us * 1e-6) / 1e12, 3) if us else None,
                  "lds_bytes_per_cu_cycle": round(mean["lds_bytes"] / (mean["gui"] * cus), 2) if mean["gui"] else None,
                  "lds_bank_conflict_pct": round(mean["conflict"], 2),
                  "waves_per_cu": round(mean["waves"], 2), "occupancy_pct": round(100 * mean["waves"] / slots_cu, 1),
                  "workgroup_size": rs[0]["wg"]}
    return {"cu_count": cus, "wave_slots_per_cu": slots_cu, "num_xcc": xcc, "kernels": out}


def lds_markdown(t: dict, title: str, command: str) -> str:
    lines = _header(title, command) + [
        f"{t['cu_count']} CUs, {t['wave_slots_per_cu']} wave slots per CU.  LDS bytes = 64 x "
        "(SQ_INSTS_LDS_LOAD_BANDWIDTH + SQ_INSTS_LDS_STORE_BANDWIDTH); conflicts = SQ_LDS_BANK_CONFLICT / "
        "(GRBM_GUI_ACTIVE x CUs); resident waves per CU = 4 x SQ_WAVE_CYCLES / (GRBM_GUI_ACTIVE x CUs).", "",
        "| kernel | dispatches | us | LDS bytes | LDS TB/s | LDS B/CU/cycle | bank conflict % | waves/CU | "
        "occupancy % | WG size |",
        "|---|---|---|---|---|---|---|---|---|---|"]
    for k, v in sorted(t["kernels"].items(), key=lambda kv: -kv[1]["duration_us"]):
        lines.append(f"| `{k}` | {v['dispatches']} | {v['duration_us']} | {v['lds_bytes']} | {v['lds_TBps']} | "
                     f"{v['lds_bytes_per_cu_cycle']} | {v['lds_bank_conflict_pct']} | {v['waves_per_cu']} | "
                     f"{v['occupancy_pct']} | {v['workgroup_size']} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("stats", "counters", "trace"):
        p = sub.add_parser(name)
        p.add_argument("--csv", required=True)
        p.add_argument("--command", required=True)
        p.add_argument("--title", required=True)
        if name == "trace":
            p.add_argument("--per-dispatch", default=None, help="kernel whose dispatches are listed one by one")
            p.add_argument("--first", type=int, default=0, help="dispatches of the plain sequence (averaged apart)")
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
    p = sub.add_parser("lds")
    p.add_argument("--csv", required=True)
    p.add_argument("--trace")
    p.add_argument("--command", required=True)
    p.add_argument("--title", required=True)
    p.add_argument("--out-json")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.title, a.command))
    elif a.cmd == "counters":
        sys.stdout.write(counters_markdown(a.csv, a.title, a.command))
    elif a.cmd == "trace":
        sys.stdout.write(trace_markdown(a.csv, a.title, a.command, a.per_dispatch, a.first))
    elif a.cmd == "lds":
        t = lds_occupancy(a.csv, a.trace)
        t["command"] = a.command
        if a.out_json:
            Path(a.out_json).write_text(json.dumps(t, indent=1) + "\n")
        sys.stdout.write(lds_markdown(t, a.title, a.command))
    else:
        t = traffic(a.fetch, a.write, a.copy_bytes, a.copy_match)
        t["events_per_launch"] = a.events_per_launch
        t["command"] = a.command
        Path(a.out_json).write_text(json.dumps(t, indent=1) + "\n")
        Path(a.out_md).write_text(traffic_markdown(t, a.title, a.command))


if __name__ == "__main__":
    main()
