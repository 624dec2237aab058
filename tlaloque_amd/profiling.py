"""Summaries of rocprofv3 CSV output for the ``profiles/`` folder.

Two reports, both markdown with the command that produced the data:

``stats``    per-kernel launch count, total/average/min/max duration and the
             launch resources (VGPR, AGPR, SGPR, LDS, scratch, grid, block),
             from a ``rocprofv3 --kernel-trace --stats --output-format csv`` run.

``counters`` the mean per launch of every counter of one or more
             ``rocprofv3 --pmc`` runs for the kernels whose names match a
             pattern, and each counter per work unit (zones, or queued
             cells) given on the command line: wave instructions per unit
             times 64 for the SQ_INSTS_* counters (lane instructions per
             unit, likewise SQ_ACTIVE_INST_* quad-cycles), the raw ratio
             for the others.

``traffic``  measured HBM-side bytes per launch of each kernel from two
             ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs.  The
             counters report KiB per dispatch; each is scaled by what it reads
             for the streaming copy kernel ``k_stream_copy`` of a known size
             (bytes moved / counter bytes), which calibrates the counter on
             this GPU.  The report ends with one JSON line (prefix
             ``traffic-json:``) that bench.py reads for ``roofline.traffic``.

Usage (from the repository root):
    python -m tlaloque_amd.profiling stats TRACE.csv --cmd "..." --title "..."
    python -m tlaloque_amd.profiling counters PMC.csv [PMC2.csv ...] --match k_stage \
        --units N --cmd "..." --title "..."
    python -m tlaloque_amd.profiling traffic FETCH.csv WRITE.csv --copy-bytes B \
        --zones N --cmd-fetch "..." --cmd-write "..." --title "..."
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys
from typing import Dict, List

COPY_KERNEL = "k_stream_copy"


def short_name(name: str) -> str:
    s = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()
    return s


def _rows(path: str) -> List[dict]:
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


def kernel_stats(trace_csv: str) -> List[dict]:
    """Per-kernel statistics from a kernel-trace CSV (durations in us)."""
    agg: Dict[str, dict] = {}
    for r in _rows(trace_csv):
        name = short_name(r["Kernel_Name"])
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        a = agg.setdefault(name, {"name": name, "calls": 0, "total": 0.0, "min": 1e300,
                                  "max": 0.0, "vgpr": r.get("VGPR_Count", ""),
                                  "agpr": r.get("Accum_VGPR_Count", ""),
                                  "sgpr": r.get("SGPR_Count", ""),
                                  "lds": r.get("LDS_Block_Size", ""),
                                  "scratch": r.get("Scratch_Size", ""),
                                  "grid": r.get("Grid_Size", r.get("Grid_Size_X", "")),
                                  "block": r.get("Workgroup_Size",
                                                 r.get("Workgroup_Size_X", ""))})
        a["calls"] += 1
        a["total"] += dur
        a["min"] = min(a["min"], dur)
        a["max"] = max(a["max"], dur)
    out = sorted(agg.values(), key=lambda a: -a["total"])
    tot = sum(a["total"] for a in out) or 1.0
    for a in out:
        a["avg"] = a["total"] / a["calls"]
        a["pct"] = 100.0 * a["total"] / tot
    return out


def stats_report(trace_csv: str, cmd: str, title: str) -> str:
    lines = [f"# {title}", "", "Command (one MI355X, from the repository root):", "", "```",
             cmd, "```", "",
             "| kernel | calls | total us | avg us | min us | max us | % | VGPR | AGPR | SGPR | "
             "LDS B | scratch | grid | block |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for a in kernel_stats(trace_csv):
        lines.append(f"| `{a['name']}` | {a['calls']} | {a['total']:.1f} | {a['avg']:.1f} | "
                     f"{a['min']:.1f} | {a['max']:.1f} | {a['pct']:.1f} | {a['vgpr']} | "
                     f"{a['agpr']} | {a['sgpr']} | {a['lds']} | {a['scratch']} | {a['grid']} | "
                     f"{a['block']} |")
    return "\n".join(lines) + "\n"


def counter_means(pmc_csv: str, counter: str) -> Dict[str, float]:
    """Mean counter value per kernel (KiB per dispatch for FETCH/WRITE_SIZE)."""
    vals = collections.defaultdict(list)
    for r in _rows(pmc_csv):
        if r["Counter_Name"] == counter:
            vals[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in vals.items()}


def traffic_report(fetch_csv: str, write_csv: str, copy_bytes: int, zones: int,
                   cmd_fetch: str, cmd_write: str, title: str, calibration: str = "") -> str:
    """With `calibration` (another traffic report), its counter factors are
    used instead of a copy kernel in these runs."""
    fetch = counter_means(fetch_csv, "FETCH_SIZE")
    write = counter_means(write_csv, "WRITE_SIZE")
    if calibration:
        cal = read_traffic(calibration)["calibration"]
        f_fetch, f_write = cal["fetch_factor"], cal["write_factor"]
        cal_line = (f"Calibration: the factors of `{calibration}` (FETCH_SIZE x {f_fetch:.4f}, "
                    f"WRITE_SIZE x {f_write:.4f}, bytes per counted byte, from a streaming copy).")
    else:
        # the plain copy kernel (not its variants of the copy-probe sweep)
        if COPY_KERNEL not in fetch or COPY_KERNEL not in write:
            raise SystemExit("no copy-kernel dispatch in the counter runs: cannot calibrate")
        ck = COPY_KERNEL
        # the copy reads copy_bytes and writes copy_bytes
        f_fetch = copy_bytes / (fetch[ck] * 1024.0)
        f_write = copy_bytes / (write[ck] * 1024.0)
        cal_line = (f"Calibration: `{COPY_KERNEL}` copies {copy_bytes} B (reads and writes that "
                    f"many); FETCH_SIZE reads {fetch[ck]:.0f} KiB -> factor {f_fetch:.4f}, "
                    f"WRITE_SIZE reads {write[ck]:.0f} KiB -> factor {f_write:.4f} (bytes per "
                    f"counted byte).")
    lines = [f"# {title}", "",
             "Commands (one MI355X, from the repository root; one counter per run):", "",
             "```", cmd_fetch, cmd_write, "```", "", cal_line, "",
             f"Per-zone figures divide by the {zones} interior zones of one launch.", "",
             "| kernel | FETCH KiB/launch | WRITE KiB/launch | scaled read B | scaled write B | "
             "total B/launch | B/zone |",
             "|---|---|---|---|---|---|---|"]
    js = {"calibration": {"fetch_factor": f_fetch, "write_factor": f_write,
                          "copy_bytes": copy_bytes}, "zones": zones, "kernels": {}}
    for k in sorted(set(fetch) & set(write), key=lambda k: -(fetch[k] + write[k])):
        rd = fetch[k] * 1024.0 * f_fetch
        wr = write[k] * 1024.0 * f_write
        per_zone = (rd + wr) / zones if zones else 0.0
        lines.append(f"| `{k}` | {fetch[k]:.0f} | {write[k]:.0f} | {rd:.4g} | {wr:.4g} | "
                     f"{rd + wr:.4g} | {per_zone:.1f} |")
        # key: the bare kernel name; a second instantiation of the same kernel
        # (listed after the one with more traffic) keeps its template arguments
        key = re.sub(r"<.*", "", k).split("::")[-1]
        if key in js["kernels"]:
            key = k.split("::")[-1]
        js["kernels"][key] = {
            "read_bytes": rd, "write_bytes": wr, "bytes": rd + wr, "bytes_per_zone": per_zone}
    lines += ["", "traffic-json: " + json.dumps(js)]
    return "\n".join(lines) + "\n"


def counters_report(pmc_csvs: List[str], match: str, units: float, unit_name: str, cmd: str,
                    title: str) -> str:
    """Mean per launch of each counter of the kernels matching `match`, and
    per work unit (SQ_INSTS_* as lane instructions: wave instructions x 64)."""
    per: Dict[str, Dict[str, List[float]]] = collections.defaultdict(
        lambda: collections.defaultdict(list))
    for path in pmc_csvs:
        for r in _rows(path):
            name = short_name(r["Kernel_Name"])
            if re.search(match, name):
                per[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    lines = [f"# {title}", "", "Commands (one MI355X, from the repository root; one pass per command):",
             "", "```", cmd, "```", "",
             f"Values are means over each kernel's launches in the runs; per-unit figures divide by "
             f"{units:.6g} {unit_name} per launch; SQ_INSTS_* (wave instructions) and SQ_ACTIVE_INST_* "
             f"(quad-cycles a wave issues) are scaled by 64, the lanes of a wave.", "",
             "| kernel | counter | per launch | per unit | launches |", "|---|---|---|---|---|"]
    for name in sorted(per):
        for cn in sorted(per[name]):
            v = per[name][cn]
            mean = sum(v) / len(v)
            lane = cn.startswith(("SQ_INSTS_", "SQ_ACTIVE_INST_"))
            pu = mean * (64.0 if lane else 1.0) / units if units else 0.0
            lines.append(f"| `{name}` | {cn} | {mean:.4g} | {pu:.4g} | {len(v)} |")
    return "\n".join(lines) + "\n"


def read_traffic(path: str) -> dict:
    """The JSON line of a traffic report."""
    with open(path) as fh:
        for ln in fh:
            if ln.startswith("traffic-json: "):
                return json.loads(ln[len("traffic-json: "):])
    raise ValueError(f"{path}: no traffic-json line")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m tlaloque_amd.profiling")
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("trace_csv")
    s.add_argument("--cmd", required=True)
    s.add_argument("--title", required=True)
    c = sub.add_parser("counters")
    c.add_argument("pmc_csv", nargs="+")
    c.add_argument("--match", required=True, help="regular expression on the kernel name")
    c.add_argument("--units", type=float, required=True, help="work units per launch")
    c.add_argument("--unit-name", default="zones")
    c.add_argument("--cmd", required=True)
    c.add_argument("--title", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--zones", type=int, required=True)
    t.add_argument("--cmd-fetch", required=True)
    t.add_argument("--cmd-write", required=True)
    t.add_argument("--title", required=True)
    t.add_argument("--calibration", default="", help="reuse this traffic report's factors")
    a = ap.parse_args(argv)
    if a.what == "stats":
        sys.stdout.write(stats_report(a.trace_csv, a.cmd, a.title))
    elif a.what == "counters":
        sys.stdout.write(counters_report(a.pmc_csv, a.match, a.units, a.unit_name, a.cmd, a.title))
    else:
        sys.stdout.write(traffic_report(a.fetch_csv, a.write_csv, a.copy_bytes, a.zones,
                                        a.cmd_fetch, a.cmd_write, a.title, a.calibration))


if __name__ == "__main__":
    main()
