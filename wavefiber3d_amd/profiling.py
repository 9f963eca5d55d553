"""Summaries of rocprofv3 output for the committed profile notes.

``python -m wavefiber3d_amd.profiling kernels STATS_CSV [--trace TRACE_CSV]``
prints a markdown table of per-kernel time (calls, total, average, share;
static LDS and scratch per block from the trace when given).

``python -m wavefiber3d_amd.profiling traffic --fetch DIR --write DIR
--fetch-copy DIR --write-copy DIR --copy-bytes B --kernel SUBSTR[|SUBSTR..]
[--units U]`` reads the counter-collection CSVs of four ``rocprofv3 --pmc``
runs (FETCH_SIZE and WRITE_SIZE of the workload, the same two counters of a
device copy of B bytes), derives the bytes-per-KiB factor of each counter from
the copy (bytes read = bytes written = B) and prints, as JSON, the HBM traffic
per dispatch of the kernels whose names contain one of the SUBSTRs -- or, with
--units, the traffic of all those dispatches divided by U (e.g. the level
transitions of the profiled solves, whose kernels are several dispatches).

``python -m wavefiber3d_amd.profiling counters DIR [DIR ...] [--kernel SUBSTR|..]``
prints a markdown table of the per-dispatch averages of every counter found
in the counter-collection CSVs of the given ``rocprofv3 --pmc`` runs, one row
per kernel.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
import sys
from collections import defaultdict


def short_name(n: str) -> str:
    """Readable kernel name: demangles the anonymous-namespace templates this
    library emits (``_ZN12_GLOBAL__N_1<len><name>I...``) to ``name<args>``."""
    m = re.match(r"_ZN12_GLOBAL__N_1(\d+)", n)
    if m:
        L = int(m.group(1))
        base = n[len(m.group(0)):len(m.group(0)) + L]
        rest = n[len(m.group(0)) + L:]
        targs = re.findall(r"L([ib])(\d+)E", rest.split("EEEv")[0] + "EE")
        args = [v if t == "i" else ("true" if v == "1" else "false") for t, v in targs]
        n = base + ("<" + ",".join(args) + ">" if args else "")
    n = re.sub(r"\(anonymous namespace\)::", "", n)
    n = re.sub(r"^void ", "", n)
    return re.sub(r"\(.*", "", n).strip()


def kernel_table(stats_csv: str, trace_csv: str | None = None, top: int = 20) -> str:
    rows = list(csv.DictReader(open(stats_csv)))
    res = {}
    if trace_csv and os.path.exists(trace_csv):
        for r in csv.DictReader(open(trace_csv)):
            res.setdefault(r["Kernel_Name"], (r.get("LDS_Block_Size", ""), r.get("Scratch_Size", "")))
    tot = sum(int(r["TotalDurationNs"]) for r in rows) or 1
    out = ["| kernel | calls | total ms | avg us | % | static LDS B | scratch B |",
           "|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -int(r["TotalDurationNs"]))[:top]:
        v = res.get(r["Name"], ("", ""))
        out.append(f"| `{short_name(r['Name'])}` | {r['Calls']} | {int(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {100 * int(r['TotalDurationNs']) / tot:.1f} | "
                   f"{v[0]} | {v[1]} |")
    return "\n".join(out)


def group_average_us(stats_csv: str, substr: str) -> tuple[int, float]:
    """(calls, average duration in us) over all kernels whose name contains substr
    (or any of its "|"-separated alternatives)."""
    calls, ns = 0, 0
    for r in csv.DictReader(open(stats_csv)):
        if _matches(short_name(r["Name"]), substr):
            calls += int(r["Calls"])
            ns += int(r["TotalDurationNs"])
    return calls, (ns / calls / 1e3 if calls else float("nan"))


def _counter(dirname: str, counter: str) -> dict[str, list[float]]:
    """kernel short name -> per-dispatch values of one counter (KiB)."""
    path = os.path.join(dirname, "run_counter_collection.csv")
    per = defaultdict(lambda: defaultdict(float))
    names = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        per[r["Dispatch_Id"]][counter] += float(r["Counter_Value"])
        names[r["Dispatch_Id"]] = short_name(r["Kernel_Name"])
    out = defaultdict(list)
    for d, v in per.items():
        out[names[d]].append(v[counter])
    return out


def _matches(name: str, kernel: str) -> bool:
    return any(k and k in name for k in kernel.split("|"))


def traffic(fetch: str, write: str, fetch_copy: str, write_copy: str, copy_bytes: float,
            kernel: str, units: float | None = None) -> dict:
    fc = _counter(fetch_copy, "FETCH_SIZE")
    wc = _counter(write_copy, "WRITE_SIZE")
    fk = max(max(v) for k, v in fc.items() if "copyBuffer" in k)
    wk = max(max(v) for k, v in wc.items() if "copyBuffer" in k)
    f_factor = copy_bytes / (fk * 1024.0)
    w_factor = copy_bytes / (wk * 1024.0)
    fw = _counter(fetch, "FETCH_SIZE")
    ww = _counter(write, "WRITE_SIZE")
    fv = [x for k, v in fw.items() if _matches(k, kernel) for x in v]
    wv = [x for k, v in ww.items() if _matches(k, kernel) for x in v]
    if not fv or not wv:
        raise SystemExit(f"no dispatches of kernels matching {kernel!r}")
    if units:
        f_kib, w_kib = sum(fv) / units, sum(wv) / units
    else:
        f_kib, w_kib = sum(fv) / len(fv), sum(wv) / len(wv)
    return {
        "kernel": kernel,
        "dispatches": [len(fv), len(wv)],
        "units": units or None,
        "fetch_kib_per_dispatch" if not units else "fetch_kib_per_unit": f_kib,
        "write_kib_per_dispatch" if not units else "write_kib_per_unit": w_kib,
        "copy_bytes": copy_bytes,
        "copy_fetch_kib": fk,
        "copy_write_kib": wk,
        "fetch_bytes_per_kib": f_factor * 1024.0,
        "write_bytes_per_kib": w_factor * 1024.0,
        "traffic_bytes_per_launch": f_kib * 1024.0 * f_factor + w_kib * 1024.0 * w_factor,
    }


def counter_table(dirs: list[str], kernel: str | None = None) -> str:
    """kernel -> per-dispatch average of every counter in the runs' CSVs."""
    vals = defaultdict(lambda: defaultdict(list))
    order = []
    for d in dirs:
        path = os.path.join(d, "run_counter_collection.csv")
        per = defaultdict(lambda: defaultdict(float))
        names = {}
        for r in csv.DictReader(open(path)):
            per[r["Dispatch_Id"]][r["Counter_Name"]] += float(r["Counter_Value"])
            names[r["Dispatch_Id"]] = short_name(r["Kernel_Name"])
        for disp, cv in per.items():
            for c, v in cv.items():
                if c not in order:
                    order.append(c)
                vals[names[disp]][c].append(v)
    rows = sorted(k for k in vals if kernel is None or _matches(k, kernel))
    out = ["| kernel | dispatches | " + " | ".join(order) + " |",
           "|---|---|" + "---|" * len(order)]
    for k in rows:
        n = max(len(v) for v in vals[k].values())
        cells = []
        for c in order:
            v = vals[k].get(c)
            cells.append(f"{sum(v) / len(v):.4g}" if v else "")
        out.append(f"| `{k}` | {n} | " + " | ".join(cells) + " |")
    return "\n".join(out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m wavefiber3d_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("stats")
    k.add_argument("--trace", default=None)
    k.add_argument("--top", type=int, default=20)
    g = sub.add_parser("group")
    g.add_argument("stats")
    g.add_argument("substr")
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--fetch-copy", "--write-copy", "--kernel"):
        t.add_argument(a, required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--units", type=float, default=None)
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--kernel", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        print(kernel_table(a.stats, a.trace, a.top))
    elif a.cmd == "group":
        calls, avg = group_average_us(a.stats, a.substr)
        print(json.dumps({"kernel": a.substr, "calls": calls, "avg_us": avg}))
    elif a.cmd == "counters":
        print(counter_table(a.dirs, a.kernel))
    else:
        print(json.dumps(traffic(a.fetch, a.write, a.fetch_copy, a.write_copy, a.copy_bytes,
                                 a.kernel, a.units), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
