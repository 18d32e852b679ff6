"""Summaries of rocprofv3 CSV output for the SPH step kernels.

    python -m nbodysimulation_experiment_amd.profiling stats  <kernel_stats.csv>
    python -m nbodysimulation_experiment_amd.profiling counters <counter_collection.csv> [--copy-bytes N]

``counters`` averages each counter per dispatch of each kernel.  With
--copy-bytes, FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are converted to
bytes with the factor that maps the copy kernel's reading onto its known
traffic (N bytes read and N bytes written per dispatch).
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys

SHORT = [
    ("k_visc_predict", "viscosity_predict"), ("k_scan_reduce", "scan_reduce"), ("k_scan_blocks", "scan_blocks"),
    ("k_scan_final", "scan_final"), ("k_scatter", "scatter"), ("k_rank_reorder", "rank_reorder"),
    ("k_density", "density"), ("k_delta", "delta"), ("k_copy", "copy"),
]


def short_name(name: str) -> str:
    for key, short in SHORT:
        if key in name:
            m = re.search(r"<(true|false), (\d+)>", name)
            return short + (f"<{m.group(1)[0]},{m.group(2)}>" if m else "")
    return name[:60]


def kernel_stats(path: str) -> list[dict]:
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                     "avg_us": float(r["AverageNs"]) / 1e3, "total_ms": float(r["TotalDurationNs"]) / 1e6,
                     "pct": float(r["Percentage"])})
    return rows


def counters(path: str) -> dict:
    """{kernel: {counter: mean per dispatch}} from a counter_collection CSV."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        name = short_name(r.get("Kernel_Name", r.get("Kernel-Name", "")))
        acc[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} | {"dispatches": max(len(v) for v in d.values())}
            for k, d in acc.items()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "counters"])
    ap.add_argument("path")
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    a = ap.parse_args(argv)
    if a.mode == "stats":
        for r in kernel_stats(a.path):
            print(f"{r['kernel']:32s} calls {r['calls']:6d}  avg {r['avg_us']:9.2f} us  total {r['total_ms']:9.3f} ms"
                  f"  {r['pct']:6.2f}%")
        return
    c = counters(a.path)
    if a.copy_bytes and "copy" in c:
        for key in ("FETCH_SIZE", "WRITE_SIZE"):
            if key in c["copy"]:
                factor = a.copy_bytes / (c["copy"][key] * 1024.0)
                for k in c:
                    if key in c[k]:
                        c[k][key + "_bytes"] = c[k][key] * 1024.0 * factor
                c["copy"][key + "_factor"] = factor
    json.dump(c, sys.stdout, indent=1)
    print()


if __name__ == "__main__":
    main()
