"""Summaries of rocprofv3 CSV output for the SPH step kernels.

    python -m nbodysimulation_experiment_amd.profiling stats  <kernel_stats.csv>
    python -m nbodysimulation_experiment_amd.profiling counters <counter_collection.csv> [--copy-bytes N]
                                                              [--workload W --steps K --json valu.json]
    python -m nbodysimulation_experiment_amd.profiling traffic <fetch.csv> <write.csv> --copy-bytes N
                                                              --workload W [--json out.json]
    python -m nbodysimulation_experiment_amd.profiling steps <counter_collection.csv> [--ranges 2,100,300,400,502]

``traffic`` turns one FETCH_SIZE run and one WRITE_SIZE run into bytes per
launch of each kernel group (dispatches of every LDS-tile variant pooled).
``counters`` averages each counter per dispatch of each kernel.  ``steps``
sums the counters of the pair kernels over ranges of steps (a step starts at
each viscosity launch) and prints, per range, the VALU issue utilisation, the
active-lane fraction of the VALU instructions (SQ_THREAD_CYCLES_VALU over 64
lanes x SQ_INSTS_VALU) and, when collected, the waves resident per busy cycle.  With
--copy-bytes, FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are converted to
bytes with the factor that maps the copy kernel's reading onto its known
traffic (N bytes read and N bytes written per dispatch).
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re
import sys

SHORT = [
    ("k_visc_predict_wide", "viscosity_predict_wide"), ("k_density_wide", "density_wide"),
    ("k_delta_wide", "delta_wide"),
    ("k_visc_predict", "viscosity_predict"), ("k_count_fill", "count_fill"),
    ("k_count_scan_small", "count_scan_small"), ("k_scan_reduce", "scan_reduce"),
    ("k_scan_small", "scan_small"),
    ("k_scan_final", "scan_final"), ("k_scatter_grad", "scatter_grad"), ("k_scatter", "scatter"), ("k_rank_reorder", "rank_reorder"),
    ("k_place", "place"),
    ("k_density", "density"), ("k_delta", "delta"), ("k_copy", "copy"), ("k_sur_forward", "sur_forward"),
    ("k_sur_back_pairs", "sur_back_pairs"), ("k_sur_back_head", "sur_back_head"),
]

# Kernel groups of the step (timed together by the library's events).  Each
# group lists its alternatives: the kernels one step launches for the phase in
# the one-lane-per-particle and in the wide form.
GROUPS = {
    "viscosity_predict": (("viscosity_predict",), ("viscosity_predict_wide", "count_fill"), ("viscosity_predict_wide",)),
    "grid": (("scan_reduce", "scan_final", "place"), ("scan_reduce", "scan_final", "scatter", "rank_reorder"),
             ("scan_small", "scatter", "rank_reorder"), ("count_scan_small", "scatter", "rank_reorder")),
    "density": (("density",), ("density_wide",)),
    "delta": (("delta",), ("delta_wide",)),
    "surrogate_forward": (("sur_forward",),),
}


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
    """{kernel: {counter: mean per dispatch}} from a counter_collection CSV.
    A kernel launched with several grid sizes (the learned relaxation's
    forward at the timed scene and at the training set) keeps its name for
    the largest grid; the others are listed as "name@grid"."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        name = short_name(r.get("Kernel_Name", r.get("Kernel-Name", "")))
        grid = int(float(r.get("Grid_Size", 0) or 0))
        acc[(name, grid)][r["Counter_Name"]].append(float(r["Counter_Value"]))
    largest = {}
    for name, grid in acc:
        largest[name] = max(largest.get(name, grid), grid)
    return {(name if grid == largest[name] else f"{name}@{grid}"):
            {c: sum(v) / len(v) for c, v in d.items()} | {"dispatches": max(len(v) for v in d.values())}
            for (name, grid), d in acc.items()}


def _per_dispatch(path: str, counter: str) -> dict:
    """{short kernel name without tile variant: [value per dispatch]} (KiB)."""
    out = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        name = short_name(r.get("Kernel_Name", r.get("Kernel-Name", ""))).split("<")[0]
        out[name].append(float(r["Counter_Value"]))
    return out


def traffic(fetch_csv: str, write_csv: str, copy_bytes: float, workload: str) -> dict:
    """HBM bytes per launch of each kernel group, each counter scaled by the
    factor that maps the copy kernel's reading onto its known traffic
    (copy_bytes read and copy_bytes written per dispatch)."""
    f = _per_dispatch(fetch_csv, "FETCH_SIZE")
    w = _per_dispatch(write_csv, "WRITE_SIZE")
    ff = copy_bytes / (sum(f["copy"]) / len(f["copy"]) * 1024.0)
    wf = copy_bytes / (sum(w["copy"]) / len(w["copy"]) * 1024.0)
    kernels, detail = {}, {}
    for k in f:
        if k in w and k != "copy":
            detail[k] = {"fetch_bytes": round(sum(f[k]) / len(f[k]) * 1024.0 * ff),
                         "write_bytes": round(sum(w[k]) / len(w[k]) * 1024.0 * wf), "dispatches": len(f[k])}
    for g, variants in GROUPS.items():
        # bytes per step of each alternative that ran, weighted by its launches
        tot = weight = 0.0
        for members in variants:
            if not all(k in detail for k in members):
                continue
            launches = detail[members[0]]["dispatches"]
            tot += launches * sum(detail[k]["fetch_bytes"] + detail[k]["write_bytes"] for k in members)
            weight += launches
        if weight:
            kernels[g] = round(tot / weight)
    return {"workload": workload, "unit": "bytes per launch", "fetch_factor": round(ff, 4),
            "write_factor": round(wf, 4), "kernels": kernels, "detail": detail}


CUS = 256  # MI355X compute units; GRBM_GUI_ACTIVE is summed over its 8 XCDs


def valu_issue(c: dict, workload: str, steps: int = 0) -> dict:
    """VALU issue utilisation of each kernel group from a counters() result:
    SQ_INSTS_VALU / (CUS * GRBM_GUI_ACTIVE / 8), a CU issuing at most one wave64
    VALU instruction per cycle; the single-scene LDS-tile variants pooled by
    dispatches (multi-scene variants, "<t,N>", belong to other workloads).
    Also the wave64 VALU instructions per launch of each group (summed over
    the group's kernels), which bench.py divides by its measured launch
    duration for the same command (``steps`` timed steps)."""
    per = collections.defaultdict(lambda: [0.0, 0.0])  # kernel: [instr, cycles] summed over dispatches
    for k, v in c.items():
        if "SQ_INSTS_VALU" not in v or "GRBM_GUI_ACTIVE" not in v or "<t," in k:
            continue
        name = k.split("<")[0]
        per[name][0] += v["SQ_INSTS_VALU"] * v["dispatches"]
        per[name][1] += v["GRBM_GUI_ACTIVE"] * v["dispatches"]
    disp = collections.Counter()
    for k, v in c.items():
        if "SQ_INSTS_VALU" in v and "<t," not in k:
            disp[k.split("<")[0]] += v["dispatches"]
    groups, per_launch = {}, {}
    for g, variants in GROUPS.items():
        for members in variants:  # the first alternative that ran, its kernels together
            if all(k in per for k in members):
                ins = sum(per[k][0] for k in members)
                cyc = sum(per[k][1] for k in members)
                groups[g] = round(ins / (CUS * cyc / 8.0), 4)
                per_launch[g] = round(sum(per[k][0] / disp[k] for k in members))
                break
    return {"workload": workload, "steps": steps, "unit": "fraction of VALU issue slots", "kernels": groups,
            "valu_insts_per_launch": per_launch}


def step_ranges(path: str, bounds: list[int]) -> list[dict]:
    """Counters of the viscosity, density and relaxation kernels summed over
    the step ranges [bounds[i], bounds[i + 1]), steps counted by viscosity
    launches (the first launch is step 0)."""
    per = collections.defaultdict(dict)
    names = {}
    for r in csv.DictReader(open(path)):
        k = int(r["Dispatch_Id"])
        per[k][r["Counter_Name"]] = float(r["Counter_Value"])
        names[k] = short_name(r.get("Kernel_Name", r.get("Kernel-Name", ""))).split("<")[0]
    step, acc = -1, collections.defaultdict(collections.Counter)
    launches = collections.Counter()
    for k in sorted(per):
        n = names[k]
        if n == "viscosity_predict":
            step += 1
        if n not in ("viscosity_predict", "density", "delta") or step < 0:
            continue
        for i in range(len(bounds) - 1):
            if bounds[i] <= step < bounds[i + 1]:
                acc[(i, n)].update(per[k])
                launches[(i, n)] += 1
    out = []
    for (i, n), c in sorted(acc.items()):
        row = {"steps": [bounds[i], bounds[i + 1]], "kernel": n, "launches": launches[(i, n)]}
        if c.get("SQ_INSTS_VALU") and c.get("GRBM_GUI_ACTIVE"):
            row["valu_issue"] = round(c["SQ_INSTS_VALU"] / (CUS * c["GRBM_GUI_ACTIVE"] / 8.0), 3)
            row["valu_instr_per_launch_M"] = round(c["SQ_INSTS_VALU"] / launches[(i, n)] / 1e6, 2)
            row["valu_instr_per_launch"] = round(c["SQ_INSTS_VALU"] / launches[(i, n)])
            row["cycles_per_launch_k"] = round(c["GRBM_GUI_ACTIVE"] / 8.0 / launches[(i, n)] / 1e3, 1)
        if c.get("SQ_THREAD_CYCLES_VALU") and c.get("SQ_INSTS_VALU"):
            row["active_lanes"] = round(c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_INSTS_VALU"]), 3)
        if c.get("SQ_WAVE_CYCLES") and c.get("SQ_BUSY_CYCLES"):
            row["wave_cycles_per_busy_cycle"] = round(c["SQ_WAVE_CYCLES"] / c["SQ_BUSY_CYCLES"], 2)
        out.append(row)
    return out


def valu_window(path: str, first: int, end: int, steps: int) -> dict:
    """The pair kernels' VALU issue and wave64 VALU instructions per launch over
    the steps [first, end) of a counter run: the timed window of a bench.py
    command of ``steps`` steps, for bench.py's roofline at that step count."""
    rows = step_ranges(path, [first, end])
    return {"steps": steps, "steps_counted": [first, end],
            "kernels": {r["kernel"]: r["valu_issue"] for r in rows if "valu_issue" in r},
            "valu_insts_per_launch": {r["kernel"]: r["valu_instr_per_launch"] for r in rows
                                      if "valu_instr_per_launch" in r}}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "counters", "traffic", "steps"])
    ap.add_argument("path")
    ap.add_argument("path2", nargs="?")
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    ap.add_argument("--workload", default="dam-break-1m")
    ap.add_argument("--json", default="")
    ap.add_argument("--ranges", default="2,100,300,400,502")
    ap.add_argument("--steps", type=int, default=0, help="timed steps of the profiled bench.py command")
    a = ap.parse_args(argv)
    if a.mode == "steps":
        bounds = [int(x) for x in a.ranges.split(",")]
        for row in step_ranges(a.path, bounds):
            print(json.dumps(row))
        if a.json and a.steps:
            # the whole range as the window of a --steps run, beside the
            # file's other windows
            d = json.load(open(a.json)) if os.path.exists(a.json) else {"workload": a.workload}
            d.setdefault("windows", {})[str(a.steps)] = valu_window(a.path, bounds[0], bounds[-1], a.steps)
            with open(a.json, "w") as fh:
                fh.write(json.dumps(d, indent=1) + "\n")
        return
    if a.mode == "traffic":
        t = traffic(a.path, a.path2, a.copy_bytes, a.workload)
        text = json.dumps(t, indent=1)
        if a.json:
            with open(a.json, "w") as fh:
                fh.write(text + "\n")
        print(text)
        return
    if a.mode == "stats":
        rows = kernel_stats(a.path)
        for r in rows:
            print(f"{r['kernel']:32s} calls {r['calls']:6d}  avg {r['avg_us']:9.2f} us  total {r['total_ms']:9.3f} ms"
                  f"  {r['pct']:6.2f}%")
        # every LDS-tile variant of a single-scene kernel pooled ("<f,*>")
        pooled = collections.defaultdict(lambda: [0, 0.0])
        for r in rows:
            m = re.match(r"(\w+)<f,\d+>$", r["kernel"])
            if m:
                pooled[m.group(1)][0] += r["calls"]
                pooled[m.group(1)][1] += r["total_ms"]
        for k, (calls, tot) in pooled.items():
            print(f"{k + '<f,*>':32s} calls {calls:6d}  avg {1e3 * tot / calls:9.2f} us  total {tot:9.3f} ms  (pooled)")
        return
    c = counters(a.path)
    if a.json:
        d = valu_issue(c, a.workload, a.steps)
        if os.path.exists(a.json):  # keep the windows of other step counts
            old = json.load(open(a.json))
            if old.get("workload") == a.workload and "windows" in old:
                d["windows"] = old["windows"]
        with open(a.json, "w") as fh:
            fh.write(json.dumps(d, indent=1) + "\n")
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
