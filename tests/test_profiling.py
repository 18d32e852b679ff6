"""The rocprofv3 CSV summaries of the profiling module, on small synthetic
counter files (no GPU)."""
from __future__ import annotations

import csv
import json

from nbodysimulation_experiment_amd import profiling

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
NAMES = {"viscosity_predict": "void nbs::k_visc_predict<false, 1024>(nbs::Buffers)",
         "density": "void nbs::k_density<false, 1024>(nbs::Buffers)",
         "delta": "void nbs::k_delta<false, 2048>(nbs::Buffers)",
         "scatter": "nbs::k_scatter(nbs::Buffers, int)"}


def _write(path, launches):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, FIELDS)
        w.writeheader()
        for d, (kernel, counters) in enumerate(launches, start=1):
            for name, value in counters.items():
                w.writerow({"Dispatch_Id": d, "Kernel_Name": NAMES[kernel], "Counter_Name": name,
                            "Counter_Value": value})


def test_step_ranges_sum_per_range_and_kernel(tmp_path):
    launches = []
    for step in range(4):
        c = {"SQ_INSTS_VALU": 1000.0 * (step + 1), "GRBM_GUI_ACTIVE": 8.0 * 10.0,
             "SQ_THREAD_CYCLES_VALU": 64.0 * 500.0 * (step + 1)}
        launches += [("viscosity_predict", c), ("scatter", c), ("density", c), ("delta", c)]
    p = tmp_path / "c.csv"
    _write(p, launches)
    rows = {(tuple(r["steps"]), r["kernel"]): r for r in profiling.step_ranges(str(p), [0, 2, 4])}
    assert set(k for _, k in rows) == {"viscosity_predict", "density", "delta"}  # scatter is not a pair kernel
    r = rows[((0, 2), "density")]
    assert r["launches"] == 2
    assert r["valu_instr_per_launch_M"] == round(3000.0 / 2 / 1e6, 2)
    # 3000 VALU instructions over 2 x 10 cycles per XCD on 256 CUs
    assert r["valu_issue"] == round(3000.0 / (256 * 160.0 / 8.0), 3)
    assert r["active_lanes"] == 0.5
    assert rows[((2, 4), "delta")]["launches"] == 2


def test_valu_window_merges_beside_the_profiled_window(tmp_path):
    launches = []
    for step in range(6):
        c = {"SQ_INSTS_VALU": 1000.0 * (step + 1), "GRBM_GUI_ACTIVE": 8.0 * 10.0}
        launches += [("viscosity_predict", c), ("density", c), ("delta", c)]
    p = tmp_path / "c.csv"
    _write(p, launches)
    j = tmp_path / "valu.json"
    j.write_text('{"workload": "w", "steps": 20, "kernels": {"delta": 1.0}, "valu_insts_per_launch": {"delta": 7}}')
    profiling.main(["steps", str(p), "--ranges", "2,4,6", "--json", str(j), "--steps", "4", "--workload", "w"])
    d = json.loads(j.read_text())
    assert d["steps"] == 20 and d["valu_insts_per_launch"]["delta"] == 7  # the profiled window is kept
    w = d["windows"]["4"]
    assert w["steps_counted"] == [2, 6]
    # steps 2..5 issue 3000..6000 instructions: 4500 per launch on average
    assert w["valu_insts_per_launch"]["delta"] == 4500
    assert w["kernels"]["density"] == round(18000.0 / (256 * 4 * 80.0 / 8.0), 3)


def test_short_names_keep_tile_variants():
    assert profiling.short_name(NAMES["delta"]) == "delta<f,2048>"
    assert profiling.short_name(NAMES["scatter"]) == "scatter"
