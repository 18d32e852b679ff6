"""bench.py's fixed parts, checked without a GPU: the metric is the one the
baseline names, the per-kernel algorithmic bytes are the per-particle records
each kernel must read and write, and runs on more than one GPU are refused."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_metric_is_the_baseline_metric():
    with open(os.path.join(ROOT, "BASELINE.json")) as fh:
        assert bench.METRIC == json.load(fh)["metric"]


def test_kernel_bytes_are_the_particle_records():
    f2, u4 = 8, 4  # float2, uint32 per particle
    assert bench.KERNEL_BYTES["viscosity_predict"] == (f2 + f2 + u4) + (f2 + u4 + u4)  # R x, v, key; W xp, key, fill
    assert bench.KERNEL_BYTES["density"] == (f2 + u4) + (f2 + f2)  # R x, key; W dens, pres
    assert bench.KERNEL_BYTES["delta"] == (f2 + u4 + f2 + f2) + (f2 + f2)  # R x, key, pres, prev; W x, v


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
