"""bench.py's fixed parts, checked without a GPU: the headline metric, the
per-kernel algorithmic bytes (the per-particle records each kernel must read
and write), and runs on more than one GPU are refused."""
from __future__ import annotations

import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_metric_is_the_headline_metric():
    assert bench.METRIC == "particle-steps/s @1M particles: plain vs surrogate at \u22641% density error"


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


def test_host_cpu_reports_usable_threads_and_model():
    h = bench.host_cpu()
    assert 1 <= h["threads"] <= h["affinity_cpus"] <= (os.cpu_count() or h["affinity_cpus"])
    assert h["threads"] == len(os.sched_getaffinity(0)) or h["cgroup_cpu_quota"] == h["threads"]
    assert h["model"]


def test_first_tuning_steps_cover_two_rounds_of_every_variant():
    # one binning step, then 2 rounds over 0/1024/2048/4096 (+ the wide kernels up to 131,072 particles)
    assert bench.first_tuning_steps(1 << 20) == 1 + 2 * 4
    assert bench.first_tuning_steps(4096) == 1 + 2 * 5


def test_compare_states():
    import numpy as np
    rng = np.random.default_rng(0)
    a = {"x": rng.normal(size=(50, 2)).astype(np.float32), "v": rng.normal(size=(50, 2)).astype(np.float32),
         "density": rng.uniform(1, 3, (50, 2)).astype(np.float32)}
    z = bench.compare_states(a, a)
    assert all(v == 0 for v in z.values())
    b = dict(a, x=a["x"] + np.float32(0.5))
    assert abs(bench.compare_states(b, a)["position_max_h"] - 0.5 * 2 ** 0.5) < 1e-5
