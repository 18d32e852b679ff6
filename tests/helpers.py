"""Shared helpers: paired GPU / oracle runs and bitwise comparison."""
from __future__ import annotations

import numpy as np

from nbodysimulation_experiment_amd.scenes import SPHParams

FIELDS = ("x", "prev", "v", "density", "pressure")
DT = 0.005
GRAVITY = (0.0, -9.8)


def oracle_for(specs, params=None, precision=32, mode=None, threads=16, gravity=GRAVITY):
    from oracle.oracle import JACOBI, OracleSim
    params = params or SPHParams()
    o = OracleSim(params, precision, JACOBI if mode is None else mode, threads)
    for s in specs:
        o.add_scene_spec(s)
    o.set_gravity(*gravity)
    return o


def gpu_for(specs, params=None, gravity=GRAVITY):
    from nbodysimulation_experiment_amd.sim import ParticleSimulation
    params = params or SPHParams()
    if len(specs) == 1:
        return ParticleSimulation.from_scene(specs[0], params, gravity)
    return ParticleSimulation.from_scenes(specs, params, gravity)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bit_exact(got: dict, want: dict, what: str = ""):
    for k in FIELDS:
        g, w = bits(got[k]), bits(want[k])
        if not np.array_equal(g, w):
            bad = np.argwhere(g != w)
            i = int(bad[0][0])
            raise AssertionError(
                f"{what}{k}: {len(bad)} words differ; first particle {i}: gpu {got[k][i].tolist()} "
                f"oracle {want[k][i].tolist()}")
    if not np.array_equal(got["key"], want["key"]):
        bad = np.argwhere(got["key"] != want["key"])
        raise AssertionError(f"{what}cell keys differ at {bad[:5].ravel().tolist()}")


def max_pos_err(a, b) -> float:
    return float(np.abs(np.asarray(a["x"], np.float64) - np.asarray(b["x"], np.float64)).max())
