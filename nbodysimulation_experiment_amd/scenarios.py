"""The reference's built-in scenarios, scenario loader and benchmark.

Python side of ``host/scenarios.cpp`` and ``host/benchmark.cpp``:

* ``SCENARIOS``: the table ``SPHScenarios`` (NBodySimulation/sph.h:262-437)
  built from the compile-time constants of sph.h:17-63, every value in float32
  in the reference's operand order;
* ``load_scenario``: ``DemoApplication::LoadScenario`` (app.cpp:477-534);
* ``run_benchmark``: the application's built-in benchmark (app.cpp:88-161,
  215-275): iterations x frames, per-phase min / max / avg milliseconds.

The scenes live in the reference's fixed domain, a 10 x 5.625 box centred on
the origin with a 33 x 18 grid of 0.3 cells (sph.h:17-22, 58-63).  The
eight-argument ``SPHParameters`` constructor keeps the compile-time kernel
height and rest density whatever it is passed (sph.h:112-122).  ``sinf`` and
``cosf`` of the rotated boxes are the float32 rounding of the double-precision
sine and cosine.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .scenes import SPHParams

__all__ = ["CONSTANTS", "SCENARIOS", "Scenario", "resolve_body", "volume_counts", "scenario_params",
           "make_simulation", "load_scenario", "run_benchmark", "BENCHMARK_FRAMES", "BENCHMARK_ITERATIONS",
           "PHASES"]

f32 = np.float32
BENCHMARK_FRAMES = 64       # app.h:24
BENCHMARK_ITERATIONS = 16   # app.h:25
PHASES = ("emitters", "integration", "viscosityForces", "predict", "updateGrid", "neighborSearch",
          "densityAndPressure", "deltaPositions", "collisions")


def _constants() -> dict:
    aspect = f32(16.0) / f32(9.0)
    c = {"boundary_width": f32(10.0)}
    c["boundary_height"] = c["boundary_width"] / aspect
    c["boundary_half_width"] = c["boundary_width"] * f32(0.5)
    c["boundary_half_height"] = c["boundary_height"] * f32(0.5)
    c["delta_time"] = (f32(1.0) / f32(60.0)) / f32(1.0)
    c["particle_radius"] = f32(0.05)
    c["kernel_height"] = f32(6.0) * c["particle_radius"]
    c["particle_spacing"] = c["kernel_height"] * f32(0.5)
    c["collision_radius"] = c["particle_radius"]
    c["collision_margin"] = f32(0.005) * f32(2.0)
    c["rest_density"] = f32(20.0)
    c["stiffness"] = f32(0.6)
    c["near_stiffness"] = c["stiffness"] * f32(10.0)
    c["linear_viscosity"] = f32(0.5)
    c["quadratic_viscosity"] = f32(0.3)
    c["grid_cell_size"] = c["kernel_height"]
    c["grid_count_x"] = int(c["boundary_width"] / c["grid_cell_size"])
    c["grid_count_y"] = int(c["boundary_height"] / c["grid_cell_size"])
    return c


CONSTANTS = _constants()


@dataclass
class Body:
    """SPHScenarioBody (sph.h:162-222): kind, position, orientation columns,
    radius and local vertices."""

    kind: str  # "plane", "circle", "segment", "polygon"
    position: tuple
    col1: tuple = (f32(1.0), f32(0.0))
    col2: tuple = (f32(0.0), f32(1.0))
    radius: np.float32 = f32(0.0)
    local_verts: list = field(default_factory=list)


@dataclass
class Volume:
    position: tuple
    size: tuple
    force: tuple = (f32(0.0), f32(0.0))


@dataclass
class Emitter:
    position: tuple
    direction: tuple
    radius: np.float32
    speed: np.float32
    rate: np.float32
    duration: np.float32


@dataclass
class Scenario:
    name: str
    gravity: tuple
    volumes: list
    emitters: list
    bodies: list
    params: dict


def _v(x, y):
    return (f32(x), f32(y))


def _plane(position, normal):
    n = _v(*normal)
    return Body("plane", _v(*position), n, (f32(-1.0) * n[1], f32(1.0) * n[0]))


def _circle(position, radius):
    return Body("circle", _v(*position), radius=f32(radius))


def _box(position, angle, ext):
    s = f32(math.sin(float(f32(angle))))
    c = f32(math.cos(float(f32(angle))))
    ex, ey = f32(ext[0]), f32(ext[1])
    return Body("polygon", _v(*position), (c, s), (-s, c), f32(0.0), [(ex, ey), (-ex, ey), (-ex, -ey), (ex, -ey)])


def _walls(top=True):
    c = CONSTANTS
    w = [_plane((0.0, -c["boundary_half_height"]), (0.0, 1.0))]
    if top:
        w.append(_plane((0.0, c["boundary_half_height"]), (0.0, -1.0)))
    w.append(_plane((-c["boundary_half_width"], 0.0), (1.0, 0.0)))
    w.append(_plane((c["boundary_half_width"], 0.0), (-1.0, 0.0)))
    return w


def _params(spacing, stiffness, near_stiffness) -> dict:
    """The eight-argument SPHParameters constructor (sph.h:112-122)."""
    c = CONSTANTS
    return {"kernel_height": c["kernel_height"], "cell_size": c["grid_cell_size"], "particle_spacing": f32(spacing),
            "inv_kernel_height": f32(1.0) / c["kernel_height"], "rest_density": c["rest_density"],
            "stiffness": f32(stiffness), "near_stiffness": f32(near_stiffness),
            "linear_viscosity": c["linear_viscosity"], "quadratic_viscosity": c["quadratic_viscosity"]}


def _scenarios() -> list:
    c = CONSTANTS
    h, k = c["kernel_height"], c["stiffness"]
    hw, hh = c["boundary_half_width"], c["boundary_half_height"]
    wall_w, wall_h = c["boundary_width"] * f32(0.05), c["boundary_height"] * f32(0.85)   # sph.h:306-307
    vol_w, vol_h = c["boundary_width"] * f32(0.25), c["boundary_height"] * f32(0.95)     # sph.h:308-309
    blob_w, blob_h = c["boundary_width"] * f32(0.5), c["boundary_height"] * f32(0.5)     # sph.h:311-312
    deg2rad = f32(math.pi) / f32(180.0)                                                  # vecmath.h:9
    s = []
    # sph.h:315-328
    d = Scenario("Dambreak", _v(0, -10), [Volume(_v(-hw + vol_w * f32(0.5), 0), (vol_w, vol_h))], [], _walls(),
                 _params(h / f32(6.0), k, c["near_stiffness"]))
    d.bodies.append(_box(_v(((-hw + vol_w) + wall_w * f32(0.5)) + c["collision_radius"], c["boundary_height"] * f32(0.1)),
                         0.0, (wall_w * f32(0.5), wall_h * f32(0.5))))
    s.append(d)
    # sph.h:330-343
    s.append(Scenario("Dambreak x 2", _v(0, -10),
                      [Volume(_v(-hw + vol_w * f32(0.5), 0), (vol_w, vol_h)),
                       Volume(_v(hw - vol_w * f32(0.5), 0), (vol_w, vol_h))],
                      [], _walls(), _params(h / f32(3.0), k, k * f32(20.0))))
    # sph.h:345-357
    s.append(Scenario("Blob", _v(0, 0), [Volume(_v(0, 0), (blob_w, blob_h))], [], _walls(),
                      _params(h / f32(3.0), k, c["near_stiffness"])))
    # sph.h:359-372
    q = blob_h * f32(0.75)
    s.append(Scenario("Blob x 2", _v(0, 0),
                      [Volume(_v(-q, 0), (q, q), _v(10, 0)), Volume(_v(q, 0), (q, q), _v(-10, 0))],
                      [], _walls(), _params(h / f32(3.0), k, c["near_stiffness"])))
    # sph.h:374-386
    s.append(Scenario("Liquid", _v(0, -2), [],
                      [Emitter(_v(-3.5, 0.0), _v(1, 0), h * f32(3.0), f32(2.5), f32(15.0), f32(30.0))],
                      _walls(), _params(h / f32(4.0), k, k * f32(10.0))))
    # sph.h:388-403
    g = Scenario("Glass", _v(0, -10), [],
                 [Emitter(_v(-1.5, 2.0), _v(1, 0), h * f32(3.0), f32(2.5), f32(15.0), f32(25.0))],
                 _walls(), _params(h / f32(4.0), k, k * f32(6.0)))
    g.bodies += [_box((0.0, -2.0), 0.0, (1.0, 0.2)), _box((-1.0, -0.5), 0.0, (0.2, 1.5)),
                 _box((1.0, -0.5), 0.0, (0.2, 1.5))]
    s.append(g)
    # sph.h:405-417
    s.append(Scenario("Fontain", _v(0, -10), [],
                      [Emitter(_v(0.0, -hh + f32(1.0)), _v(0, 1), h * f32(4.0), f32(8.0), f32(15.0), f32(25.0))],
                      _walls(), _params(h / f32(4.0), k, k * f32(2.0))))
    # sph.h:419-436 (no top wall)
    fun = Scenario("Fun", _v(0, -10), [],
                   [Emitter(_v(-4, 2), _v(1, 0), h * f32(4.0), f32(3.5), f32(15.0), f32(20.0))],
                   _walls(top=False), _params(h / f32(4.0), k, k * f32(6.0)))
    fun.bodies += [_box((-1.5, 1.0), deg2rad * f32(-2.5), (3.5, 0.1)), _box((1.5, -0.25), deg2rad * f32(2.5), (3.5, 0.1)),
                   _circle((-4.0, -1.5), 0.5), _box((0.0, -hh + f32(0.5)), 0.0, (0.3, 1.0))]
    s.append(fun)
    return s


SCENARIOS = _scenarios()


def resolve_body(b: Body) -> dict:
    """The body as LoadScenario hands it over (app.cpp:487-515): planes as
    normal and distance = Vec2Dot(col1, position), polygons and segments as
    Vec2MultMat2(orientation, v) + position per vertex."""
    if b.kind == "plane":
        return {"kind": "plane", "normal": b.col1,
                "distance": b.col1[0] * b.position[0] + b.col1[1] * b.position[1]}
    if b.kind == "circle":
        return {"kind": "circle", "center": b.position, "radius": b.radius}
    verts = []
    for vx, vy in b.local_verts:
        rx = b.col1[0] * vx + b.col2[0] * vy
        ry = b.col1[1] * vx + b.col2[1] * vy
        verts.append((rx + b.position[0], ry + b.position[1]))
    return {"kind": b.kind, "verts": verts}


def volume_counts(v: Volume, spacing) -> tuple:
    """floor(size / spacing) per axis in float32 (app.cpp:523-524)."""
    return (int(math.floor(float(v.size[0] / f32(spacing)))), int(math.floor(float(v.size[1] / f32(spacing)))))


def scenario_params(s: Scenario) -> SPHParams:
    c = CONSTANTS
    p = s.params
    return SPHParams(h=float(p["kernel_height"]), rho0=float(p["rest_density"]), k=float(p["stiffness"]),
                     k_near=float(p["near_stiffness"]), sigma=float(p["linear_viscosity"]),
                     beta=float(p["quadratic_viscosity"]), collision_radius=float(c["collision_radius"]),
                     collision_margin=float(c["collision_margin"]))


def _get(s) -> Scenario:
    if isinstance(s, Scenario):
        return s
    if isinstance(s, str):
        for sc in SCENARIOS:
            if sc.name == s:
                return sc
        raise KeyError(s)
    return SCENARIOS[int(s)]


def make_simulation(scenario=0, capacity: int = 0):
    """A GPU simulation in the reference's domain with the scenario's parameters."""
    from .sim import ParticleSimulation
    s = _get(scenario)
    c = CONSTANTS
    return ParticleSimulation(scenario_params(s), float(c["boundary_width"]), float(c["boundary_height"]),
                              float(c["grid_cell_size"]), capacity, float(s.params["particle_spacing"]))


def load_scenario(sim, scenario, seed: int = 0) -> None:
    """DemoApplication::LoadScenario (app.cpp:477-534) on a ParticleSimulation:
    statistics reset, bodies, particles and emitters cleared, gravity and
    parameters set, then bodies, volumes and emitters in table order.  ``seed``
    selects the jitter of the volumes and emitters."""
    from .sim import params_to_c
    s = _get(scenario)
    sim.ResetStats()
    sim.ClearBodies()
    sim.ClearParticles()
    sim.ClearEmitters()
    sim.SetGravity(s.gravity)
    spacing = s.params["particle_spacing"]
    sim.SetParams(params_to_c(scenario_params(s), float(s.params["cell_size"]), float(spacing)))
    for b in s.bodies:
        r = resolve_body(b)
        if r["kind"] == "plane":
            sim.AddPlane(r["normal"], r["distance"])
        elif r["kind"] == "circle":
            sim.AddCircle(r["center"], r["radius"])
        elif r["kind"] == "segment":
            sim.AddLineSegment(r["verts"][0], r["verts"][1])
        else:
            sim.AddPolygon(np.array(r["verts"], np.float32))
    for v in s.volumes:
        nx, ny = volume_counts(v, spacing)
        sim.AddVolume(v.position, v.force, nx, ny, float(spacing), seed=seed)
    for e in s.emitters:
        sim.AddEmitter(e.position, e.direction, e.radius, e.speed, e.rate, e.duration, seed=seed)


def run_benchmark(sim, scenario, frames: int = BENCHMARK_FRAMES, iterations: int = BENCHMARK_ITERATIONS,
                  seed: int = 0, substeps: int = 1) -> dict:
    """The built-in benchmark (app.cpp:88-161, 215-275): per iteration the
    scenario is reloaded and stepped ``frames`` times with the frame dt; the
    wall time of each frame's Update calls and the per-phase statistics of
    every frame are reduced to min / max / avg.  The per-phase minimums start
    from +inf (the reference starts them from 0, so its minimum reads 0)."""
    s = _get(scenario)
    dt = float(CONSTANTS["delta_time"])
    sim.set_timing(True)
    sim_ms, phases, counts = [], {p: [] for p in PHASES}, []
    for _ in range(iterations):
        load_scenario(sim, s, seed)
        for _ in range(frames):
            t0 = time.perf_counter()
            for _ in range(substeps):
                sim.Update(dt)
            sim_ms.append((time.perf_counter() - t0) * 1e3)
            st = sim.GetStats()
            for p in PHASES:
                phases[p].append(st["time"][p])
            counts.append(sim.GetParticleCount())

    def mma(a):
        return [float(min(a)), float(max(a)), float(sum(a) / len(a))] if a else [0.0, 0.0, 0.0]

    return {"scenario": s.name, "frames": frames, "iterations": iterations, "particles": mma(counts),
            "simulation_ms": mma(sim_ms), "phases_ms": {p: mma(v) for p, v in phases.items()}}
