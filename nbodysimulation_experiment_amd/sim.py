"""Python mirror of the reference's simulation interface over the C ABI.

``ParticleSimulation`` keeps the method names, argument meaning and call order
of ``class BaseSimulation`` (NBodySimulation/base.h:8-39) as implemented by
``Demo4::ParticleSimulation`` (demo4.h:137-226), so a test written against the
reference's interface reads the same here.  Every call goes to the HIP library
(libnbs_sph.so); errors raise ``NbsError`` with the library's status code.

Extensions beyond the reference's interface: several independent scenes per
object (``AddScene`` and the ``scene`` keyword), bulk ``AddParticles``,
``get_state``/``set_state``, ``get_grid``, FP64 ``diagnostics``, per-kernel
timing, HIP-graph replay, and the learned relaxation (``surrogate_*``).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import NbsDiagnostics, NbsDomain, NbsParams, NbsStats, NbsSurrogateReport, NbsTuning, call
from .scenes import SPHParams

__all__ = ["ParticleSimulation", "NbsError", "params_to_c", "selftest_math", "selftest_mfma", "bench_copy"]

NbsError = _lib.NbsError


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, cols=2):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return a.reshape(-1, cols) if cols else a


def params_to_c(p: SPHParams, cell_size: float = 1.0, particle_spacing: float | None = None) -> NbsParams:
    spacing = p.h * 0.5 if particle_spacing is None else particle_spacing
    return NbsParams(p.h, cell_size, spacing, p.rho0, p.k, p.k_near, p.sigma, p.beta, p.collision_radius,
                     p.collision_margin)


class ParticleSimulation:
    """BaseSimulation over one MI355X (the GPU current when created)."""

    STAT_TIME_FIELDS = ("emitters", "integration", "viscosityForces", "predict", "updateGrid", "neighborSearch",
                        "densityAndPressure", "deltaPositions", "collisions")

    def __init__(self, params: SPHParams | None = None, width: float = 80.0, height: float = 60.0,
                 cell_size: float | None = None, capacity: int = 0, particle_spacing: float | None = None):
        self._lib = _lib.load()
        self.params = params or SPHParams()
        cs = self.params.h if cell_size is None else cell_size
        self._cparams = params_to_c(self.params, cs, particle_spacing)
        dom = NbsDomain(width, height, cs)
        h = ctypes.c_void_p()
        call("nbs_create", ctypes.byref(self._cparams), ctypes.byref(dom), int(capacity), ctypes.byref(h))
        self._h = h
        self._multithreading = False

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            call("nbs_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- base.h:10-13 --------------------------------------------------------
    def ResetStats(self):
        call("nbs_reset_stats", self._h)

    def ClearBodies(self):
        call("nbs_clear_bodies", self._h)

    def ClearParticles(self):
        call("nbs_clear_particles", self._h)

    def ClearEmitters(self):
        call("nbs_clear_emitters", self._h)

    # -- base.h:15-18 --------------------------------------------------------
    def AddPlane(self, normal, distance, scene: int = 0):
        call("nbs_add_plane", self._h, scene, float(normal[0]), float(normal[1]), float(distance))

    def AddCircle(self, pos, radius, scene: int = 0):
        call("nbs_add_circle", self._h, scene, float(pos[0]), float(pos[1]), float(radius))

    def AddLineSegment(self, a, b, scene: int = 0):
        call("nbs_add_line_segment", self._h, scene, float(a[0]), float(a[1]), float(b[0]), float(b[1]))

    def AddPolygon(self, verts, scene: int = 0):
        v = _f32(verts)
        call("nbs_add_polygon", self._h, scene, v.shape[0], _ptr(v))

    # -- base.h:20-22 --------------------------------------------------------
    def AddParticle(self, position, force=(0.0, 0.0), scene: int = 0) -> int:
        return self.AddParticles([position], [force], scene)

    def AddParticles(self, positions, forces=None, scene: int = 0) -> int:
        pos = _f32(positions)
        frc = None if forces is None else _f32(forces)
        if frc is not None and frc.shape != pos.shape:
            raise ValueError("forces must match positions")
        first = ctypes.c_uint64()
        call("nbs_add_particles", self._h, scene, _ptr(pos), _ptr(frc), pos.shape[0], ctypes.byref(first))
        return int(first.value)

    def AddVolume(self, center, force, countX: int, countY: int, spacing: float, scene: int = 0, seed: int = 0):
        call("nbs_add_volume", self._h, scene, float(center[0]), float(center[1]), float(force[0]), float(force[1]),
             int(countX), int(countY), float(spacing), int(seed))

    def AddEmitter(self, position, direction, radius, speed, rate, duration, scene: int = 0, seed: int = 0):
        call("nbs_add_emitter", self._h, scene, float(position[0]), float(position[1]), float(direction[0]),
             float(direction[1]), float(radius), float(speed), float(rate), float(duration), int(seed))

    # -- base.h:24 -------------------------------------------------------------
    def Update(self, deltaTime: float, steps: int = 1):
        call("nbs_update", self._h, float(deltaTime), int(steps))

    # -- base.h:27-28 ----------------------------------------------------------
    def AddExternalForces(self, force):
        call("nbs_add_external_force", self._h, float(force[0]), float(force[1]))

    def ClearExternalForce(self):
        call("nbs_clear_external_force", self._h)

    # -- base.h:30-38 ----------------------------------------------------------
    def GetParticleCount(self) -> int:
        c = ctypes.c_uint64()
        call("nbs_get_particle_count", self._h, ctypes.byref(c))
        return int(c.value)

    def SetGravity(self, gravity):
        call("nbs_set_gravity", self._h, float(gravity[0]), float(gravity[1]))

    def GetParams(self) -> NbsParams:
        p = NbsParams()
        call("nbs_get_params", self._h, ctypes.byref(p))
        return p

    def SetParams(self, params):
        if isinstance(params, SPHParams):
            params = params_to_c(params, self._cparams.cell_size, self._cparams.particle_spacing)
        call("nbs_set_params", self._h, ctypes.byref(params))
        self._cparams = params

    def GetStats(self) -> dict:
        s = NbsStats()
        call("nbs_get_stats", self._h, ctypes.byref(s))
        return {
            "minParticleNeighborCount": int(s.min_particle_neighbor_count),
            "maxParticleNeighborCount": int(s.max_particle_neighbor_count),
            "minCellParticleCount": int(s.min_cell_particle_count),
            "maxCellParticleCount": int(s.max_cell_particle_count),
            "time": {
                "emitters": s.time_emitters, "integration": s.time_integration,
                "viscosityForces": s.time_viscosity_forces, "predict": s.time_predict,
                "updateGrid": s.time_update_grid, "neighborSearch": s.time_neighbor_search,
                "densityAndPressure": s.time_density_and_pressure, "deltaPositions": s.time_delta_positions,
                "collisions": s.time_collisions,
            },
        }

    def SetMultiThreading(self, value: bool):
        # The GPU is the parallelism; the switch is recorded and ignored.
        self._multithreading = bool(value)

    def IsMultiThreadingSupported(self) -> bool:
        s = ctypes.c_int32()
        call("nbs_is_multithreading_supported", self._h, ctypes.byref(s))
        return bool(s.value)

    def IsMultiThreading(self) -> bool:
        return False

    def GetWorkerThreadCount(self) -> int:
        return 1

    # -- extensions ----------------------------------------------------------
    def AddScene(self, width: float, height: float, cell_size: float | None = None) -> int:
        cs = self._cparams.cell_size if cell_size is None else cell_size
        s = ctypes.c_int32()
        call("nbs_add_scene", self._h, ctypes.byref(NbsDomain(width, height, cs)), ctypes.byref(s))
        return int(s.value)

    def GetSceneCount(self) -> int:
        s = ctypes.c_int32()
        call("nbs_get_scene_count", self._h, ctypes.byref(s))
        return int(s.value)

    def add_scene_spec(self, spec, scene: int | None = None) -> int:
        """Adds a SceneSpec: a new scene unless ``scene`` is given."""
        if scene is None:
            scene = self.AddScene(spec.width, spec.height, spec.cell_size)
        for nx, ny, d in spec.planes:
            self.AddPlane((nx, ny), d, scene)
        for cx, cy, r in spec.circles:
            self.AddCircle((cx, cy), r, scene)
        self.AddParticles(spec.positions, None, scene)
        return scene

    @classmethod
    def from_scene(cls, spec, params: SPHParams | None = None, gravity=(0.0, -9.8), capacity: int = 0):
        sim = cls(params, spec.width, spec.height, spec.cell_size, capacity=max(capacity, spec.n))
        sim.add_scene_spec(spec, scene=0)
        sim.SetGravity(gravity)
        return sim

    @classmethod
    def from_scenes(cls, specs, params: SPHParams | None = None, gravity=(0.0, -9.8)):
        total = sum(s.n for s in specs)
        sim = cls(params, specs[0].width, specs[0].height, specs[0].cell_size, capacity=total)
        sim.add_scene_spec(specs[0], scene=0)
        for spec in specs[1:]:
            sim.add_scene_spec(spec)
        sim.SetGravity(gravity)
        return sim

    def get_state(self) -> dict:
        n = self.GetParticleCount()
        out = {k: np.zeros((n, 2), np.float32) for k in ("x", "prev", "v", "density", "pressure")}
        out["key"] = np.zeros(n, np.uint32)
        out["scene"] = np.zeros(n, np.int32)
        call("nbs_get_state", self._h, _ptr(out["x"]), _ptr(out["prev"]), _ptr(out["v"]), _ptr(out["density"]),
             _ptr(out["pressure"]), _ptr(out["key"]), _ptr(out["scene"]))
        return out

    def get_colors(self) -> np.ndarray:
        """Per-particle RGBA of the reference's Render (sph.h:683-695), particle-index order."""
        out = np.zeros((self.GetParticleCount(), 4), np.float32)
        call("nbs_get_colors", self._h, _ptr(out))
        return out

    def set_state(self, x, v):
        x, v = _f32(x), _f32(v)
        n = self.GetParticleCount()
        if x.shape != (n, 2) or v.shape != (n, 2):
            raise ValueError(f"state must be ({n}, 2)")
        call("nbs_set_state", self._h, _ptr(x), _ptr(v))

    @property
    def total_cells(self) -> int:
        c = ctypes.c_uint64()
        call("nbs_get_total_cells", self._h, ctypes.byref(c))
        return int(c.value)

    def get_grid(self):
        ids = np.zeros(self.GetParticleCount(), np.uint32)
        start = np.zeros(self.total_cells + 1, np.uint32)
        call("nbs_get_grid", self._h, _ptr(ids), _ptr(start))
        return ids, start

    def diagnostics(self) -> dict:
        d = NbsDiagnostics()
        call("nbs_get_diagnostics", self._h, ctypes.byref(d))
        return {"particle_count": int(d.particle_count), "momentum": (d.momentum[0], d.momentum[1]),
                "kinetic_energy": d.kinetic_energy, "mean_density": d.mean_density, "max_density": d.max_density,
                "mean_abs_density_error": d.mean_abs_density_error, "rms_density_error": d.rms_density_error,
                "centroid": (d.centroid[0], d.centroid[1]), "non_finite_count": int(d.non_finite_count)}

    def set_timing(self, enabled: bool):
        call("nbs_set_timing", self._h, 1 if enabled else 0)

    def set_options(self, hip_graph: bool = False, check_finite: bool = False):
        flags = (_lib.OPT_HIP_GRAPH if hip_graph else 0) | (_lib.OPT_CHECK_FINITE if check_finite else 0)
        call("nbs_set_options", self._h, flags)

    def get_tuning(self) -> dict:
        t = NbsTuning()
        call("nbs_get_tuning", self._h, ctypes.byref(t))
        return {"autotune": bool(t.autotune), "retune_interval": int(t.retune_interval),
                "tile_cap": [int(c) for c in t.tile_cap], "wide_max_particles": int(t.wide_max_particles)}

    def set_tuning(self, autotune: bool = True, retune_interval: int = 256, tile_cap=(2048, 2048, 2048),
                   wide_max_particles: int = 0, tuned: bool = False):
        """autotune: time the kernel variants on real steps (first all of them,
        then every retune_interval steps the current choice and its
        neighbours); tuned=True takes tile_cap as an earlier tuning's result,
        so only the periodic re-tuning runs."""
        mode = (2 if tuned else 1) if autotune else 0
        t = NbsTuning(mode, int(retune_interval), (ctypes.c_int32 * 3)(*tile_cap), int(wide_max_particles))
        call("nbs_set_tuning", self._h, ctypes.byref(t))

    def kernel_times(self) -> dict:
        avg = np.zeros(len(_lib.KERNEL_NAMES), np.float64)
        cnt = np.zeros(len(_lib.KERNEL_NAMES), np.uint64)
        call("nbs_get_kernel_times", self._h, _ptr(avg), _ptr(cnt))
        return {k: (float(a), int(c)) for k, a, c in zip(_lib.KERNEL_NAMES, avg, cnt)}

    # -- learned relaxation (include/nbs_surrogate.h) ---------------------------
    def surrogate_init(self, seed: int = 0):
        call("nbs_surrogate_init", self._h, int(seed))

    def surrogate_params(self) -> np.ndarray:
        p = np.zeros(_lib.SURROGATE_PARAM_COUNT, np.float32)
        call("nbs_surrogate_get_params", self._h, _ptr(p), p.size)
        return p

    def surrogate_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        call("nbs_surrogate_set_params", self._h, _ptr(p), p.size)

    def surrogate_calibrate(self, dt: float):
        call("nbs_surrogate_calibrate", self._h, float(dt))

    def surrogate_train(self, dt: float, steps: int, learning_rate: float = 1e-3, relative: bool = False):
        """Per-step losses (mean squared error); with relative=True also the
        relative errors sum|o - t/os|^2 / sum|t/os|^2 as a second array."""
        losses = np.zeros(max(int(steps), 1), np.float32)
        rel = np.zeros(max(int(steps), 1), np.float32)
        call("nbs_surrogate_train", self._h, float(dt), int(steps), float(learning_rate), _ptr(losses), _ptr(rel))
        return (losses[:int(steps)], rel[:int(steps)]) if relative else losses[:int(steps)]

    def surrogate_set_mode(self, surrogate: bool, check_interval: int = 8, max_density_error: float = 0.01,
                           carried_density: bool = False):
        """The network relaxes every step (surrogate=True), checked every
        check_interval steps; with carried_density the steps between checks
        skip the density pass and take the last check's density."""
        mode = _lib.RELAX_PLAIN
        if surrogate:
            mode = _lib.RELAX_SURROGATE_CARRIED_DENSITY if carried_density else _lib.RELAX_SURROGATE
        call("nbs_surrogate_set_mode", self._h, mode, int(check_interval), float(max_density_error))

    def surrogate_set_online_training(self, learning_rate: float):
        """Check steps in surrogate mode also train on their plain frame (0: off)."""
        call("nbs_surrogate_set_online_training", self._h, float(learning_rate))

    def surrogate_scene_errors(self) -> np.ndarray:
        """Per-scene density errors of the last check step."""
        out = np.zeros(self.GetSceneCount(), np.float64)
        call("nbs_surrogate_get_scene_errors", self._h, _ptr(out), out.size)
        return out

    def surrogate_report(self) -> dict:
        r = NbsSurrogateReport()
        call("nbs_surrogate_get_report", self._h, ctypes.byref(r))
        return {n: (getattr(r, n) if t is ctypes.c_double else int(getattr(r, n))) for n, t in r._fields_}

    def surrogate_frame(self, dt: float) -> dict:
        """One plain step; the frame the network saw (slot order) and the
        gradient of the loss on it, without a weight update."""
        n0 = self.GetParticleCount()
        cap = n0 + 4096  # room for particles an emitter adds during the step
        grad = np.zeros(_lib.SURROGATE_WEIGHT_COUNT, np.float32)
        loss = np.zeros(1, np.float32)
        arrs = {k: np.zeros((cap, 2), np.float32) for k in ("out", "x", "prev", "density", "target")}
        scene = np.zeros(cap, np.int32)
        call("nbs_surrogate_frame", self._h, float(dt), _ptr(grad), _ptr(loss), _ptr(arrs["out"]), _ptr(arrs["x"]),
             _ptr(arrs["prev"]), _ptr(arrs["density"]), _ptr(arrs["target"]), _ptr(scene))
        n = self.GetParticleCount()
        if n > cap:
            raise RuntimeError("an emitter added more particles than the frame buffers hold")
        res = {k: v[:n].copy() for k, v in arrs.items()}
        res.update(grad=grad, loss=float(loss[0]), scene=scene[:n].copy())
        return res


def selftest_math(a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros((a.size, 6), np.float32)
    call("nbs_selftest_math", _ptr(a), _ptr(b), a.size, _ptr(out))
    return out


def selftest_mfma(a, b) -> np.ndarray:
    """C = A (32 x 8) B (8 x 32) with BF16-rounded operands by one MFMA."""
    a = np.ascontiguousarray(a, np.float32).reshape(32, 8)
    b = np.ascontiguousarray(b, np.float32).reshape(8, 32)
    c = np.zeros((32, 32), np.float32)
    call("nbs_selftest_mfma", _ptr(a), _ptr(b), _ptr(c))
    return c


def selftest_mfma16(a, b) -> np.ndarray:
    """C = A (32 x 16) B (16 x 32) with BF16-rounded operands by one K = 16 MFMA."""
    a = np.ascontiguousarray(a, np.float32).reshape(32, 16)
    b = np.ascontiguousarray(b, np.float32).reshape(16, 32)
    c = np.zeros((32, 32), np.float32)
    call("nbs_selftest_mfma16", _ptr(a), _ptr(b), _ptr(c))
    return c


def bench_copy(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Device copy bandwidth (read + write GB/s) of the library's copy kernel."""
    g = ctypes.c_double()
    call("nbs_bench_copy", int(nbytes), int(reps), ctypes.byref(g))
    return float(g.value)
