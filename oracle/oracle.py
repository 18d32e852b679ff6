"""ctypes wrapper of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference results the GPU path is
compared with or as the CPU baseline, never as the measured GPU path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsph_oracle.so")

JACOBI, GS, GS_MT = 0, 1, 2
BODY_PLANE, BODY_CIRCLE, BODY_SEGMENT, BODY_POLYGON = 1, 2, 3, 4

# Planted-mistake switches (oracle/sph_oracle.hpp enum Mutation).
MUTATIONS = {
    "drop_near_density": 1,
    "delta_sign": 2,
    "key_swap_xy": 3,
    "cutoff_le": 4,
    "drop_visc_quad": 5,
    "plane_sign": 6,
    "pressure_no_rest": 7,
    "velocity_no_inv_dt": 8,
    "circle_no_radius": 9,
    "skip_self": 10,
    "grid_by_index": 11,
}


class ParamsF(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("h", "rho0", "k", "k_near", "sigma", "beta", "collision_radius", "collision_margin")]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, f, i32, i64 = ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_int64
        P = ctypes.POINTER
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i32, i32, i32, P(ParamsF)]
        L.orc_destroy.argtypes = [vp]
        L.orc_add_scene.restype = i32
        L.orc_add_scene.argtypes = [vp, f, f, f]
        L.orc_add_body.restype = i32
        L.orc_add_body.argtypes = [vp, i32, i32, vp, i32]
        L.orc_add_particles.restype = i64
        L.orc_add_particles.argtypes = [vp, i32, vp, vp, i64]
        L.orc_set_gravity.argtypes = [vp, f, f]
        L.orc_set_spacing.argtypes = [vp, f]
        L.orc_set_external_force.argtypes = [vp, f, f]
        L.orc_step.argtypes = [vp, f, i32]
        L.orc_add_emitter.argtypes = [vp, i32, f, f, f, f, f, f, f, f, ctypes.c_uint64]
        L.orc_count.restype = i64
        L.orc_count.argtypes = [vp]
        L.orc_total_cells.restype = i64
        L.orc_total_cells.argtypes = [vp]
        L.orc_get_state.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_set_state.argtypes = [vp, vp, vp]
        L.orc_neighbors.restype = i64
        L.orc_neighbors.argtypes = [vp, i64, vp, i64]
        L.orc_grid.argtypes = [vp, vp, vp]
        L.orc_get_stats.argtypes = [vp, vp, vp]
        L.orc_reset_stats.argtypes = [vp]
        L.orc_set_mutation.argtypes = [i32]
        L.orc_get_mutation.restype = i32
        L.orc_kat_density.argtypes = [i32, P(ParamsF), vp, vp, i64, vp]
        L.orc_kat_pressure.argtypes = [i32, P(ParamsF), vp, i64, vp]
        L.orc_kat_delta.argtypes = [i32, P(ParamsF), vp, vp, vp, f, i64, vp]
        L.orc_kat_viscosity.argtypes = [i32, P(ParamsF), vp, vp, vp, vp, i64, vp]
        L.orc_kat_collide.argtypes = [i32, P(ParamsF), i32, vp, i32, vp, i64, vp]
        L.orc_kat_cell.argtypes = [i32, f, f, f, vp, i64, vp]
        L.orc_kat_color.argtypes = [f, vp, vp, vp, i64, vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def params_struct(params) -> ParamsF:
    return ParamsF(*[float(v) for v in params.as_tuple()])


class set_mutation:
    """Context manager planting one mistake in the oracle (tests only)."""

    def __init__(self, name_or_id):
        self.m = MUTATIONS[name_or_id] if isinstance(name_or_id, str) else int(name_or_id)

    def __enter__(self):
        lib().orc_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutation(0)
        return False


class OracleSim:
    """One oracle simulation.  mode: JACOBI (GPU order), GS (reference order),
    GS_MT (reference order with pooled phases).  precision: 32 or 64."""

    def __init__(self, params, precision=32, mode=JACOBI, threads=0):
        self._params = params_struct(params)
        self.h = lib().orc_create(int(precision), int(mode), int(threads), ctypes.byref(self._params))
        self.precision = precision
        self.mode = mode

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, width, height, cell_size=1.0) -> int:
        return lib().orc_add_scene(self.h, float(width), float(height), float(cell_size))

    def add_plane(self, scene, nx, ny, d):
        a = _f32([nx, ny, d])
        assert lib().orc_add_body(self.h, scene, BODY_PLANE, _ptr(a), 0) == 0

    def add_circle(self, scene, cx, cy, r):
        a = _f32([cx, cy, r])
        assert lib().orc_add_body(self.h, scene, BODY_CIRCLE, _ptr(a), 0) == 0

    def add_segment(self, scene, ax, ay, bx, by):
        a = _f32([ax, ay, bx, by])
        assert lib().orc_add_body(self.h, scene, BODY_SEGMENT, _ptr(a), 0) == 0

    def add_polygon(self, scene, verts):
        a = _f32(verts).reshape(-1)
        assert lib().orc_add_body(self.h, scene, BODY_POLYGON, _ptr(a), a.size // 2) == 0

    def add_particles(self, scene, pos, force=None) -> int:
        pos = _f32(pos, (-1, 2))
        force = None if force is None else _f32(force, (-1, 2))
        r = lib().orc_add_particles(self.h, scene, _ptr(pos), _ptr(force), pos.shape[0])
        assert r >= 0
        return r

    def add_scene_spec(self, spec) -> int:
        s = self.add_scene(spec.width, spec.height, spec.cell_size)
        for nx, ny, d in spec.planes:
            self.add_plane(s, nx, ny, d)
        for cx, cy, r in spec.circles:
            self.add_circle(s, cx, cy, r)
        self.add_particles(s, spec.positions)
        return s

    def add_emitter(self, scene, pos, direction, radius, speed, rate, duration, seed=0):
        lib().orc_add_emitter(self.h, scene, float(pos[0]), float(pos[1]), float(direction[0]), float(direction[1]),
                              float(radius), float(speed), float(rate), float(duration), int(seed))

    def set_spacing(self, spacing):
        """SPHParameters::particleSpacing used by the emitters (demo4.cpp:258)."""
        lib().orc_set_spacing(self.h, float(spacing))

    def set_gravity(self, gx, gy):
        lib().orc_set_gravity(self.h, float(gx), float(gy))

    def set_external_force(self, fx, fy):
        lib().orc_set_external_force(self.h, float(fx), float(fy))

    def step(self, dt, nsteps=1):
        lib().orc_step(self.h, float(dt), int(nsteps))

    @property
    def n(self) -> int:
        return lib().orc_count(self.h)

    @property
    def total_cells(self) -> int:
        return lib().orc_total_cells(self.h)

    def state(self) -> dict:
        n = self.n
        out = {k: np.zeros((n, 2), np.float64) for k in ("x", "prev", "v", "density", "pressure")}
        key = np.zeros(n, np.uint32)
        lib().orc_get_state(self.h, _ptr(out["x"]), _ptr(out["prev"]), _ptr(out["v"]), _ptr(out["density"]),
                            _ptr(out["pressure"]), _ptr(key))
        out["key"] = key
        if self.precision == 32:
            for k in ("x", "prev", "v", "density", "pressure"):
                out[k] = out[k].astype(np.float32)
        return out

    def set_state(self, x, v):
        x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, 2))
        v = np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1, 2))
        lib().orc_set_state(self.h, _ptr(x), _ptr(v))

    def neighbors(self, i) -> np.ndarray:
        cap = 4096
        while True:
            buf = np.zeros(cap, np.uint32)
            c = lib().orc_neighbors(self.h, int(i), _ptr(buf), cap)
            if c <= cap:
                return buf[:c].copy()
            cap = c

    def grid(self):
        order = np.zeros(self.n, np.uint32)
        start = np.zeros(self.total_cells + 1, np.uint32)
        lib().orc_grid(self.h, _ptr(order), _ptr(start))
        return order, start

    def stats(self) -> dict:
        c = np.zeros(4, np.uint64)
        t = np.zeros(9, np.float32)
        lib().orc_get_stats(self.h, _ptr(c), _ptr(t))
        names = ("emitters", "integration", "viscosityForces", "predict", "updateGrid", "neighborSearch",
                 "densityAndPressure", "deltaPositions", "collisions")
        return {"minParticleNeighborCount": int(c[0]), "maxParticleNeighborCount": int(c[1]),
                "minCellParticleCount": int(c[2]), "maxCellParticleCount": int(c[3]),
                "time": dict(zip(names, [float(v) for v in t]))}


# ---------------------------------------------------------------------------
# Single-function entry points (known-answer checks)
# ---------------------------------------------------------------------------
def kat_density(params, xi, xj, precision=64):
    xi, xj = _f32(xi, (-1, 2)), _f32(xj, (-1, 2))
    out = np.zeros((xi.shape[0], 2), np.float64)
    lib().orc_kat_density(precision, ctypes.byref(params_struct(params)), _ptr(xi), _ptr(xj), xi.shape[0], _ptr(out))
    return out


def kat_pressure(params, dens, precision=64):
    dens = _f32(dens, (-1, 2))
    out = np.zeros((dens.shape[0], 2), np.float64)
    lib().orc_kat_pressure(precision, ctypes.byref(params_struct(params)), _ptr(dens), dens.shape[0], _ptr(out))
    return out


def kat_delta(params, xi, xj, pres, dt, precision=64):
    xi, xj, pres = _f32(xi, (-1, 2)), _f32(xj, (-1, 2)), _f32(pres, (-1, 2))
    out = np.zeros((xi.shape[0], 2), np.float64)
    lib().orc_kat_delta(precision, ctypes.byref(params_struct(params)), _ptr(xi), _ptr(xj), _ptr(pres), float(dt),
                        xi.shape[0], _ptr(out))
    return out


def kat_viscosity(params, xi, xj, vi, vj, precision=64):
    xi, xj, vi, vj = (_f32(a, (-1, 2)) for a in (xi, xj, vi, vj))
    out = np.zeros((xi.shape[0], 2), np.float64)
    lib().orc_kat_viscosity(precision, ctypes.byref(params_struct(params)), _ptr(xi), _ptr(xj), _ptr(vi), _ptr(vj),
                            xi.shape[0], _ptr(out))
    return out


def kat_collide(params, body_type, data, pos, precision=64):
    data = _f32(data).reshape(-1)
    pos = _f32(pos, (-1, 2))
    nverts = data.size // 2 if body_type == BODY_POLYGON else 0
    out = np.zeros((pos.shape[0], 2), np.float64)
    lib().orc_kat_collide(precision, ctypes.byref(params_struct(params)), int(body_type), _ptr(data), nverts,
                          _ptr(pos), pos.shape[0], _ptr(out))
    return out


def kat_cell(width, height, cell_size, pos, precision=32):
    pos = _f32(pos, (-1, 2))
    out = np.zeros((pos.shape[0], 3), np.int32)
    lib().orc_kat_cell(precision, float(width), float(height), float(cell_size), _ptr(pos), pos.shape[0], _ptr(out))
    return out


def kat_color(rest_density, density, pressure, vel):
    density, pressure = _f32(density).reshape(-1), _f32(pressure).reshape(-1)
    vel = _f32(vel, (-1, 2))
    out = np.zeros((density.size, 4), np.float32)
    lib().orc_kat_color(float(rest_density), _ptr(density), _ptr(pressure), _ptr(vel), density.size, _ptr(out))
    return out
