"""ctypes binding of libnbs_sph.so (the C ABI declared in include/nbs_sph.h and
include/nbs_surrogate.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libnbs_sph.so")
HEADER_PATH = os.path.join(ROOT, "include", "nbs_sph.h")
HEADER_PATHS = (HEADER_PATH, os.path.join(ROOT, "include", "nbs_surrogate.h"))
SURROGATE_WEIGHT_COUNT = 9474
SURROGATE_PARAM_COUNT = 9482
RELAX_PLAIN = 0
RELAX_SURROGATE = 1
RELAX_SURROGATE_CARRIED_DENSITY = 2

c_float, c_int32, c_uint32, c_uint64, c_double = (ctypes.c_float, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64,
                                                  ctypes.c_double)
vp = ctypes.c_void_p

STATUS_NAMES = {
    0: "NBS_OK",
    1: "NBS_ERR_INVALID_ARGUMENT",
    2: "NBS_ERR_OUT_OF_MEMORY",
    3: "NBS_ERR_HIP",
    4: "NBS_ERR_CAPACITY",
    5: "NBS_ERR_NON_FINITE",
    6: "NBS_ERR_NO_DEVICE",
    7: "NBS_ERR_UNSUPPORTED",
}

KERNEL_NAMES = ("viscosity_predict", "grid", "density", "delta")
OPT_HIP_GRAPH = 1
OPT_CHECK_FINITE = 2


class NbsParams(ctypes.Structure):
    _fields_ = [(n, c_float) for n in ("kernel_height", "cell_size", "particle_spacing", "rest_density", "stiffness",
                                       "near_stiffness", "linear_viscosity", "quadratic_viscosity",
                                       "collision_radius", "collision_margin")]


class NbsDomain(ctypes.Structure):
    _fields_ = [("width", c_float), ("height", c_float), ("cell_size", c_float)]


class NbsStats(ctypes.Structure):
    _fields_ = [("min_particle_neighbor_count", c_uint64), ("max_particle_neighbor_count", c_uint64),
                ("min_cell_particle_count", c_uint64), ("max_cell_particle_count", c_uint64)] + \
               [(n, c_float) for n in ("time_emitters", "time_integration", "time_viscosity_forces", "time_predict",
                                       "time_update_grid", "time_neighbor_search", "time_density_and_pressure",
                                       "time_delta_positions", "time_collisions")]


class NbsDiagnostics(ctypes.Structure):
    _fields_ = [("particle_count", c_uint64), ("momentum", c_double * 2), ("kinetic_energy", c_double),
                ("mean_density", c_double), ("max_density", c_double), ("mean_abs_density_error", c_double),
                ("rms_density_error", c_double), ("centroid", c_double * 2), ("non_finite_count", c_uint64)]


class NbsTuning(ctypes.Structure):
    _fields_ = [("autotune", c_int32), ("retune_interval", c_uint32), ("tile_cap", c_int32 * 3),
                ("wide_max_particles", c_int32)]


class NbsSurrogateReport(ctypes.Structure):
    _fields_ = [(n, c_uint64) for n in ("surrogate_steps", "plain_steps", "check_steps", "fallbacks",
                                        "train_steps")] + \
               [(n, c_double) for n in ("last_density_error", "max_density_error", "mean_density_error")] + \
               [("last_pair_count", c_uint64), ("last_max_scene_error", c_double),
                ("last_scenes_over_limit", c_uint64)]


class NbsError(RuntimeError):
    def __init__(self, code: int, func: str, message: str):
        self.code = code
        self.status = STATUS_NAMES.get(code, f"status {code}")
        super().__init__(f"{func}: {self.status}: {message}")


P = ctypes.POINTER
SIGNATURES = {
    "nbs_create": [P(NbsParams), P(NbsDomain), c_uint64, P(vp)],
    "nbs_destroy": [vp],
    "nbs_add_scene": [vp, P(NbsDomain), P(c_int32)],
    "nbs_get_scene_count": [vp, P(c_int32)],
    "nbs_set_params": [vp, P(NbsParams)],
    "nbs_get_params": [vp, P(NbsParams)],
    "nbs_set_gravity": [vp, c_float, c_float],
    "nbs_add_external_force": [vp, c_float, c_float],
    "nbs_clear_external_force": [vp],
    "nbs_reset_stats": [vp],
    "nbs_clear_bodies": [vp],
    "nbs_clear_particles": [vp],
    "nbs_clear_emitters": [vp],
    "nbs_add_plane": [vp, c_int32, c_float, c_float, c_float],
    "nbs_add_circle": [vp, c_int32, c_float, c_float, c_float],
    "nbs_add_line_segment": [vp, c_int32, c_float, c_float, c_float, c_float],
    "nbs_add_polygon": [vp, c_int32, c_uint64, vp],
    "nbs_add_particles": [vp, c_int32, vp, vp, c_uint64, P(c_uint64)],
    "nbs_add_volume": [vp, c_int32, c_float, c_float, c_float, c_float, c_int32, c_int32, c_float, c_uint64],
    "nbs_add_emitter": [vp, c_int32, c_float, c_float, c_float, c_float, c_float, c_float, c_float, c_float,
                        c_uint64],
    "nbs_update": [vp, c_float, c_uint32],
    "nbs_get_particle_count": [vp, P(c_uint64)],
    "nbs_get_state": [vp, vp, vp, vp, vp, vp, vp, vp],
    "nbs_set_state": [vp, vp, vp],
    "nbs_get_colors": [vp, vp],
    "nbs_get_stats": [vp, P(NbsStats)],
    "nbs_get_kernel_times": [vp, vp, vp],
    "nbs_set_timing": [vp, c_int32],
    "nbs_set_options": [vp, c_uint32],
    "nbs_get_diagnostics": [vp, P(NbsDiagnostics)],
    "nbs_diagnostics_noalloc": [vp, vp, vp, c_uint64, c_float, vp, vp],
    "nbs_set_tuning": [vp, P(NbsTuning)],
    "nbs_get_tuning": [vp, P(NbsTuning)],
    "nbs_get_total_cells": [vp, P(c_uint64)],
    "nbs_get_grid": [vp, vp, vp],
    "nbs_is_multithreading_supported": [vp, P(c_int32)],
    "nbs_selftest_math": [vp, vp, c_uint64, vp],
    "nbs_bench_copy": [c_uint64, c_int32, P(c_double)],
    "nbs_surrogate_init": [vp, c_uint64],
    "nbs_surrogate_set_params": [vp, vp, c_uint64],
    "nbs_surrogate_get_params": [vp, vp, c_uint64],
    "nbs_surrogate_calibrate": [vp, c_float],
    "nbs_surrogate_train": [vp, c_float, c_uint32, c_float, vp, vp],
    "nbs_surrogate_set_mode": [vp, c_int32, c_uint32, c_float],
    "nbs_surrogate_get_report": [vp, P(NbsSurrogateReport)],
    "nbs_surrogate_set_online_training": [vp, c_float],
    "nbs_surrogate_get_scene_errors": [vp, vp, c_uint64],
    "nbs_surrogate_frame": [vp, c_float, vp, vp, vp, vp, vp, vp, vp, vp],
    "nbs_selftest_mfma": [vp, vp, vp],
    "nbs_selftest_mfma16": [vp, vp, vp],
}


def declared_symbols(headers=HEADER_PATHS) -> list[str]:
    """Every function the C headers declare."""
    if isinstance(headers, str):
        headers = (headers,)
    text = "".join(open(h).read() for h in headers)
    return sorted(set(re.findall(r"^\s*(?:nbs_status|const char \*)\s*\**\s*(nbs_\w+)\s*\(", text, re.M)))


_lib = None


def load():
    """Loads libnbs_sph.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = c_int32
    L.nbs_status_string.argtypes = [c_int32]
    L.nbs_status_string.restype = ctypes.c_char_p
    L.nbs_last_error.argtypes = []
    L.nbs_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(code: int, func: str):
    if code != 0:
        msg = load().nbs_last_error().decode(errors="replace")
        raise NbsError(code, func, msg)


def call(func: str, *args):
    check(getattr(load(), func)(*args), func)
