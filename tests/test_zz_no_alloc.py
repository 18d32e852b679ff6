"""nbs_diagnostics_noalloc, the entry point that allocates nothing, against
nbs_get_diagnostics on the same particle states.

The caller owns the device memory: this test allocates the input arrays, the
output record and a stream through the HIP runtime the library is linked
against, as a caller that cannot let the library allocate would.  Two calls
in a row, the second on fewer particles (fewer partial-sum rows of the shared
scratch array than the first left behind), the second on the null stream.

Tolerance: the counts and the maximum density are exact; the sums are FP64
over the same values in another order (particle order here, the library's
cell-sorted slot order in nbs_get_diagnostics), within 1e-12 relative.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from nbodysimulation_experiment_amd import _lib, scenes
from nbodysimulation_experiment_amd._lib import NbsDiagnostics
from tests.helpers import DT, gpu_for

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.load()  # the library's runtime is loaded; this resolves to it
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


def _noalloc(hip, state, rho0, stream):
    n = state["x"].shape[0]
    host = [np.ascontiguousarray(state[k], np.float32) for k in ("x", "v", "density")]
    bufs = []
    try:
        for a in host:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            bufs.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        out = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(out), ctypes.sizeof(NbsDiagnostics)) == 0
        bufs.append(out)
        rc = _lib.load().nbs_diagnostics_noalloc(bufs[0], bufs[1], bufs[2], n, ctypes.c_float(rho0), out, stream)
        assert rc == 0, _lib.load().nbs_last_error()
        assert hip.hipStreamSynchronize(stream) == 0
        rec = NbsDiagnostics()
        assert hip.hipMemcpy(ctypes.addressof(rec), out, ctypes.sizeof(rec), D2H) == 0
        return rec
    finally:
        for p in bufs:
            hip.hipFree(p)


def _compare(got: NbsDiagnostics, want: dict):
    assert got.particle_count == want["particle_count"]
    assert got.non_finite_count == want["non_finite_count"]
    assert got.max_density == want["max_density"]
    for k in ("kinetic_energy", "mean_density", "mean_abs_density_error", "rms_density_error"):
        assert got.__getattribute__(k) == pytest.approx(want[k], rel=1e-12), k
    for k in ("momentum", "centroid"):
        assert list(getattr(got, k)) == pytest.approx(list(want[k]), rel=1e-12, abs=1e-9), k


def test_noalloc_diagnostics_match_the_allocating_entry_point():
    hip = _hip()
    big = gpu_for(scenes.batched_scenes(seed=5, count=6, n_min=2048, n_max=8000))
    small = gpu_for([scenes.dam_break_4k()])
    big.Update(DT, 40)
    small.Update(DT, 25)
    rho0 = float(big.GetParams().rest_density)
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        nb = big.GetParticleCount()
        assert nb > 2 * 4096  # several partial-sum rows; the second call uses one
        _compare(_noalloc(hip, big.get_state(), rho0, stream), big.diagnostics())
        _compare(_noalloc(hip, small.get_state(), rho0, None), small.diagnostics())
    finally:
        hip.hipStreamDestroy(stream)
        big.close()
        small.close()
