"""CPU-only checks of the C ABI library: it loads, exports every declared
symbol, and host-side argument checks fail cleanly without a GPU."""
import ctypes
import os
import subprocess

import pytest

from nbodysimulation_experiment_amd import _lib


def test_library_exports_every_declared_symbol():
    declared = _lib.declared_symbols()
    assert len(declared) >= 30
    L = _lib.load()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    # and the binding covers every declaration
    bound = set(_lib.SIGNATURES) | {"nbs_status_string", "nbs_last_error"}
    assert set(declared) == bound, set(declared) ^ bound


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for s in _lib.declared_symbols():
        assert s in names, s


def test_status_strings():
    L = _lib.load()
    for code, name in _lib.STATUS_NAMES.items():
        assert L.nbs_status_string(code).decode() == name


def test_null_arguments_are_rejected_without_a_device():
    L = _lib.load()
    out = ctypes.c_void_p()
    assert L.nbs_create(None, None, 0, ctypes.byref(out)) == 1
    assert L.nbs_update(None, ctypes.c_float(0.005), 1) == 1
    assert L.nbs_destroy(None) == 0
    assert L.nbs_get_particle_count(None, None) == 1


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "nbs_sph.h"\nint main(void){nbs_stats s; (void)s; return (int)sizeof(nbs_params);}\n')
    inc = os.path.join(_lib.ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-c", str(src), "-o", str(tmp_path / "t.o")],
                   check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++", "-I", inc, "-c", str(src), "-o",
                    str(tmp_path / "t2.o")], check=True)


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_lib.NbsParams) == 40
    assert ctypes.sizeof(_lib.NbsDomain) == 12
    assert ctypes.sizeof(_lib.NbsStats) == 4 * 8 + 9 * 4 + 4  # padded to 8
    assert ctypes.sizeof(_lib.NbsDiagnostics) == 8 + 16 + 8 * 5 + 16 + 8
    assert ctypes.sizeof(_lib.NbsTuning) == 24
    assert ctypes.sizeof(_lib.NbsSurrogateReport) == 5 * 8 + 3 * 8 + 8 + 8 + 8


def test_noalloc_diagnostics_argument_checks_without_a_device():
    """nbs_diagnostics_noalloc rejects null arrays and sizes beyond its fixed
    scratch array before it touches the GPU."""
    L = _lib.load()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: the checks come first
    assert L.nbs_diagnostics_noalloc(fake, fake, fake, 10, ctypes.c_float(10.0), None, None) == 1
    assert L.nbs_diagnostics_noalloc(None, fake, fake, 10, ctypes.c_float(10.0), fake, None) == 1
    too_many = 16777216 + 1
    assert L.nbs_diagnostics_noalloc(fake, fake, fake, too_many, ctypes.c_float(10.0), fake, None) == 4
    assert "16,777,216" in L.nbs_last_error().decode()
    assert L.nbs_diagnostics_noalloc(fake, fake, fake, 10, ctypes.c_float(float("nan")), fake, None) == 1
