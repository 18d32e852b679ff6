"""Builds libnbs_sph.so in-tree for gfx950 with hipcc.

Flags:
  -ffp-contract=off                       no a*b+c fusion: FP32 results are those
                                          of the expressions as written
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE-correct FP32 sqrt and divide
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                          no packed FP32 VALU instructions
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnbs_sph.so")
ARCH = os.environ.get("NBS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", f"-I{os.path.join(ROOT, 'include')}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> str:
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-o", tmp, *sources]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, LIB)
    return LIB


DRIVER = os.path.join(ROOT, "tests", "cpp", "demo5_driver")
BENCHMARK = os.path.join(HERE, "nbs_benchmark")
HOST = os.path.join(HERE, "host")
HOST_SOURCES = [os.path.join(HOST, f) for f in ("demo5.cpp", "scenarios.cpp", "benchmark.cpp")]


def _build_host_program(main_source: str, out: str, rpath_rel: str, verbose: bool) -> str:
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{HOST}", *HOST_SOURCES, main_source, f"-L{HERE}", "-lnbs_sph",
           f"-Wl,-rpath,{HERE}", f"-Wl,-rpath,$ORIGIN/{rpath_rel}", "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"host build of {os.path.basename(out)} failed ({r.returncode})")
    return out


def build_host_driver(verbose: bool = False) -> str:
    """Builds the C++ BaseSimulation wrapper (host/demo5.cpp) into the test
    driver tests/cpp/demo5_driver, linked against libnbs_sph.so."""
    return _build_host_program(DRIVER + ".cpp", DRIVER, "../../nbodysimulation_experiment_amd", verbose)


def build_benchmark(verbose: bool = False) -> str:
    """Builds the scenario table and built-in benchmark front end
    (host/nbs_benchmark.cpp) into nbodysimulation_experiment_amd/nbs_benchmark."""
    return _build_host_program(os.path.join(HOST, "nbs_benchmark.cpp"), BENCHMARK, ".", verbose)


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_host_driver(verbose=True))
    print(build_benchmark(verbose=True))
