"""MI355X-native data-oriented SPH step (the reference's "Demo 4" path).

The compute runs in libnbs_sph.so (hand-written HIP kernels for gfx950) behind
the C ABI in include/nbs_sph.h.  This package holds the Python mirror of the
simulation interface (``sim.ParticleSimulation``), the seeded workload builders
(``scenes``) and the build script (``build``).
"""
from .scenes import SPHParams, SceneSpec, batched_scenes, dam_break_1m, dam_break_4k  # noqa: F401

__all__ = ["SPHParams", "SceneSpec", "batched_scenes", "dam_break_1m", "dam_break_4k", "ParticleSimulation"]


def __getattr__(name):
    if name == "ParticleSimulation":
        from .sim import ParticleSimulation
        return ParticleSimulation
    raise AttributeError(name)
