"""Seeded scene builders for the SPH step's workloads.

Particle placement follows the reference's ``AddVolume``
(NBodySimulation/demo4.cpp:169-181) in float32 with the same operand order:

    offset = (countX*spacing, countY*spacing) * 0.5
    p      = (ix, iy) * spacing
    p      = (spacing*0.5) + p
    p      = (center - offset) + p
    p      = jitter + p

The reference draws its jitter from unseeded ``rand()`` (vecmath.h:317-322);
here the jitter is per-axis uniform in [-0.01h, +0.01h) from SplitMix64 so that
every consumer (the GPU library, the oracle, the bench) sees identical inputs.
Planes use the scenario loader's convention d = n . p (app.cpp:493-494) and
the Dambreak wall order bottom, top, left, right (sph.h:323-326).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "SPHParams",
    "SceneSpec",
    "splitmix64_uniform",
    "volume_positions",
    "wall_planes",
    "dam_break_4k",
    "dam_break_1m",
    "batched_scenes",
    "training_scenes",
]

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def _splitmix64(seed: int, counters: np.ndarray) -> np.ndarray:
    """SplitMix64 output for stream positions ``counters`` (0-based) of ``seed``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (counters.astype(np.uint64) + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def splitmix64_uniform(seed: int, n: int, start: int = 0) -> np.ndarray:
    """n float32 values uniform in [-1, 1) with 24-bit resolution (exact in float32)."""
    z = _splitmix64(seed, np.arange(start, start + n, dtype=np.uint64))
    k = (z >> np.uint64(40)).astype(np.float32)  # < 2^24, exact
    return k * np.float32(2.0 ** -23) - np.float32(1.0)


@dataclass
class SPHParams:
    """Runtime parameters (NBodySimulation/sph.h:77-123) plus the two collision
    constants the reference keeps at compile time (sph.h:35-38, 53-54)."""

    h: float = 1.0
    rho0: float = 10.0
    k: float = 0.004
    k_near: float = 0.01
    sigma: float = 0.0
    beta: float = 0.3
    collision_radius: float = 1.0 / 6.0
    collision_margin: float = 1.0 / 30.0

    def as_tuple(self):
        return (self.h, self.rho0, self.k, self.k_near, self.sigma, self.beta,
                self.collision_radius, self.collision_margin)


@dataclass
class SceneSpec:
    """One independent scene: an origin-centred box with its own uniform grid."""

    width: float
    height: float
    cell_size: float
    positions: np.ndarray  # (n, 2) float32
    planes: list = field(default_factory=list)   # (nx, ny, d)
    circles: list = field(default_factory=list)  # (cx, cy, r)

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def volume_positions(center, count_x: int, count_y: int, spacing: float, h: float,
                     seed: int, limit: int | None = None) -> np.ndarray:
    """Particle positions of ``AddVolume`` (demo4.cpp:169-181) with seeded jitter.

    Particles are generated row by row (y outer, x inner) like the reference;
    ``limit`` stops after that many particles (a partial last row).
    """
    f32 = np.float32
    s = f32(spacing)
    cx, cy = f32(center[0]), f32(center[1])
    off_x = f32(f32(count_x) * s) * f32(0.5)
    off_y = f32(f32(count_y) * s) * f32(0.5)
    iy, ix = np.meshgrid(np.arange(count_y, dtype=np.float32), np.arange(count_x, dtype=np.float32), indexing="ij")
    ix = ix.reshape(-1)
    iy = iy.reshape(-1)
    if limit is not None:
        ix = ix[:limit]
        iy = iy[:limit]
    n = ix.shape[0]
    px = ix * s
    py = iy * s
    half = s * f32(0.5)
    px = half + px
    py = half + py
    px = f32(cx - off_x) + px
    py = f32(cy - off_y) + py
    u = splitmix64_uniform(seed, 2 * n)
    jscale = f32(h)
    jx = (u[0::2] * jscale) * f32(0.01)
    jy = (u[1::2] * jscale) * f32(0.01)
    px = jx + px
    py = jy + py
    return np.stack([px, py], axis=1).astype(np.float32)


def wall_planes(width: float, height: float):
    """Four inward-facing walls of an origin-centred box (sph.h:323-326 order),
    with d = n . p as the scenario loader computes it (app.cpp:493-494)."""
    f32 = np.float32
    hw = f32(width) * f32(0.5)
    hh = f32(height) * f32(0.5)
    walls = [((0.0, 1.0), (0.0, -hh)), ((0.0, -1.0), (0.0, hh)), ((1.0, 0.0), (-hw, 0.0)), ((-1.0, 0.0), (hw, 0.0))]
    out = []
    for (nx, ny), (px, py) in walls:
        d = f32(f32(nx) * f32(px)) + f32(f32(ny) * f32(py))
        out.append((float(nx), float(ny), float(d)))
    return out


def dam_break_4k(seed: int = 0) -> SceneSpec:
    """4,096 particles (64x64 at 0.5h) resting on the floor against the left
    wall of an 80h x 60h box; 4 walls."""
    pos = volume_positions((-24.0, -14.0), 64, 64, 0.5, 1.0, seed)
    return SceneSpec(80.0, 60.0, 1.0, pos, planes=wall_planes(80.0, 60.0))


def _place_circles(rng_seed: int, count: int, width: float, height: float, block_lo, block_hi,
                   rmin: float, rmax: float, gap: float):
    """Seeded rejection sampling of non-overlapping circles that stay clear of
    the walls and of the initial particle block."""
    circles = []
    hw, hh = width * 0.5, height * 0.5
    draw = 0
    while len(circles) < count:
        if draw > 3_000_000:
            raise RuntimeError("circle placement did not converge")
        u = splitmix64_uniform(rng_seed, 3, start=draw)
        draw += 3
        r = rmin + (float(u[0]) * 0.5 + 0.5) * (rmax - rmin)
        cx = (float(u[1]) * 0.5 + 0.5) * (width - 2 * (r + gap)) - hw + r + gap
        cy = (float(u[2]) * 0.5 + 0.5) * (height - 2 * (r + gap)) - hh + r + gap
        # clear of the block's bounding box
        qx = min(max(cx, block_lo[0]), block_hi[0])
        qy = min(max(cy, block_lo[1]), block_hi[1])
        if (cx - qx) ** 2 + (cy - qy) ** 2 < (r + gap) ** 2:
            continue
        if any((cx - ox) ** 2 + (cy - oy) ** 2 < (r + orr + gap) ** 2 for ox, oy, orr in circles):
            continue
        circles.append((float(np.float32(cx)), float(np.float32(cy)), float(np.float32(r))))
    return circles


def dam_break_1m(seed: int = 1, n_side: int = 1024, n_circles: int = 64) -> SceneSpec:
    """1,048,576 particles (1024x1024 at 0.5h) against the left wall and floor of
    a 600h x 600h box, 4 walls plus 64 seeded non-overlapping circles with radius
    U[2h, 8h] placed clear of the initial block."""
    width = height = 600.0
    side = n_side * 0.5
    center = (-300.0 + side * 0.5, -300.0 + side * 0.5)
    pos = volume_positions(center, n_side, n_side, 0.5, 1.0, seed)
    lo = (center[0] - side * 0.5 - 0.5, center[1] - side * 0.5 - 0.5)
    hi = (center[0] + side * 0.5 + 0.5, center[1] + side * 0.5 + 0.5)
    circles = _place_circles(seed * 7919 + 17, n_circles, width, height, lo, hi, 2.0, 8.0, 1.0)
    return SceneSpec(width, height, 1.0, pos, planes=wall_planes(width, height), circles=circles)


def dam_break_4m(seed: int = 4, n_side: int = 2048) -> SceneSpec:
    """4,194,304 particles (2048x2048 at 0.5h) against the left wall and floor of
    a 1536h x 1536h box with 4 walls."""
    width = height = 1536.0
    side = n_side * 0.5
    center = (-width * 0.5 + side * 0.5, -height * 0.5 + side * 0.5)
    pos = volume_positions(center, n_side, n_side, 0.5, 1.0, seed)
    return SceneSpec(width, height, 1.0, pos, planes=wall_planes(width, height))


def batched_scenes(seed: int = 2, count: int = 256, n_min: int = 2048, n_max: int = 65536,
                   box_lo: float = 1.5, box_hi: float = 3.0) -> list[SceneSpec]:
    """Independent scenes with particle counts log-uniform in [n_min, n_max];
    each block (0.5h spacing, ceil(sqrt(n)) columns) rests in the lower-left
    corner of a square box whose side is U[box_lo, box_hi] x the block side,
    rounded up to whole cells."""
    u = splitmix64_uniform(seed, 2 * count)
    scenes = []
    for s in range(count):
        t = float(u[2 * s]) * 0.5 + 0.5
        n = int(round(math.exp(math.log(n_min) + t * (math.log(n_max) - math.log(n_min)))))
        n = max(n_min, min(n_max, n))
        cols = int(math.ceil(math.sqrt(n)))
        rows = int(math.ceil(n / cols))
        block = max(cols, rows) * 0.5
        scale = box_lo + (float(u[2 * s + 1]) * 0.5 + 0.5) * (box_hi - box_lo)
        side = float(math.ceil(block * scale))
        if int(side) % 2:
            side += 1.0
        center = (-side * 0.5 + cols * 0.25, -side * 0.5 + rows * 0.25)
        pos = volume_positions(center, cols, rows, 0.5, 1.0, seed * 1_000_003 + s, limit=n)
        scenes.append(SceneSpec(side, side, 1.0, pos, planes=wall_planes(side, side)))
    return scenes


def training_scenes(seed: int = 3, count: int = 200, n_side: int = 128, box_lo: float = 1.5,
                    box_hi: float = 3.0, tall: int = 0, tall_shape: tuple[int, int] = (64, 1024),
                    tall_box: tuple[float, float] = (160.0, 560.0)) -> list[SceneSpec]:
    """Frames for the learned relaxation: ``count`` dam breaks of n_side^2
    particles (16,384 by default, 0.5h spacing) resting in the lower-left
    corner of a box of width and height U[box_lo, box_hi] x the block side
    each (rounded up to an even number of cells), 4 walls; then ``tall``
    narrow columns of tall_shape particles (64 x 1,024 by default: 32h wide,
    512h tall, as tall as the 1M dam break's column, whose bottom rows reach
    densities that shallow columns never do) in the lower-left corner of a
    tall_box box, 4 walls."""
    u = splitmix64_uniform(seed, 2 * count)
    block = n_side * 0.5
    scenes = []
    for s in range(count):
        dims = []
        for k in range(2):
            side = float(math.ceil(block * (box_lo + (float(u[2 * s + k]) * 0.5 + 0.5) * (box_hi - box_lo))))
            if int(side) % 2:
                side += 1.0
            dims.append(side)
        w, h = dims
        center = (-w * 0.5 + block * 0.5, -h * 0.5 + block * 0.5)
        pos = volume_positions(center, n_side, n_side, 0.5, 1.0, seed * 1_000_003 + s)
        scenes.append(SceneSpec(w, h, 1.0, pos, planes=wall_planes(w, h)))
    cx, cy = tall_shape
    w, h = tall_box
    for s in range(tall):
        center = (-w * 0.5 + cx * 0.25, -h * 0.5 + cy * 0.25)
        pos = volume_positions(center, cx, cy, 0.5, 1.0, seed * 1_000_003 + count + s)
        scenes.append(SceneSpec(w, h, 1.0, pos, planes=wall_planes(w, h)))
    return scenes
