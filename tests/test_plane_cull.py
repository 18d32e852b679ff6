"""The wave-level plane cull of the collision pass (body_candidates in
csrc/nbs_device.hpp), in FP32 on the CPU: the projection of the reference's
plane test (sph.h:515-519: pp = n d, proj = (x - pp) . n, pushed when
proj <= R) computed at the box corner on the plane's side is at most the
projection of every position in the box, so a plane whose corner projection
exceeds R moves no position of the wave.  Each step is a rounded subtraction,
product or sum, monotone in the position for a fixed normal; this checks the
claim on random normals (signs, zeros and axis-aligned walls included), boxes
and positions, with numpy's IEEE FP32 arithmetic."""
from __future__ import annotations

import numpy as np

f32 = np.float32


def proj(x, y, nx, ny, d):
    # solve_plane's operations in its order, each rounded to FP32
    ppx, ppy = f32(nx * d), f32(ny * d)
    dx, dy = (x - ppx).astype(f32), (y - ppy).astype(f32)
    return (dx * nx).astype(f32) + (dy * ny).astype(f32)


def corner(box, nx, ny):
    x0, y0, x1, y1 = box
    return (x0 if nx >= 0 else x1), (y0 if ny >= 0 else y1)


def test_corner_projection_bounds_every_position_in_the_box():
    rng = np.random.default_rng(7)
    normals = [(f32(1), f32(0)), (f32(-1), f32(0)), (f32(0), f32(1)), (f32(0), f32(-1)),
               (f32(-0.0), f32(1)), (f32(1), f32(-0.0))]
    for _ in range(300):
        a = rng.uniform(0, 2 * np.pi)
        normals.append((f32(np.cos(a)), f32(np.sin(a))))
    culled = 0
    for nx, ny in normals:
        for _ in range(20):
            scale = 10.0 ** rng.uniform(-3, 3)
            cx0, cy0 = rng.uniform(-scale, scale, 2)
            w, h = rng.uniform(0, scale, 2) * rng.choice([1e-6, 1e-3, 1.0])
            x = f32(cx0) + (rng.uniform(0, 1, 64) * w).astype(f32)
            y = f32(cy0) + (rng.uniform(0, 1, 64) * h).astype(f32)
            x, y = x.astype(f32), y.astype(f32)
            box = (x.min(), y.min(), x.max(), y.max())
            d = f32(rng.uniform(-2 * scale, 2 * scale))
            px, py = corner(box, nx, ny)
            pc = proj(np.array([px], f32), np.array([py], f32), nx, ny, d)[0]
            pa = proj(x, y, nx, ny, d)
            assert pc <= pa.min(), (nx, ny, d, box)
            # a cull threshold just below the corner value: culling it is right
            R = f32(np.nextafter(pc, f32(-np.inf)))
            if pc > R:
                culled += 1
                assert (pa > R).all()
    assert culled > 1000


def test_nan_and_empty_boxes_keep_the_plane():
    # a NaN corner gives a NaN projection, and NaN > R is false: the plane stays
    p = proj(np.array([np.nan], f32), np.array([0.0], f32), f32(1), f32(0), f32(0))[0]
    assert not (p > f32(0.5))
