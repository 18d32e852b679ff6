"""Writes tests/golden/kat.json: known answers fixed by mathematics, not by any
implementation.  Every input is exactly representable in FP32 so the
expected values are those of the real-number formulas.

The reference ships no tests, fixtures or golden data (its repository has no
test directory), so every vector here is a closed form:

  cell_keys        integer floor / clamp / y-major key of positions whose sum
                   with the half extent is exact in FP32
  lattice_density  density and near density of a particle on an unjittered
                   square lattice, summed exactly over lattice points with
                   r < h (strict cutoff)
  pair_density     isolated pair at distance r: (1 + q^2, 1 + q^3), q = 1 - r/h
  pair_delta       D = dt^2 (P q + Pn q^2) n for one pair
  pair_viscosity   f = (1 - r/h)(sigma u + beta u^2) n if u > 0 else 0
  free_fall        semi-implicit Euler: v_n = n g dt, x_n = x_0 + g dt^2 n(n+1)/2

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def cell_keys():
    # 80 x 60 grid, cell size 1, domain centred on the origin (half extents 40, 30)
    W, H, nx, ny = 80.0, 60.0, 80, 60
    xs = [-45.0, -40.0, -39.75, -39.0, -0.25, 0.0, 0.5, 17.25, 38.999755859375, 39.0, 39.5, 40.0, 41.0, 1e6, -1e6]
    ys = [-31.0, -30.0, -29.5, 0.0, 12.75, 29.0, 29.75, 30.0, 77.0]
    out = []
    for x in xs:
        for y in ys:
            cx = min(max(math.floor(x + W / 2), 0), nx - 1)
            cy = min(max(math.floor(y + H / 2), 0), ny - 1)
            out.append({"pos": [x, y], "cx": cx, "cy": cy, "key": cy * nx + cx})
    return {"width": W, "height": H, "cell_size": 1.0, "cases": out}


def lattice_density():
    cases = []
    for s in (0.5, 0.375, 0.3125, 0.25):
        h = 1.0
        m = int(math.ceil(h / s)) + 1
        d0 = d1 = 0.0
        for a in range(-m, m + 1):
            for b in range(-m, m + 1):
                r2 = (a * a + b * b) * s * s
                if r2 < h * h:
                    q = 1.0 - math.sqrt(r2) / h
                    d0 += q * q
                    d1 += q * q * q
        cases.append({"spacing": s, "h": h, "density": d0, "near_density": d1, "reach": m})
    return cases


def pair_density():
    out = []
    for r in (0.0, 0.125, 0.25, 0.5, 0.75, 0.9921875, 1.0, 1.5):
        q = 1.0 - r if r < 1.0 else None
        out.append({"r": r, "density": 1.0 + (q * q if q is not None else 0.0),
                    "near_density": 1.0 + (q ** 3 if q is not None else 0.0)})
    return out


def pair_delta():
    out = []
    dt = 0.0078125
    for (dx, dy), (p0, p1) in [((0.375, 0.5), (0.25, 0.0625)), ((-0.625, 0.0), (-0.03125, 0.015625)),
                               ((0.0, 0.875), (1.0, 2.0)), ((0.0, 0.0), (1.0, 1.0)), ((0.75, 0.75), (1.0, 1.0))]:
        r = math.hypot(dx, dy)
        if r < 1.0 and r > 0.0:
            q = 1.0 - r
            d = dt * dt * (p0 * q + p1 * q * q)
            D = [d * dx / r, d * dy / r]
        else:
            D = [0.0, 0.0]
        out.append({"xi": [0.0, 0.0], "xj": [dx, dy], "pressure": [p0, p1], "dt": dt, "delta": D})
    return out


def pair_viscosity():
    out = []
    sigma, beta = 0.5, 0.25
    for (dx, dy), (ux, uy) in [((0.375, 0.5), (1.0, 0.0)), ((0.375, 0.5), (-1.0, 0.0)), ((0.5, 0.0), (2.0, 1.0)),
                               ((0.0, 0.0), (1.0, 1.0)), ((0.875, 0.0), (0.25, 0.0)), ((1.25, 0.0), (3.0, 0.0))]:
        r = math.hypot(dx, dy)
        f = [0.0, 0.0]
        if 0.0 < r < 1.0:
            n = (dx / r, dy / r)
            u = ux * n[0] + uy * n[1]  # vi - vj with vj = 0
            if u > 0:
                mag = (1.0 - r) * (sigma * u + beta * u * u)
                f = [mag * n[0], mag * n[1]]
        out.append({"xi": [0.0, 0.0], "xj": [dx, dy], "vi": [ux, uy], "vj": [0.0, 0.0], "sigma": sigma,
                    "beta": beta, "force": f})
    return out


def free_fall():
    g, dt, x0 = -9.75, 0.0078125, 10.0
    return [{"steps": n, "g": g, "dt": dt, "y0": x0, "y": x0 + g * dt * dt * n * (n + 1) / 2.0, "v": n * g * dt}
            for n in (1, 2, 10, 100)]


def main():
    data = {"cell_keys": cell_keys(), "lattice_density": lattice_density(), "pair_density": pair_density(),
            "pair_delta": pair_delta(), "pair_viscosity": pair_viscosity(), "free_fall": free_fall()}
    path = os.path.join(HERE, "kat.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
