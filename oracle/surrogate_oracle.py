"""CPU oracle of the learned relaxation (test infrastructure only).

The product path is the HIP library (nbodysimulation_experiment_amd/csrc/
nbs_surrogate.hip); this module computes the reference results the tests
compare it with and is never called by the product.  The network has no counterpart in the
reference: it stands in for the double-density relaxation of
NBodySimulation/demo4.cpp:239-255 (sph.h:483-495), whose displacement is its
training target.

Given a frame (predicted positions x, previous positions prev, density and
near density per particle, scene per particle) and the parameter vector:

  pairs     every (i, j != i) of one scene with r^2 < h^2, r^2 computed in
            float32 as (xj - xi)^2 + (yj - yi)^2 (the GPU's in-range test);
  features  f_ij (16, float32 arithmetic in the GPU's operation order):
            R/h, (v_j - v_i) sv, (rho_i - mr) sr, (rhon_i - mn) sn,
            (rho_j - mr) sr, (rhon_j - mn) sn, 1, q, q^2, 0 x 5
            with v = (x - prev) * (1/dt), q = 1 - sqrt(r^2) * (1/h);
  network   H_i = sum_j relu(W1 f_ij); M_i = W2 H_i + n_i b2;
            G_i = relu(W3 M_i + b3); o_i = W4 G_i + b4;
  loss      L = (1/N) sum_i |o_i - target_i / os|^2.

With ``bf16=True`` the operands the GPU feeds to its BF16 matrix cores are
rounded the same way, everything else is float64:
  forward   features, W1..W3, each pair's relu(z) (the GPU sums the BF16
            values on the matrix core), H, M;
  backward  dG and dM (the operands of dM = dG W3, dH = dM W2 and of the W3,
            W2 gradients), W2, W3, dO and G in the W4 gradient, and the W1
            gradient's masked dH; the bias gradients sum the unrounded values.
With ``bf16=False`` the whole network is float64, which is what the
finite-difference checks of the gradient use.
"""
from __future__ import annotations

import numpy as np

HID = 64
FEAT = 16
OFF_W1 = 0
OFF_W2 = OFF_W1 + HID * FEAT
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID * HID
OFF_W4 = OFF_B3 + HID
OFF_B4 = OFF_W4 + 2 * HID
NUM_WEIGHTS = OFF_B4 + 2
OFF_NORM = NUM_WEIGHTS
NUM_PARAMS = OFF_NORM + 8

f32 = np.float32


def bf16(a) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def unpack(params):
    p = np.asarray(params)
    if p.dtype != np.float64:
        p = p.astype(np.float32)
    return {
        "W1": p[OFF_W1:OFF_W2].reshape(HID, FEAT), "W2": p[OFF_W2:OFF_B2].reshape(HID, HID), "b2": p[OFF_B2:OFF_W3],
        "W3": p[OFF_W3:OFF_B3].reshape(HID, HID), "b3": p[OFF_B3:OFF_W4], "W4": p[OFF_W4:OFF_B4].reshape(2, HID),
        "b4": p[OFF_B4:NUM_WEIGHTS], "norm": p[OFF_NORM:NUM_PARAMS],
    }


def pairs(x, scene, h):
    """(i, j) of every in-range pair, i != j, same scene; float32 test r^2 < h^2."""
    from scipy.spatial import cKDTree
    x = np.asarray(x, np.float32)
    scene = np.asarray(scene)
    h2 = f32(h) * f32(h)
    out_i, out_j = [], []
    for s in np.unique(scene):
        idx = np.nonzero(scene == s)[0]
        tree = cKDTree(x[idx].astype(np.float64))
        cand = tree.query_pairs(float(h) * (1.0 + 1e-5), output_type="ndarray")
        if cand.size == 0:
            continue
        a, b = idx[cand[:, 0]], idx[cand[:, 1]]
        rx = x[b, 0] - x[a, 0]
        ry = x[b, 1] - x[a, 1]
        r2 = rx * rx + ry * ry
        keep = r2 < h2
        a, b = a[keep], b[keep]
        out_i += [a, b]
        out_j += [b, a]
    if not out_i:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    return np.concatenate(out_i), np.concatenate(out_j)


def features(frame, pi, pj, h, dt, norm):
    """16 float32 features per pair in the GPU's operation order."""
    x = np.asarray(frame["x"], np.float32)
    prev = np.asarray(frame["prev"], np.float32)
    dens = np.asarray(frame["density"], np.float32)
    invdt = f32(1.0) / f32(dt)
    invh = f32(1.0) / f32(h)
    v = (x - prev) * invdt
    sv, mr, sr, mn, sn = (f32(norm[k]) for k in range(5))
    rx = x[pj, 0] - x[pi, 0]
    ry = x[pj, 1] - x[pi, 1]
    r = np.sqrt(rx * rx + ry * ry)
    q = f32(1.0) - r * invh
    f = np.zeros((pi.size, FEAT), np.float32)
    f[:, 0] = rx * invh
    f[:, 1] = ry * invh
    f[:, 2] = (v[pj, 0] - v[pi, 0]) * sv
    f[:, 3] = (v[pj, 1] - v[pi, 1]) * sv
    f[:, 4] = (dens[pi, 0] - mr) * sr
    f[:, 5] = (dens[pi, 1] - mn) * sn
    f[:, 6] = (dens[pj, 0] - mr) * sr
    f[:, 7] = (dens[pj, 1] - mn) * sn
    f[:, 8] = 1.0
    f[:, 9] = q
    f[:, 10] = q * q
    return f


def forward(params, frame, h, dt, bf16_ops=True):
    """Network outputs o (n x 2, float64) and the activations the backward
    pass needs."""
    w = unpack(params)
    n = np.asarray(frame["x"]).shape[0]
    pi, pj = pairs(frame["x"], frame.get("scene", np.zeros(n, np.int32)), h)
    F = features(frame, pi, pj, h, dt, w["norm"])
    rnd = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    Fb = rnd(F).astype(np.float64)
    W1 = rnd(w["W1"]).astype(np.float64)
    W2 = rnd(w["W2"]).astype(np.float64)
    W3 = rnd(w["W3"]).astype(np.float64)
    Z = Fb @ W1.T                                        # pairs x 64
    R = np.maximum(Z, 0.0)
    if bf16_ops:
        R = bf16(R.astype(np.float32)).astype(np.float64)
    H = np.zeros((n, HID))
    np.add.at(H, pi, R)
    cnt = np.bincount(pi, minlength=n).astype(np.float64)
    Hin = rnd(H.astype(np.float32)).astype(np.float64) if bf16_ops else H
    M = Hin @ W2.T + cnt[:, None] * w["b2"].astype(np.float64)
    Min = rnd(M.astype(np.float32)).astype(np.float64) if bf16_ops else M
    G = np.maximum(Min @ W3.T + w["b3"].astype(np.float64), 0.0)
    o = G @ w["W4"].astype(np.float64).T + w["b4"].astype(np.float64)
    return o, {"pi": pi, "pj": pj, "F": Fb, "Z": Z, "Hin": Hin, "Min": Min, "G": G, "cnt": cnt, "bf16": bf16_ops}


def forward_rows(params, frame, h, dt, rows, bf16_ops=True):
    """Network outputs o (len(rows) x 2, float64) of the particles ``rows``
    only, with their neighbours taken from the whole frame: the forward pass
    of a slice of a frame too large for the all-pairs matrices."""
    w = unpack(params)
    rows = np.asarray(rows, np.int64)
    n = np.asarray(frame["x"]).shape[0]
    pi, pj = pairs(frame["x"], frame.get("scene", np.zeros(n, np.int32)), h)
    local = np.full(n, -1, np.int64)
    local[rows] = np.arange(rows.size)
    keep = local[pi] >= 0
    pi, pj = pi[keep], pj[keep]
    F = features(frame, pi, pj, h, dt, w["norm"])
    rnd = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    Z = rnd(F).astype(np.float64) @ rnd(w["W1"]).astype(np.float64).T
    R = np.maximum(Z, 0.0)
    if bf16_ops:
        R = bf16(R.astype(np.float32)).astype(np.float64)
    li = local[pi]
    H = np.zeros((rows.size, HID))
    np.add.at(H, li, R)
    cnt = np.bincount(li, minlength=rows.size).astype(np.float64)
    Hin = rnd(H.astype(np.float32)).astype(np.float64) if bf16_ops else H
    M = Hin @ rnd(w["W2"]).astype(np.float64).T + cnt[:, None] * w["b2"].astype(np.float64)
    Min = rnd(M.astype(np.float32)).astype(np.float64) if bf16_ops else M
    G = np.maximum(Min @ rnd(w["W3"]).astype(np.float64).T + w["b3"].astype(np.float64), 0.0)
    return G @ w["W4"].astype(np.float64).T + w["b4"].astype(np.float64)


def loss_and_grad(params, frame, h, dt, bf16_ops=True):
    """Loss and the gradient of the NUM_WEIGHTS trained parameters (float64),
    with the GPU's BF16 operand roundings when bf16_ops (module docstring)."""
    w = unpack(params)
    o, c = forward(params, frame, h, dt, bf16_ops)
    n = o.shape[0]
    os_ = float(w["norm"][5])
    t = np.asarray(frame["target"], np.float64) / os_
    e = o - t
    L = float((e * e).sum() / n)
    dO = 2.0 * e / n
    rnd = (lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)) if bf16_ops else \
        (lambda a: np.asarray(a, np.float64))
    W4 = w["W4"].astype(np.float64)
    W3 = rnd(w["W3"])
    W2 = rnd(w["W2"])
    dGpre = (dO @ W4) * (c["G"] > 0.0)
    dGr = rnd(dGpre)
    dM = dGr @ W3
    dMr = rnd(dM)
    dH = dMr @ W2
    g = np.zeros(NUM_WEIGHTS)
    g[OFF_W4:OFF_B4] = (rnd(dO).T @ rnd(c["G"])).reshape(-1)
    g[OFF_B4:NUM_WEIGHTS] = dO.sum(0)
    g[OFF_W3:OFF_B3] = (dGr.T @ c["Min"]).reshape(-1)
    g[OFF_B3:OFF_W4] = dGpre.sum(0)
    g[OFF_W2:OFF_B2] = (dMr.T @ c["Hin"]).reshape(-1)
    g[OFF_B2:OFF_W3] = (dM * c["cnt"][:, None]).sum(0)
    A = dH[c["pi"]] * (c["Z"] > 0.0)                     # pairs x 64
    if c["bf16"]:
        A = bf16(A.astype(np.float32)).astype(np.float64)
    g[OFF_W1:OFF_W2] = (A.T @ c["F"]).reshape(-1)
    return L, g, o
