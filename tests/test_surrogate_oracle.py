"""Pins of the learned relaxation's CPU oracle (oracle/surrogate_oracle.py).

* its gradient against central finite differences of its own float64 loss;
* its forward pass against a brute-force double loop for networks whose
  weights select single features, so the output is a closed-form pair sum;
* its BF16 rounding against hand-computed ties.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import surrogate_oracle as so

H = 1.0
DT = 0.005


def random_frame(n=48, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.6, 1.6, (n, 2)).astype(np.float32)
    prev = (x - rng.normal(0, 0.01, (n, 2))).astype(np.float32)
    dens = np.stack([rng.uniform(1, 6, n), rng.uniform(0.5, 3, n)], 1).astype(np.float32)
    target = rng.normal(0, 1e-3, (n, 2)).astype(np.float32)
    return {"x": x, "prev": prev, "density": dens, "target": target, "scene": np.zeros(n, np.int32)}


def random_params(seed=1, dtype=np.float64):
    rng = np.random.default_rng(seed)
    p = np.zeros(so.NUM_PARAMS, dtype)
    p[:so.NUM_WEIGHTS] = rng.normal(0, 0.3, so.NUM_WEIGHTS)
    p[so.OFF_W2:so.OFF_B2] *= 0.2
    p[so.OFF_NORM:so.NUM_PARAMS] = [2.0, 3.0, 0.8, 1.5, 1.1, 1e-3, 0.0, 0.0]
    return p


def test_gradient_matches_finite_differences():
    frame = random_frame()
    p = random_params()
    L, g, _ = so.loss_and_grad(p, frame, H, DT, bf16_ops=False)
    assert L > 0
    rng = np.random.default_rng(7)
    blocks = [(so.OFF_W1, so.OFF_W2), (so.OFF_W2, so.OFF_B2), (so.OFF_B2, so.OFF_W3), (so.OFF_W3, so.OFF_B3),
              (so.OFF_B3, so.OFF_W4), (so.OFF_W4, so.OFF_B4), (so.OFF_B4, so.NUM_WEIGHTS)]
    eps = 1e-6
    checked = 0
    for lo, hi in blocks:
        for k in rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False):
            if lo == so.OFF_W1 and (k - so.OFF_W1) % so.FEAT >= 11:
                continue  # features 11..15 are always zero: their gradient is exactly 0
            pp, pm = p.copy(), p.copy()
            pp[k] += eps
            pm[k] -= eps
            fd = (so.loss_and_grad(pp, frame, H, DT, False)[0] - so.loss_and_grad(pm, frame, H, DT, False)[0]) / (2 * eps)
            assert math.isclose(g[k], fd, rel_tol=1e-4, abs_tol=1e-9 * max(1.0, abs(L))), (k, g[k], fd)
            checked += 1
    assert checked > 60
    # the always-zero feature columns get no gradient
    w1 = g[so.OFF_W1:so.OFF_W2].reshape(so.HID, so.FEAT)
    assert np.all(w1[:, 11:] == 0.0)


def brute_pairs(x, h):
    n = x.shape[0]
    out = []
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            rx = np.float32(x[j, 0] - x[i, 0])
            ry = np.float32(x[j, 1] - x[i, 1])
            if np.float32(rx * rx + ry * ry) < np.float32(h) * np.float32(h):
                out.append((i, j, float(rx), float(ry)))
    return out


def selection_params(pick0, pick1):
    """W1 row k = e_k (k < 16); W2 = W3 = I, zero biases: G = H = sum_j relu(f_ij);
    W4 picks hidden units pick0 and pick1."""
    p = np.zeros(so.NUM_PARAMS, np.float64)
    w1 = np.zeros((so.HID, so.FEAT))
    w1[np.arange(so.FEAT), np.arange(so.FEAT)] = 1.0
    p[so.OFF_W1:so.OFF_W2] = w1.reshape(-1)
    p[so.OFF_W2:so.OFF_B2] = np.eye(so.HID).reshape(-1)
    p[so.OFF_W3:so.OFF_B3] = np.eye(so.HID).reshape(-1)
    w4 = np.zeros((2, so.HID))
    w4[0, pick0] = 1.0
    w4[1, pick1] = 1.0
    p[so.OFF_W4:so.OFF_B4] = w4.reshape(-1)
    p[so.OFF_NORM:so.NUM_PARAMS] = [1, 0, 1, 0, 1, 1, 0, 0]
    return p


def test_forward_closed_form_pair_sums():
    frame = random_frame(n=40, seed=3)
    x = frame["x"]
    pairs = brute_pairs(x, H)
    n = x.shape[0]
    # o = (sum_j relu(Rx/h), sum_j relu(Ry/h))
    o, _ = so.forward(selection_params(0, 1), frame, H, DT, bf16_ops=False)
    want = np.zeros((n, 2))
    for i, j, rx, ry in pairs:
        want[i, 0] += max(rx, 0.0)
        want[i, 1] += max(ry, 0.0)
    np.testing.assert_allclose(o, want, rtol=1e-6, atol=1e-9)
    # o = (number of pairs via the constant feature, sum_j q_ij)
    o, _ = so.forward(selection_params(8, 9), frame, H, DT, bf16_ops=False)
    cnt = np.zeros(n)
    qs = np.zeros(n)
    for i, j, rx, ry in pairs:
        cnt[i] += 1
        qs[i] += 1.0 - math.sqrt(rx * rx + ry * ry) / H
    np.testing.assert_allclose(o[:, 0], cnt, rtol=0, atol=0)
    np.testing.assert_allclose(o[:, 1], qs, rtol=1e-5, atol=1e-6)
    # o = (sum_j relu(rho_j - mr), n_i * relu(rho_i))   (selection of features 6 and 4)
    o, _ = so.forward(selection_params(6, 4), frame, H, DT, bf16_ops=False)
    rho = frame["density"][:, 0].astype(np.float64)
    want0 = np.zeros(n)
    for i, j, rx, ry in pairs:
        want0[i] += rho[j]
    np.testing.assert_allclose(o[:, 0], want0, rtol=1e-6)
    np.testing.assert_allclose(o[:, 1], cnt * rho, rtol=1e-6)


def test_scenes_do_not_interact():
    frame = random_frame(n=40, seed=4)
    frame["scene"] = np.arange(40, dtype=np.int32) % 2
    pi, pj = so.pairs(frame["x"], frame["scene"], H)
    assert pi.size > 0 and np.all(frame["scene"][pi] == frame["scene"][pj])
    want = {(i, j) for i, j, _, _ in brute_pairs(frame["x"], H) if frame["scene"][i] == frame["scene"][j]}
    assert set(zip(pi.tolist(), pj.tolist())) == want


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -9), 1.0 + 2 ** -8 + 2 ** -20, np.finfo(np.float32).max,
                     np.inf, 0.0, -0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, -1.0, 1.0 + 2 ** -7, np.inf, np.inf, 0.0, -0.0], np.float32)
    got = so.bf16(vals)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (got, want)
    assert np.isnan(so.bf16(np.array([np.nan], np.float32)))[0]


def test_forward_rows_equals_the_full_forward_on_those_rows():
    """The slice forward (used on frames too large for the all-pairs matrices)
    gives the full forward's outputs for the chosen rows, bit for bit in
    float64, with and without the BF16 operand rounding; two scenes, so pairs
    must not cross scenes."""
    frame = random_frame(n=120, seed=3)
    frame["scene"] = (np.arange(120) >= 70).astype(np.int32)
    p = random_params(seed=4, dtype=np.float32)
    rows = np.array([0, 5, 17, 69, 70, 71, 118, 119])
    for bf in (True, False):
        full, _ = so.forward(p, frame, H, DT, bf16_ops=bf)
        part = so.forward_rows(p, frame, H, DT, rows, bf16_ops=bf)
        np.testing.assert_allclose(part, full[rows], rtol=1e-12, atol=1e-14)
