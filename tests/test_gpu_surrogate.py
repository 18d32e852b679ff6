"""The learned relaxation on the GPU against its CPU oracle, and what training
buys.

* the MFMA operand layouts the kernels assume (one v_mfma_f32_32x32x8_bf16 and
  one v_mfma_f32_32x32x16_bf16 against host products);
* the forward output, the loss and every block of the weight gradient of one
  frame against oracle/surrogate_oracle.py, which rounds the same operands
  to BF16 and computes the rest in float64 (single scene and batched scenes);
* training lowers the loss, and the trained network's density error is far
  below that of the untrained network and of dropping the relaxation;
* a density check that always fails reproduces the plain method bit for bit.

Tolerances: the GPU accumulates in FP32 in its own order and rounds each
pair's relu(z), H, M, dG and dM to BF16 from FP32 sums, the oracle from
float64 sums; an element whose two sums straddle a BF16 rounding boundary
moves by one BF16 ulp (2^-8 relative).  Measured relative L2 errors are about 1e-6 (outputs, loss and
every gradient block); the bound is 1e-3.
"""
from __future__ import annotations

import numpy as np
import pytest

from nbodysimulation_experiment_amd.scenes import SPHParams, SceneSpec, dam_break_4k, volume_positions, wall_planes
from nbodysimulation_experiment_amd.sim import ParticleSimulation, selftest_mfma

pytestmark = pytest.mark.gpu

DT = 0.005
GRAVITY = (0.0, -9.8)
REL_TOL = 1e-3


def so():
    from oracle import surrogate_oracle
    return surrogate_oracle


def test_mfma_operand_layout():
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 8, (32, 8)).astype(np.float32)
    b = rng.integers(-8, 8, (8, 32)).astype(np.float32)
    c = selftest_mfma(a, b)
    assert np.array_equal(c, a @ b)  # small integers: exact in BF16 and in the FP32 sums
    a = rng.normal(size=(32, 8)).astype(np.float32)
    b = rng.normal(size=(8, 32)).astype(np.float32)
    want = so().bf16(a).astype(np.float64) @ so().bf16(b).astype(np.float64)
    np.testing.assert_allclose(selftest_mfma(a, b), want, rtol=1e-5, atol=1e-5)


def test_mfma16_operand_layout():
    """The K = 16 form the network's kernels use: A[c][8g + t], B[8g + t][c]."""
    from nbodysimulation_experiment_amd.sim import selftest_mfma16
    rng = np.random.default_rng(1)
    a = rng.integers(-8, 8, (32, 16)).astype(np.float32)
    b = rng.integers(-8, 8, (16, 32)).astype(np.float32)
    assert np.array_equal(selftest_mfma16(a, b), a @ b)
    a = rng.normal(size=(32, 16)).astype(np.float32)
    b = rng.normal(size=(16, 32)).astype(np.float32)
    want = so().bf16(a).astype(np.float64) @ so().bf16(b).astype(np.float64)
    np.testing.assert_allclose(selftest_mfma16(a, b), want, rtol=1e-5, atol=1e-5)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_frame(sim, h=1.0):
    fr = sim.surrogate_frame(DT)
    p = sim.surrogate_params()
    L, g, o = so().loss_and_grad(p, fr, h, DT, bf16_ops=True)
    print(f"n={fr['x'].shape[0]} loss gpu {fr['loss']:.6g} oracle {L:.6g} out rel {rel(fr['out'], o):.2e}", flush=True)
    assert rel(fr["out"], o) < REL_TOL
    assert abs(fr["loss"] - L) <= REL_TOL * L
    s = so()
    blocks = {"W1": (s.OFF_W1, s.OFF_W2), "W2": (s.OFF_W2, s.OFF_B2), "b2": (s.OFF_B2, s.OFF_W3),
              "W3": (s.OFF_W3, s.OFF_B3), "b3": (s.OFF_B3, s.OFF_W4), "W4": (s.OFF_W4, s.OFF_B4),
              "b4": (s.OFF_B4, s.NUM_WEIGHTS)}
    for name, (lo, hi) in blocks.items():
        e = rel(fr["grad"][lo:hi], g[lo:hi])
        print(f"  grad {name}: rel {e:.2e}", flush=True)
        assert e < REL_TOL, name
    return fr


def developed_4k(steps=60):
    sim = ParticleSimulation.from_scene(dam_break_4k(0), SPHParams(), GRAVITY)
    sim.Update(DT, steps)
    return sim


def test_forward_and_gradient_match_oracle_single_scene():
    with developed_4k() as sim:
        sim.surrogate_init(5)
        sim.surrogate_calibrate(DT)
        check_frame(sim)
        # a second frame after some training: other weights, same agreement
        sim.surrogate_train(DT, 20, 1e-3)
        check_frame(sim)


def test_forward_and_gradient_match_oracle_batched_scenes():
    specs = []
    for k, (nx, ny) in enumerate([(20, 24), (33, 17), (12, 40)]):
        pos = volume_positions((-4.0, -2.0), nx, ny, 0.5, 1.0, seed=10 + k)
        specs.append(SceneSpec(40.0, 30.0, 1.0, pos, planes=wall_planes(40.0, 30.0)))
    with ParticleSimulation.from_scenes(specs, SPHParams(), GRAVITY) as sim:
        sim.Update(DT, 40)
        sim.surrogate_init(9)
        sim.surrogate_calibrate(DT)
        fr = check_frame(sim)
        assert set(np.unique(fr["scene"]).tolist()) == {0, 1, 2}


def from_state(state):
    sim = ParticleSimulation.from_scene(dam_break_4k(0), SPHParams(), GRAVITY)
    sim.set_state(state["x"], state["v"])
    sim.Update(DT, 2)
    return sim


def frame_loss(state, params):
    with from_state(state) as sim:
        sim.surrogate_set_params(params)
        return sim.surrogate_frame(DT)["loss"]


def density_error(state, params, steps=16):
    with from_state(state) as sim:
        sim.surrogate_set_params(params)
        sim.surrogate_set_mode(True, 1, 1e30)  # check every step, never fall back
        sim.Update(DT, steps)
        r = sim.surrogate_report()
        assert r["check_steps"] == steps and r["fallbacks"] == 0
        st = sim.get_state()
        assert np.isfinite(st["x"]).all()
        return r["mean_density_error"]


def test_training_lowers_loss_and_beats_untrained_and_no_relaxation():
    with developed_4k(100) as sim:
        sim.surrogate_init(1)
        sim.surrogate_calibrate(DT)
        untrained = sim.surrogate_params()
        losses = sim.surrogate_train(DT, 400, 1e-3)
        trained = sim.surrogate_params()
        state = sim.get_state()
    assert np.isfinite(losses).all()
    zero = trained.copy()
    zero[:so().NUM_WEIGHTS] = 0.0  # output 0: the step without relaxation
    l_trained, l_untrained, l_zero = (frame_loss(state, w) for w in (trained, untrained, zero))
    print(f"held-out frame loss: trained {l_trained:.4g} untrained {l_untrained:.4g} no relaxation {l_zero:.4g}",
          flush=True)
    assert l_trained < 0.3 * l_untrained and l_trained < 0.3 * l_zero
    e_trained, e_untrained, e_zero = (density_error(state, w) for w in (trained, untrained, zero))
    print(f"density error: trained {e_trained:.3e} untrained {e_untrained:.3e} no relaxation {e_zero:.3e}",
          flush=True)
    assert e_trained < 0.5 * e_zero
    assert e_trained < 0.5 * e_untrained


def test_failing_checks_reproduce_the_plain_method():
    with developed_4k(20) as plain, developed_4k(20) as fb:
        fb.surrogate_init(3)
        fb.surrogate_calibrate(DT)
        plain.Update(DT, 1)  # the calibration step is a plain step
        fb.surrogate_set_mode(True, 1, 0.0)  # every step is checked; every check falls back
        plain.Update(DT, 9)
        fb.Update(DT, 9)
        a, b = plain.get_state(), fb.get_state()
        for k in ("x", "prev", "v", "density", "pressure"):
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
        r = fb.surrogate_report()
        assert r["check_steps"] >= 4 and r["fallbacks"] == r["check_steps"]


def _trained_4k(steps=100, train=200):
    sim = developed_4k(steps)
    sim.surrogate_init(1)
    sim.surrogate_calibrate(DT)
    sim.surrogate_train(DT, train, 1e-3)
    return sim, sim.surrogate_params()


def test_carried_density_mode():
    """NBS_RELAX_SURROGATE_CARRIED_DENSITY: between checks the density pass
    is skipped and the network's density features are the last accepted
    check's, carried with each particle through the binning.
    * between two checks the state's density is constant per particle and its
      pressure is the reference's formula of it (sph.h:478-481), bit for bit;
    * with a check on every step (no step in between) it is the per-step
      surrogate bit for bit;
    * over 64 steps it stays close to the per-step surrogate run (both from
      the same state with the same network)."""
    p = SPHParams()
    sim, params = _trained_4k()
    start = sim.get_state()
    with sim:
        sim.surrogate_set_mode(True, 8, 0.01, carried_density=True)
        sim.Update(DT, 8)  # ends with a check step
        sim.Update(DT, 2)  # two network steps on the carried density
        a = sim.get_state()
        sim.Update(DT, 3)
        b = sim.get_state()
        assert np.array_equal(a["density"].view(np.uint32), b["density"].view(np.uint32))
        f32 = np.float32
        d = b["density"]
        pres = np.stack([f32(p.k) * (d[:, 0] - f32(p.rho0)), f32(p.k_near) * d[:, 1]], 1)
        assert np.array_equal(b["pressure"].view(np.uint32), pres.astype(np.float32).view(np.uint32))
        assert np.isfinite(b["x"]).all() and not np.array_equal(a["x"], b["x"])
        r = sim.surrogate_report()
        assert r["check_steps"] == 1 and r["fallbacks"] == 0

    def run(carried, interval, steps):
        s = ParticleSimulation.from_scene(dam_break_4k(0), SPHParams(), GRAVITY)
        s.set_state(start["x"], start["v"])
        s.Update(DT, 1)  # a restored state's first step rebuilds the lists
        s.surrogate_set_params(params)
        s.surrogate_set_mode(True, interval, 0.01, carried_density=carried)
        s.Update(DT, steps)
        st = s.get_state()
        s.close()
        return st

    e1, c1 = run(False, 1, 12), run(True, 1, 12)
    for k in ("x", "v", "density", "pressure"):
        assert np.array_equal(e1[k].view(np.uint32), c1[k].view(np.uint32)), k
    e8, c8 = run(False, 8, 64), run(True, 8, 64)  # 64 = 8 checks: the states' densities are fresh
    dx = np.linalg.norm(e8["x"].astype(np.float64) - c8["x"], axis=1)
    rel = np.abs(e8["density"][:, 0].astype(np.float64) - c8["density"][:, 0]).sum() / np.abs(e8["density"][:, 0]).sum()
    print(f"carried vs per-step density, 64 steps: position mean {dx.mean():.2e} h max {dx.max():.2e} h, "
          f"density {rel:.2e}", flush=True)
    assert dx.mean() < 0.01 and rel < 0.01


def test_surrogate_mode_argument_errors():
    from nbodysimulation_experiment_amd._lib import NbsError
    with developed_4k(2) as sim:
        with pytest.raises(NbsError):
            sim.surrogate_set_mode(True)  # not initialised
        sim.surrogate_init(0)
        with pytest.raises(NbsError):
            sim.surrogate_set_mode(True, 0, 0.01)
        with pytest.raises(NbsError):
            sim.surrogate_set_params(np.zeros(10, np.float32))
        with pytest.raises(NbsError):
            sim.surrogate_train(DT, 1, -1.0)


def test_batched_surrogate_with_in_run_training_and_scene_errors():
    """Uneven batched scenes plus one larger scene (the batched configuration
    at reduced size): surrogate inference with fine-tuning on each check
    step's plain frame, and per-scene density errors."""
    from nbodysimulation_experiment_amd import scenes
    specs = scenes.batched_scenes(seed=5, count=24, n_min=2048, n_max=16384)
    big = volume_positions((-30.0, -30.0), 256, 256, 0.5, 1.0, seed=77)
    specs.append(SceneSpec(200.0, 200.0, 1.0, big, planes=wall_planes(200.0, 200.0)))
    with ParticleSimulation.from_scenes(specs, SPHParams(), GRAVITY) as sim:
        sim.Update(DT, 40)
        sim.surrogate_init(2)
        sim.surrogate_calibrate(DT)
        sim.surrogate_train(DT, 30, 1e-3)
        p0 = sim.surrogate_params()
        sim.surrogate_set_mode(True, 4, 1e30)  # check every 4th step, never fall back
        sim.surrogate_set_online_training(1e-3)
        before = sim.surrogate_report()["train_steps"]
        sim.Update(DT, 32)
        r = sim.surrogate_report()
        errs = sim.surrogate_scene_errors()
        st = sim.get_state()
        p1 = sim.surrogate_params()
    print(f"checks {r['check_steps']} train steps {r['train_steps'] - before} last error {r['last_density_error']:.3e} "
          f"scene errors {errs.min():.3e}..{errs.max():.3e}", flush=True)
    assert r["check_steps"] == 8 and r["fallbacks"] == 0
    assert r["train_steps"] - before == 8          # one Adam step per check
    assert not np.array_equal(p0[:so().NUM_WEIGHTS], p1[:so().NUM_WEIGHTS])
    assert np.array_equal(p0[so().OFF_NORM:], p1[so().OFF_NORM:])
    assert errs.shape == (len(specs),) and np.isfinite(errs).all() and (errs > 0).all()
    # the global error is the plain-density-weighted mean of the scene errors
    assert errs.min() * (1 - 1e-9) <= r["last_density_error"] <= errs.max() * (1 + 1e-9)
    assert np.isfinite(st["x"]).all()


def test_single_scene_error_equals_global_error():
    with developed_4k(40) as sim:
        sim.surrogate_init(4)
        sim.surrogate_calibrate(DT)
        sim.surrogate_set_mode(True, 2, 1e30)
        sim.Update(DT, 4)
        r = sim.surrogate_report()
        errs = sim.surrogate_scene_errors()
    assert errs.shape == (1,)
    assert abs(errs[0] - r["last_density_error"]) <= 1e-9 * r["last_density_error"]
