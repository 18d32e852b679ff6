"""GPU parity on the rest of the path and its edge cases, bit-exact against
the oracle's FP32 Jacobi-order run (zero-ULP tolerance)."""
import numpy as np
import pytest

from nbodysimulation_experiment_amd.scenes import SPHParams, SceneSpec, volume_positions, wall_planes
from nbodysimulation_experiment_amd.sim import NbsError, ParticleSimulation
from tests.helpers import DT, assert_bit_exact, gpu_for, oracle_for

pytestmark = pytest.mark.gpu


def _box_scene(seed=1, n=24):
    pos = volume_positions((-3.0, -2.0), n, n, 0.5, 1.0, seed=seed)
    return SceneSpec(30.0, 20.0, 1.0, pos, planes=wall_planes(30.0, 20.0))


def _pair(spec, params=None, gravity=(0.0, -9.8)):
    return gpu_for([spec], params, gravity), oracle_for([spec], params, gravity=gravity)


def test_segments_and_polygons_bit_exact():
    # the reference's Dambreak wall box and Fun-scenario slopes are polygons and segments (sph.h:327, 431-434)
    spec = _box_scene()
    sim, ref = _pair(spec)
    box = [(4.0, 6.0), (2.0, 6.0), (2.0, -9.0), (4.0, -9.0)]
    tri = [(-9.0, -9.0), (-5.0, -9.0), (-9.0, -5.0)]
    for s in (sim,):
        s.AddPolygon(box)
        s.AddPolygon(tri)
        s.AddLineSegment((-14.0, 2.0), (-6.0, -3.0))
        s.AddLineSegment((5.0, -5.0), (13.0, -8.0))
    ref.add_polygon(0, box)
    ref.add_polygon(0, tri)
    ref.add_segment(0, -14.0, 2.0, -6.0, -3.0)
    ref.add_segment(0, 5.0, -5.0, 13.0, -8.0)
    sim.Update(DT, 300)
    ref.step(DT, 300)
    assert_bit_exact(sim.get_state(), ref.state(), "segments/polygons: ")


def test_circles_and_viscosity_bit_exact():
    spec = _box_scene(seed=2)
    spec.circles = [(0.5, -6.0, 2.5), (-8.0, 4.0, 1.0), (9.0, -7.0, 3.0)]
    params = SPHParams(sigma=0.5, beta=0.3, k=0.02, k_near=0.05)
    sim, ref = _pair(spec, params)
    sim.Update(DT, 200)
    ref.step(DT, 200)
    assert_bit_exact(sim.get_state(), ref.state(), "circles+sigma: ")


def test_empty_simulation_and_zero_steps():
    sim = ParticleSimulation(SPHParams(), 10.0, 10.0)
    sim.Update(DT, 5)
    assert sim.GetParticleCount() == 0
    st = sim.get_state()
    assert st["x"].shape == (0, 2)
    sim.AddParticle((0.0, 0.0))
    sim.Update(DT, 0)
    assert sim.get_state()["x"].tolist() == [[0.0, 0.0]]


def test_particles_outside_the_grid_clamp_into_border_cells():
    # sph.h:456-463: positions beyond the domain bin into the border cells
    rng = np.random.default_rng(4)
    pos = np.concatenate([rng.uniform(-14, 14, (300, 2)), rng.uniform(-40, 40, (200, 2))]).astype(np.float32)
    spec = SceneSpec(20.0, 20.0, 1.0, pos)  # no walls: particles outside stay outside
    sim, ref = _pair(spec)
    sim.Update(DT, 40)
    ref.step(DT, 40)
    assert_bit_exact(sim.get_state(), ref.state(), "outside grid: ")


def test_coincident_particles():
    # zero-length normalise (vecmath.h:272-280): coincident particles add density, exert no push
    pos = np.array([[0.25, 0.25]] * 6 + [[0.75, 0.25]] * 3 + [[3.0, 3.0]], np.float32)
    spec = SceneSpec(10.0, 10.0, 1.0, pos, planes=wall_planes(10.0, 10.0))
    sim, ref = _pair(spec, SPHParams(sigma=0.4))
    sim.Update(DT, 30)
    ref.step(DT, 30)
    got = sim.get_state()
    assert_bit_exact(got, ref.state(), "coincident: ")
    assert got["density"][9, 0] == 1.0


def test_particles_added_mid_run_and_forces():
    spec = _box_scene(seed=5, n=16)
    sim, ref = _pair(spec)
    sim.Update(DT, 10)
    ref.step(DT, 10)
    extra = volume_positions((6.0, 4.0), 6, 6, 0.5, 1.0, seed=9)
    force = np.tile(np.array([[30.0, -5.0]], np.float32), (extra.shape[0], 1))
    first = sim.AddParticles(extra, force)
    assert first == spec.n and ref.add_particles(0, extra, force) == spec.n
    sim.AddExternalForces((2.0, 0.5))
    ref.set_external_force(2.0, 0.5)
    sim.Update(DT, 25)
    ref.step(DT, 25)
    assert_bit_exact(sim.get_state(), ref.state(), "added mid-run: ")
    sim.ClearExternalForce()
    ref.set_external_force(0.0, 0.0)
    sim.Update(DT, 5)
    ref.step(DT, 5)
    assert_bit_exact(sim.get_state(), ref.state(), "cleared force: ")


def test_set_state_restores_and_continues():
    spec = _box_scene(seed=6)
    sim, ref = _pair(spec)
    sim.Update(DT, 20)
    st = sim.get_state()
    ref.set_state(st["x"].astype(np.float64), st["v"].astype(np.float64))
    sim.set_state(st["x"], st["v"])
    sim.Update(DT, 30)
    ref.step(DT, 30)
    assert_bit_exact(sim.get_state(), ref.state(), "restored: ")


def test_emitters_bit_exact():
    # demo4.cpp:257-284 emitters feeding a dam break, against the oracle's emitter
    spec = _box_scene(seed=8, n=12)
    sim, ref = _pair(spec)
    for s in (sim,):
        s.AddEmitter((-12.0, 6.0), (1.0, 0.0), 2.0, 0.02, 20.0, 0.4, seed=3)
        s.AddEmitter((10.0, 8.0), (-0.6, -0.8), 1.5, 0.05, 10.0, 1.0, seed=4)
    ref.add_emitter(0, (-12.0, 6.0), (1.0, 0.0), 2.0, 0.02, 20.0, 0.4, seed=3)
    ref.add_emitter(0, (10.0, 8.0), (-0.6, -0.8), 1.5, 0.05, 10.0, 1.0, seed=4)
    sim.Update(DT, 150)
    ref.step(DT, 150)
    assert sim.GetParticleCount() == ref.n > spec.n
    assert_bit_exact(sim.get_state(), ref.state(), "emitters: ")


def test_invalid_arguments_raise():
    sim = ParticleSimulation(SPHParams(), 10.0, 10.0)
    with pytest.raises(NbsError) as e:
        sim.Update(-1.0)
    assert e.value.status == "NBS_ERR_INVALID_ARGUMENT"
    with pytest.raises(NbsError):
        sim.AddPlane((0.0, 1.0), 0.0, scene=3)
    with pytest.raises(NbsError):
        sim.AddPolygon([(0.0, 0.0), (1.0, 0.0)])
    with pytest.raises(NbsError):
        sim.set_tuning(tile_cap=(7, 0, 0))
    with pytest.raises(NbsError):
        ParticleSimulation(SPHParams(h=-1.0), 10.0, 10.0)
    with pytest.raises(NbsError):  # beyond the pair kernels' operand range r^2 < 2^80
        ParticleSimulation(SPHParams(h=2.0 ** 41), 10.0, 10.0)


def test_check_finite_option_reports_non_finite_state():
    sim = ParticleSimulation(SPHParams(), 10.0, 10.0)
    sim.AddParticles([[0.0, 0.0], [1.0, 1.0]])
    sim.set_options(check_finite=True)
    sim.Update(DT, 2)
    st = sim.get_state()
    x = st["x"].copy()
    x[1, 0] = np.nan
    sim.set_state(x, st["v"])
    with pytest.raises(NbsError) as e:
        sim.Update(DT, 1)
    assert e.value.status == "NBS_ERR_NON_FINITE"
    assert sim.diagnostics()["non_finite_count"] >= 1


def test_fp64_diagnostics_match_host_sums():
    spec = _box_scene(seed=7)
    sim = gpu_for([spec])
    sim.Update(DT, 50)
    st = sim.get_state()
    d = sim.diagnostics()
    v = st["v"].astype(np.float64)
    rho = st["density"][:, 0].astype(np.float64)
    assert d["particle_count"] == spec.n
    np.testing.assert_allclose(d["momentum"], v.sum(0), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(d["kinetic_energy"], 0.5 * (v ** 2).sum(), rtol=1e-12)
    np.testing.assert_allclose(d["mean_density"], rho.mean(), rtol=1e-12)
    np.testing.assert_allclose(d["max_density"], rho.max(), rtol=0)
    np.testing.assert_allclose(d["mean_abs_density_error"], np.abs(rho - 10.0).mean() / 10.0, rtol=1e-12)


def test_particle_colors_bit_exact():
    from oracle.oracle import kat_color
    spec = _box_scene(seed=10)
    sim = gpu_for([spec])
    sim.Update(DT, 80)
    st = sim.get_state()
    got = sim.get_colors()
    want = kat_color(10.0, st["density"][:, 0], st["pressure"][:, 0], st["v"])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _circle_field(seed, count, lo, hi, rmin, rmax):
    rng = np.random.default_rng(seed)
    return [(float(np.float32(rng.uniform(lo[0], hi[0]))), float(np.float32(rng.uniform(lo[1], hi[1]))),
             float(np.float32(rng.uniform(rmin, rmax)))) for _ in range(count)]


@pytest.mark.parametrize("caps", [(2048, 2048, 2048), (-1, -1, -1)])
def test_many_circles_and_mixed_body_lists_bit_exact(caps):
    """Circles are skipped per wave when the wave's bounding box lies outside
    a circle's cull box.  A field of circles the fluid falls through, one
    scene without circles and one with a few touching circles, stepped
    together (waves that span scenes take the per-lane body walk), in the
    narrow and the wide relaxation kernels, against the oracle bit for bit."""
    a = SceneSpec(40.0, 30.0, 1.0, volume_positions((-6.0, 4.0), 28, 20, 0.5, 1.0, seed=21),
                  planes=wall_planes(40.0, 30.0))
    a.circles = _circle_field(5, 40, (-15.0, -13.0), (15.0, 3.0), 0.3, 1.6)
    b = SceneSpec(20.0, 20.0, 1.0, volume_positions((-3.0, -3.0), 9, 13, 0.5, 1.0, seed=22),
                  planes=wall_planes(20.0, 20.0))
    c = SceneSpec(24.0, 24.0, 1.0, volume_positions((-4.0, 0.0), 16, 10, 0.5, 1.0, seed=23),
                  planes=wall_planes(24.0, 24.0))
    c.circles = [(-2.0, -4.0, 1.0), (-1.0, -4.5, 1.2), (3.0, -6.0, 2.0)]  # overlapping pair
    params = SPHParams(sigma=0.3, beta=0.1)
    for specs in ([a], [a, b, c]):
        sim, ref = gpu_for(specs, params), oracle_for(specs, params)
        sim.set_tuning(autotune=False, tile_cap=caps, wide_max_particles=0)
        sim.Update(DT, 160)
        ref.step(DT, 160)
        assert_bit_exact(sim.get_state(), ref.state(), f"{len(specs)} scene(s), caps {caps}: ")


@pytest.mark.parametrize("caps", [(1024, 1024, 1024), (-1, -1, -1)])
def test_scene_added_mid_run_bit_exact(caps):
    """A scene added after steps have run grows the cell tables: the per-cell
    scene table of the multi-scene kernels is rebuilt (the old one no longer
    covers the grid), and the next step starts without neighbour lists, as in
    the oracle.  Two scenes over 19,200 cells (placement binning with the
    narrow kernels), then a third; state and grid bit-exact at every mark."""
    a = SceneSpec(120.0, 100.0, 1.0, volume_positions((-20.0, -10.0), 40, 30, 0.5, 1.0, seed=51),
                  planes=wall_planes(120.0, 100.0))
    b = SceneSpec(90.0, 80.0, 1.0, volume_positions((-8.0, -5.0), 24, 20, 0.5, 1.0, seed=52),
                  planes=wall_planes(90.0, 80.0))
    c = SceneSpec(60.0, 50.0, 1.0, volume_positions((-4.0, 0.0), 20, 16, 0.5, 1.0, seed=53),
                  planes=wall_planes(60.0, 50.0))
    sim, ref = gpu_for([a, b]), oracle_for([a, b])
    sim.set_tuning(autotune=False, tile_cap=caps, wide_max_particles=0)

    def check(what):
        assert_bit_exact(sim.get_state(), ref.state(), what)
        ids, start = sim.get_grid()
        oids, ostart = ref.grid()
        np.testing.assert_array_equal(start, ostart)
        np.testing.assert_array_equal(ids, oids)

    sim.Update(DT, 8)
    ref.step(DT, 8)
    check(f"two scenes, caps {caps}: ")
    assert sim.add_scene_spec(c) == 2 and ref.add_scene_spec(c) == 2
    assert sim.total_cells == 12000 + 7200 + 3000
    sim.Update(DT, 8)
    ref.step(DT, 8)
    check(f"third scene added, caps {caps}: ")
    sim.Update(DT, 8)
    ref.step(DT, 8)
    check(f"eight steps on, caps {caps}: ")


@pytest.mark.parametrize("caps", [(1024, 1024, 1024), (0, 0, 2048), (1024, -1, 1024)])
def test_dense_cells_bit_exact(caps):
    """Particles at 0.12h spacing put about 70 in a cell: rows of 200+
    candidates walked in 32-candidate chunks (more than 16 per particle), in
    LDS-tile, global-memory and wide kernels, against the oracle bit for bit."""
    pos = volume_positions((-2.0, -2.0), 34, 34, 0.12, 1.0, seed=31)
    spec = SceneSpec(20.0, 20.0, 1.0, pos, planes=wall_planes(20.0, 20.0))
    sim, ref = gpu_for([spec]), oracle_for([spec])
    sim.set_tuning(autotune=False, tile_cap=caps, wide_max_particles=0)
    sim.Update(DT, 12)
    ref.step(DT, 12)
    assert_bit_exact(sim.get_state(), ref.state(), f"dense, caps {caps}: ")
    assert sim.GetStats()["maxParticleNeighborCount"] > 16 * 32


@pytest.mark.parametrize("caps", [(4096, 4096, 4096), (0, 2048, 0), (-1, -1, -1)])
def test_crowded_cells_bit_exact(caps):
    """Particles at 0.05h spacing put about 400 in a cell: rows of 1,000+
    candidates, dozens of 32-candidate chunks per row, and chunks with every
    count of marked pairs, so the walk's four-pairs-per-iteration slots run
    full, partly used and past the last marked bit.  1,156 particles leave a
    ragged last workgroup; a second scene runs the multi-scene kernels.
    Against the oracle bit for bit."""
    pos = volume_positions((-0.8, -0.8), 34, 34, 0.05, 1.0, seed=41)
    a = SceneSpec(20.0, 20.0, 1.0, pos, planes=wall_planes(20.0, 20.0))
    b = SceneSpec(16.0, 16.0, 1.0, volume_positions((-0.5, -0.5), 20, 21, 0.05, 1.0, seed=42),
                  planes=wall_planes(16.0, 16.0))
    for specs in ([a], [a, b]):
        sim, ref = gpu_for(specs), oracle_for(specs)
        sim.set_tuning(autotune=False, tile_cap=caps, wide_max_particles=0)
        sim.Update(DT, 8)
        ref.step(DT, 8)
        assert_bit_exact(sim.get_state(), ref.state(), f"crowded, {len(specs)} scene(s), caps {caps}: ")
        assert sim.GetStats()["maxParticleNeighborCount"] > 1000
