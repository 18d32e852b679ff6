"""Workload builders: determinism, AddVolume placement arithmetic, circle layout."""
import numpy as np

from nbodysimulation_experiment_amd import scenes


def test_splitmix_uniform_range_and_determinism():
    a = scenes.splitmix64_uniform(7, 10000)
    b = scenes.splitmix64_uniform(7, 10000)
    assert np.array_equal(a, b)
    assert a.dtype == np.float32 and a.min() >= -1.0 and a.max() < 1.0
    assert abs(float(a.mean())) < 0.03
    # SplitMix64 reference value for seed 0, first output (0xE220A8397B1DCDAF)
    z = scenes._splitmix64(0, np.array([0], np.uint64))[0]
    assert int(z) == 0xE220A8397B1DCDAF


def test_volume_positions_follow_add_volume_without_jitter_bias():
    pos = scenes.volume_positions((-24.0, -14.0), 64, 64, 0.5, 1.0, seed=0)
    assert pos.shape == (4096, 2) and pos.dtype == np.float32
    # lattice point (ix, iy) = ix*s + s/2 + centre - 64*s/2, jitter within 0.01h
    ix = np.tile(np.arange(64), 64)
    iy = np.repeat(np.arange(64), 64)
    lat = np.stack([ix * 0.5 + 0.25 - 24.0 - 16.0, iy * 0.5 + 0.25 - 14.0 - 16.0], 1)
    d = pos - lat
    assert np.abs(d).max() <= 0.01 + 1e-6
    assert np.abs(d).max() > 0.009


def test_dam_break_4k_matches_its_box():
    s = scenes.dam_break_4k()
    assert s.n == 4096 and s.width == 80.0 and s.height == 60.0
    assert s.planes == [(0.0, 1.0, -30.0), (0.0, -1.0, -30.0), (1.0, 0.0, -40.0), (-1.0, 0.0, -40.0)]
    assert pos_inside(s)


def pos_inside(s):
    hw, hh = s.width / 2, s.height / 2
    return bool((np.abs(s.positions[:, 0]) < hw).all() and (np.abs(s.positions[:, 1]) < hh).all())


def test_dam_break_1m_circles_disjoint_and_clear_of_block():
    s = scenes.dam_break_1m()
    assert s.n == 1 << 20 and len(s.circles) == 64 and pos_inside(s)
    c = np.array(s.circles)
    assert (c[:, 2] >= 2.0).all() and (c[:, 2] <= 8.0).all()
    d = np.hypot(c[:, None, 0] - c[None, :, 0], c[:, None, 1] - c[None, :, 1])
    np.fill_diagonal(d, np.inf)
    assert (d > c[:, None, 2] + c[None, :, 2]).all()
    lo, hi = s.positions.min(0), s.positions.max(0)
    qx = np.clip(c[:, 0], lo[0], hi[0])
    qy = np.clip(c[:, 1], lo[1], hi[1])
    assert (np.hypot(c[:, 0] - qx, c[:, 1] - qy) > c[:, 2]).all()
    assert (np.abs(c[:, :2]) + c[:, 2:3] < 300).all()


def test_batched_scenes_counts_and_boxes():
    sp = scenes.batched_scenes()
    assert len(sp) == 256
    n = np.array([s.n for s in sp])
    assert n.min() >= 2048 and n.max() <= 65536
    assert 3.5e6 < n.sum() < 6.5e6
    for s in sp:
        assert pos_inside(s) and s.width == s.height and s.width % 2 == 0


def test_training_scenes_with_tall_columns():
    """168 shallow dam breaks plus 8 columns as tall as the 1M scene's: the
    bench's training set of 3,276,800 particles, every block inside its box."""
    specs = scenes.training_scenes(count=168, tall=8)
    assert len(specs) == 176
    assert sum(s.n for s in specs) == 3_276_800
    for s in specs[:168]:
        assert s.n == 128 * 128
    for s in specs[168:]:
        assert s.n == 64 * 1024
        lo, hi = s.positions.min(0), s.positions.max(0)
        assert hi[1] - lo[1] > 511.0 and hi[0] - lo[0] < 32.0  # 512h tall, 32h wide
        assert lo[0] > -s.width / 2 and lo[1] > -s.height / 2 and hi[0] < s.width / 2 and hi[1] < s.height / 2
    # the shallow scenes are those of the plain set, and the columns differ by seed
    plain = scenes.training_scenes(count=168)
    assert all(np.array_equal(a.positions, b.positions) for a, b in zip(specs[:168], plain))
    assert not np.array_equal(specs[168].positions, specs[169].positions)
