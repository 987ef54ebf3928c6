"""Scene generators: counts, bounds and the reference initDambreak loop."""
import numpy as np

from mlsmpm import scenes


def test_config0_shape_and_bounds():
    rec, box = scenes.config0()
    assert rec.shape == (40000, 20) and box == [64.0, 64.0, 64.0]
    p = rec[:, 0:3]
    assert p[:, 0].min() >= 2.9 and p[:, 0].max() < 35.1
    assert p[:, 2].max() < 35.1 and p[:, 1].max() < 45.1
    # y outer loop: 59 x 59 sites per layer
    assert np.allclose(rec[:3481, 1], rec[0, 1], atol=0.2)
    assert np.all(rec[:, 3:] [:, [0, 1, 2, 3]] == 0)


def test_config1_density_keeps_mass_below_int32():
    rec, box = scenes.config1()
    assert rec.shape == (300000, 20) and box == [80.0, 80.0, 80.0]
    p = rec[:, 0:3]
    assert p.min() > 2.8 and p[:, 0].max() < 3 + 0.45 * 80 + 0.2
    cells = np.floor(p).astype(np.int64)
    counts = np.bincount((cells[:, 0] * 80 + cells[:, 1]) * 80 + cells[:, 2])
    assert counts.max() * 1e7 < 2**31 / 4


def test_init_dambreak_matches_reference_loop():
    # mls-mpm/mls-mpm.ts:195-223 with a deterministic rand()
    seq = iter(np.linspace(0, 0.99, 100000))
    rec = scenes.init_dambreak([40, 30, 60], 70000, rand=lambda: next(seq))
    assert rec.shape[0] == 70000
    # first particle: i=3, j=0, k=3 plus the shared jitter 2*rand() = 0
    assert list(rec[0, 0:3]) == [3.0, 0.0, 3.0]
    # one scalar jitter on all three axes
    d = rec[5, 0:3] - np.array([3.0, 0.0, 3.0 + 5 * 0.65], dtype=np.float32)
    assert np.allclose(d, d[0], atol=1e-5)
    # z runs 3 .. < 60/2
    assert rec[:, 2].max() < 30 + 2.0


def test_small_scene_is_seeded():
    a = scenes.small_scene(1000, [24, 24, 24], seed=4, velocity=1.0, affine=0.1)
    b = scenes.small_scene(1000, [24, 24, 24], seed=4, velocity=1.0, affine=0.1)
    assert np.array_equal(a, b)
