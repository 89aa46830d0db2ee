"""The oracle's Surface::ray against the reference's raycasting checks (test_Surface.cpp:
1278-1535) and against geometry the reference's checks leave open."""
import numpy as np

import oracle_lib
import ray_kats
from felt_amd._binding import speed_constant
from felt_amd.workloads import spheres_phi


def test_oracle_ray_matches_reference_checks():
    fails = ray_kats.check_ray_kats(oracle_lib.surface)
    assert not fails, fails


def test_oracle_ray_hits_sphere_surface():
    """Rays aimed through the centre of a loaded sphere of radius r hit at distance r from the
    centre, to the accuracy of the level set (a fraction of a voxel)."""
    s = oracle_lib.surface(3, 2, (40, 40, 40), (8, 8, 8))
    s.load_dense(spheres_phi((40, 40, 40), 2, [(1.5, -2.0, 0.5)], [9.0]))
    rng = np.random.default_rng(7)
    d = rng.normal(size=(64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = np.array([1.5, -2.0, 0.5])
    o = (c - 25.0 * d).astype(np.float32)
    hits = s.ray(o, d.astype(np.float32))
    assert not np.any(hits == np.float32(3.4028235e38))
    r = np.linalg.norm(hits - c, axis=1)
    assert np.all(np.abs(r - 9.0) < 0.15), r
    # the hit lies on the near side: between origin and centre
    assert np.all(np.sum((hits - o) * d, axis=1) < 25.0)


def test_oracle_ray_origin_on_ray_line_invariance():
    """Moving an outside origin back along its own ray does not change the hit.  The seeded
    surface grown by three unit steps is the city-block ball |x|+|y|+|z| = 3 (the reference's
    diagonal check hits it at (-1.5, -1.5, 0))."""
    s = ray_kats.sphere3(oracle_lib.surface)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(32, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o1 = (-12.0 * d).astype(np.float32)
    o2 = (-14.0 * d).astype(np.float32)
    h1, h2 = s.ray(o1, d), s.ray(o2, d)
    assert np.allclose(h1, h2, atol=2e-3)
    assert np.allclose(np.abs(h1).sum(axis=1), 3.0, atol=0.05), np.abs(h1).sum(axis=1)


def test_ray_kats_detect_planted_mistake():
    """A surface whose hits are shifted by a tenth of a voxel fails the reference's checks."""
    class Shifted:
        def __init__(self, s):
            self.s = s

        def __getattr__(self, k):
            return getattr(self.s, k)

        def ray(self, o, d):
            h = self.s.ray(o, d)
            miss = np.all(h == np.float32(3.4028235e38), axis=1)
            h[~miss, 0] += 0.1
            return h

    fails = ray_kats.check_ray_kats(lambda *a: Shifted(oracle_lib.surface(*a)))
    assert fails


def test_oracle_ray_misses_outside_band():
    s = ray_kats.sphere3(oracle_lib.surface)
    # parallel to the x axis, far above the sphere
    h = s.ray([(-20.0, 10.0, 0.0)], [(1.0, 0.0, 0.0)])
    assert np.all(h == np.float32(3.4028235e38))
    s.update(speed_constant(0.0))
