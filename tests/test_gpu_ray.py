"""Surface::ray on the device against the reference's raycasting checks (test_Surface.cpp:
1278-1535) and against the oracle on seeded random rays.  Hits and misses must agree ray for ray;
hit positions within 1e-5."""
import numpy as np
import pytest

import oracle_lib
import ray_kats
from felt_amd import workloads
from felt_amd._binding import FeltError, speed_const_curv, speed_constant

pytestmark = pytest.mark.gpu
MISS = np.float32(3.4028235e38)


def _gpu_surface(D, L, size, part):
    import felt_amd
    return felt_amd.Surface(D, L, size, part)


def ray_diff(got, want, tol=1e-5):
    gm = np.all(got == MISS, axis=1)
    wm = np.all(want == MISS, axis=1)
    out = []
    if np.any(gm != wm):
        i = np.nonzero(gm != wm)[0]
        out.append(f"{len(i)} rays differ in hit/miss, first {i[:5]}: {got[i[:2]]} vs {want[i[:2]]}")
    both = ~gm & ~wm
    if np.any(both):
        err = np.max(np.abs(got[both].astype(np.float64) - want[both]))
        if err > tol:
            out.append(f"hit positions differ by {err}")
    return out


def test_gpu_ray_matches_reference_checks():
    print("[ray] reference checks", flush=True)
    fails = ray_kats.check_ray_kats(_gpu_surface)
    assert not fails, fails


@pytest.mark.parametrize("case", ["seed3", "spheres3", "circle2", "noisy3"])
def test_gpu_ray_matches_oracle(case):
    print(f"[ray] parity {case}", flush=True)
    rng = np.random.default_rng(11)
    if case == "seed3":
        g, c = ray_kats.sphere3(_gpu_surface), ray_kats.sphere3(oracle_lib.surface)
        size = (32, 32, 32)
    elif case == "circle2":
        g, c = ray_kats.circle2(_gpu_surface), ray_kats.circle2(oracle_lib.surface)
        size = (16, 16)
    else:
        size, part, L = (48, 40, 56), (8, 8, 8), 2
        cs = [(-6.0, 3.0, 2.5), (9.0, -7.0, -8.0), (0.5, 10.0, 12.0)]
        rs = [9.5, 6.0, 5.0]
        phi = workloads.spheres_phi(size, L, cs, rs)
        g, c = _gpu_surface(3, L, size, part), oracle_lib.surface(3, L, size, part)
        g.load_dense(phi)
        c.load_dense(phi)
        if case == "noisy3":
            for s in (g, c):
                for _ in range(4):
                    s.update(speed_const_curv(-0.3, 0.2))
            assert ray_kats.approx_vec(g.snapshot(), c.snapshot())
    D = len(size)
    o, d = ray_kats.random_rays(rng, D, 2000, size)
    got, want = g.ray(o, d), c.ray(o, d)
    hits = int(np.sum(~np.all(want == MISS, axis=1)))
    print(f"[ray] {case}: {hits} of {len(o)} rays hit", flush=True)
    assert hits > len(o) // 4
    fails = ray_diff(got, want)
    assert not fails, fails


def test_gpu_batch_ray_matches_oracle():
    print("[ray] batch members", flush=True)
    from test_gpu_batch import make_batch
    sizes = [(24, 24, 24), (40, 32, 24), (16, 48, 32), (32, 32, 32)]
    bt, oracles = make_batch(3, 2, sizes, (8, 8, 8), seed=5)
    rng = np.random.default_rng(2)
    mem, org, dirs, want = [], [], [], []
    for i, size in enumerate(sizes):
        o, d = ray_kats.random_rays(rng, 3, 300, size)
        mem += [i] * len(o)
        org.append(o)
        dirs.append(d)
        want.append(oracles[i].ray(o, d))
    perm = rng.permutation(len(mem))  # rays of all members interleaved in one launch
    mem = np.array(mem)[perm]
    org, dirs, want = (np.concatenate(x)[perm] for x in (org, dirs, want))
    got = bt.ray(mem, org, dirs)
    fails = ray_diff(got, want)
    assert not fails, fails


def test_gpu_ray_rejects_bad_rays():
    g = ray_kats.circle2(_gpu_surface)
    with pytest.raises(FeltError):
        g.ray([(0.0, 0.0)], [(0.0, 0.0)])      # no direction
    with pytest.raises(FeltError):
        g.ray([(np.nan, 0.0)], [(1.0, 0.0)])
    with pytest.raises(FeltError):
        g.ray([(0.0, 0.0)], [(3.0, 0.0)])      # not a unit vector
    assert g.ray(np.zeros((0, 2)), np.zeros((0, 2))).shape == (0, 2)
