"""The large configurations at 1/8 scale (every length divided by 8, counts and steps kept),
checked against the oracle: layers, active partitions and partition lists bit-exact, phi within
1e-5, meshes as unordered sets within 1e-5.

- 1024^3 in 32^3, 64 spheres r in [20, 60], -0.5 until two fronts meet, then 0.2 * curvature,
  300 steps, 16 random 64^3 bounded updates per step, incremental marching cubes
  -> 128^3 in 4^3, r in [2.5, 7.5], 8^3 boxes.  The switch step is when two separate
  components first touch: spheres that overlap at the start form one component, and every
  radius grows by 0.5 per step, so after ceil(smallest gap between components) steps.  The
  speeds are kept, so at this scale the curvature flow shrinks the small spheres away; meshes
  are compared while the surface still exists and at the end.
- 2D 16384^2 in 64^2, 512 circles r in [4, 64], 300 steps at -0.5, marching squares
  -> 2048^2 in 8^2, r in [0.5, 8]."""
import math
import time

import numpy as np
import pytest

import oracle_lib
import parity
from felt_amd import workloads
from felt_amd._binding import speed_constant, speed_curvature

pytestmark = pytest.mark.gpu


def check(g, c, where, snapshot=True):
    d = parity.state_diff(g, c, snapshot=snapshot)
    assert not d, f"{where}: {d[:4]}"


def merge_step(cs, rs):
    n = len(cs)
    comp = list(range(n))

    def root(i):
        while comp[i] != i:
            comp[i] = comp[comp[i]]
            i = comp[i]
        return i

    gaps = [(math.dist(cs[i], cs[j]) - rs[i] - rs[j], i, j)
            for i in range(n) for j in range(i + 1, n)]
    for g, i, j in gaps:
        if g < 0:
            comp[root(i)] = root(j)
    between = [g for g, i, j in gaps if root(i) != root(j)]
    return max(0, math.ceil(min(between))) if between else 0


def test_large_3d_eighth_scale():
    import felt_amd
    N, part, L = 128, 4, 3
    rng = workloads.SplitMix64(0xFE1705)
    cs = [tuple(rng.uniform(-N / 2, N / 2) for _ in range(3)) for _ in range(64)]
    rs = [rng.uniform(20, 60) / 8 for _ in range(64)]
    k0 = merge_step(cs, rs)
    g = felt_amd.Surface(3, L, (N,) * 3, (part,) * 3)
    c = oracle_lib.surface(3, L, (N,) * 3, (part,) * 3)
    g.load_spheres(cs, rs)
    c.load_dense(workloads.spheres_phi((N,) * 3, L, cs, rs))
    check(g, c, "load")
    pg, pc = felt_amd.Polys(g), oracle_lib.polys(c)
    for p in (pg, pc):
        p.invalidate()
        p.march()
    t0 = time.time()
    meshed = 0
    for step in range(300):
        sp = speed_constant(-0.5) if step < k0 else speed_curvature(0.2)
        g.update(sp)
        c.update(sp)
        pg.notify()
        pc.notify()
        for _ in range(16):
            lo = [rng.randint(-N // 2, N // 2 - 8) for _ in range(3)]
            hi = [x + 7 for x in lo]
            g.update_box(lo, hi, sp)
            c.update_box(lo, hi, sp)
            pg.notify()
            pc.notify()
        pg.march()
        pc.march()
        if step % 25 == 24 or step == k0:
            check(g, c, f"step {step}", snapshot=step % 100 == 99)
            print(f"[c5a/8] step {step + 1} ok (switch at {k0}), band "
                  f"{g.stats().band_voxels}, {time.time() - t0:.0f}s", flush=True)
        if step in (k0, k0 + 20, 299):
            d = parity.mesh_diff(pg, g, pc, c)
            assert not d, f"step {step}: {d[:4]}"
            meshed = max(meshed, pg.totals()[1])
    assert meshed > 0


def test_large_2d_eighth_scale():
    import felt_amd
    N, part, L = 2048, 8, 2
    rng = workloads.SplitMix64(0xFE1706)
    cs = [tuple(rng.uniform(-N / 2, N / 2) for _ in range(2)) for _ in range(512)]
    rs = [rng.uniform(4, 64) / 8 for _ in range(512)]
    g = felt_amd.Surface(2, L, (N,) * 2, (part,) * 2)
    c = oracle_lib.surface(2, L, (N,) * 2, (part,) * 2)
    g.load_spheres(cs, rs)
    c.load_dense(workloads.spheres_phi((N,) * 2, L, cs, rs))
    check(g, c, "load")
    t0 = time.time()
    sp = speed_constant(-0.5)
    for step in range(300):
        g.update(sp)
        c.update(sp)
        if step % 30 == 29:
            check(g, c, f"step {step}", snapshot=step % 150 == 149)
            print(f"[c5b/8] step {step + 1} ok, band {g.stats().band_voxels}, "
                  f"{time.time() - t0:.0f}s", flush=True)
    pg, pc = felt_amd.Polys(g), oracle_lib.polys(c)
    for p in (pg, pc):
        p.invalidate()
        p.march()
    d = parity.mesh_diff(pg, g, pc, c)
    assert not d, d[:4]
    assert pg.totals()[1] > 0
