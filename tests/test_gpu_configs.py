"""The large configurations at 1/8 scale (every length divided by 8, counts and steps kept),
checked against the oracle: layers, active partitions and partition lists bit-exact, phi within
1e-5, meshes as unordered sets within 1e-5.

- 1024^3 in 32^3, 64 spheres r in [20, 60], -0.5 until the fronts have merged, then
  0.2 * curvature, 300 steps, 16 random 64^3 bounded updates per step, incremental marching
  cubes -> 128^3 in 4^3, r in [2.5, 7.5], 8^3 boxes.  "Merged" is pinned as the first step
  after which the inside region (phi < 0) is one face-connected component: every initial sphere
  has joined one front (step 35 of 300 at this scale, 56 components at the start).  Layers,
  active partitions and partition lists are compared after every step, phi every 50 steps,
  meshes at the switch and at steps 101 and 300.
- 2D 16384^2 in 64^2, 512 circles r in [4, 64], 300 steps at -0.5, marching squares
  -> 2048^2 in 8^2, r in [0.5, 8]; layers and lists compared after every step, phi every 50."""
import time
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np
import pytest

import oracle_lib
import progress
import oracle_pool
import parity
from felt_amd import workloads
from felt_amd._binding import speed_constant

pytestmark = pytest.mark.gpu


def check(g, c, where, snapshot=True):
    d = parity.state_diff(g, c, snapshot=snapshot)
    assert not d, f"{where}: {d[:4]}"


@pytest.mark.timeout(900)
def test_large_3d_eighth_scale():
    import felt_amd
    cfg = oracle_pool.C5AQ
    cs, rs, _ = oracle_pool.c5aq_setup()
    size = (cfg["N"],) * 3
    with ProcessPoolExecutor(1, mp_context=mp.get_context("spawn")) as ex:
        fut = ex.submit(oracle_pool.c5aq_job, None)
        g = felt_amd.Surface(3, cfg["L"], size, (cfg["part"],) * 3)
        g.load_spheres(cs, rs)
        pg = felt_amd.Polys(g)
        pg.invalidate()
        pg.march()
        states, phis, meshes, bands = [], {}, {}, []
        t0 = time.time()

        def on_step(step, switch):
            states.append(oracle_pool.state_digest(g))
            bands.append(int(g.stats().band_voxels))
            if step % 50 == 49:
                phis[step] = g.snapshot()
            if step == switch or step in (100, 299):
                meshes[step] = parity.mesh_sets(pg, g)
            if step % 10 == 9:
                progress.say(f"[c5a/8] gpu step {step + 1} (switch {switch}), band {bands[-1]}, "
                             f"{time.time() - t0:.0f}s")

        switch = oracle_pool.c5aq_run(g, pg, on_step)
        want = progress.wait(fut, "c5a/8")
    assert switch == want["switch"] and switch is not None and switch > 10
    for step, (a, b) in enumerate(zip(states, want["state"])):
        assert a[0] == b[0], f"step {step}: layers differ ({[x[0] for x in a[0]]} vs " \
                             f"{[x[0] for x in b[0]]})"
        assert a[1] == b[1], f"step {step}: active partitions or partition lists differ"
    assert bands == want["band"]
    for step, phi in phis.items():
        d = np.abs(phi.astype(np.float64) - want["phi"][step].astype(np.float64)).max()
        assert d <= parity.PHI_TOL, f"step {step}: phi max diff {d}"
    for step, m in meshes.items():
        d = parity.mesh_sets_diff(m, want["mesh"][step], 3)
        assert not d, f"mesh at step {step}: {d[:4]}"
    assert min(bands) > 0 and max(len(m) for m in meshes.values()) > 0
    progress.say(f"[c5a/8] 300 steps equal the oracle after every step; fronts merged at "
                 f"step {switch + 1}; band {min(bands)}..{max(bands)} voxels")


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
        check(g, c, f"step {step}", snapshot=step % 50 == 49)
        if step % 50 == 49:
            progress.say(f"[c5b/8] step {step + 1} ok, band {g.stats().band_voxels}, "
                         f"{time.time() - t0:.0f}s")
    pg, pc = felt_amd.Polys(g), oracle_lib.polys(c)
    for p in (pg, pc):
        p.invalidate()
        p.march()
    d = parity.mesh_diff(pg, g, pc, c)
    assert not d, d[:4]
    assert pg.totals()[1] > 0
