"""HIP surface vs the CPU oracle on seeded workloads, compared after every step: layer
membership, active partitions, isogrid/delta/status-change partition lists bit-exact, phi
within 1e-5 (tolerance PHI_TOL in parity.py), meshes as unordered sets within 1e-5."""
import time

import numpy as np
import pytest

import oracle_lib
import parity
from felt_amd import workloads
from felt_amd._binding import speed_const_curv, speed_constant, speed_curvature, speed_noise_curv

pytestmark = pytest.mark.gpu


def pair(D, L, size, part):
    import felt_amd
    return felt_amd.Surface(D, L, size, part), oracle_lib.surface(D, L, size, part)


def step_both(g, c, how, *args):
    getattr(g, how)(*args)
    getattr(c, how)(*args)


def check(g, c, where, snapshot=True, exact=False):
    d = parity.state_diff(g, c, snapshot=snapshot, exact=exact)
    assert not d, f"{where}: {d}"


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "launches"])
def test_c1_plain_config(graph):
    """128^3 isogrid in 16^3 partitions, singularity seed, 40 steps of -0.5 then 40 of
    0.2*curvature, compared after every step, phi bit for bit (both sides round every operation
    the same way, so the distances are bit-identical, not just within PHI_TOL); then a full march
    compared as meshes.  Run with the step captured as one graph and with kernels launched one by
    one."""
    import felt_amd
    g, c = pair(3, 3, (128,) * 3, (16,) * 3)
    # the FP64 oracle beside it (graph run only): the largest |phi - phi_fp64| is reported
    c64 = oracle_lib.surface(3, 3, (128,) * 3, (16,) * 3, fp64=True) if graph else None
    g.set_graph(graph)
    step_both(g, c, "seed", (0, 0, 0))
    if c64 is not None:
        c64.seed((0, 0, 0))
    check(g, c, "seed")
    t0 = time.time()
    worst64 = 0.0
    for i in range(80):
        sp = speed_constant(-0.5) if i < 40 else speed_curvature(0.2)
        step_both(g, c, "update", sp)
        check(g, c, f"step {i}", exact=True)
        if c64 is not None:
            c64.update(sp)
            if i % 10 == 9:
                worst64 = max(worst64, float(np.abs(g.snapshot().astype(np.float64) -
                                                    c64.snapshot().astype(np.float64)).max()))
        if i % 10 == 9:
            print(f"[c1] step {i + 1} ok, band {g.stats().band_voxels}, "
                  f"{time.time() - t0:.1f}s", flush=True)
    assert c.stats().nan_events == g.stats().nan_events
    if c64 is not None:
        assert worst64 < 1e-4
        print(f"[c1] max |phi - phi_fp64| over the run = {worst64:.3g}", flush=True)
    pg, pc = felt_amd.Polys(g), oracle_lib.polys(c)
    pg.invalidate(); pc.invalidate()
    pg.march(); pc.march()
    d = parity.mesh_diff(pg, g, pc, c)
    assert not d, d
    assert pg.totals()[1] > 0


@pytest.mark.parametrize("seed", [1, 2])
def test_2d_circles_global_and_bounded(seed):
    rng = workloads.SplitMix64(seed)
    size, part, L = (96, 80), (8, 8), 2
    cs = [(rng.uniform(-30, 30), rng.uniform(-25, 25)) for _ in range(4)]
    rs = [rng.uniform(4, 14) for _ in range(4)]
    phi = workloads.spheres_phi(size, L, cs, rs)
    g, c = pair(2, L, size, part)
    step_both(g, c, "load_dense", phi)
    check(g, c, "load")
    speeds = [speed_constant(-0.5), speed_curvature(0.3), speed_const_curv(-0.3, 0.2),
              speed_constant(0.4), speed_noise_curv(-0.3, 0.1, 0.5, 99, 16)]
    for i in range(30):
        sp = speeds[rng.randint(0, len(speeds) - 1)]
        if rng.uniform() < 0.5:
            step_both(g, c, "update", sp)
        else:
            lo = [rng.randint(-50, 40) for _ in range(2)]
            hi = [x + rng.randint(0, 60) for x in lo]
            step_both(g, c, "update_box", lo, hi, sp)
        check(g, c, f"seed {seed} step {i}")
    print(f"[2d] seed {seed} ok, band {g.stats().band_voxels}", flush=True)


def test_3d_spheres_bounded_noise():
    rng = workloads.SplitMix64(7)
    size, part, L = (48, 48, 40), (8, 8, 8), 3
    cs = [tuple(rng.uniform(-15, 15) for _ in range(3)) for _ in range(3)]
    rs = [rng.uniform(4, 10) for _ in range(3)]
    g, c = pair(3, L, size, part)
    step_both(g, c, "load_dense", workloads.spheres_phi(size, L, cs, rs))
    check(g, c, "load")
    for i in range(24):
        sp = speed_noise_curv(-0.3, 0.1, 0.5, 1234, 16) if i % 3 else speed_curvature(0.25)
        if i % 2:
            lo = [rng.randint(-24, 10) for _ in range(3)]
            hi = [x + rng.randint(4, 30) for x in lo]
            step_both(g, c, "update_box", lo, hi, sp)
        else:
            step_both(g, c, "update", sp)
        check(g, c, f"step {i}")
    print(f"[3d] ok, band {g.stats().band_voxels}", flush=True)


def test_uneven_partitions_and_grid_edge():
    """Grid sizes not divisible by the partition size (children grid +1 on every axis,
    PartitionedMixin.hpp:131-141), with the surface growing into the grid boundary."""
    g, c = pair(3, 2, (20, 17, 13), (6, 6, 6))
    step_both(g, c, "seed", (3, 2, 1))
    for i in range(14):
        step_both(g, c, "update", speed_constant(-0.7 if i < 10 else 0.35))
        check(g, c, f"step {i}")


@pytest.mark.parametrize("D,size,part", [(3, (14, 12, 10), (2, 2, 2)),
                                         (3, (13, 11, 9), (1, 4, 2)),
                                         (3, (12, 13, 14), (3, 1, 5)),
                                         (2, (21, 17), (1, 3)),
                                         (2, (20, 18), (2, 5))])
def test_thin_partitions(D, size, part):
    """Partition edges of 1 and 2 (every voxel on a brick face; with an edge of 1 on both faces
    along that axis): the curvature stencil and the expansion reach across bricks on every step."""
    print(f"[thin] D={D} size={size} part={part}", flush=True)
    g, c = pair(D, 2, size, part)
    step_both(g, c, "seed", (1,) * D)
    for i in range(9):
        speed = speed_const_curv(-0.9, 0.1) if i < 6 else speed_curvature(0.3)
        step_both(g, c, "update", speed)
        check(g, c, f"step {i}")


def test_collapse_and_reexpand():
    g, c = pair(3, 3, (40,) * 3, (8,) * 3)
    step_both(g, c, "load_dense", workloads.spheres_phi((40,) * 3, 3, [(0.5, -1.0, 2.0)], [9.0]))
    for i in range(16):
        step_both(g, c, "update", speed_constant(0.9))
        check(g, c, f"contract {i}")
    assert g.layer_size(0) == 0
    step_both(g, c, "seed", (0, 0, 0))
    for i in range(6):
        step_both(g, c, "update", speed_constant(-1.0))
        check(g, c, f"expand {i}")


def test_host_functor_and_low_level_api():
    g, c = pair(2, 3, (24, 24), (4, 4))
    step_both(g, c, "seed", (-3, -3))
    step_both(g, c, "seed", (4, 5))
    pts = {(-3, -2): -1.0, (4, 4): 1.0}
    step_both(g, c, "update", lambda p: pts.get(tuple(p), 0.0))
    check(g, c, "functor")
    for i in range(5):
        z = [tuple(int(x) for x in p) for p in c.layer_positions(0)]
        pick = sorted(z)[i % len(z)]
        for s in (g, c):
            s.update_start()
            s.delta([pick], [-0.8])
            s.update_end()
        check(g, c, f"local {i}")
        for k in range(7):
            a = {tuple(int(x) for x in p) for p in g.affected_positions(k)}
            b = {tuple(int(x) for x in p) for p in c.affected_positions(k)}
            assert a == b, (i, k)


def test_polys_incremental_parity():
    import felt_amd
    rng = workloads.SplitMix64(11)
    size, part, L = (40, 36, 32), (8, 6, 8), 3
    cs = [tuple(rng.uniform(-10, 10) for _ in range(3)) for _ in range(2)]
    g, c = pair(3, L, size, part)
    step_both(g, c, "load_dense", workloads.spheres_phi(size, L, cs, [6.0, 8.5]))
    pg, pc = felt_amd.Polys(g), oracle_lib.polys(c)
    for p in (pg, pc):
        p.invalidate()
        p.march()
    assert not parity.mesh_diff(pg, g, pc, c)
    for i in range(12):
        sp = speed_const_curv(-0.4, 0.2) if i < 8 else speed_constant(0.6)
        if i % 3 == 2:
            step_both(g, c, "update_box", [-10, -10, -10], [5, 8, 3], sp)
        else:
            step_both(g, c, "update", sp)
        pg.notify(); pc.notify()
        if i % 2:
            pg.march(); pc.march()
            d = parity.mesh_diff(pg, g, pc, c)
            assert not d, (i, d)
    print(f"[polys] ok, triangles {pg.totals()[1]}", flush=True)


def test_mesh_order_deterministic():
    """Two surfaces stepped and marched the same way hold the same meshes in the same order:
    every child's vertex and simplex arrays equal element by element, and changes() equal in
    order (the march lists and numbers its work with scans, not atomics)."""
    import felt_amd
    rng = workloads.SplitMix64(23)
    size, part, L = (64, 56, 48), (8, 8, 8), 3
    cs = [tuple(rng.uniform(-15, 15) for _ in range(3)) for _ in range(3)]
    phi = workloads.spheres_phi(size, L, cs, [9.0, 11.5, 7.0])
    gs = []
    for _ in range(2):
        g = felt_amd.Surface(3, L, size, part)
        g.load_dense(phi)
        pg = felt_amd.Polys(g)
        pg.invalidate()
        pg.march()
        gs.append((g, pg))
    for i in range(8):
        sp = speed_const_curv(-0.4, 0.2) if i % 2 else speed_noise_curv(-0.3, 0.1, 0.5, 5, 8)
        for g, pg in gs:
            g.update(sp)
            pg.notify()
            pg.march()
        (ga, pa), (gb, pb) = gs
        assert [int(x) for x in pa.changes()] == [int(x) for x in pb.changes()], f"step {i}"
        for p in range(ga.num_partitions):
            ia, ib = pa.child_info(p), pb.child_info(p)
            assert ia == ib, (i, p)
            if not ia[0]:
                continue
            (va, ta), (vb, tb) = pa.child_mesh(p), pb.child_mesh(p)
            assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)), (i, p)
            assert np.array_equal(ta, tb), (i, p)
    assert pa.totals()[1] > 0


def test_polys_2d_parity():
    import felt_amd
    size, part, L = (64, 48), (8, 8), 2
    g, c = pair(2, L, size, part)
    step_both(g, c, "load_dense", workloads.spheres_phi(size, L, [(-10, 3), (12, -5)], [7.5, 9.0]))
    pg, pc = felt_amd.Polys(g), oracle_lib.polys(c)
    for i in range(6):
        step_both(g, c, "update", speed_const_curv(-0.5, 0.1))
        pg.notify(); pc.notify()
        pg.march(); pc.march()
        d = parity.mesh_diff(pg, g, pc, c)
        assert not d, (i, d)


@pytest.mark.parametrize("D,size,part", [(3, (96, 96, 64), (32, 32, 32)),
                                         (3, (96, 96, 64), (16, 16, 32)),
                                         (2, (384, 256), (128, 128))],
                         ids=["3d-32^3", "3d-16x16x32", "2d-128^2"])
def test_large_partitions(D, size, part):
    """Bricks large enough that several wavefronts share one in the band compaction (8, 2 and 4
    per brick): global and bounded steps compared after every other step, then a full march
    compared as meshes."""
    import felt_amd
    L = 3 if D == 3 else 2
    g, c = pair(D, L, size, part)
    if D == 3:
        cs, rs = [(-20.0, 10.0, 3.0), (18.0, -12.0, -6.0)], [22.0, 15.0]
    else:
        cs, rs = [(-60.0, 20.0), (70.0, -30.0)], [70.0, 45.0]
    step_both(g, c, "load_dense", workloads.spheres_phi(size, L, cs, rs))
    check(g, c, "load")
    for i in range(12):
        sp = speed_const_curv(-0.4, 0.2) if i % 3 else speed_constant(0.5)
        if i % 4 == 3:
            lo = [-s // 4 for s in size]
            hi = [s // 8 for s in size]
            step_both(g, c, "update_box", lo, hi, sp)
        else:
            step_both(g, c, "update", sp)
        if i % 2:
            check(g, c, f"step {i}", snapshot=i >= 9)
    pg, pc = felt_amd.Polys(g), oracle_lib.polys(c)
    pg.invalidate(); pc.invalidate()
    pg.march(); pc.march()
    d = parity.mesh_diff(pg, g, pc, c)
    assert not d, d
    assert pg.totals()[1] > 0
    print(f"[large] {D}D {part}: band {g.stats().band_voxels}, simplices {pg.totals()[1]}",
          flush=True)


@pytest.mark.parametrize("D,L,size,part,cs,rs", [
    (2, 2, (384, 256), (32, 32), [(-60.0, 20.0), (70.0, -30.0)], [70.0, 45.0]),
    (3, 2, (160, 160, 32), (16, 16, 16), [(-30.0, 10.0, 0.0)], [70.0])],
    ids=["2d", "3d"])
def test_expansion_reclaims_half_values(D, L, size, part, cs, rs):
    """A voxel claimed by the expansion with exactly -(L+0.5) is tracked in layer -L, but its
    value rounds outside the band (iround(-(L+0.5) + FLT_EPSILON) = -(L+1) in fp32), so a later
    expansion next to it recomputes its distance (Surface.hpp:1249-1355 tests the value, not
    the lists).  These spheres produce such voxels in the first step."""
    g, c = pair(D, L, size, part)
    step_both(g, c, "load_dense", workloads.spheres_phi(size, L, cs, rs))
    for i in range(3):
        step_both(g, c, "update", speed_constant(0.5))
        check(g, c, f"step {i}")


@pytest.mark.parametrize("D,L,size,part", [(2, 1, (192, 128), (16, 16)),
                                           (3, 1, (96, 96, 48), (16, 16, 16)),
                                           (2, 4, (192, 128), (16, 16)),
                                           (3, 4, (96, 96, 64), (32, 32, 32)),
                                           (2, 5, (192, 128), (8, 8))],
                         ids=["2d-L1", "3d-L1", "2d-L4", "3d-L4-32^3", "2d-L5"])
def test_layer_counts(D, L, size, part):
    """Other narrow-band widths: L=1, and L=4, 5 whose 2L+1 > 8 layers take the band
    compaction's per-layer path; global and bounded steps compared after every step."""
    g, c = pair(D, L, size, part)
    if D == 3:
        cs, rs = [(-10.0, 5.0, 0.0)], [30.0]
    else:
        cs, rs = [(-30.0, 10.0), (35.0, -15.0)], [35.0, 22.0]
    step_both(g, c, "load_dense", workloads.spheres_phi(size, L, cs, rs))
    for i in range(8):
        sp = speed_const_curv(-0.4, 0.2) if i % 2 else speed_constant(0.5)
        if i % 3 == 2:
            step_both(g, c, "update_box", [-s // 4 for s in size], [s // 8 for s in size], sp)
        else:
            step_both(g, c, "update", sp)
        check(g, c, f"L={L} step {i}")


def test_bounded_steps_over_large_affected_sets():
    """Bounded steps whose affected set is large enough (over 65536 voxels beyond the delta list)
    for the device to sort it by layer before the first round's distance passes."""
    size, part, L = (128, 128, 96), (16, 16, 16), 3
    cs, rs = [(-8.0, 5.0, 2.0), (20.0, -12.0, -6.0)], [34.0, 22.0]
    g, c = pair(3, L, size, part)
    step_both(g, c, "load_dense", workloads.spheres_phi(size, L, cs, rs))
    lo, hi = [-64, -64, -48], [63, 63, 47]
    for i in range(4):
        sp = speed_const_curv(-0.4, 0.2) if i % 2 == 0 else speed_curvature(0.3)
        step_both(g, c, "update_box", lo, hi, sp)
        check(g, c, f"bounded step {i}")
    assert g.step_info()[0] > 70000  # voxel updates of the last step: the affected set
    print(f"[bounded-large] ok, affected {g.step_info()[0]}", flush=True)
