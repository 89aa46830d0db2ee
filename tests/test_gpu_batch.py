"""Batched surfaces (felt_batch_*) against the CPU oracle: every member of a batch, stepped
together with the others, must match a separate oracle surface of its size stepped alone on
the same inputs — layers bit-exact, phi within 1e-5 (PHI_TOL), active partition counts equal,
meshes as unordered sets within 1e-5."""
import numpy as np
import pytest

import oracle_lib
import parity
from felt_amd import workloads
from felt_amd._binding import speed_const_curv, speed_constant, speed_curvature, speed_noise_curv

pytestmark = pytest.mark.gpu


def make_batch(D, L, sizes, part, seed, max_spheres=3):
    import felt_amd
    from felt_amd.batch import Batch
    rng = workloads.SplitMix64(seed)
    bt = Batch(felt_amd.lib().lib, D, L, sizes, part)
    oracles = []
    for i, size in enumerate(sizes):
        ns = rng.randint(1, max_spheres)
        e = min(size)
        cs = [tuple(rng.uniform(-s / 2, s / 2) for s in size) for _ in range(ns)]
        rs = [rng.uniform(0.1, 0.3) * e for _ in range(ns)]
        bt.load_spheres(i, cs, rs)
        c = oracle_lib.surface(D, L, size, part)
        c.load_dense(workloads.spheres_phi(size, L, cs, rs))
        oracles.append(c)
    return bt, oracles


def member_diff(bt, i, c, snapshot=True):
    out = []
    for k in range(-bt.L, bt.L + 1):
        gs = {tuple(int(x) for x in p) for p in bt.layer_positions(i, k)}
        cs = {tuple(int(x) for x in p) for p in c.layer_positions(k)}
        if gs != cs:
            out.append(f"member {i} layer {k}: gpu-only {sorted(gs - cs)[:4]} "
                       f"cpu-only {sorted(cs - gs)[:4]} ({len(gs)} vs {len(cs)})")
    if snapshot:
        sg, sc = bt.snapshot(i), c.snapshot()
        d = np.abs(sg.astype(np.float64) - sc.astype(np.float64))
        if d.size and float(d.max()) > parity.PHI_TOL:
            out.append(f"member {i} phi max diff {float(d.max())}")
    if bt.stats(i).active_isogrid_partitions != c.stats().active_isogrid_partitions:
        out.append(f"member {i} active partitions differ")
    return out


def check_all(bt, oracles, where, snapshot=True):
    d = []
    for i, c in enumerate(oracles):
        d += member_diff(bt, i, c, snapshot)
    assert not d, f"{where}: {d[:6]}"


SPEEDS = [speed_constant(-0.4), speed_curvature(0.15), speed_const_curv(-0.3, 0.1),
          speed_noise_curv(-0.3, 0.1, 0.5, 7, 8), speed_constant(0.3)]


def test_batch_global_3d_uneven_sizes():
    """Five 3D members of different sizes (one not a multiple of the partition size), each
    with its own speed, advanced by batch global steps."""
    sizes = [(32, 32, 32), (40, 24, 32), (20, 17, 13), (48, 48, 48), (24, 24, 24)]
    bt, cs = make_batch(3, 3, sizes, (8, 8, 8), 11)
    check_all(bt, cs, "load")
    speeds = [SPEEDS[i % len(SPEEDS)] for i in range(len(sizes))]
    for step in range(10):
        bt.update(speeds)
        for c, sp in zip(cs, speeds):
            c.update(sp)
        check_all(bt, cs, f"step {step}", snapshot=step % 3 == 2)
    check_all(bt, cs, "end")
    ups = sum(c.stats().last_voxel_updates for c in cs)
    assert bt.step_info()[0] == ups
    print(f"[batch3d] ok, voxel updates last step {ups}", flush=True)


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "launches"])
def test_batch_global_2d(graph):
    sizes = [(64, 48), (96, 96), (40, 72), (30, 50)]
    bt, cs = make_batch(2, 2, sizes, (8, 8), 5, max_spheres=4)
    bt.set_graph(graph)
    speeds = [SPEEDS[(i + 1) % len(SPEEDS)] for i in range(len(sizes))]
    for step in range(16):
        bt.update(speeds)
        for c, sp in zip(cs, speeds):
            c.update(sp)
        check_all(bt, cs, f"step {step}")
    print("[batch2d] ok", flush=True)


def test_batch_boxes_subsets():
    """Bounded steps of random member subsets over random boxes in one call; members not in
    a step must stay unchanged."""
    sizes = [(32, 32, 32), (48, 40, 32), (24, 24, 24), (40, 40, 40)]
    bt, cs = make_batch(3, 3, sizes, (8, 8, 8), 23)
    rng = workloads.SplitMix64(99)
    for step in range(12):
        ids, lo, hi, sp = [], [], [], []
        for i, size in enumerate(sizes):
            if rng.uniform() < 0.6:
                ids.append(i)
                l = [rng.randint(-s // 2 - 4, s // 2) for s in size]
                h = [x + rng.randint(0, s) for x, s in zip(l, size)]
                lo.append(l)
                hi.append(h)
                sp.append(SPEEDS[rng.randint(0, 2)])
        bt.update_boxes(ids, lo, hi, sp)
        for i, l, h, s in zip(ids, lo, hi, sp):
            cs[i].update_box(l, h, s)
        check_all(bt, cs, f"step {step} ids {ids}")
    print("[batch boxes] ok", flush=True)


class _MemberPolys:
    """One member's view of a BatchPolys with the Polys query interface."""

    def __init__(self, bp, i):
        self.bp, self.i = bp, i

    def child_info(self, p):
        return self.bp.child_info(self.i, p)

    def child_mesh(self, p):
        return self.bp.child_mesh(self.i, p)

    def changes(self):
        return self.bp.changes(self.i)

    def totals(self):
        nv = ns = 0
        for p in range(self.nparts):
            a, v, s = self.child_info(p)
            nv += v
            ns += s
        return nv, ns


class _MemberGeo:
    def __init__(self, D, num_partitions):
        self.D = D
        self.num_partitions = num_partitions


def test_batch_polys():
    """Full then incremental polygonisation of every member at once, against each member's
    oracle polys."""
    from felt_amd.batch import BatchPolys
    sizes = [(32, 32, 32), (40, 24, 32), (20, 17, 13), (24, 24, 24)]
    part = (8, 8, 8)
    bt, cs = make_batch(3, 3, sizes, part, 31)
    bp = BatchPolys(bt)
    pcs = [oracle_lib.polys(c) for c in cs]
    speeds = [SPEEDS[i % 3] for i in range(len(sizes))]

    def compare(where):
        for i, (c, pc) in enumerate(zip(cs, pcs)):
            mp = _MemberPolys(bp, i)
            mp.nparts = c.num_partitions
            d = parity.mesh_diff(mp, _MemberGeo(3, c.num_partitions), pc, c)
            assert not d, f"{where} member {i}: {d[:4]}"

    bp.invalidate()
    bp.march()
    for pc in pcs:
        pc.invalidate()
        pc.march()
    compare("full march")
    for step in range(6):
        bt.update(speeds)
        bp.notify()
        for c, pc, sp in zip(cs, pcs, speeds):
            c.update(sp)
            pc.notify()
        if step % 2 == 1:
            bp.march()
            for pc in pcs:
                pc.march()
            compare(f"step {step}")
    assert bp.totals()[1] > 0
    print("[batch polys] ok", flush=True)


def test_two_batches_async_overlap():
    """Steps of two batches queued together (two streams) before waiting: every member still
    matches its oracle, and calls between queue and wait are refused."""
    from felt_amd._binding import FeltError
    b8, c8 = make_batch(3, 3, [(32, 32, 32), (24, 40, 24), (40, 40, 40)], (8, 8, 8), 41)
    b16, c16 = make_batch(3, 3, [(48, 48, 48), (64, 32, 48)], (16, 16, 16), 42)
    s8 = [SPEEDS[i % 3] for i in range(3)]
    s16 = [SPEEDS[(i + 1) % 3] for i in range(2)]
    for step in range(8):
        b8.update_async(s8)
        b16.update_async(s16)
        if step == 0:
            with pytest.raises(FeltError):
                b8.snapshot(0)
        b8.wait()
        b16.wait()
        for c, sp in zip(c8, s8):
            c.update(sp)
        for c, sp in zip(c16, s16):
            c.update(sp)
        check_all(b8, c8, f"8^3 step {step}", snapshot=step % 4 == 3)
        check_all(b16, c16, f"16^3 step {step}", snapshot=step % 4 == 3)
    print("[batch async] ok", flush=True)
