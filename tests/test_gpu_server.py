"""Served surfaces (felt_server_*): requests arriving while others run are batched into the
same steps, and every surface must end exactly as it would alone — against the oracle running
that surface's requests in order (layers bit-exact, phi within 1e-5, meshes as unordered sets
within 1e-5) and against the plain GPU surface doing the same (phi bit-identical)."""
import numpy as np
import pytest

import oracle_lib
import parity
from felt_amd import workloads
from test_gpu_batch import SPEEDS, _MemberGeo, _MemberPolys, member_diff

pytestmark = pytest.mark.gpu


def setup(sizes, part, seed, budget, max_active):
    import felt_amd
    from felt_amd.server import Server
    rng = workloads.SplitMix64(seed)
    srv = Server(felt_amd.lib().lib, 3, 3, sizes, part, budget, max_active, polygonise=True)
    fields = []
    for i, size in enumerate(sizes):
        ns = rng.randint(1, 3)
        e = min(size)
        cs = [tuple(rng.uniform(-s / 2, s / 2) for s in size) for _ in range(ns)]
        rs = [rng.uniform(0.1, 0.3) * e for _ in range(ns)]
        srv.batch.load_spheres(i, cs, rs)
        fields.append((cs, rs))
    srv.start()
    return srv, fields


def make_requests(sizes, n, seed, max_steps):
    from felt_amd.server import request
    rng = workloads.SplitMix64(seed)
    out = []
    for k in range(n):
        i = rng.randint(0, len(sizes) - 1)
        lo, hi = [], []
        for s in sizes[i]:
            side = max(1, int(rng.uniform(0.05, 1.0) * s))
            a = rng.randint(-(s // 2), s - s // 2 - side)
            lo.append(a)
            hi.append(a + side - 1)
        out.append(request(i, lo, hi, SPEEDS[rng.randint(0, 2)], rng.randint(1, max_steps), k))
    return out


def run_mixed(srv, reqs, seed):
    """Submit requests a few at a time between ticks until all have completed."""
    rng = workloads.SplitMix64(seed)
    todo = list(reqs)
    done = set()
    ticks = 0
    while len(done) < len(reqs):
        k = rng.randint(0, 3)
        srv.submit(todo[:k])
        todo = todo[k:]
        srv.tick()
        ticks += 1
        done |= set(srv.poll())
        assert ticks < 10000
    return ticks


def alone(sizes, part, fields, reqs):
    """Each surface's requests in submission order, on the oracle and on a plain GPU surface."""
    import felt_amd
    res = []
    for i, size in enumerate(sizes):
        cs, rs = fields[i]
        c = oracle_lib.surface(3, 3, size, part)
        c.load_dense(workloads.spheres_phi(size, 3, cs, rs))
        g = felt_amd.Surface(3, 3, size, part)
        g.load_spheres(cs, rs)
        pc = oracle_lib.polys(c)
        pc.invalidate()
        pc.march()
        for r in reqs:
            if r.surface != i:
                continue
            lo, hi = list(r.lo)[:3], list(r.hi)[:3]
            for _ in range(r.steps):
                c.update_box(lo, hi, r.speed)
                g.update_box(lo, hi, r.speed)
                pc.notify()
                pc.march()
        res.append((c, pc, g))
    return res


def compare(srv, sizes, res, where):
    for i, (c, pc, g) in enumerate(res):
        d = member_diff(srv.batch, i, c)
        assert not d, f"{where}: {d[:4]}"
        assert np.array_equal(srv.batch.snapshot(i), g.snapshot()), f"{where}: member {i} vs plain"
        mp = _MemberPolys(srv.polys, i)
        mp.nparts = c.num_partitions
        d = parity.mesh_diff(mp, _MemberGeo(3, c.num_partitions), pc, c, changes=False)
        assert not d, f"{where} member {i} mesh: {d[:4]}"


def test_served_mix_equals_alone():
    sizes = [(32, 32, 32), (48, 48, 48), (40, 32, 48), (64, 64, 64), (32, 32, 32), (48, 48, 48)]
    part = (16, 16, 16)
    srv, fields = setup(sizes, part, 3, 4 << 30, 4)
    reqs = make_requests(sizes, 24, 8, 4)
    ticks = run_mixed(srv, reqs, 5)
    st = srv.stats()
    assert st["requests_done"] == len(reqs) and st["requests_queued"] == 0
    assert st["ticks"] == ticks or st["ticks"] <= ticks
    compare(srv, sizes, alone(sizes, part, fields, reqs), "mix")
    print(f"[served] {len(reqs)} requests in {ticks} ticks, p99 step {st['step_p99_ms']:.2f} ms",
          flush=True)


def test_served_budget_defers_admission():
    """A budget that holds the resident bricks but the step lists of only a few requests:
    admissions wait for memory, and every result is still the one it has alone."""
    import felt_amd
    from felt_amd.server import Server
    sizes = [(32, 32, 32), (48, 48, 48), (32, 48, 32), (48, 48, 48)]
    part = (16, 16, 16)
    # a server admitting one request at a time reserves step lists for the largest band: its
    # device bytes are the resident bytes plus those lists; half a band more is the budget
    probe, _ = setup(sizes, part, 4, 4 << 30, 1)
    resident = probe.stats()["device_bytes"]
    bands = [sum(probe.batch.layer_size(i, k) for k in range(-3, 4)) for i in range(len(sizes))]
    probe.close()
    budget = resident + int(max(bands) * 0.5 * 45)
    srv, fields = setup(sizes, part, 4, budget, 8)
    reqs = make_requests(sizes, 16, 12, 3)
    srv.submit(reqs)
    ticks = 0
    while srv.stats()["requests_done"] < len(reqs):
        srv.tick()
        ticks += 1
        assert ticks < 10000
    st = srv.stats()
    assert st["admissions_deferred"] > 0
    assert st["device_bytes"] <= budget
    compare(srv, sizes, alone(sizes, part, fields, reqs), "budget")
    print(f"[served budget] deferred {st['admissions_deferred']} times", flush=True)
