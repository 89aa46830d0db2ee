"""The benchmark workloads against the oracle.

- Single large surface (512^3 in 32^3, harmonic sphere r = 150, noise + curvature speed, 200
  steps, incremental marching cubes every 10) at a quarter of its lengths: layers, active
  partitions and partition lists compared after every step, phi every 10 steps, meshes at steps
  10, 100, 200; the FP64 oracle runs the same steps and the largest |phi - phi_fp64| is reported.
- Batched: the real 256-surface mixed workload (mixed_surfaces(256, 0xFE1703, 32, 256), both
  partition sizes) for 10 steps, every member against its own oracle: layers at steps 1, 2, 4, 7,
  10 and phi after step 10.
- Served: the first 16 surfaces of the served stream (mixed_surfaces(1000, 0xFE1704, 64, 384,
  16^3 partitions)) with the config's Poisson requests (2000/s), run through the server with
  requests joining while others are in flight; every surface must equal the oracle and the plain
  GPU surface running its requests alone.

Oracles of independent surfaces run in spawned worker processes (tests/oracle_pool.py) and report
digests of sorted layer keys and of snapshots; the GPU side computes the same digests."""
import time

import numpy as np
import pytest

import oracle_pool
import progress
import parity
from felt_amd import workloads

pytestmark = pytest.mark.gpu


def _fail_digest(where, g, c, L):
    """Which layers differ between two layer_digests tuples."""
    bad = [k - L for k in range(2 * L + 1) if g[k] != c[k]]
    return f"{where}: layers {bad} differ (gpu counts {[x[0] for x in g]}, cpu {[x[0] for x in c]})"


@pytest.mark.timeout(600)
def test_single_large_surface_quarter_scale():
    import felt_amd
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    cfg = oracle_pool.C2Q
    with ProcessPoolExecutor(2, mp_context=mp.get_context("spawn")) as ex:
        f32, f64 = ex.submit(oracle_pool.c2q_job, False), ex.submit(oracle_pool.c2q_job, True)
        g = felt_amd.Surface(3, 3, cfg["size"], cfg["part"])
        g.load_dense(oracle_pool.c2q_field())
        pg = felt_amd.Polys(g)
        pg.invalidate()
        pg.march()
        sp = oracle_pool.c2q_speed()
        states, phis, meshes = [], {}, {}
        t0 = time.time()
        for step in range(1, cfg["steps"] + 1):
            g.update(sp)
            states.append(oracle_pool.state_digest(g))
            pg.notify()
            if step % cfg["march_every"] == 0:
                pg.march()
                phis[step] = g.snapshot()
            if step in cfg["mesh_steps"]:
                meshes[step] = parity.mesh_sets(pg, g)
        progress.say(f"[c2/4] gpu done in {time.time() - t0:.0f}s, band {g.stats().band_voxels}")
        c32, c64 = progress.wait(f32, "c2/4"), progress.wait(f64, "c2/4 fp64")
    for step, (gs, cs) in enumerate(zip(states, c32["state"]), 1):
        assert gs[0] == cs[0], _fail_digest(f"step {step}", gs[0], cs[0], 3)
        assert gs[1] == cs[1], f"step {step}: active partitions or partition lists differ"
    worst64, identical = 0.0, 0
    for step, phi in phis.items():
        d32 = np.abs(phi.astype(np.float64) - c32["phi"][step].astype(np.float64)).max()
        assert d32 <= parity.PHI_TOL, f"step {step}: phi max diff {d32}"
        identical += int(np.array_equal(phi, c32["phi"][step]))
        worst64 = max(worst64, float(np.abs(phi.astype(np.float64) -
                                            c64["phi"][step].astype(np.float64)).max()))
    for step, m in meshes.items():
        d = parity.mesh_sets_diff(m, c32["mesh"][step], 3)
        assert not d, f"mesh at step {step}: {d[:4]}"
    assert worst64 < 1e-4
    # both sides round every operation the same way: the distances are bit-identical
    assert identical == len(phis), f"phi bit-identical at {identical}/{len(phis)} checks only"
    print(f"[c2/4] 200 steps: layers and lists equal every step, phi bit-identical at "
          f"{identical}/{len(phis)} checks, max |phi - phi_fp64| = {worst64:.3g}")


@pytest.mark.timeout(600)
def test_batched_mixed_workload_256():
    import felt_amd
    from felt_amd.batch import Batch
    n, seed, lo, hi, steps = 256, 0xFE1703, 32, 256, 10
    checkpoints = (1, 2, 4, 7, 10)
    specs = workloads.mixed_surfaces(n, seed, lo, hi)
    # the oracle members, largest first so the workers stay balanced
    order = sorted(range(n), key=lambda i: -int(np.prod(specs[i].size)))
    chunks = [order[w::oracle_pool.workers()] for w in range(oracle_pool.workers())]
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    with ProcessPoolExecutor(len(chunks), mp_context=mp.get_context("spawn")) as ex:
        futs = [ex.submit(oracle_pool.mixed_member_job, (n, seed, lo, hi, ch, steps, checkpoints))
                for ch in chunks]
        groups = {}
        for i, sp in enumerate(specs):
            groups.setdefault(sp.part, []).append(i)
        batches = []
        for part, ids in sorted(groups.items()):
            bt = Batch(felt_amd.lib().lib, 3, 3, [specs[i].size for i in ids], part)
            for j, i in enumerate(ids):
                bt.load_spheres(j, specs[i].centres, specs[i].radii)
            batches.append((bt, ids, [workloads.speed_of(specs[i].speed) for i in ids]))
        got = {}
        t0 = time.time()
        for step in range(1, steps + 1):
            for bt, ids, sps in batches:
                bt.update(sps)
            if step in checkpoints:
                for bt, ids, _ in batches:
                    for j, i in enumerate(ids):
                        got[(i, step)] = oracle_pool.layer_digests(
                            lambda k: bt.layer_positions(j, k), specs[i].size, 3)
        for bt, ids, _ in batches:
            for j, i in enumerate(ids):
                got[(i, "phi")] = oracle_pool.phi_digest(bt.snapshot(j))
        progress.say(f"[c3] gpu {steps} steps of {n} surfaces in {len(batches)} batches, "
                     f"{time.time() - t0:.0f}s with digests")
        want = {}
        for f in futs:
            want.update(progress.wait(f, "c3"))
    bad = [key for key in want if got[key] != want[key]]
    assert not bad, f"{len(bad)} member checkpoints differ, first {sorted(bad, key=str)[:6]}"
    assert len(want) == n * (len(checkpoints) + 1)
    progress.say(f"[c3] all {n} members equal their oracles (layers at {checkpoints}, "
                 f"phi at {steps})")


@pytest.mark.timeout(600)
def test_served_stream_subset():
    import felt_amd
    from felt_amd.server import Server, request
    specs = workloads.mixed_surfaces(1000, 0xFE1704, 64, 384, fixed_part=16)[:16]
    arrivals = workloads.served_requests(specs, 0.012, 2000.0, 0xFE1704 ^ 0x5EED)
    assert len(arrivals) >= 12
    per_surface = [[] for _ in specs]
    for _, i, lo, hi, k, steps in arrivals:
        per_surface[i].append((lo, hi, k, steps))
    jobs = [(specs[i], per_surface[i]) for i in range(len(specs))]
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    with ProcessPoolExecutor(oracle_pool.workers(), mp_context=mp.get_context("spawn")) as ex:
        futs = [ex.submit(oracle_pool.served_surface_job, j) for j in jobs]
        part = specs[0].part
        srv = Server(felt_amd.lib().lib, 3, 3, [s.size for s in specs], part, 64 << 30,
                     max_active=8, polygonise=True)
        for i, s in enumerate(specs):
            srv.batch.load_spheres(i, s.centres, s.radii)
        srv.start()
        reqs = [request(i, lo, hi, workloads.speed_of(k), steps, rid)
                for rid, (_, i, lo, hi, k, steps) in enumerate(arrivals)]
        # arrivals in order, two per tick, so requests join while others are in flight
        todo, done, ticks = list(reqs), set(), 0
        while len(done) < len(reqs):
            srv.submit(todo[:2])
            todo = todo[2:]
            srv.tick()
            ticks += 1
            done |= set(srv.poll())
            assert ticks < 100000 and not srv.poll_failed()
        st = srv.stats()
        # each surface alone on a plain GPU surface, with its own polygoniser
        for i, s in enumerate(specs):
            g = felt_amd.Surface(3, 3, s.size, s.part)
            g.load_spheres(s.centres, s.radii)
            pg = felt_amd.Polys(g)
            pg.invalidate()
            pg.march()
            for lo, hi, k, steps in per_surface[i]:
                for _ in range(steps):
                    g.update_box(lo, hi, workloads.speed_of(k))
                    pg.notify()
                    pg.march()
            assert np.array_equal(srv.batch.snapshot(i), g.snapshot()), f"surface {i} vs alone"
            from test_gpu_batch import _MemberGeo, _MemberPolys
            mp_ = _MemberPolys(srv.polys, i)
            d = parity.mesh_sets_diff(parity.mesh_sets(mp_, _MemberGeo(3, g.num_partitions)),
                                      parity.mesh_sets(pg, g), 3)
            assert not d, f"surface {i} mesh vs alone: {d[:4]}"
            g.close()
        want = [progress.wait(f, "c4") for f in futs]
    for i, s in enumerate(specs):
        layers = oracle_pool.layer_digests(lambda k: srv.batch.layer_positions(i, k), s.size, 3)
        assert layers == want[i][0], _fail_digest(f"surface {i}", layers, want[i][0], 3)
        assert oracle_pool.phi_digest(srv.batch.snapshot(i)) == want[i][1], f"surface {i} phi"
    print(f"[c4] {len(reqs)} requests over {len(specs)} surfaces in {ticks} ticks "
          f"(p99 tick {st['step_p99_ms']:.1f} ms): every surface equals its oracle and its "
          f"plain run alone")
