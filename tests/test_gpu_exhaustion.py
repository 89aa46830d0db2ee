"""Brick-pool exhaustion and request failure (PartitionedMixin.hpp:201-213; GridMixin.hpp:75-81).

A step whose worst-case partition activations exceed the free bricks is refused before it
writes anything: FELT_E_NOMEM, and every member's values, layers and active partitions are what
they were (the oracle, never stepped, still matches).  A served request that cannot run fails
alone; the rest of its tick runs, and the server keeps serving."""
import numpy as np
import pytest

import felt_amd
import oracle_lib
from felt_amd import workloads
from felt_amd._binding import FELT_E_ARG, FELT_E_NOMEM, FeltError, speed_constant
from test_gpu_batch import check_all

pytestmark = pytest.mark.gpu

L = 3
PART = (8, 8, 8)


def band_partitions(phi, part):
    """Partitions of a dense field that hold a band voxel (load activates exactly these)."""
    t = phi.astype(np.float32) + np.float32(np.finfo(np.float32).eps)
    k = np.sign(t) * np.floor(np.abs(t) + 0.5)  # round half away from zero
    band = np.abs(k) <= L
    s = [n // p for n, p in zip(phi.shape, part)]
    b = band.reshape(s[0], part[0], s[1], part[1], s[2], part[2]).any(axis=(1, 3, 5))
    return int(b.sum())


def loaded_batch(sizes, centres, radii, max_bricks):
    from felt_amd.batch import Batch
    bt = Batch(felt_amd.lib().lib, 3, L, sizes, PART, max_bricks=max_bricks)
    oracles = []
    for i, size in enumerate(sizes):
        bt.load_spheres(i, [centres[i]], [radii[i]])
        c = oracle_lib.surface(3, L, size, PART)
        c.load_dense(workloads.spheres_phi(size, L, [centres[i]], [radii[i]]))
        oracles.append(c)
    return bt, oracles


def test_global_step_refused_when_pool_is_short():
    sizes = [(32, 32, 32), (40, 32, 32)]
    centres, radii = [(0.5, -1.0, 0.25), (-3.0, 2.0, 1.5)], [6.3, 5.2]
    need = sum(band_partitions(workloads.spheres_phi(s, L, [c], [r]), PART)
               for s, c, r in zip(sizes, centres, radii))
    bt, oracles = loaded_batch(sizes, centres, radii, max_bricks=need + 1)
    _, used, total = bt.memory()
    assert (used, total) == (need, need + 1)
    before = [bt.snapshot(i) for i in range(2)]
    for _ in range(2):  # refused every time, never half-applied
        with pytest.raises(FeltError) as e:
            bt.update([speed_constant(-0.5), speed_constant(0.3)])
        assert e.value.code == FELT_E_NOMEM
        for i in range(2):
            assert np.array_equal(bt.snapshot(i), before[i])
        check_all(bt, oracles, "after a refused global step")
    assert bt.memory()[1:] == (need, need + 1)


def test_bounded_step_refused_when_pool_is_short():
    sizes = [(32, 32, 32), (32, 32, 40)]
    centres, radii = [(0.5, -1.0, 0.25), (2.0, 2.0, -3.5)], [6.3, 7.1]
    need = sum(band_partitions(workloads.spheres_phi(s, L, [c], [r]), PART)
               for s, c, r in zip(sizes, centres, radii))
    bt, oracles = loaded_batch(sizes, centres, radii, max_bricks=need + 1)
    before = [bt.snapshot(i) for i in range(2)]
    with pytest.raises(FeltError) as e:
        bt.update_boxes([0, 1], [-16, -16, -16, -16, -16, -20], [15, 15, 15, 15, 15, 19],
                        [speed_constant(-0.5), speed_constant(-0.5)])
    assert e.value.code == FELT_E_NOMEM
    for i in range(2):
        assert np.array_equal(bt.snapshot(i), before[i])
    check_all(bt, oracles, "after a refused bounded step")
    # a box with no zero-layer voxel cannot activate anything: that step runs (and changes
    # nothing), as the oracle's does
    bt.update_boxes([0], [12, 12, 12], [15, 15, 15], [speed_constant(-0.5)])
    oracles[0].update_box([12, 12, 12], [15, 15, 15], speed_constant(-0.5))
    check_all(bt, oracles, "after an empty bounded step")


def _brick_budget(bricks):
    """Smallest server budget that gives exactly `bricks` bricks (three quarters of the budget
    in bricks of 8^3 voxels: 4 B value, 1 B layer, 1 B mark, 32 B of slot tables)."""
    brick_bytes = 512 * 6 + 32
    b = bricks * brick_bytes * 4 // 3
    while b // 4 * 3 // brick_bytes < bricks:
        b += 1
    return b


def test_server_fails_a_request_alone_and_keeps_serving():
    from felt_amd.server import Server, request
    n = 40
    sizes = [(32, 32, 32)] * n
    rng = workloads.SplitMix64(0xFE1799)
    centres = [tuple(rng.uniform(-3, 3) for _ in range(3)) for _ in range(n)]
    radii = [rng.uniform(3.5, 4.5) for _ in range(n)]
    need = sum(band_partitions(workloads.spheres_phi(s, L, [c], [r]), PART)
               for s, c, r in zip(sizes, centres, radii))
    v = Server(felt_amd.lib().lib, 3, L, sizes, PART, _brick_budget(need + 2), max_active=8,
               polygonise=False)
    for i in range(n):
        v.batch.load_spheres(i, [centres[i]], [radii[i]])
    v.start()
    st = v.stats()
    assert st["bricks_total"] == need + 2 and st["bricks_used"] == need
    before = v.batch.snapshot(0)
    # surface 0 grows: its step may activate more partitions than the two free bricks; surface 1
    # is stepped over a corner box without zero-layer voxels, which activates nothing
    v.submit([request(0, [-16] * 3, [15] * 3, speed_constant(-0.5), 3, 100),
              request(1, [12] * 3, [15] * 3, speed_constant(-0.5), 2, 101)])
    assert v.tick() == 2
    assert v.poll_failed() == [100]
    assert np.array_equal(v.batch.snapshot(0), before)
    assert v.tick() == 1
    assert v.poll() == [101]
    assert v.stats()["requests_failed"] == 1
    # surface 0 is free again: a new request on it is admitted (and fails the same way)
    v.submit([request(0, [-16] * 3, [15] * 3, speed_constant(-0.5), 1, 102)])
    assert v.tick() == 1
    assert v.poll_failed() == [102]
    v.close()


def test_server_submit_is_all_or_nothing_and_cancel():
    from felt_amd.server import Server, request
    sizes = [(32, 32, 32)] * 3
    v = Server(felt_amd.lib().lib, 3, L, sizes, PART, 1 << 30, max_active=2, polygonise=False)
    for i in range(3):
        v.batch.load_spheres(i, [(0.5, 0.5, 0.5)], [5.0 + i])
    v.start()
    good = request(0, [-16] * 3, [15] * 3, speed_constant(-0.3), 2, 1)
    bad = request(7, [-16] * 3, [15] * 3, speed_constant(-0.3), 2, 2)  # no surface 7
    with pytest.raises(FeltError) as e:
        v.submit([good, bad])
    assert e.value.code == FELT_E_ARG
    assert v.stats()["requests_queued"] == 0
    v.submit([good, request(1, [-16] * 3, [15] * 3, speed_constant(-0.3), 3, 3),
              request(2, [-16] * 3, [15] * 3, speed_constant(-0.3), 3, 4)])
    v.cancel(4)  # still queued
    assert v.tick() == 2
    v.cancel(3)  # in flight after one step
    assert v.stats()["requests_active"] == 1
    with pytest.raises(FeltError):
        v.cancel(99)
    assert v.tick() == 1
    assert v.poll() == [1]
    assert v.poll_failed() == []
    assert v.tick() == 0
    v.close()
