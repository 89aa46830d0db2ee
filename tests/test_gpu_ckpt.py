"""Checkpoint / resume of device surfaces in the reference's on-disk format (Surface::save /
load, Surface.hpp:180-196): device archives against the oracle's, archives crossing between
the two, resumed evolution, and batch members evicted and restored."""
import numpy as np
import pytest

import cereal_isogrid as ci
import kat_replay
import oracle_lib
from felt_amd import workloads
from felt_amd._binding import FeltError, Surface, speed_const_curv, speed_constant
from parity import state_diff

pytestmark = pytest.mark.gpu
KATS = [k for k in kat_replay.load_kats() if k["ops"] and k["name"] != "save_load_update"]


def _gpu(D, L, size, part):
    import felt_amd
    return felt_amd.Surface(D, L, size, part)


def _gpu_load(data):
    import felt_amd
    return felt_amd.Surface.load(felt_amd.lib(), data)


def same_archive(a: bytes, b: bytes):
    """Equal once list order (the one free choice of a writer) is made canonical."""
    ga, gb = ci.canonical(ci.parse(a)), ci.canonical(ci.parse(b))
    return ci.write(ga) == ci.write(gb)


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_gpu_archive_equals_oracle_archive(kat):
    print(f"[ckpt] {kat['name']}", flush=True)
    g = kat_replay.replay(_gpu(kat["D"], kat["L"], kat["size"], kat["part"]), kat["ops"])
    c = kat_replay.replay(oracle_lib.surface(kat["D"], kat["L"], kat["size"], kat["part"]),
                          kat["ops"])
    gb, cb = g.save(), c.save()
    assert len(gb) == len(cb)
    assert not ci.check_consistent(ci.parse(gb))
    assert same_archive(gb, cb)


def _spheres(make, size=(48, 40, 32), part=(8, 8, 8), L=3):
    s = make(3, L, size, part)
    s.load_dense(workloads.spheres_phi(size, L, [(-7.0, 3.0, 1.5), (10.0, -6.0, -4.0)],
                                       [9.0, 6.5]))
    return s


def test_gpu_archive_crosses_to_oracle_and_back():
    print("[ckpt] cross", flush=True)
    g, c = _spheres(_gpu), _spheres(oracle_lib.surface)
    for _ in range(3):
        for s in (g, c):
            s.update(speed_const_curv(-0.3, 0.2))
    assert same_archive(g.save(), c.save())
    # oracle archive resumed on the device, device archive resumed on the oracle
    g2 = _gpu_load(c.save())
    c2 = Surface.load(oracle_lib.binding(), g.save())
    for step in range(5):
        sp = speed_const_curv(-0.25, 0.15) if step % 2 else speed_constant(0.4)
        for s in (g, c, g2, c2):
            s.update(sp)
        for x, y in ((g, c), (g2, c), (g, c2)):
            fails = state_diff(x, y)
            assert not fails, (step, fails)
    g2.update_box((-8, -6, -6), (5, 5, 3), speed_constant(-0.4))
    c.update_box((-8, -6, -6), (5, 5, 3), speed_constant(-0.4))
    assert not state_diff(g2, c)


def test_gpu_load_rejects_bad_archives():
    g = _gpu(2, 2, (9, 9), (3, 3))
    g.seed((0, 0))
    g.update(speed_constant(-1.0))
    b = g.save()
    with pytest.raises(FeltError):
        _gpu_load(b[:-5])                                   # truncated
    gp = ci.parse(b)
    k = next(k for k, lst in enumerate(gp["clists"]) if lst)
    gp["clists"][k] = gp["clists"][k][:-1]                  # parent list missing a partition
    with pytest.raises(FeltError):
        _gpu_load(ci.write(gp))
    gp = ci.parse(b)
    gp["bg"] = 7.0                                          # background other than L+1
    with pytest.raises(FeltError):
        _gpu_load(ci.write(gp))
    assert np.array_equal(_gpu_load(b).snapshot(), g.snapshot())


def test_gpu_batch_member_evict_and_restore():
    """Save a member, release its bricks, restore it: the batch then evolves exactly as one that
    was never evicted, and the released bricks return to the pool."""
    print("[ckpt] batch evict/restore", flush=True)
    from test_gpu_batch import make_batch, member_diff
    sizes = [(24, 24, 24), (40, 32, 24), (32, 32, 32)]
    a, oracles = make_batch(3, 3, sizes, (8, 8, 8), seed=9)
    b, _ = make_batch(3, 3, sizes, (8, 8, 8), seed=9)
    speeds = [speed_const_curv(-0.3, 0.1), speed_constant(-0.4), speed_const_curv(0.2, 0.3)]
    for bt in (a, b):
        bt.update(speeds)
    used0 = b.memory()[1]
    data = b.save(1)
    b.clear(1)
    assert b.memory()[1] < used0
    assert np.all(b.snapshot(1) == 4.0) and b.layer_size(1, 0) == 0
    with pytest.raises(FeltError):
        b.load(0, data)                                     # member 0 has another size
    b.load(1, data)
    assert b.memory()[1] == used0
    assert same_archive(b.save(1), a.save(1))
    for _ in range(3):
        for bt in (a, b):
            bt.update(speeds)
    for i in range(len(sizes)):
        assert np.array_equal(a.snapshot(i), b.snapshot(i))
        assert a.layers(i) == b.layers(i)
    # the restored member against the oracle resumed from the same archive
    c = Surface.load(oracle_lib.binding(), data)
    for _ in range(3):
        c.update(speeds[1])
    assert not member_diff(b, 1, c)
