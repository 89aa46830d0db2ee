"""Launch-shape tuning (felt_surface_autotune / felt_batch_autotune): the tuner's choice is one of
its candidates, a cached choice is applied without timing, and every launch shape steps a
surface to bit-identical results, checked against the oracle."""
import numpy as np
import pytest

import oracle_lib
import parity
from felt_amd import workloads
from felt_amd._binding import FeltError, speed_const_curv, speed_constant

pytestmark = pytest.mark.gpu

SHAPES = [(128, 2048, 8), (512, 16384, 64), (64, 1, 16), (1024, 3, 32)]


def spheres(seed, size, n=3):
    rng = workloads.SplitMix64(seed)
    cs = [tuple(rng.uniform(-s / 3, s / 3) for s in size) for _ in range(n)]
    rs = [rng.uniform(0.1, 0.25) * min(size) for _ in range(n)]
    return cs, rs


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{t}x{g}v{v}" for t, g, v in SHAPES])
def test_launch_shapes_step_identically(shape):
    """A surface stepped with each shape equals the oracle bit for bit after every step."""
    import felt_amd
    size, part, L = (64, 56, 48), (16, 16, 16), 3
    cs, rs = spheres(0x7E57, size)
    g = felt_amd.Surface(3, L, size, part)
    c = oracle_lib.surface(3, L, size, part)
    g.load_spheres(cs, rs)
    c.load_dense(workloads.spheres_phi(size, L, cs, rs))
    g.set_launch_config(*shape)
    for step in range(12):
        sp = speed_constant(-0.5) if step < 6 else speed_const_curv(0.2, 0.3)
        g.update(sp)
        c.update(sp)
        d = parity.state_diff(g, c, exact=True)
        assert not d, f"shape {shape} step {step}: {d[:4]}"


def test_set_launch_config_rejects_bad_shapes():
    import felt_amd
    g = felt_amd.Surface(3, 2, (32,) * 3, (8,) * 3)
    for bad in [(96 + 1, 4096, 8), (2048, 4096, 8), (256, 0, 8), (256, 4096, 12)]:
        with pytest.raises(FeltError):
            g.set_launch_config(*bad)


def test_batch_autotune_choice_and_cache():
    """felt_batch_autotune picks one of its candidates and never reports a slower time than
    the default it measured; a second batch of the same shape and band picks the cached choice
    up with reps == 0; both batches then step to identical snapshots."""
    import felt_amd
    from felt_amd.batch import Batch
    sizes = [(48, 48, 48), (64, 40, 56), (40, 40, 40)]
    part = (8, 8, 8)
    bts = []
    for _ in range(2):
        bt = Batch(felt_amd.lib().lib, 3, 3, sizes, part)
        for i, size in enumerate(sizes):
            cs, rs = spheres(100 + i, size)
            bt.load_spheres(i, cs, rs)
        bts.append(bt)
    sps = [speed_const_curv(-0.3, 0.1)] * len(sizes)
    for bt in bts:
        for _ in range(3):
            bt.update(sps)
        bt.sync()
    cfg = bts[0].autotune(reps=3)
    assert cfg["threads_per_block"] in (128, 256, 512)
    assert cfg["grid_cap"] in (2048, 4096, 8192, 16384)
    assert cfg["voxels_per_lane"] in (8, 16, 32, 64)
    # 8^3 bricks: one wavefront scans 64 * voxels_per_lane voxels, at most 8 per brick
    assert cfg["threads_per_brick"] == 64 * max(1, min(8, 512 // (64 * cfg["voxels_per_lane"])))
    assert 0 < cfg["us_tuned"] <= cfg["us_default"]
    cached = bts[1].autotune(reps=0)
    for k in ("threads_per_block", "grid_cap", "voxels_per_lane", "threads_per_brick"):
        assert cached[k] == cfg[k], k
    assert cached["us_default"] == 0 and cached["us_tuned"] == 0
    bts[1].set_launch_config(256, 4096, 8)  # the default shape on the second batch
    for _ in range(8):
        for bt in bts:
            bt.update(sps)
    for i in range(len(sizes)):
        a, b = bts[0].snapshot(i), bts[1].snapshot(i)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"member {i}"
    for bt in bts:
        bt.close()
