"""The device's numeric operators against the reference's own answers (test_Grid.cpp, fixtures in
tests/golden/grid_kats.json), and against the oracle bit for bit.

Three device paths are checked: the grid view a caller's device functor is compiled against
(felt_surface_stencil*), the update kernel's own speed code (felt_surface_speed_eval, whose
curvature feeds every curvature-driven step), and get() for the index convention."""
import itertools

import numpy as np
import pytest

import felt_amd
import grid_kats
import oracle_lib

pytestmark = pytest.mark.gpu

DATA = grid_kats.load()
KATS = DATA["kats"]
OPS = ["get", "curv", "divergence", "grad", "gradE", "gradF", "gradB", "gradC"]


def same_bits(a, b) -> bool:
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_device_stencils_match_reference_kat(kat):
    fails = grid_kats.check_kat(felt_amd.Surface, kat)
    assert not fails, f"{kat['name']} ({kat['cite']}): {fails}"


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_device_stencils_match_oracle_everywhere(kat):
    """Every operator at every grid position (and a ring outside it, which reads the grid
    background), plus real positions, with dx 1 and 2: identical bits to the oracle."""
    g = grid_kats.make_surface(felt_amd.Surface, kat)
    c = grid_kats.make_surface(oracle_lib.surface, kat)
    D, size = kat["D"], kat["size"]
    off = [-(s // 2) for s in size]
    pos = np.array(list(itertools.product(*[range(o - 1, o + s + 1) for o, s in zip(off, size)])),
                   np.int32)
    rng = np.random.default_rng(17)
    rpos = (rng.random((64, D)) * (np.array(size) + 1) + np.array(off) - 0.5).astype(np.float32)
    rpos[:8] = np.round(rpos[:8] * 2) / 2  # half-voxel positions
    for op in OPS:
        for dx in (1.0, 2.0):
            assert same_bits(g.stencil(op, pos, dx), c.stencil(op, pos, dx)), (op, dx)
            assert same_bits(g.stencil_real(op, rpos, dx), c.stencil_real(op, rpos, dx)), (op, dx)


def test_update_kernel_curvature_matches_reference_kats():
    """The speed code the update kernel runs per zero-layer voxel, with fn = b * curv, gives b
    times the reference's curvature answers (2 and 1 in 2D; 3, 1 and 1.5 in 3D)."""
    n = 0
    for kat in KATS:
        for c in kat["checks"]:
            if c["op"] != "curv":
                continue
            g = grid_kats.make_surface(felt_amd.Surface, kat)
            for b in (1.0, 0.25):
                got = g.speed_eval(felt_amd.speed_curvature(b), [c["pos"]])[0]
                assert got == np.float32(b) * np.float32(c["expect"]), (kat["name"], b, got)
            n += 1
    assert n == 5


def test_update_kernel_speed_matches_oracle_on_a_sphere():
    """All four device functors at every zero-layer voxel of a loaded sphere: the update
    kernel's speed code and the oracle agree bit for bit."""
    from felt_amd import workloads
    size, part = (40, 36, 32), (8, 8, 8)
    phi = workloads.spheres_phi(size, 3, [(2.0, -1.5, 0.5)], [11.3])
    g = felt_amd.Surface(3, 3, size, part)
    c = oracle_lib.surface(3, 3, size, part)
    g.load_dense(phi)
    c.load_dense(phi)
    pos = c.layer_positions(0)
    for fn in (felt_amd.speed_constant(-0.4), felt_amd.speed_curvature(0.15),
               felt_amd.speed_const_curv(-0.3, 0.1),
               felt_amd.speed_noise_curv(-0.3, 0.1, 0.5, 0xFE1702, 16)):
        assert same_bits(g.speed_eval(fn, pos), c.speed_eval(fn, pos)), fn.kind
    # the grid view's curvature (what a device functor calls) is the same number
    assert same_bits(g.stencil("curv", pos), c.speed_eval(felt_amd.speed_curvature(1.0), pos))


def test_index_convention_on_device():
    """Row-major, last axis fastest (test_Grid.cpp:72-89, index 613 pinned in the oracle test):
    a grid of the KAT's size loaded with each voxel's index reads back that index."""
    k = DATA["index"]
    size = k["size"]
    vals = np.arange(np.prod(size), dtype=np.float32).reshape(size)
    g = felt_amd.Surface(3, 3, size, size)
    g.load_dense(vals)
    off = [-(s // 2) for s in size]
    rng = np.random.default_rng(3)
    pos = np.stack([rng.integers(o, o + s, 200) for o, s in zip(off, size)], 1).astype(np.int32)
    want = [((p[0] - off[0]) * size[1] + (p[1] - off[1])) * size[2] + (p[2] - off[2]) for p in pos]
    assert np.array_equal(g.get(pos), np.asarray(want, np.float32))
    assert np.array_equal(g.snapshot(), vals)
