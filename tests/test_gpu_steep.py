"""Fields steep enough that voxels of opposite sides of the zero layer are face neighbours,
against the oracle: the reference updates every inner layer before any outer one
(Surface.hpp:993-1001), and with such neighbours the order is observable."""
import numpy as np
import pytest

import oracle_lib
from felt_amd import workloads
from felt_amd._binding import speed_const_curv, speed_constant
from parity import state_diff

pytestmark = pytest.mark.gpu


def _gpu(D, L, size, part):
    import felt_amd
    return felt_amd.Surface(D, L, size, part)


def steep_field(size, L, slope):
    """A tilted plane whose values jump by `slope` per voxel along x: with slope > 1 a voxel of
    layer -1 sits next to one of layer +1."""
    grids = workloads.grid_coords(size)
    phi = slope * (grids[0] + 0.5) + 0.13 * grids[1]
    return np.clip(phi, -(L + 1), L + 1).astype(np.float32)


@pytest.mark.parametrize("D,size,part,slope", [(2, (24, 20), (4, 4), 1.2),
                                               (3, (24, 20, 16), (8, 8, 8), 1.3),
                                               (3, (24, 20, 16), (8, 8, 8), 2.1)])
def test_cross_side_neighbours_follow_reference_order(D, size, part, slope):
    print(f"[pairs] D={D} slope={slope}", flush=True)
    L = 3
    phi = steep_field(size, L, slope)
    g, c = _gpu(D, L, size, part), oracle_lib.surface(D, L, size, part)
    g.load_dense(phi)
    c.load_dense(phi)
    for step in range(5):
        sp = speed_const_curv(-0.2, 0.1) if step % 2 else speed_constant(0.3)
        for s in (g, c):
            s.update(sp)
        fails = state_diff(g, c)
        assert not fails, (step, fails)


def test_smooth_fields_follow_reference_order():
    """Ordinary workloads (no cross-side neighbours) over several steps, 3D, batch-sized."""
    size, part, L = (40, 40, 32), (8, 8, 8), 3
    phi = workloads.spheres_phi(size, L, [(-6.0, 4.0, 2.0), (8.0, -7.0, -5.0)], [10.0, 7.0])
    g, c = _gpu(3, L, size, part), oracle_lib.surface(3, L, size, part)
    g.load_dense(phi)
    c.load_dense(phi)
    for step in range(6):
        for s in (g, c):
            s.update(speed_const_curv(-0.35, 0.15))
        fails = state_diff(g, c)
        assert not fails, (step, fails)
