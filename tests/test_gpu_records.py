"""Every active brick's face-neighbour record agrees with the page table after every operation that
activates or deactivates partitions: seeding, loading, global and bounded steps that expand and
collapse the band, member eviction and restore, on grids the partitions divide and on grids they do
not (the +1 rule of PartitionedMixin.hpp:131-141 adds a partition past the grid's end)."""
import numpy as np
import pytest

from felt_amd import workloads
from felt_amd._binding import speed_const_curv, speed_constant

pytestmark = pytest.mark.gpu


def _surface(D, L, size, part):
    import felt_amd
    return felt_amd.Surface(D, L, size, part)


@pytest.mark.parametrize("D,size,part", [(3, (24, 20, 16), (8, 8, 8)),
                                         (3, (40, 40, 40), (8, 8, 8)),
                                         (2, (50, 36), (8, 8)),
                                         (3, (33, 17, 25), (4, 8, 4))])
def test_records_follow_activation_and_deactivation(D, size, part):
    print(f"[records] D={D} size={size} part={part}", flush=True)
    L = 3
    g = _surface(D, L, size, part)
    g.seed([0] * D)
    assert g.check_records() == 0
    # grow until the band reaches the grid's faces, then shrink until it collapses: partitions are
    # activated and deactivated along the way
    for step in range(14):
        g.update(speed_constant(-1.0))
        assert g.check_records() == 0, f"grow step {step}"
    for step in range(40):
        g.update(speed_constant(1.0))
        assert g.check_records() == 0, f"shrink step {step}"
    assert g.stats().band_voxels == 0


def test_records_after_loads_and_bounded_steps():
    print("[records] loads and bounded steps", flush=True)
    L = 3
    size, part = (40, 32, 24), (8, 8, 8)
    cs = [(-6.0, 3.0, 1.0), (9.0, -4.0, -2.0)]
    rs = [7.0, 5.0]
    g = _surface(3, L, size, part)
    g.load_dense(workloads.spheres_phi(size, L, cs, rs))
    assert g.check_records() == 0
    rng = np.random.default_rng(5)
    for step in range(12):
        lo = [int(rng.integers(-s // 2, s // 2 - 8)) for s in size]
        hi = [x + 12 for x in lo]
        g.update_box(lo, hi, speed_const_curv(-0.6 if step < 8 else 0.6, 0.2))
        assert g.check_records() == 0, f"bounded step {step}"


def test_records_in_a_batch_across_evict_and_restore():
    print("[records] batch evict/restore", flush=True)
    from test_gpu_batch import make_batch
    sizes = [(24, 24, 24), (40, 32, 24), (33, 20, 17)]
    bt, _ = make_batch(3, 3, sizes, (8, 8, 8), seed=11)
    speeds = [speed_const_curv(-0.3, 0.1), speed_constant(-0.4), speed_const_curv(0.2, 0.3)]
    for _ in range(3):
        bt.update(speeds)
        assert bt.check_records() == 0
    data = bt.save(1)
    bt.clear(1)
    assert bt.check_records() == 0
    bt.update(speeds)
    assert bt.check_records() == 0
    bt.load(1, data)
    assert bt.check_records() == 0
    for _ in range(3):
        bt.update(speeds)
        assert bt.check_records() == 0
