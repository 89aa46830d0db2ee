"""Groups of batches (felt_group_*): batches of two partition sizes stepped by one graph launch
per step end bit-identical to the same batches stepped one by one, and to the oracle; the group
rebuilds its parent graph when a batch recaptures its step, and keeps the batches' state rules."""
import numpy as np
import pytest

import oracle_lib
import parity
from felt_amd import workloads
from felt_amd._binding import FeltError, speed_const_curv, speed_constant, speed_curvature

pytestmark = pytest.mark.gpu

SIZES = {8: [(40, 48, 40), (56, 40, 48), (48, 48, 48)], 16: [(96, 80, 96), (112, 96, 96)]}
SPEEDS = [speed_constant(-0.4), speed_curvature(0.15), speed_const_curv(-0.3, 0.1)]


def make(seed):
    import felt_amd
    from felt_amd.batch import Batch
    rng = workloads.SplitMix64(seed)
    out, members = [], []
    for p, sizes in sorted(SIZES.items()):
        bt = Batch(felt_amd.lib().lib, 3, 3, sizes, (p,) * 3)
        for i, size in enumerate(sizes):
            ns = rng.randint(1, 3)
            cs = [tuple(rng.uniform(-s / 3, s / 3) for s in size) for _ in range(ns)]
            rs = [rng.uniform(0.1, 0.25) * min(size) for _ in range(ns)]
            bt.load_spheres(i, cs, rs)
            members.append((p, size, cs, rs, SPEEDS[(i + p) % len(SPEEDS)]))
        out.append(bt)
    return out, members


def speeds_of(bts, members):
    k, out = 0, []
    for bt in bts:
        out.append([m[4] for m in members[k:k + bt.n]])
        k += bt.n
    return out


def test_group_equals_separate_batches_and_oracle():
    from felt_amd.batch import Group
    ga, members = make(0x6A0)
    gb, _ = make(0x6A0)
    sp = speeds_of(ga, members)
    grp = Group(ga)
    oracles = []
    for p, size, cs, rs, f in members:
        c = oracle_lib.surface(3, 3, size, (p,) * 3)
        c.load_dense(workloads.spheres_phi(size, 3, cs, rs))
        oracles.append((c, f))
    for step in range(10):
        grp.update(sp)
        for bt, s in zip(gb, sp):
            bt.update(s)
        for c, f in oracles:
            c.update(f)
    k = 0
    for bt_g, bt_s in zip(ga, gb):
        for i in range(bt_g.n):
            a, b = bt_g.snapshot(i), bt_s.snapshot(i)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"member {k}"
            c = oracles[k][0]
            assert np.array_equal(a.view(np.uint32), c.snapshot().view(np.uint32)), \
                f"member {k} vs oracle"
            for layer in range(-3, 4):
                got = parity.position_keys(bt_g.layer_positions(i, layer), bt_g.sizes[i])
                want = parity.position_keys(c.layer_positions(layer), c.size)
                assert np.array_equal(got, want), f"member {k} layer {layer}"
            k += 1
    grp.close()
    for bt in ga + gb:
        bt.close()


def test_group_recaptures_and_state_rules():
    """Changing a batch's launch shape invalidates its step graph; the group's next step must
    pick the new capture up (results unchanged).  While a group step is in flight the batches
    and the group refuse other calls."""
    from felt_amd.batch import Group
    ga, members = make(0x6A1)
    gb, _ = make(0x6A1)
    sp = speeds_of(ga, members)
    grp = Group(ga)
    for step in range(6):
        if step == 3:
            ga[1].set_launch_config(128, 2048, 16)
        grp.update_async(sp)
        with pytest.raises(FeltError):
            ga[0].update(sp[0])
        with pytest.raises(FeltError):
            grp.update_async(sp)
        grp.wait()
        for bt, s in zip(gb, sp):
            bt.update(s)
    for bt_g, bt_s in zip(ga, gb):
        for i in range(bt_g.n):
            assert np.array_equal(bt_g.snapshot(i).view(np.uint32),
                                  bt_s.snapshot(i).view(np.uint32))
    grp.close()
    for bt in ga + gb:
        bt.close()


def test_group_rejects_duplicates():
    from felt_amd.batch import Group
    ga, _ = make(0x6A2)
    with pytest.raises(FeltError):
        Group([ga[0], ga[0]])
    for bt in ga:
        bt.close()
