"""The oracle's Surface::save / load against the reference's on-disk format, read independently
from its serialisation code (tests/cereal_isogrid.py), over every known-answer state."""
import numpy as np
import pytest

import cereal_isogrid as ci
import kat_replay
import oracle_lib
from felt_amd._binding import Surface, speed_const_curv, speed_constant
from felt_amd.workloads import spheres_phi

KATS = [k for k in kat_replay.load_kats() if k["ops"]]


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_oracle_archive_matches_format(kat):
    s = oracle_lib.surface(kat["D"], kat["L"], kat["size"], kat["part"])
    s = kat_replay.replay(s, kat["ops"])
    b = s.save()
    g = ci.parse(b)                     # consumes every byte
    assert ci.write(g) == b             # the independent writer agrees byte for byte
    assert not ci.check_consistent(g)
    assert g["sizes"][0] == list(kat["size"]) and g["child_size"] == list(kat["part"])
    assert g["sizes"][1] == [-(x // 2) for x in kat["size"]]
    assert len(g["clists"]) == 2 * kat["L"] + 1
    t = Surface.load(oracle_lib.binding(), b)
    assert t.save() == b
    assert np.array_equal(t.snapshot(), s.snapshot())


def test_oracle_archive_size_closed_form():
    """Byte count of an unseeded 7x7 L=2 surface with 3x3 partitions from the layout: 9
    inactive children (children grid 3x3), no list entries."""
    s = oracle_lib.surface(2, 2, (7, 7), (3, 3))
    b = s.save()
    V, N, np_ = 16 + 8, 5, 9
    lists0 = 16 + 8 * N
    child = 4 + 8 + 8 + 8 + lists0 + 6 * V
    total = 4 + V + child + 8 + np_ * child + (16 + 8 * N) + 8 + np_ * (16 + 8 * N) + lists0 \
        + 6 * V + 3 * V + 4 + 2 * V
    assert len(b) == total


def test_oracle_resume_continues_identically():
    """A surface loaded from an archive evolves exactly as the original."""
    size, part, L = (40, 32, 24), (8, 8, 8), 3
    phi = spheres_phi(size, L, [(-5.0, 2.0, 1.0), (9.0, -6.0, -3.0)], [8.0, 5.5])
    a = oracle_lib.surface(3, L, size, part)
    a.load_dense(phi)
    for _ in range(3):
        a.update(speed_const_curv(-0.2, 0.1))
    b = Surface.load(oracle_lib.binding(), a.save())
    for _ in range(4):
        for s in (a, b):
            s.update(speed_const_curv(-0.25, 0.15))
        assert np.array_equal(a.snapshot(), b.snapshot())
        assert a.layers() == b.layers()
    b.update_box((-6, -6, -6), (4, 4, 4), speed_constant(0.3))
    a.update_box((-6, -6, -6), (4, 4, 4), speed_constant(0.3))
    assert np.array_equal(a.snapshot(), b.snapshot())


def test_consistency_check_detects_planted_mistake():
    s = kat_replay.replay(oracle_lib.surface(2, 2, (9, 9), (3, 3)),
                          [["seed", [0, 0]], ["update", -1.0]])
    g = ci.parse(s.save())
    p, k = next((i, k) for i, c in enumerate(g["children"])
                for k, lst in enumerate(c["lists"]) if len(lst) >= 2)
    g["children"][p]["lists"][k] = g["children"][p]["lists"][k][1:] + [g["children"][p]["lists"][k][0]]
    assert ci.check_consistent(g)
