"""Polygonisation known answers of the reference (test_Poly.cpp) on the GPU and the oracle."""
import numpy as np
import pytest

import oracle_lib
from felt_amd._binding import speed_constant

pytestmark = pytest.mark.gpu


def both():
    import felt_amd
    return [("gpu", felt_amd.Surface, felt_amd.Polys),
            ("cpu", oracle_lib.surface, oracle_lib.polys)]


@pytest.mark.parametrize("which", ["gpu", "cpu"])
def test_single_partition_octahedron(which):
    """test_Poly.cpp:49-103: 6 vertices at +-0.4 with axis normals, 8 triangles."""
    mk_s, mk_p = {n: (s, p) for n, s, p in both()}[which]
    s = mk_s(3, 3, (9, 9, 9), (3, 3, 3))
    s.seed((0, 0, 0))
    s.update(speed_constant(-0.4))
    p = mk_p(s)
    centre = s.child_index((0, 0, 0))
    p.mark(centre)
    p.march()
    vtx, spx = p.child_mesh(centre)
    assert len(vtx) == 6 and len(spx) == 8
    want = {(0.4, 0, 0), (0, 0, 0.4), (0, 0.4, 0), (-0.4, 0, 0), (0, -0.4, 0), (0, 0, -0.4)}
    got = {tuple(round(float(x), 5) + 0.0 for x in v[:3]) for v in vtx}
    assert got == {tuple(round(x, 5) + 0.0 for x in w) for w in want}
    for v in vtx:
        n = v[3:]
        axis = int(np.argmax(np.abs(v[:3])))
        assert abs(abs(n[axis]) - 1) < 1e-6 and np.sign(n[axis]) == np.sign(v[axis])
    if which == "cpu":
        # exact vertex and index order of the reference (test_Poly.cpp:75-103)
        order = [(0.4, 0, 0), (0, 0, 0.4), (0, 0.4, 0), (-0.4, 0, 0), (0, -0.4, 0), (0, 0, -0.4)]
        assert np.allclose(vtx[:, :3], np.array(order), atol=1e-6)
        assert spx.tolist() == [[1, 0, 2], [1, 2, 3], [1, 4, 0], [0, 5, 2], [4, 1, 3],
                                [3, 2, 5], [0, 4, 5], [4, 3, 5]]


@pytest.mark.parametrize("which", ["gpu", "cpu"])
def test_single_partition_diamond_2d(which):
    """test_Poly.cpp:167-211: 4 vertices, 4 segments."""
    mk_s, mk_p = {n: (s, p) for n, s, p in both()}[which]
    s = mk_s(2, 3, (9, 9), (3, 3))
    s.seed((0, 0))
    s.update(speed_constant(-0.4))
    p = mk_p(s)
    p.mark(s.child_index((0, 0)))
    p.march()
    vtx, spx = p.child_mesh(s.child_index((0, 0)))
    assert len(vtx) == 4 and len(spx) == 4
    if which == "cpu":
        assert np.allclose(vtx, [[0.4, 0], [0, 0.4], [-0.4, 0], [0, -0.4]], atol=1e-6)
        assert spx.tolist() == [[0, 1], [1, 2], [3, 0], [2, 3]]


@pytest.mark.parametrize("which", ["gpu", "cpu"])
def test_polys_counts_and_changes(which):
    """test_Poly.cpp:278-442: 30 vertices / 56 simplices in the centre, changed partitions."""
    mk_s, mk_p = {n: (s, p) for n, s, p in both()}[which]
    s = mk_s(3, 3, (15, 15, 15), (5, 5, 5))
    p = mk_p(s)
    s.seed((0, 0, 0))
    s.update(speed_constant(-1.0))
    p.notify()
    p.march()
    c = s.child_index((0, 0, 0))
    assert p.child_info(c)[1:] == (30, 56)
    seven = {s.child_index(x) for x in [(0, 0, -1), (0, -1, 0), (-1, 0, 0), (0, 0, 0), (0, 0, 1),
                                         (0, 1, 0), (1, 0, 0)]}
    assert set(int(x) for x in p.changes()) == seven
    s.update_start()
    s.delta([(0, 1, 0)], [-0.3])
    s.update_end()
    p.notify()
    assert set(int(x) for x in p.changes()) == seven
    p.march()
    six = {s.child_index(x) for x in [(-1, 0, 0), (0, 0, 0), (0, 1, 0), (1, 0, 0), (0, 0, -1),
                                       (0, 0, 1)]}
    assert set(int(x) for x in p.changes()) == six


@pytest.mark.parametrize("which", ["gpu", "cpu"])
def test_polys_expand_contract_roundtrip(which):
    """test_Poly.cpp:341-358, 474-527: re-expanded surface has the same counts; contracted
    to nothing leaves no active children."""
    mk_s, mk_p = {n: (s, p) for n, s, p in both()}[which]
    s = mk_s(3, 3, (15, 15, 15), (5, 5, 5))
    p = mk_p(s)
    s.seed((0, 0, 0))
    s.update(speed_constant(-1.0))
    p.notify(); p.march()
    s.update(speed_constant(1.0)); p.notify()
    s.update(speed_constant(-1.0)); p.notify(); p.march()
    assert p.child_info(s.child_index((0, 0, 0)))[1:] == (30, 56)
    s.update(speed_constant(-1.0)); p.notify(); p.march()
    lower = [s.child_index(x) for x in [(-1, 0, 0), (0, -1, 0), (0, 0, -1)]]
    assert [p.child_info(i)[1] for i in lower] == [5, 5, 5]
    s.update(speed_constant(1.0)); p.notify(); p.march()
    assert p.totals()[0] == 30
    act = s.partition_active()
    assert all(p.child_info(i)[0] == bool(act[i]) for i in range(s.num_partitions))
    for _ in range(2):
        s.update(speed_constant(1.0)); p.notify()
    p.march()
    assert p.totals() == (0, 0)
    act = s.partition_active()
    assert all(p.child_info(i)[0] == bool(act[i]) for i in range(s.num_partitions))
