"""CPU-side checks of the HIP library boundary: it loads, exports every entry point that
include/felt_amd.h declares, and its marching-cubes table equals the oracle's (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "felt_amd.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(felt_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import felt_amd
    lib = ctypes.CDLL(felt_amd.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared()) >= 30


def test_ms_table_matches_oracle_and_topology():
    """Marching-squares segments: same edge lists as the oracle's table, and every segment
    joins two edges whose endpoints differ in sign (Felt's square corners and edges,
    PolyMixin.hpp:221-265)."""
    import felt_amd
    lib = ctypes.CDLL(felt_amd.LIB_PATH)
    out = np.zeros(16, dtype=np.uint32)
    assert lib.felt_ms_table(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 16) == 0
    src = open(os.path.join(ROOT, "oracle", "felt_oracle.hpp")).read()
    blk = src[src.index("kSeg2[16] = {"):]
    rows = re.findall(r'"([0-9a-f]*)"', blk[:blk.index("};")])
    assert len(rows) == 16
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
    edges = [(0, 1), (1, 2), (3, 2), (0, 3)]  # corner pairs of edges 0..3
    for m in range(16):
        t = int(out[m])
        n = t >> 28
        got = "".join("%x" % ((t >> (4 * k)) & 0xF) for k in range(2 * n))
        assert got == rows[m], (m, got, rows[m])
        crossed = {e for e, (a, b) in enumerate(edges) if ((m >> a) & 1) != ((m >> b) & 1)}
        assert set(int(c, 16) for c in got) == crossed, m


def test_mc_table_matches_oracle_and_topology():
    import felt_amd
    lib = ctypes.CDLL(felt_amd.LIB_PATH)
    out = np.zeros(256, dtype=np.uint64)
    assert lib.felt_mc_table(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 256) == 0
    src = open(os.path.join(ROOT, "oracle", "felt_oracle.hpp")).read()
    blk = src[src.index("kTri3[256] = {"):]
    rows = re.findall(r'"([0-9a-f]*)"', blk[:blk.index("};")])
    assert len(rows) == 256
    # corners and edges of Felt's cube (PolyMixin.hpp:380-406)
    corners = [(0, 0, 0), (1, 0, 0), (1, 0, -1), (0, 0, -1), (0, 1, 0), (1, 1, 0), (1, 1, -1),
               (0, 1, -1)]
    edges = [((0, 0, 0), 0), ((1, 0, -1), 2), ((0, 0, -1), 0), ((0, 0, -1), 2), ((0, 1, 0), 0),
             ((1, 1, -1), 2), ((0, 1, -1), 0), ((0, 1, -1), 2), ((0, 0, 0), 1), ((1, 0, 0), 1),
             ((1, 0, -1), 1), ((0, 0, -1), 1)]
    cidx = {c: i for i, c in enumerate(corners)}
    for m in range(256):
        packed = int(out[m])
        n = packed >> 60
        es = [(packed >> (4 * i)) & 0xF for i in range(3 * n)]
        assert es == [int(c, 16) for c in rows[m]], m
        crossed = set()
        for (o, ax) in edges:
            b = list(o)
            b[ax] += 1
            if ((m >> cidx[tuple(o)]) & 1) != ((m >> cidx[tuple(b)]) & 1):
                crossed.add(edges.index((o, ax)))
        # a case's triangles use exactly the edges whose endpoints differ in sign
        assert set(es) == crossed, m
        # complementary cases cross the same edges
        n2 = int(out[255 - m]) >> 60
        es2 = {(int(out[255 - m]) >> (4 * i)) & 0xF for i in range(3 * n2)}
        assert es2 == crossed or m in (0, 255)


def test_product_refuses_to_run_without_library(tmp_path, monkeypatch):
    import felt_amd
    monkeypatch.setattr(felt_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(felt_amd, "_lib", None)
    with pytest.raises(felt_amd.FeltError):
        felt_amd.lib()
