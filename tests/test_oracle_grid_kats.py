"""The CPU oracle's numeric operators (curvature, interpolation, gradients, divergence) and index
convention against the known answers of the reference's test_Grid.cpp
(tests/golden/grid_kats.json)."""
import ctypes as C

import numpy as np
import pytest

import grid_kats
import oracle_lib

DATA = grid_kats.load()
KATS = DATA["kats"]


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_oracle_matches_grid_kat(kat):
    fails = grid_kats.check_kat(oracle_lib.surface, kat)
    assert not fails, f"{kat['name']} ({kat['cite']}): {fails}"


def test_oracle_index_613():
    """Row-major index with the last axis fastest, and its inverse (test_Grid.cpp:72-89)."""
    k = DATA["index"]
    lib = oracle_lib.binding().lib
    size = np.asarray(k["size"], np.int32)
    off = np.asarray(k["offset"], np.int32)
    pos = np.asarray(k["pos"], np.int32)
    back = np.zeros(3, np.int32)
    idx = C.c_uint64()
    P = C.POINTER(C.c_int32)
    assert lib.felt_oracle_index(3, size.ctypes.data_as(P), off.ctypes.data_as(P),
                                 pos.ctypes.data_as(P), C.byref(idx), back.ctypes.data_as(P)) == 0
    assert idx.value == k["index"]
    assert back.tolist() == k["pos"]


def test_curvature_kats_pin_the_speed_functor():
    """fn = b * curv evaluated by the oracle's speed functor equals b times the reference's
    curvature answers: the same stencil the update step uses."""
    import felt_amd
    for kat in KATS:
        for c in kat["checks"]:
            if c["op"] != "curv":
                continue
            s = grid_kats.make_surface(oracle_lib.surface, kat)
            got = s.speed_eval(felt_amd.speed_curvature(0.5), [c["pos"]])[0]
            assert got == np.float32(0.5) * np.float32(c["expect"]), (kat["name"], got)


def test_transposed_literal_is_caught():
    """A grid read with its axes reversed (a plausible orientation mistake) fails the
    reference's gradient answers."""
    kat = next(k for k in KATS if k["name"] == "grad_3d")
    g = np.asarray(kat["grid"], np.float32).reshape(kat["size"]).transpose(2, 1, 0).reshape(-1)
    assert grid_kats.check_kat(oracle_lib.surface, kat, grid=g)


def test_swapped_difference_is_caught():
    """Forward and backward differences swapped (a plausible sign mistake) fail the answers."""
    kat = next(k for k in KATS if k["name"] == "grad_3d")
    s = grid_kats.make_surface(oracle_lib.surface, kat)
    checks = [c for c in kat["checks"] if c["op"] in ("gradF", "gradB") and c["dx"] == 1.0]
    for c in checks:
        other = "gradB" if c["op"] == "gradF" else "gradF"
        assert not grid_kats.matches(s.stencil(other, [c["pos"]])[0], c)
