"""Replays the grid known answers of the reference's test_Grid.cpp (tests/golden/grid_kats.json)
on any implementation of the felt ABI: each grid is loaded into a Surface<D,3> whose single
partition covers it, and each check evaluates the operator through felt_surface_stencil[_real]."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "grid_kats.json")
L = 3
# Catch's Approx: |a - b| < 100 * FLT_EPSILON * (1 + max(|a|, |b|))
APPROX_EPS = 100 * float(np.finfo(np.float32).eps)


def load():
    with open(FIXTURE) as f:
        return json.load(f)


def make_surface(factory, kat, grid=None):
    """A Surface<D,3> of the KAT's size in one partition, loaded with its grid."""
    s = factory(kat["D"], L, kat["size"], kat["size"])
    g = np.asarray(kat["grid"] if grid is None else grid, dtype=np.float32)
    s.load_dense(g.reshape(kat["size"]))
    return s


def evaluate(s, check):
    if "pos" in check:
        return s.stencil(check["op"], [check["pos"]], dx=check["dx"])[0]
    return s.stencil_real(check["op"], [check["rpos"]], dx=check["dx"])[0]


def matches(got, check) -> bool:
    got = np.atleast_1d(np.asarray(got, dtype=np.float32))
    want = np.atleast_1d(np.asarray(check["expect"], dtype=np.float32))
    if got.shape != want.shape:
        return False
    if check["exact"]:
        return bool(np.all(got == want))
    tol = APPROX_EPS * (1.0 + np.maximum(np.abs(got), np.abs(want)))
    return bool(np.all(np.abs(got - want) < tol))


def check_kat(factory, kat, grid=None):
    """Failures of one KAT as a list of strings (empty when every check holds)."""
    s = make_surface(factory, kat, grid)
    fails = []
    for c in kat["checks"]:
        got = evaluate(s, c)
        if not matches(got, c):
            where = c.get("pos", c.get("rpos"))
            fails.append(f"{c['op']}{tuple(where)} dx={c['dx']}: got {np.asarray(got).tolist()}, "
                         f"reference expects {c['expect']} (test_Grid.cpp:{c['line']})")
    s.close()
    return fails
