"""felt_surface_ray_noalloc through the C ABI: the same hits as felt_surface_ray, call after call
on the one shared device array (a second, smaller call must not see the first's rays), and sizes
beyond the array refused."""
import ctypes as C

import numpy as np
import pytest

import felt_amd
from felt_amd import workloads
from felt_amd._binding import FELT_E_ARG, FELT_OK

pytestmark = pytest.mark.gpu
MAX_RAYS = 65536
F32P = C.POINTER(C.c_float)


def _lib():
    lib = felt_amd.lib().lib
    f = lib.felt_surface_ray_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_size_t, F32P, F32P, F32P]
    return lib


def _rays(D, n, size, seed):
    """n rays from points around the grid centre towards random nearby targets (unit
    directions), some from inside the surface."""
    rng = np.random.default_rng(seed)
    half = np.array(size[:D], dtype=np.float64) / 2
    org = (rng.random((n, D)) * 2 - 1) * half * 0.9
    tgt = (rng.random((n, D)) * 2 - 1) * half * 0.2
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.ascontiguousarray(org, np.float32), np.ascontiguousarray(d, np.float32)


def _noalloc(lib, s, org, dirs):
    out = np.full_like(org, np.nan)
    rc = lib.felt_surface_ray_noalloc(s.h, len(org), org.ctypes.data_as(F32P),
                                      dirs.ctypes.data_as(F32P), out.ctypes.data_as(F32P))
    return rc, out


@pytest.mark.parametrize("D", [3, 2])
def test_ray_noalloc_matches_ray(D):
    lib = _lib()
    size = (96, 80, 72)[:D]
    part = (16,) * D
    s = felt_amd.Surface(D, 3, size, part)
    centres = [(-9.0, 4.0, 2.0)[:D], (12.0, -6.0, -3.0)[:D]]
    s.load_dense(workloads.spheres_phi(size, 3, centres, [17.0, 13.0]))
    for _ in range(3):
        s.update(felt_amd.speed_const_curv(-0.3, 0.1))
    # a full call, then a smaller one: each must equal felt_surface_ray on its own rays
    for n, seed in ((MAX_RAYS, 1), (1000, 2), (MAX_RAYS // 3, 3)):
        org, dirs = _rays(D, n, size, seed)
        want = s.ray(org, dirs)
        rc, got = _noalloc(lib, s, org, dirs)
        assert rc == FELT_OK, felt_amd.lib().fn["last_error"]()
        assert np.array_equal(got, want), n
        hit = want[:, 0] < 1e30
        assert hit.sum() > n // 4  # the rays do reach the surface
    org, dirs = _rays(D, MAX_RAYS + 1, size, 4)
    rc, _ = _noalloc(lib, s, org, dirs)
    assert rc == FELT_E_ARG
    rc, _ = _noalloc(lib, s, org[:0], dirs[:0])
    assert rc == FELT_OK
