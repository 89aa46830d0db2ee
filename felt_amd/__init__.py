"""felt_amd — Felt's sparse-field level-set surface and polygonisation on MI355X (gfx950).

The compute path is the HIP library libfelt_amd.so built from felt_amd/csrc; this package only
loads it and mirrors Felt's API (``Surface``, ``Polys``) over its C ABI.  There is no host
fallback: importing ``lib()`` without a built library raises.
"""
import os

from ._binding import (Binding, FeltError, Polys as _Polys, Surface as _Surface, felt_speed,
                       felt_stats, speed_const_curv, speed_constant, speed_curvature,
                       speed_noise_curv, LIST_AFFECTED, LIST_DELTA, LIST_ISOGRID,
                       LIST_STATUS_CHANGE)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfelt_amd.so")
_lib = None


def lib() -> Binding:
    """The HIP library binding; raises if the library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FeltError(-2, f"{LIB_PATH} is not built; run python felt_amd/build.py")
        _lib = Binding(LIB_PATH, prefix="felt_")
    return _lib


def device_count() -> int:
    import ctypes
    n = ctypes.c_int(0)
    lib().call("device_count", ctypes.byref(n)) if "device_count" in lib().fn else None
    return n.value


class Surface(_Surface):
    """``Felt::Surface<D,L>`` on the GPU."""

    def __init__(self, D, L, size, part=None):
        super().__init__(lib(), D, L, size, part)


class Polys(_Polys):
    """``Felt::Polys<Surface>`` on the GPU."""


__all__ = ["Surface", "Polys", "FeltError", "felt_speed", "felt_stats", "speed_constant",
           "speed_curvature", "speed_const_curv", "speed_noise_curv", "lib", "LIB_PATH",
           "LIST_AFFECTED", "LIST_DELTA", "LIST_ISOGRID", "LIST_STATUS_CHANGE"]
