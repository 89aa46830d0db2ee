"""Loads the CPU restatement (oracle/libfelt_oracle.so) behind the same Python mirror as the
HIP library.  Test infrastructure: only tests, smoke() and bench.py's CPU baseline use it."""
import ctypes
import os
import subprocess

from felt_amd._binding import Binding, Polys, Surface

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "libfelt_oracle.so")
_binding = None


def binding() -> Binding:
    global _binding
    if _binding is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _binding = Binding(LIB, prefix="felt_oracle_", polys_on_surface=True,
                           create_has_precision=True)
        _binding.lib.felt_oracle_set_threads(1)
        f = _binding.lib.felt_oracle_set_round_diag
        f.argtypes = [ctypes.c_longlong] * 3
        f.restype = ctypes.c_int
    return _binding


def surface(D, L, size, part, fp64=False) -> Surface:
    return Surface(binding(), D, L, size, part, fp64=fp64)


def polys(s: Surface) -> Polys:
    return Polys(s)
