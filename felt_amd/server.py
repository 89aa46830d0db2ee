"""Served surfaces: a request loop over one batch of resident surfaces (felt_server_*).

Requests are (surface, world box, speed, steps); each tick admits queued requests (one per
surface in flight, at most ``max_active``, within the HBM budget) and advances every in-flight
request by one bounded update plus repolygonisation, all in one batched step.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from ._binding import FeltError, felt_speed
from .batch import Batch, BatchPolys, _Api, _arr, _ptr, _I32P

_P = C.c_void_p


class felt_request(C.Structure):
    _fields_ = [("surface", C.c_int32), ("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3),
                ("speed", felt_speed), ("steps", C.c_int32), ("reserved", C.c_int32),
                ("id", C.c_uint64)]


class felt_server_stats(C.Structure):
    _fields_ = [("ticks", C.c_uint64), ("requests_done", C.c_uint64),
                ("requests_queued", C.c_uint64), ("requests_active", C.c_uint64),
                ("admissions_deferred", C.c_uint64), ("voxel_updates", C.c_uint64),
                ("busy_s", C.c_double), ("step_p50_ms", C.c_double), ("step_p99_ms", C.c_double),
                ("request_p50_ms", C.c_double), ("request_p99_ms", C.c_double),
                ("device_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("bricks_used", C.c_uint64), ("bricks_total", C.c_uint64),
                ("phase_ms", C.c_double * 4), ("requests_failed", C.c_uint64)]


_SIGS = {
    "felt_server_create": [C.c_int, C.c_int, _I32P, C.c_int32, _I32P, C.c_uint64, C.c_int32,
                           C.c_int32, C.POINTER(_P)],
    "felt_server_destroy": [_P],
    "felt_server_batch": [_P, C.POINTER(_P)],
    "felt_server_polys": [_P, C.POINTER(_P)],
    "felt_server_start": [_P],
    "felt_server_set_target": [_P, C.c_double],
    "felt_server_submit": [_P, C.POINTER(felt_request), C.c_size_t],
    "felt_server_tick": [_P, C.POINTER(C.c_int32)],
    "felt_server_poll": [_P, C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(C.c_size_t)],
    "felt_server_poll_failed": [_P, C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(C.c_size_t)],
    "felt_server_cancel": [_P, C.c_uint64],
    "felt_server_read_stats": [_P, C.POINTER(felt_server_stats)],
}


def request(surface: int, lo: Sequence[int], hi: Sequence[int], speed: felt_speed, steps: int,
            rid: int) -> felt_request:
    r = felt_request()
    r.surface = int(surface)
    for a in range(len(lo)):
        r.lo[a] = int(lo[a])
        r.hi[a] = int(hi[a])
    r.speed = speed
    r.steps = int(steps)
    r.id = int(rid)
    return r


class Server:
    def __init__(self, lib: C.CDLL, D: int, L: int, sizes, part, budget_bytes: int,
                 max_active: int, polygonise: bool = True):
        self.api = _Api(lib)
        for name, args in _SIGS.items():
            f = getattr(lib, name)
            f.restype = C.c_int
            f.argtypes = args
        self.D, self.L = D, L
        sizes = [tuple(int(x) for x in s) for s in sizes]
        h = _P()
        sz = _arr(sizes, np.int32)
        pt = _arr(part, np.int32)
        self.api("felt_server_create", D, L, _ptr(pt, _I32P), len(sizes), _ptr(sz, _I32P),
                 int(budget_bytes), int(max_active), int(bool(polygonise)), C.byref(h))
        self.h = h
        bh = _P()
        self.api("felt_server_batch", self.h, C.byref(bh))
        self.batch = Batch(lib, D, L, sizes, part, handle=bh)
        self.polys = None
        if polygonise:
            ph = _P()
            self.api("felt_server_polys", self.h, C.byref(ph))
            self.polys = BatchPolys(self.batch, handle=ph)

    def close(self):
        if self.h:
            self.api("felt_server_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def start(self) -> None:
        self.api("felt_server_start", self.h)

    def set_target(self, tick_ms: float) -> None:
        """Admit requests only while the predicted tick time stays within tick_ms (0: off)."""
        self.api("felt_server_set_target", self.h, float(tick_ms))

    def submit(self, reqs: Sequence[felt_request]) -> None:
        if not reqs:
            return
        arr = (felt_request * len(reqs))(*reqs)
        self.api("felt_server_submit", self.h, arr, len(reqs))

    def tick(self) -> int:
        n = C.c_int32()
        self.api("felt_server_tick", self.h, C.byref(n))
        return int(n.value)

    def poll(self, cap: int = 4096):
        """Ids of the requests completed since the last poll."""
        return self._drain("felt_server_poll", cap)

    def poll_failed(self, cap: int = 4096):
        """Ids of the requests that failed since the last poll_failed (felt_server_tick)."""
        return self._drain("felt_server_poll_failed", cap)

    def _drain(self, name, cap):
        out = np.empty(max(cap, 1), dtype=np.uint64)
        n = C.c_size_t()
        self.api(name, self.h, out.ctypes.data_as(C.POINTER(C.c_uint64)), cap, C.byref(n))
        return [int(x) for x in out[: n.value]]

    def cancel(self, rid: int) -> None:
        """Withdraw a queued or in-flight request."""
        self.api("felt_server_cancel", self.h, int(rid))

    def stats(self) -> dict:
        st = felt_server_stats()
        self.api("felt_server_read_stats", self.h, C.byref(st))
        out = {name: getattr(st, name) for name, _ in felt_server_stats._fields_}
        out["phase_ms"] = [float(x) for x in st.phase_ms]
        return out
