"""ctypes binding of the felt C ABI (include/felt_amd.h) and a Python mirror of Felt's API.

``Surface`` mirrors ``Felt::Surface<D,L>`` (Surface.hpp:49-696) and ``Polys`` mirrors
``Felt::Polys<S>`` (Polys.hpp:29-201).  The same classes drive any library that exports the
ABI under a symbol prefix, so the classes are written against a ``Binding`` rather than one
shared object.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Iterable, Optional, Sequence

import numpy as np

FELT_OK = 0
FELT_E_ARG = -1
FELT_E_DEVICE = -2
FELT_E_NOMEM = -3
FELT_E_CAPACITY = -4
FELT_E_STATE = -5
FELT_E_DELTA = -6
RAY_MISS = float(np.finfo(np.float32).max)  # Surface::ray_miss, every component

SPEED_CONSTANT = 0
SPEED_CURVATURE = 1
SPEED_CONST_CURV = 2
SPEED_NOISE_CURV = 3

LIST_ISOGRID = 0
LIST_DELTA = 1
LIST_STATUS_CHANGE = 2
LIST_AFFECTED = 3

PASS_SPEED = 0
PASS_DIST1 = 1
PASS_BAND = 2

# Grid operators of felt_surface_stencil (NumericMixin.hpp:72-500)
STENCIL_GET = 0
STENCIL_CURV = 1
STENCIL_DIVERGENCE = 2
STENCIL_GRAD = 3
STENCIL_GRAD_E = 4
STENCIL_GRAD_F = 5
STENCIL_GRAD_B = 6
STENCIL_GRAD_C = 7
STENCIL_NAMES = {"get": STENCIL_GET, "curv": STENCIL_CURV, "divergence": STENCIL_DIVERGENCE,
                 "grad": STENCIL_GRAD, "gradE": STENCIL_GRAD_E, "gradF": STENCIL_GRAD_F,
                 "gradB": STENCIL_GRAD_B, "gradC": STENCIL_GRAD_C}


class FeltError(RuntimeError):
    """Raised for a non-zero status; the reference throws std::domain_error there."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"felt error {code}: {msg}")
        self.code = code


class felt_speed(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_float), ("b", C.c_float), ("c", C.c_float),
                ("seed", C.c_uint64), ("cell", C.c_int32), ("reserved", C.c_int32)]


class felt_stats(C.Structure):
    _fields_ = [("active_isogrid_partitions", C.c_uint64),
                ("active_delta_partitions", C.c_uint64),
                ("band_voxels", C.c_uint64),
                ("last_voxel_updates", C.c_uint64),
                ("last_rounds", C.c_uint64),
                ("nan_events", C.c_uint64)]


class felt_launch_config(C.Structure):
    _fields_ = [("threads_per_block", C.c_int32), ("grid_cap", C.c_int32),
                ("voxels_per_lane", C.c_int32), ("threads_per_brick", C.c_int32),
                ("us_default", C.c_double), ("us_tuned", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def speed_constant(a: float) -> felt_speed:
    return felt_speed(SPEED_CONSTANT, a, 0.0, 0.0, 0, 0, 0)


def speed_curvature(b: float) -> felt_speed:
    return felt_speed(SPEED_CURVATURE, 0.0, b, 0.0, 0, 0, 0)


def speed_const_curv(a: float, b: float) -> felt_speed:
    return felt_speed(SPEED_CONST_CURV, a, b, 0.0, 0, 0, 0)


def speed_noise_curv(a: float, b: float, c: float, seed: int, cell: int) -> felt_speed:
    return felt_speed(SPEED_NOISE_CURV, a, b, c, seed, cell, 0)


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_F32P = C.POINTER(C.c_float)
_U64P = C.POINTER(C.c_uint64)
_U32P = C.POINTER(C.c_uint32)
_U8P = C.POINTER(C.c_uint8)
_SZ = C.c_size_t
_SZP = C.POINTER(C.c_size_t)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Binding:
    """Loads a shared object exporting the felt ABI under ``prefix``.

    ``polys_on_surface`` selects the call shape of the polys entry points: the HIP library
    has its own polys handle; the CPU restatement keys polys by the surface handle.
    """

    def __init__(self, path: str, prefix: str = "felt_", polys_on_surface: bool = False,
                 create_has_precision: bool = False):
        self.path = path
        self.lib = C.CDLL(path)
        self.prefix = prefix
        self.polys_on_surface = polys_on_surface
        self.create_has_precision = create_has_precision
        L = self.lib
        sig = {
            "last_error": (C.c_char_p, []),
            "device_count": (C.c_int, [C.POINTER(C.c_int)]),
            "surface_destroy": (C.c_int, [_P]),
            "surface_geometry": (C.c_int, [_P, _I32P, _U64P]),
            "surface_seed": (C.c_int, [_P, _I32P]),
            "surface_load_dense": (C.c_int, [_P, _F32P, _SZ]),
            "surface_update": (C.c_int, [_P, C.POINTER(felt_speed)]),
            "surface_update_box": (C.c_int, [_P, _I32P, _I32P, C.POINTER(felt_speed)]),
            "surface_update_deltas": (C.c_int, [_P, _SZ, _I32P, _F32P]),
            "surface_update_start": (C.c_int, [_P]),
            "surface_delta": (C.c_int, [_P, _SZ, _I32P, _F32P]),
            "surface_update_end": (C.c_int, [_P]),
            "surface_snapshot": (C.c_int, [_P, _F32P, _SZ]),
            "surface_get": (C.c_int, [_P, _SZ, _I32P, _F32P]),
            "surface_layer_size": (C.c_int, [_P, C.c_int, _U64P]),
            "surface_layer_positions": (C.c_int, [_P, C.c_int, _I32P, _SZ, _SZP]),
            "surface_zero_layer_in_box": (C.c_int, [_P, _I32P, _I32P, _I32P, _SZ, _SZP]),
            "surface_partition_list": (C.c_int, [_P, C.c_int, C.c_int, _U64P, _SZ, _SZP]),
            "surface_affected_positions": (C.c_int, [_P, C.c_int, _I32P, _SZ, _SZP]),
            "surface_partition_active": (C.c_int, [_P, _U8P, _SZ]),
            "surface_partition_layer_size": (C.c_int, [_P, C.c_uint64, C.c_int, _U64P]),
            "surface_stats": (C.c_int, [_P, C.POINTER(felt_stats)]),
            "surface_check_records": (C.c_int, [_P, _U64P]),
            "surface_sync": (C.c_int, [_P]),
            "surface_step_info": (C.c_int, [_P, _U64P, _U64P]),
            "surface_set_graph": (C.c_int, [_P, C.c_int]),
            "surface_autotune": (C.c_int, [_P, C.c_int, C.POINTER(felt_launch_config)]),
            "surface_set_round_limit": (C.c_int, [_P, C.c_int32]),
            "surface_set_launch_config": (C.c_int, [_P, C.POINTER(felt_launch_config)]),
            "surface_load_spheres": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double)]),
            "surface_time_pass": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.POINTER(C.c_int)]),
            "surface_ray": (C.c_int, [_P, _SZ, _F32P, _F32P, _F32P]),
            "surface_save": (C.c_int, [_P, C.POINTER(C.c_uint8), _SZ, _SZP]),
            "surface_load": (C.c_int, [C.POINTER(C.c_uint8), _SZ, C.POINTER(_P)]),
            "surface_time_ray": (C.c_int, [_P, _SZ, _F32P, _F32P, C.c_int,
                                           C.POINTER(C.c_double)]),
            "surface_stencil": (C.c_int, [_P, C.c_int, C.c_float, _SZ, _I32P, _F32P]),
            "surface_stencil_real": (C.c_int, [_P, C.c_int, C.c_float, _SZ, _F32P, _F32P]),
            "surface_speed_eval": (C.c_int, [_P, C.POINTER(felt_speed), _SZ, _I32P, _F32P]),
            "polys_notify": (C.c_int, [_P]),
            "polys_march": (C.c_int, [_P]),
            "polys_invalidate": (C.c_int, [_P]),
            "polys_mark": (C.c_int, [_P, C.c_uint64]),
            "polys_changes": (C.c_int, [_P, _U64P, _SZ, _SZP]),
            "polys_child_info": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int), _U64P, _U64P]),
            "polys_child_mesh": (C.c_int, [_P, C.c_uint64, _F32P, _SZ, _U32P, _SZ]),
            "polys_totals": (C.c_int, [_P, _U64P, _U64P]),
        }
        if create_has_precision:
            sig["surface_create"] = (C.c_int, [C.c_int, C.c_int, _I32P, _I32P, C.c_int,
                                               C.POINTER(_P)])
        else:
            sig["surface_create"] = (C.c_int, [C.c_int, C.c_int, _I32P, _I32P, C.POINTER(_P)])
        if polys_on_surface:
            sig["polys_create"] = (C.c_int, [_P])
        else:
            sig["polys_create"] = (C.c_int, [_P, C.POINTER(_P)])
            sig["polys_destroy"] = (C.c_int, [_P])
        self.fn = {}
        for name, (res, args) in sig.items():
            if not hasattr(L, prefix + name) and name in ("device_count", "surface_step_info",
                                                            "surface_set_graph",
                                                            "surface_autotune",
                                                            "surface_set_round_limit",
                                                            "surface_set_launch_config",
                                                            "surface_time_pass",
                                                            "surface_load_spheres",
                                                            "surface_ray",
                                                            "surface_time_ray",
                                                            "surface_save",
                                                            "surface_load",
                                                            "surface_check_records"):
                continue
            f = getattr(L, prefix + name)
            f.restype = res
            f.argtypes = args
            self.fn[name] = f

    def call(self, name: str, *args) -> int:
        rc = self.fn[name](*args)
        if rc != FELT_OK:
            msg = self.fn["last_error"]()
            raise FeltError(rc, msg.decode() if msg else "")
        return rc


def _vec(v: Sequence[int], D: int) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(-1))
    if a.size != D:
        raise ValueError(f"expected {D} coordinates, got {a.size}")
    return a


def archive_geometry(data: bytes):
    """(D, L, size, part) of a surface archive: the child size is its first Eigen vector, the
    background child's list tuple holds 2L+1 lists, and the grid size is the isogrid's Size just
    before the trailing Spatial fields (dx and two vectors)."""
    import struct
    if len(data) < 12:
        raise FeltError(FELT_E_ARG, "not a surface archive")
    D = struct.unpack_from("<i", data, 4)[0]
    if D not in (2, 3):
        raise FeltError(FELT_E_ARG, "not a surface archive")
    V = 16 + 4 * D
    part = struct.unpack_from(f"<{D}i", data, 4 + 16)
    at = 4 + V
    at += 4                                                   # background child: bg
    at += 8 + 4 * struct.unpack_from("<Q", data, at)[0]        # its data
    at += 8                                                   # its lookup's background
    at += 8 + 8 * struct.unpack_from("<Q", data, at)[0]        # its lookup data
    nl = struct.unpack_from("<i", data, at)[0]
    end = len(data) - (4 + 2 * V) - 3 * V
    size = struct.unpack_from(f"<{D}i", data, end + 16)
    if nl < 3 or nl % 2 == 0:
        raise FeltError(FELT_E_ARG, "not a surface archive")
    return D, (nl - 1) // 2, tuple(size), tuple(part)


class Surface:
    """Mirror of ``Felt::Surface<D,L>`` (Surface.hpp:49-696) over a felt ABI binding."""

    def __init__(self, binding: Binding, D: int, L: int, size: Sequence[int],
                 part: Optional[Sequence[int]] = None, fp64: bool = False):
        self.b = binding
        self.D, self.L = D, L
        self.NL = 2 * L + 1
        self.size = tuple(int(x) for x in size)
        self.part = tuple(int(x) for x in (part if part is not None else [8] * D))
        h = C.c_void_p()
        sz = _vec(self.size, D)
        pt = _vec(self.part, D)
        if binding.create_has_precision:
            binding.call("surface_create", D, L, _ptr(sz, _I32P), _ptr(pt, _I32P), int(fp64),
                         C.byref(h))
        else:
            binding.call("surface_create", D, L, _ptr(sz, _I32P), _ptr(pt, _I32P), C.byref(h))
        self._setup(h)

    def _setup(self, h):
        binding = self.b
        self.h = h
        self.offset = tuple(-(s // 2) for s in self.size)
        cs = np.zeros(3, np.int32)
        npart = C.c_uint64()
        binding.call("surface_geometry", self.h, _ptr(cs, _I32P), C.byref(npart))
        self.children_size = tuple(int(x) for x in cs[:self.D])
        self.num_partitions = int(npart.value)
        self.children_offset = tuple(int(np.fix(o / p)) for o, p in zip(self.offset, self.part))
        self._polys = None

    def close(self):
        if self.h:
            self.b.call("surface_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- layers / indices (Surface.hpp:672-675) --
    def layer_idx(self, layer: int) -> int:
        return layer + self.L

    def child_index(self, pos_child: Sequence[int]) -> int:
        """PosIdx of a children-grid position (with children offset)."""
        idx = 0
        for a in range(self.D):
            idx = idx * self.children_size[a] + (pos_child[a] - self.children_offset[a])
        return idx

    # -- stepping --
    def seed(self, pos: Sequence[int]) -> None:
        p = _vec(pos, self.D)
        self.b.call("surface_seed", self.h, _ptr(p, _I32P))

    def load_dense(self, phi: np.ndarray) -> None:
        a = np.ascontiguousarray(phi, dtype=np.float32).reshape(-1)
        self.b.call("surface_load_dense", self.h, _ptr(a, _F32P), a.size)

    def load_spheres(self, centres, radii) -> None:
        """Device-side load of min_k(|x - c_k| - r_k) (felt_surface_load_spheres)."""
        c = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1))
        r = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).reshape(-1))
        self.b.call("surface_load_spheres", self.h, int(r.size),
                    c.ctypes.data_as(C.POINTER(C.c_double)), r.ctypes.data_as(C.POINTER(C.c_double)))

    def update(self, fn) -> None:
        """``update(fn)``: fn is a felt_speed descriptor (device functor) or a Python callable
        ``fn(pos) -> float`` evaluated on the host for each zero-layer voxel."""
        if isinstance(fn, felt_speed):
            self.b.call("surface_update", self.h, C.byref(fn))
            return
        pos = self.layer_positions(0)
        vals = np.array([fn(tuple(int(x) for x in p)) for p in pos], dtype=np.float32)
        pos = np.ascontiguousarray(pos, dtype=np.int32)
        self.b.call("surface_update_deltas", self.h, len(vals), _ptr(pos, _I32P),
                    _ptr(vals, _F32P))

    def update_box(self, lo: Sequence[int], hi: Sequence[int], fn) -> None:
        """``update(pos_lo, pos_hi, fn)`` (Surface.hpp:310-386)."""
        lo_a, hi_a = _vec(lo, self.D), _vec(hi, self.D)
        if isinstance(fn, felt_speed):
            self.b.call("surface_update_box", self.h, _ptr(lo_a, _I32P), _ptr(hi_a, _I32P),
                        C.byref(fn))
            return
        pos = self.zero_layer_in_box(lo, hi)
        vals = np.array([fn(tuple(int(x) for x in p)) for p in pos], dtype=np.float32)
        self.update_start()
        if len(vals):
            self.delta(pos, vals)
        self.update_end()

    def update_start(self) -> None:
        self.b.call("surface_update_start", self.h)

    def delta(self, pos, val) -> None:
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.int32).reshape(-1, self.D))
        v = np.ascontiguousarray(np.asarray(val, dtype=np.float32).reshape(-1))
        if len(p) != len(v):
            raise ValueError("positions and values differ in length")
        self.b.call("surface_delta", self.h, len(v), _ptr(p, _I32P), _ptr(v, _F32P))

    def update_end(self) -> None:
        self.b.call("surface_update_end", self.h)

    # -- read-back --
    def snapshot(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.float32)
        self.b.call("surface_snapshot", self.h, _ptr(out, _F32P), out.size)
        return out

    def get(self, pos) -> np.ndarray:
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.int32).reshape(-1, self.D))
        out = np.empty(len(p), dtype=np.float32)
        self.b.call("surface_get", self.h, len(p), _ptr(p, _I32P), _ptr(out, _F32P))
        return out

    def save(self) -> bytes:
        """Surface::save: the isogrid as the reference's cereal binary archive."""
        n = C.c_size_t()
        self.b.call("surface_save", self.h, None, 0, C.byref(n))
        # numpy buffers: ctypes.string_at takes a C int length and cuts archives of 2 GiB or more
        buf = np.empty(max(n.value, 1), dtype=np.uint8)
        self.b.call("surface_save", self.h, buf.ctypes.data_as(C.POINTER(C.c_uint8)), n.value,
                    C.byref(n))
        return buf[:n.value].tobytes()

    @classmethod
    def load(cls, binding: "Binding", data: bytes) -> "Surface":
        """Surface::load: a new surface from a cereal archive of an isogrid."""
        D, L, size, part = archive_geometry(data)
        buf = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
        h = _P()
        binding.call("surface_load", buf.ctypes.data_as(C.POINTER(C.c_uint8)), len(data),
                     C.byref(h))
        obj = Surface.__new__(cls)
        obj.b = binding
        obj.D, obj.L, obj.NL = D, L, 2 * L + 1
        obj.size, obj.part = size, part
        obj._setup(h)
        return obj

    def ray(self, origins, dirs) -> np.ndarray:
        """Surface::ray for each (origin, direction) row; a miss reads RAY_MISS (FLT_MAX)."""
        o = np.ascontiguousarray(np.asarray(origins, dtype=np.float32).reshape(-1, self.D))
        d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float32).reshape(-1, self.D))
        if o.shape != d.shape:
            raise ValueError("origins and directions differ in count")
        out = np.empty_like(o)
        self.b.call("surface_ray", self.h, len(o), _ptr(o, _F32P), _ptr(d, _F32P),
                    _ptr(out, _F32P))
        return out

    def time_ray(self, origins, dirs, reps: int = 10) -> float:
        """Average device microseconds of one ray launch over these rays."""
        o = np.ascontiguousarray(np.asarray(origins, dtype=np.float32).reshape(-1, self.D))
        d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float32).reshape(-1, self.D))
        us = C.c_double()
        self.b.call("surface_time_ray", self.h, len(o), _ptr(o, _F32P), _ptr(d, _F32P),
                    int(reps), C.byref(us))
        return float(us.value)

    def stencil(self, op, pos, dx: float = 1.0) -> np.ndarray:
        """A grid operator (name or STENCIL_*) at integer positions: (n,) for get, curv and
        divergence, (n, D) for the gradients (felt_surface_stencil)."""
        return self._stencil("surface_stencil", op, np.int32, _I32P, pos, dx)

    def stencil_real(self, op, pos, dx: float = 1.0) -> np.ndarray:
        """The same at real positions, every sample interpolated (felt_surface_stencil_real)."""
        return self._stencil("surface_stencil_real", op, np.float32, _F32P, pos, dx)

    def _stencil(self, name, op, dtype, ptype, pos, dx):
        op = STENCIL_NAMES.get(op, op)
        p = np.ascontiguousarray(np.asarray(pos, dtype=dtype).reshape(-1, self.D))
        width = self.D if op >= STENCIL_GRAD else 1
        out = np.empty((len(p), width), dtype=np.float32)
        self.b.call(name, self.h, int(op), float(dx), len(p), _ptr(p, ptype), _ptr(out, _F32P))
        return out if width > 1 else out.reshape(-1)

    def speed_eval(self, fn: felt_speed, pos) -> np.ndarray:
        """fn(pos) of a device functor at positions of active partitions, computed by the
        update kernel's own speed code (felt_surface_speed_eval)."""
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.int32).reshape(-1, self.D))
        out = np.empty(len(p), dtype=np.float32)
        self.b.call("surface_speed_eval", self.h, C.byref(fn), len(p), _ptr(p, _I32P),
                    _ptr(out, _F32P))
        return out

    def layer_size(self, layer: int) -> int:
        n = C.c_uint64()
        self.b.call("surface_layer_size", self.h, layer, C.byref(n))
        return int(n.value)

    def _positions(self, name: str, *args) -> np.ndarray:
        n = C.c_size_t()
        cap = 1024
        while True:
            out = np.empty((cap, self.D), dtype=np.int32)
            try:
                self.b.call(name, self.h, *args, _ptr(out, _I32P), cap, C.byref(n))
                return out[: n.value].copy()
            except FeltError as e:
                if e.code != FELT_E_CAPACITY:
                    raise
                cap = int(n.value) + 1

    def layer_positions(self, layer: int) -> np.ndarray:
        return self._positions("surface_layer_positions", layer)

    def zero_layer_in_box(self, lo, hi) -> np.ndarray:
        lo_a, hi_a = _vec(lo, self.D), _vec(hi, self.D)
        return self._positions("surface_zero_layer_in_box", _ptr(lo_a, _I32P), _ptr(hi_a, _I32P))

    def affected_positions(self, k: int) -> np.ndarray:
        return self._positions("surface_affected_positions", k)

    def partition_list(self, kind: int, k: int) -> np.ndarray:
        n = C.c_size_t()
        cap = 256
        while True:
            out = np.empty(cap, dtype=np.uint64)
            try:
                self.b.call("surface_partition_list", self.h, kind, k, _ptr(out, _U64P), cap,
                            C.byref(n))
                return out[: n.value].copy()
            except FeltError as e:
                if e.code != FELT_E_CAPACITY:
                    raise
                cap = int(n.value) + 1

    def delta_list(self, k: int) -> np.ndarray:
        return self.partition_list(LIST_DELTA, k)

    def status_change_list(self, k: int) -> np.ndarray:
        return self.partition_list(LIST_STATUS_CHANGE, k)

    def partition_active(self) -> np.ndarray:
        out = np.empty(self.num_partitions, dtype=np.uint8)
        self.b.call("surface_partition_active", self.h, _ptr(out, _U8P), out.size)
        return out.astype(bool)

    def partition_layer_size(self, pos_idx: int, layer: int) -> int:
        n = C.c_uint64()
        self.b.call("surface_partition_layer_size", self.h, int(pos_idx), layer, C.byref(n))
        return int(n.value)

    def stats(self) -> felt_stats:
        st = felt_stats()
        self.b.call("surface_stats", self.h, C.byref(st))
        return st

    def check_records(self) -> int:
        """Active bricks whose face-neighbour record disagrees with the page table."""
        n = C.c_uint64()
        self.b.call("surface_check_records", self.h, C.byref(n))
        return int(n.value)

    def set_graph(self, enable: bool) -> None:
        """Capture global device-functor steps as one HIP graph (default) or launch kernels
        one by one."""
        if "surface_set_graph" in self.b.fn:
            self.b.call("surface_set_graph", self.h, int(bool(enable)))

    def autotune(self, reps: int = 10) -> dict:
        """Times launch shapes on the current state and keeps the fastest
        (felt_surface_autotune); returns the chosen felt_launch_config as a dict."""
        out = felt_launch_config()
        self.b.call("surface_autotune", self.h, int(reps), C.byref(out))
        return out.as_dict()

    def set_launch_config(self, threads_per_block: int, grid_cap: int,
                          voxels_per_lane: int) -> None:
        cfg = felt_launch_config(threads_per_block, grid_cap, voxels_per_lane, 0, 0.0, 0.0)
        self.b.call("surface_set_launch_config", self.h, C.byref(cfg))

    def set_round_limit(self, max_rounds: int) -> None:
        """Most rounds one update may take before it fails (felt_surface_set_round_limit)."""
        self.b.call("surface_set_round_limit", self.h, int(max_rounds))

    def step_info(self):
        """(voxel updates, rounds) of the last update call."""
        u, r = C.c_uint64(), C.c_uint64()
        if "surface_step_info" in self.b.fn:
            self.b.call("surface_step_info", self.h, C.byref(u), C.byref(r))
            return int(u.value), int(r.value)
        st = self.stats()
        return int(st.last_voxel_updates), int(st.last_rounds)

    def time_pass(self, pass_id: int, reps: int = 20):
        """(microseconds per launch, entries per launch, launches) of one pass of a global step
        replayed on the current state (felt_surface_time_pass)."""
        us, items, n = C.c_double(), C.c_double(), C.c_int()
        self.b.call("surface_time_pass", self.h, int(pass_id), int(reps), C.byref(us),
                    C.byref(items), C.byref(n))
        return float(us.value), float(items.value), int(n.value)

    def sync(self) -> None:
        self.b.call("surface_sync", self.h)

    def layers(self) -> dict:
        """All tracked positions as {layer: set of position tuples}."""
        return {k: {tuple(int(x) for x in p) for p in self.layer_positions(k)}
                for k in range(-self.L, self.L + 1)}


class Polys:
    """Mirror of ``Felt::Polys<S>`` (Polys.hpp:29-201)."""

    def __init__(self, surface: Surface):
        self.s = surface
        self.b = surface.b
        if self.b.polys_on_surface:
            self.b.call("polys_create", surface.h)
            self.h = surface.h
            self.own = False
        else:
            h = C.c_void_p()
            self.b.call("polys_create", surface.h, C.byref(h))
            self.h = h
            self.own = True

    def close(self):
        if self.own and self.h:
            self.b.call("polys_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def notify(self) -> None:
        self.b.call("polys_notify", self.h)

    def march(self) -> None:
        self.b.call("polys_march", self.h)

    def invalidate(self) -> None:
        self.b.call("polys_invalidate", self.h)

    def mark(self, pos_idx: int) -> None:
        """Queue one partition for the next march (Impl::Poly::Single use)."""
        self.b.call("polys_mark", self.h, int(pos_idx))

    def changes(self) -> np.ndarray:
        n = C.c_size_t()
        cap = max(16, self.s.num_partitions)
        out = np.empty(cap, dtype=np.uint64)
        self.b.call("polys_changes", self.h, _ptr(out, _U64P), cap, C.byref(n))
        return out[: n.value].copy()

    def child_info(self, pos_idx: int):
        act = C.c_int()
        nv, ns = C.c_uint64(), C.c_uint64()
        self.b.call("polys_child_info", self.h, int(pos_idx), C.byref(act), C.byref(nv),
                    C.byref(ns))
        return bool(act.value), int(nv.value), int(ns.value)

    def child_mesh(self, pos_idx: int):
        """(vertices, simplices): vertices are (n, 2D) pos+normal in 3D, (n, 2) in 2D."""
        _, nv, ns = self.child_info(pos_idx)
        D = self.s.D
        vw = 6 if D == 3 else 2
        vtx = np.empty((max(nv, 1), vw), dtype=np.float32)
        spx = np.empty((max(ns, 1), D), dtype=np.uint32)
        self.b.call("polys_child_mesh", self.h, int(pos_idx), _ptr(vtx, _F32P), vtx.size,
                    _ptr(spx, _U32P), spx.size)
        return vtx[:nv].copy(), spx[:ns].copy()

    def totals(self):
        nv, ns = C.c_uint64(), C.c_uint64()
        self.b.call("polys_totals", self.h, C.byref(nv), C.byref(ns))
        return int(nv.value), int(ns.value)
