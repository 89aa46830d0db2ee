"""Batched surfaces: many independent ``Felt::Surface<D,L>`` instances stepped together.

``Batch`` wraps the ``felt_batch_*`` entry points of include/felt_amd.h.  Member ``i`` behaves
exactly as a ``felt_amd.Surface`` of its own grid size would; a step of the batch advances
every member (``update``) or a chosen set of members over boxes (``update_boxes``) with one
launch sequence over all their bricks.  ``BatchPolys`` polygonises every member at once.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from ._binding import FeltError, felt_launch_config, felt_speed, felt_stats

_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_F32P = C.POINTER(C.c_float)
_F64P = C.POINTER(C.c_double)
_U64P = C.POINTER(C.c_uint64)
_U32P = C.POINTER(C.c_uint32)
_U8P = C.POINTER(C.c_uint8)
_SZ = C.c_size_t
_SZP = C.POINTER(C.c_size_t)

_SIGS = {
    "felt_batch_create": [C.c_int, C.c_int, _I32P, C.c_int32, _I32P, C.c_uint64, C.POINTER(_P)],
    "felt_batch_destroy": [_P],
    "felt_batch_count": [_P, _I32P],
    "felt_batch_memory": [_P, _U64P, _U64P, _U64P],
    "felt_batch_seed": [_P, C.c_int32, _I32P],
    "felt_batch_load_dense": [_P, C.c_int32, _F32P, _SZ],
    "felt_batch_load_spheres": [_P, C.c_int32, C.c_int32, _F64P, _F64P],
    "felt_batch_update": [_P, C.POINTER(felt_speed)],
    "felt_batch_update_boxes": [_P, C.c_int32, _I32P, _I32P, _I32P, C.POINTER(felt_speed)],
    "felt_batch_update_async": [_P, C.POINTER(felt_speed)],
    "felt_batch_wait": [_P],
    "felt_batch_snapshot": [_P, C.c_int32, _F32P, _SZ],
    "felt_batch_layer_size": [_P, C.c_int32, C.c_int, _U64P],
    "felt_batch_layer_positions": [_P, C.c_int32, C.c_int, _I32P, _SZ, _SZP],
    "felt_batch_stats": [_P, C.c_int32, C.POINTER(felt_stats)],
    "felt_batch_check_records": [_P, _U64P],
    "felt_batch_ray": [_P, _SZ, _I32P, _F32P, _F32P, _F32P],
    "felt_batch_save": [_P, C.c_int32, C.POINTER(C.c_uint8), _SZ, _SZP],
    "felt_batch_load": [_P, C.c_int32, C.POINTER(C.c_uint8), _SZ],
    "felt_batch_clear": [_P, C.c_int32],
    "felt_batch_step_info": [_P, _U64P, _U64P],
    "felt_batch_set_graph": [_P, C.c_int],
    "felt_batch_time_pass": [_P, C.c_int, C.c_int, C.POINTER(C.c_double),
                             C.POINTER(C.c_double), C.POINTER(C.c_int)],
    "felt_batch_sync": [_P],
    "felt_group_create": [C.POINTER(_P), C.c_int32, C.POINTER(_P)],
    "felt_group_destroy": [_P],
    "felt_group_update_async": [_P, C.POINTER(C.POINTER(felt_speed))],
    "felt_group_wait": [_P],
    "felt_batch_autotune": [_P, C.c_int, C.POINTER(felt_launch_config)],
    "felt_batch_set_round_limit": [_P, C.c_int32],
    "felt_batch_set_launch_config": [_P, C.POINTER(felt_launch_config)],
    "felt_batch_polys_create": [_P, C.POINTER(_P)],
    "felt_batch_polys_changes": [_P, C.c_int32, _U64P, _SZ, _SZP],
    "felt_batch_polys_child_info": [_P, C.c_int32, C.c_uint64, C.POINTER(C.c_int), _U64P, _U64P],
    "felt_batch_polys_child_mesh": [_P, C.c_int32, C.c_uint64, _F32P, _SZ, _U32P, _SZ],
    "felt_polys_destroy": [_P],
    "felt_polys_notify": [_P],
    "felt_polys_march": [_P],
    "felt_polys_invalidate": [_P],
    "felt_polys_totals": [_P, _U64P, _U64P],
}


def _arr(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1))


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class _Api:
    def __init__(self, lib: C.CDLL):
        self.lib = lib
        for name, args in _SIGS.items():
            f = getattr(lib, name)
            f.restype = C.c_int
            f.argtypes = args
        lib.felt_last_error.restype = C.c_char_p

    def __call__(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != 0:
            msg = self.lib.felt_last_error()
            raise FeltError(rc, msg.decode() if msg else "")
        return rc


class Batch:
    """n independent surfaces sharing D, L and the partition size."""

    def __init__(self, lib: C.CDLL, D: int, L: int, sizes: Sequence[Sequence[int]],
                 part: Sequence[int], max_bricks: int = 0, handle=None):
        self.api = _Api(lib)
        self.D, self.L = D, L
        self.sizes = [tuple(int(x) for x in s) for s in sizes]
        self.part = tuple(int(x) for x in part)
        self.n = len(self.sizes)
        self.own = handle is None
        if handle is not None:  # borrowed from a server
            self.h = handle
            return
        h = _P()
        sz = _arr(self.sizes, np.int32)
        pt = _arr(self.part, np.int32)
        self.api("felt_batch_create", D, L, _ptr(pt, _I32P), self.n, _ptr(sz, _I32P),
                 int(max_bricks), C.byref(h))
        self.h = h

    def close(self):
        if self.h and self.own:
            self.api("felt_batch_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def offset(self, i: int):
        return tuple(-(s // 2) for s in self.sizes[i])

    def memory(self):
        """(device bytes, brick slots in use, brick slots in the pool)."""
        b, u, t = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.api("felt_batch_memory", self.h, C.byref(b), C.byref(u), C.byref(t))
        return int(b.value), int(u.value), int(t.value)

    def seed(self, i: int, pos: Sequence[int]) -> None:
        p = _arr(pos, np.int32)
        self.api("felt_batch_seed", self.h, int(i), _ptr(p, _I32P))

    def load_dense(self, i: int, phi: np.ndarray) -> None:
        a = _arr(phi, np.float32)
        self.api("felt_batch_load_dense", self.h, int(i), _ptr(a, _F32P), a.size)

    def load_spheres(self, i: int, centres, radii) -> None:
        c = _arr(centres, np.float64)
        r = _arr(radii, np.float64)
        self.api("felt_batch_load_spheres", self.h, int(i), int(r.size), _ptr(c, _F64P),
                 _ptr(r, _F64P))

    def update(self, speeds: Sequence[felt_speed]) -> None:
        """update(fn) of every member, member i with speeds[i]."""
        arr = (felt_speed * self.n)(*speeds)
        self.api("felt_batch_update", self.h, arr)

    def update_async(self, speeds: Sequence[felt_speed]) -> None:
        """Queue update(fn) of every member; finish with wait()."""
        self._speeds = (felt_speed * self.n)(*speeds)  # kept alive until the launch returns
        self.api("felt_batch_update_async", self.h, self._speeds)

    def wait(self) -> None:
        self.api("felt_batch_wait", self.h)

    def update_boxes(self, ids: Sequence[int], lo, hi, speeds: Sequence[felt_speed]) -> None:
        """update(lo, hi, fn) of the listed members as one step."""
        m = len(ids)
        idv = _arr(ids, np.int32)
        lov = _arr(lo, np.int32)
        hiv = _arr(hi, np.int32)
        arr = (felt_speed * max(m, 1))(*speeds)
        self.api("felt_batch_update_boxes", self.h, m, _ptr(idv, _I32P), _ptr(lov, _I32P),
                 _ptr(hiv, _I32P), arr)

    def snapshot(self, i: int) -> np.ndarray:
        out = np.empty(self.sizes[i], dtype=np.float32)
        self.api("felt_batch_snapshot", self.h, int(i), _ptr(out, _F32P), out.size)
        return out

    def layer_size(self, i: int, layer: int) -> int:
        n = C.c_uint64()
        self.api("felt_batch_layer_size", self.h, int(i), int(layer), C.byref(n))
        return int(n.value)

    def layer_positions(self, i: int, layer: int) -> np.ndarray:
        n = self.layer_size(i, layer)
        out = np.empty((max(n, 1), self.D), dtype=np.int32)
        got = C.c_size_t()
        self.api("felt_batch_layer_positions", self.h, int(i), int(layer), _ptr(out, _I32P),
                 out.shape[0], C.byref(got))
        return out[: got.value].copy()

    def layers(self, i: int) -> dict:
        return {k: {tuple(int(x) for x in p) for p in self.layer_positions(i, k)}
                for k in range(-self.L, self.L + 1)}

    def stats(self, i: int) -> felt_stats:
        st = felt_stats()
        self.api("felt_batch_stats", self.h, int(i), C.byref(st))
        return st

    def check_records(self) -> int:
        """Active bricks whose face-neighbour record disagrees with the page table."""
        n = C.c_uint64()
        self.api("felt_batch_check_records", self.h, C.byref(n))
        return int(n.value)

    def ray(self, members, origins, dirs) -> np.ndarray:
        """Surface::ray of ray i against member members[i]; a miss reads FLT_MAX."""
        mv = _arr(members, np.int32)
        o = np.ascontiguousarray(np.asarray(origins, dtype=np.float32).reshape(-1, self.D))
        d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float32).reshape(-1, self.D))
        if not (len(mv) == len(o) == len(d)):
            raise ValueError("members, origins and directions differ in count")
        out = np.empty_like(o)
        self.api("felt_batch_ray", self.h, len(o), _ptr(mv, _I32P), _ptr(o, _F32P),
                 _ptr(d, _F32P), _ptr(out, _F32P))
        return out

    def save(self, i: int) -> bytes:
        """Surface::save of member i: the reference's cereal archive of its isogrid."""
        n = C.c_size_t()
        self.api("felt_batch_save", self.h, int(i), None, 0, C.byref(n))
        # numpy buffers: ctypes.string_at takes a C int length and cuts archives of 2 GiB or more
        buf = np.empty(max(n.value, 1), dtype=np.uint8)
        self.api("felt_batch_save", self.h, int(i), _ptr(buf, _U8P), n.value, C.byref(n))
        return buf[:n.value].tobytes()

    def load(self, i: int, data: bytes) -> None:
        """Replace member i's state with an archive of the same geometry."""
        buf = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
        self.api("felt_batch_load", self.h, int(i), _ptr(buf, _U8P), len(data))

    def clear(self, i: int) -> None:
        """Release member i's bricks; it becomes an empty grid."""
        self.api("felt_batch_clear", self.h, int(i))

    def step_info(self):
        u, r = C.c_uint64(), C.c_uint64()
        self.api("felt_batch_step_info", self.h, C.byref(u), C.byref(r))
        return int(u.value), int(r.value)

    def set_graph(self, enable: bool) -> None:
        self.api("felt_batch_set_graph", self.h, int(bool(enable)))

    def time_pass(self, pass_id: int, reps: int = 20):
        us, items, n = C.c_double(), C.c_double(), C.c_int()
        self.api("felt_batch_time_pass", self.h, int(pass_id), int(reps), C.byref(us),
                 C.byref(items), C.byref(n))
        return float(us.value), float(items.value), int(n.value)

    def autotune(self, reps: int = 10) -> dict:
        """felt_batch_autotune: times launch shapes on the members' current state and keeps the
        fastest; returns the chosen felt_launch_config as a dict."""
        out = felt_launch_config()
        self.api("felt_batch_autotune", self.h, int(reps), C.byref(out))
        return out.as_dict()

    def set_launch_config(self, threads_per_block: int, grid_cap: int,
                          voxels_per_lane: int) -> None:
        cfg = felt_launch_config(threads_per_block, grid_cap, voxels_per_lane, 0, 0.0, 0.0)
        self.api("felt_batch_set_launch_config", self.h, C.byref(cfg))

    def set_round_limit(self, max_rounds: int) -> None:
        self.api("felt_batch_set_round_limit", self.h, int(max_rounds))

    def sync(self) -> None:
        self.api("felt_batch_sync", self.h)


class Group:
    """Batches stepped by one graph launch per step (felt_group_*).  The batches must outlive
    the group."""

    def __init__(self, batches: Sequence[Batch]):
        self.batches = list(batches)
        self.api = self.batches[0].api
        arr = (_P * len(self.batches))(*[b.h for b in self.batches])
        h = _P()
        self.api("felt_group_create", arr, len(self.batches), C.byref(h))
        self.h = h
        self._keep = None

    def update_async(self, speeds: Sequence[Sequence[felt_speed]]) -> None:
        """Queue update(fn) of every member of every batch; speeds[k] are batch k's."""
        rows = [(felt_speed * len(sp))(*sp) for sp in speeds]
        ptrs = (C.POINTER(felt_speed) * len(rows))(*[C.cast(r, C.POINTER(felt_speed))
                                                     for r in rows])
        self._keep = (rows, ptrs)
        self.api("felt_group_update_async", self.h, ptrs)

    def wait(self) -> None:
        self.api("felt_group_wait", self.h)
        self._keep = None

    def update(self, speeds) -> None:
        self.update_async(speeds)
        self.wait()

    def close(self):
        if self.h:
            self.api("felt_group_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPolys:
    """Polygonisation of every member of a batch (Polys.hpp:29-201 per member)."""

    def __init__(self, batch: Batch, handle=None):
        self.bt = batch
        self.api = batch.api
        self.own = handle is None
        if handle is not None:  # borrowed from a server
            self.h = handle
            return
        h = _P()
        self.api("felt_batch_polys_create", batch.h, C.byref(h))
        self.h = h

    def close(self):
        if self.h and self.own:
            self.api("felt_polys_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def notify(self) -> None:
        self.api("felt_polys_notify", self.h)

    def march(self) -> None:
        self.api("felt_polys_march", self.h)

    def invalidate(self) -> None:
        self.api("felt_polys_invalidate", self.h)

    def totals(self):
        nv, ns = C.c_uint64(), C.c_uint64()
        self.api("felt_polys_totals", self.h, C.byref(nv), C.byref(ns))
        return int(nv.value), int(ns.value)

    def changes(self, i: int) -> np.ndarray:
        n = C.c_size_t()
        cap = 1
        for c, p in zip(self.bt.sizes[i], self.bt.part):
            cap *= c // p + 1
        out = np.empty(max(cap, 16), dtype=np.uint64)
        self.api("felt_batch_polys_changes", self.h, int(i), _ptr(out, _U64P), out.size,
                 C.byref(n))
        return out[: n.value].copy()

    def child_info(self, i: int, pos_idx: int):
        act = C.c_int()
        nv, ns = C.c_uint64(), C.c_uint64()
        self.api("felt_batch_polys_child_info", self.h, int(i), int(pos_idx), C.byref(act),
                 C.byref(nv), C.byref(ns))
        return bool(act.value), int(nv.value), int(ns.value)

    def child_mesh(self, i: int, pos_idx: int):
        _, nv, ns = self.child_info(i, pos_idx)
        D = self.bt.D
        vw = 6 if D == 3 else 2
        vtx = np.empty((max(nv, 1), vw), dtype=np.float32)
        spx = np.empty((max(ns, 1), D), dtype=np.uint32)
        self.api("felt_batch_polys_child_mesh", self.h, int(i), int(pos_idx), _ptr(vtx, _F32P),
                 vtx.size, _ptr(spx, _U32P), spx.size)
        return vtx[:nv].copy(), spx[:ns].copy()
