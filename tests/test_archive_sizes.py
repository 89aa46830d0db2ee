"""Archives of 2 GiB and more pass through the Python save / load paths whole (a full-size
1024^3 surface in 32^3 partitions saves about 5 GB).  A stand-in binding writes a pattern of
2^31 + 24 bytes; save must return all of it, and load must hand the library the same length
and bytes."""
import ctypes as C

import numpy as np

from felt_amd._binding import Surface

N = (1 << 31) + 24


class _FakeBinding:
    def __init__(self):
        self.seen = None

    def call(self, name, *args):
        if name == "surface_save":
            _, dst, cap, n = args
            n._obj.value = N
            if dst is not None:
                assert cap >= N
                addr = C.cast(dst, C.c_void_p).value
                C.memset(addr, 0x5A, N)
                C.memmove(addr + N - 8, (C.c_uint8 * 8)(*range(1, 9)), 8)
            return 0
        if name == "surface_load":
            src, n, _ = args
            addr = C.cast(src, C.c_void_p).value
            self.seen = (n, bytes((C.c_uint8 * 8).from_address(addr + n - 8)))
            raise RuntimeError("stop after the call")
        raise AssertionError(name)


def test_save_returns_archives_over_2gib_whole():
    s = Surface.__new__(Surface)
    s.b = _FakeBinding()
    s.h = None
    data = s.save()
    assert len(data) == N
    assert data[:4] == b"\x5a" * 4 and data[-8:] == bytes(range(1, 9))
    del data


def test_load_passes_archives_over_2gib_whole(monkeypatch):
    import felt_amd._binding as bmod
    monkeypatch.setattr(bmod, "archive_geometry", lambda d: (3, 3, (8, 8, 8), (8, 8, 8)))
    data = np.full(N, 0x5A, np.uint8)
    data[-8:] = np.arange(1, 9, dtype=np.uint8)
    fake = _FakeBinding()
    raw = data.tobytes()
    del data
    try:
        Surface.load(fake, raw)
    except RuntimeError:
        pass
    assert fake.seen == (N, bytes(range(1, 9)))
