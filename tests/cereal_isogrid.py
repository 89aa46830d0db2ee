"""Reader and writer of the reference's on-disk surface format, written from the serialisation
code itself: Surface::save / load write the IsoGrid with a cereal BinaryOutputArchive
(Surface.hpp:180-196, PartitionedMixin.hpp:692-712).

cereal writes arithmetic values as raw little-endian bytes and container sizes as uint64
(cereal/macros.hpp:61-70, cereal/types/vector.hpp:40-85); Felt writes an Eigen vector as int32
rows, int32 cols and a std::vector of its elements (PartitionedMixin.hpp:719-735).  The IsoGrid
serialises Access, Children, Size, Spatial in that order (Partitioned.hpp:151-158); a child grid
Activate, Data, LookupInterface, Size (Tracked.hpp:53-60, 210-217); a lookup grid Activate, Data,
lists, Size (Lookup.hpp:170-177, 229-236)."""
import struct

NULL = (1 << 64) - 1


class _In:
    def __init__(self, b: bytes):
        self.b, self.at = b, 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.at)
        self.at += struct.calcsize("<" + fmt)
        return v

    def one(self, fmt):
        return self.take(fmt)[0]

    def eigen(self, elem):
        rows, cols, n = self.take("iiQ")
        assert cols == 1 and n == rows, (rows, cols, n)
        if elem == "lists":
            return [list(self.take(f"{self.one('Q')}Q")) for _ in range(n)]
        return list(self.take(f"{n}{elem}"))

    def vec(self, elem):
        n = self.one("Q")
        return list(self.take(f"{n}{elem}"))

    def sizes(self):
        return [self.eigen("i") for _ in range(3)]


def _child(r: _In):
    c = {"bg": r.one("f"), "data": r.vec("f")}
    c["lookup_bg"] = r.one("Q")
    c["look"] = r.vec("Q")
    c["lists"] = r.eigen("lists")
    c["lookup_sizes"] = r.sizes()
    c["sizes"] = r.sizes()
    return c


def parse(b: bytes) -> dict:
    """The archive as a dict; asserts that every byte is consumed."""
    r = _In(b)
    g = {"bg": r.one("f"), "child_size": r.eigen("i")}
    g["bg_child"] = _child(r)
    g["children"] = [_child(r) for _ in range(r.one("Q"))]
    g["clookup_bg"] = r.eigen("Q")
    g["ctuples"] = [r.eigen("Q") for _ in range(r.one("Q"))]
    g["clists"] = r.eigen("lists")
    g["clookup_sizes"] = r.sizes()
    g["csizes"] = r.sizes()
    g["sizes"] = r.sizes()
    g["dx"] = r.one("f")
    g["pos_min"] = r.eigen("i")
    g["pos_max"] = r.eigen("i")
    assert r.at == len(b), f"{len(b) - r.at} trailing bytes"
    return g


class _Out:
    def __init__(self):
        self.parts = []

    def put(self, fmt, *v):
        self.parts.append(struct.pack("<" + fmt, *v))

    def eigen(self, elem, v):
        self.put("iiQ", len(v), 1, len(v))
        if elem == "lists":
            for lst in v:
                self.put("Q", len(lst))
                self.put(f"{len(lst)}Q", *lst)
        else:
            self.put(f"{len(v)}{elem}", *v)

    def vec(self, elem, v):
        self.put("Q", len(v))
        self.put(f"{len(v)}{elem}", *v)

    def sizes(self, s):
        for v in s:
            self.eigen("i", v)


def _put_child(o: _Out, c):
    o.put("f", c["bg"])
    o.vec("f", c["data"])
    o.put("Q", c["lookup_bg"])
    o.vec("Q", c["look"])
    o.eigen("lists", c["lists"])
    o.sizes(c["lookup_sizes"])
    o.sizes(c["sizes"])


def write(g: dict) -> bytes:
    o = _Out()
    o.put("f", g["bg"])
    o.eigen("i", g["child_size"])
    _put_child(o, g["bg_child"])
    o.put("Q", len(g["children"]))
    for c in g["children"]:
        _put_child(o, c)
    o.eigen("Q", g["clookup_bg"])
    o.put("Q", len(g["ctuples"]))
    for t in g["ctuples"]:
        o.eigen("Q", t)
    o.eigen("lists", g["clists"])
    o.sizes(g["clookup_sizes"])
    o.sizes(g["csizes"])
    o.sizes(g["sizes"])
    o.put("f", g["dx"])
    o.eigen("i", g["pos_min"])
    o.eigen("i", g["pos_max"])
    return b"".join(o.parts)


def canonical(g: dict) -> dict:
    """The same state with every list in ascending order and the lookups re-indexed to match:
    list order is the only part of the format two correct writers may choose differently."""
    g = dict(g)
    kids = []
    for c in g["children"]:
        c = dict(c)
        c["lists"] = [sorted(lst) for lst in c["lists"]]
        if c["look"]:
            look = [NULL] * len(c["look"])
            for lst in c["lists"]:
                for i, leaf in enumerate(lst):
                    look[leaf] = i
            c["look"] = look
        kids.append(c)
    g["children"] = kids
    g["clists"] = [sorted(lst) for lst in g["clists"]]
    nl = len(g["clists"])
    tuples = [[NULL] * nl for _ in g["ctuples"]]
    for k, lst in enumerate(g["clists"]):
        for i, p in enumerate(lst):
            tuples[p][k] = i
    g["ctuples"] = tuples
    return g


def check_consistent(g: dict) -> list:
    """Internal consistency of an archive: lookups index their lists, each listed voxel's value
    rounds to its list's layer, a partition is in parent list k iff its list k is non-empty."""
    out = []
    nl = len(g["clists"])
    L = (nl - 1) // 2
    for p, c in enumerate(g["children"]):
        active = bool(c["data"])
        if active != bool(c["look"]):
            out.append(f"child {p}: data and lookup disagree on activation")
        for k, lst in enumerate(c["lists"]):
            for i, leaf in enumerate(lst):
                if c["look"][leaf] != i:
                    out.append(f"child {p}: lookup of leaf {leaf} is not {i}")
                v = c["data"][leaf]
                if round_half_away(v + 1.1920928955078125e-07) != k - L:
                    out.append(f"child {p}: leaf {leaf} value {v} not in layer {k - L}")
            listed = p in g["clists"][k]
            if listed != bool(lst):
                out.append(f"child {p}: parent list {k} membership {listed}, list length {len(lst)}")
    return out


def round_half_away(x: float) -> int:
    import math
    import numpy as np
    x = float(np.float32(x))
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)
