"""Comparison of a GPU surface with the CPU oracle after the same calls on the same inputs.
Layer membership, active partitions and partition lists must be identical; distances within
the stated tolerance (they are expected to be bit-identical: both sides build with
-ffp-contract=off and use correctly rounded fp32 division and square root)."""
import numpy as np

from felt_amd._binding import LIST_DELTA, LIST_ISOGRID, LIST_STATUS_CHANGE

PHI_TOL = 1e-5


def position_keys(pos, size) -> np.ndarray:
    """Sorted row-major indices (last axis fastest) of world positions in a grid of `size`."""
    pos = np.asarray(pos, dtype=np.int64).reshape(-1, len(size))
    key = np.zeros(len(pos), dtype=np.int64)
    for a, n in enumerate(size):
        key = key * n + (pos[:, a] + n // 2)
    return np.sort(key)


def keys_diff(name, a, b, size):
    """A message when the sorted key arrays a and b differ, with a few positions of each side."""
    if np.array_equal(a, b):
        return None

    def pos(keys):
        out = []
        for k in keys[:5].tolist():
            p = []
            for n in reversed(size):
                p.append(k % n - n // 2)
                k //= n
            out.append(tuple(reversed(p)))
        return out
    return (f"{name}: gpu-only {pos(np.setdiff1d(a, b))} cpu-only {pos(np.setdiff1d(b, a))}"
            f" ({len(a)} vs {len(b)})")


def layer_diff(g, c, k):
    return keys_diff(f"layer {k}", position_keys(g.layer_positions(k), g.size),
                     position_keys(c.layer_positions(k), c.size), g.size)


def state_diff(g, c, snapshot=True, exact=False):
    """List of differences between GPU surface g and oracle surface c (empty when equal).  With
    `exact` the snapshots must be bit-identical rather than within PHI_TOL."""
    out = []
    L = g.L
    for k in range(-L, L + 1):
        d = layer_diff(g, c, k)
        if d:
            out.append(d)
    if not np.array_equal(g.partition_active(), c.partition_active()):
        out.append("active partitions differ")
    for kind, name in ((LIST_ISOGRID, "isogrid"), (LIST_DELTA, "delta"),
                       (LIST_STATUS_CHANGE, "status_change")):
        for k in range(2 * L + 1):
            a = set(int(x) for x in g.partition_list(kind, k))
            b = set(int(x) for x in c.partition_list(kind, k))
            if a != b:
                out.append(f"{name}({k}) partitions: gpu-only {sorted(a - b)[:5]} "
                           f"cpu-only {sorted(b - a)[:5]}")
    if snapshot:
        sg, sc = g.snapshot(), c.snapshot()
        d = np.abs(sg.astype(np.float64) - sc.astype(np.float64))
        if not np.all(np.isfinite(sg) == np.isfinite(sc)) or np.nanmax(d) > PHI_TOL:
            i = int(np.nanargmax(d))
            out.append(f"phi max diff {float(np.nanmax(d))} at flat {i}: gpu {sg.reshape(-1)[i]}"
                       f" cpu {sc.reshape(-1)[i]}")
        elif exact and not np.array_equal(sg.view(np.uint32), sc.view(np.uint32)):
            i = int(np.argmax(sg.view(np.uint32) != sc.view(np.uint32)))
            out.append(f"phi not bit-identical at flat {i}: gpu {sg.reshape(-1)[i]!r} "
                       f"cpu {sc.reshape(-1)[i]!r}")
    sa, sb = g.stats(), c.stats()
    if (sa.active_isogrid_partitions, sa.active_delta_partitions) != \
            (sb.active_isogrid_partitions, sb.active_delta_partitions):
        out.append("stats differ")
    return out


def mesh_sets(polys, surface, ndigits=4):
    """Per active child: (vertex position set, simplex position-tuple set) rounded to a grid of
    10^-ndigits, plus raw arrays for tolerance checks."""
    res = {}
    D = surface.D
    for p in range(surface.num_partitions):
        act, nv, ns = polys.child_info(p)
        if not act:
            continue
        vtx, spx = polys.child_mesh(p)
        res[p] = (vtx, spx)
    return res


def mesh_diff(pg, sg, pc, sc, tol=1e-5, changes=True):
    """Unordered comparison of two Polys meshes, child by child (and of their last march's
    changed partitions when `changes`)."""
    out = mesh_sets_diff(mesh_sets(pg, sg), mesh_sets(pc, sc), sg.D, tol)
    if pg.totals() != pc.totals():
        out.append(f"totals {pg.totals()} vs {pc.totals()}")
    if changes and set(int(x) for x in pg.changes()) != set(int(x) for x in pc.changes()):
        out.append("changes() differ")
    return out


def mesh_sets_diff(a, b, D, tol=1e-5):
    """Unordered comparison of two meshes given as {child: (vertices, simplices)}."""
    out = []
    if set(a) != set(b):
        out.append(f"active children differ: gpu-only {sorted(set(a) - set(b))[:5]} "
                   f"cpu-only {sorted(set(b) - set(a))[:5]}")
    for p in sorted(set(a) & set(b)):
        (vg, tg), (vc, tc) = a[p], b[p]
        if len(vg) != len(vc) or len(tg) != len(tc):
            out.append(f"child {p}: counts {len(vg)}/{len(tg)} vs {len(vc)}/{len(tc)}")
            continue
        if len(vg) == 0:
            continue
        og = np.lexsort(vg[:, :D].T[::-1])
        oc = np.lexsort(vc[:, :D].T[::-1])
        dv = np.abs(vg[og].astype(np.float64) - vc[oc].astype(np.float64)).max()
        if dv > tol:
            out.append(f"child {p}: vertex max diff {dv}")
            continue
        # simplices as sets of rounded position tuples
        def spx_set(v, t):
            key = np.round(v[:, :D].astype(np.float64), 4)
            return {tuple(tuple(key[i]) for i in tri) for tri in t.tolist()}
        if spx_set(vg, tg) != spx_set(vc, tc):
            out.append(f"child {p}: simplex sets differ")
    return out
