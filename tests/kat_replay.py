"""Replays the reference's known-answer scenarios (tests/golden/surface_kats.json) on any
Surface built over the felt ABI, and checks the recorded answers."""
import json
import os

import numpy as np

from felt_amd._binding import Surface, speed_constant

HERE = os.path.dirname(os.path.abspath(__file__))


def load_kats():
    with open(os.path.join(HERE, "golden", "surface_kats.json")) as f:
        return json.load(f)["kats"]


def replay(surface: Surface, ops):
    """Runs the operations; returns the surface they end on (save_load replaces it)."""
    for op in ops:
        kind = op[0]
        if kind == "save_load":
            surface = type(surface).load(surface.b, surface.save())
            continue
        if kind == "seed":
            surface.seed(op[1])
        elif kind == "update":
            surface.update(speed_constant(op[1]))
        elif kind == "update_points":
            pts = {tuple(p): v for p, v in op[1]}
            surface.update(lambda pos: pts.get(tuple(pos), 0.0))
        elif kind == "local":
            surface.update_start()
            surface.delta([p for p, _ in op[1]], [v for _, v in op[1]])
            surface.update_end()
        elif kind == "box":
            surface.update_box(op[1], op[2], speed_constant(op[3]))
        else:
            raise ValueError(kind)
    return surface


def grid_matches(snap: np.ndarray, expected, tol=1e-5):
    """The reference compares summed differences only (test_Surface.cpp:94-98), which hides the
    orientation of its literal; accept the literal as [x][y] or as [y][x] but require an
    element-wise match in one of them, and the summed check."""
    exp = np.asarray(expected, dtype=np.float64)
    s = snap.astype(np.float64)
    if abs(float((exp - s.reshape(-1)).sum())) > 1e-3:
        return False, "sum"
    if np.allclose(exp, s.reshape(-1), atol=tol):
        return True, "xy"
    if snap.ndim == 2 and np.allclose(exp, s.T.reshape(-1), atol=tol):
        return True, "yx"
    return False, "elementwise"


def check_kat(make_surface, kat):
    """Returns a list of failure strings (empty when the scenario matches)."""
    s = make_surface(kat["D"], kat["L"], kat["size"], kat["part"])
    fails = []
    s = replay(s, kat["ops"])
    L = kat["L"]
    if "grid" in kat:
        ok, how = grid_matches(s.snapshot(), kat["grid"])
        if not ok:
            fails.append(f"grid mismatch ({how})")
    if "layer_sizes" in kat:
        got = [s.layer_size(k) for k in range(-L, L + 1)]
        if got != kat["layer_sizes"]:
            fails.append(f"layer sizes {got} != {kat['layer_sizes']}")
    if "zero_layer" in kat:
        got = {tuple(int(x) for x in p) for p in s.layer_positions(0)}
        if got != {tuple(p) for p in kat["zero_layer"]}:
            fails.append(f"zero layer {sorted(got)}")
    if "active" in kat:
        act = s.partition_active()
        for cpos, want in kat["active"]:
            if bool(act[s.child_index(cpos)]) != want:
                fails.append(f"partition {cpos} active != {want}")
    if "affected" in kat:
        for layer, want in kat["affected"].items():
            got = {tuple(int(x) for x in p) for p in s.affected_positions(int(layer) + L)}
            if got != {tuple(p) for p in want}:
                fails.append(f"affected layer {layer}: {sorted(got)}")
    if "box_visits" in kat:
        lo, hi, want = kat["box_visits"]
        got = {tuple(int(x) for x in p) for p in s.zero_layer_in_box(lo, hi)}
        if got != {tuple(p) for p in want}:
            fails.append(f"box visits {sorted(got)}")
        # fn returning 0 visits exactly those voxels: the delta list holds them afterwards
        s.update_box(lo, hi, speed_constant(0.0))
    if "partition_layer" in kat:
        for cpos, k, n in kat["partition_layer"]:
            got = s.partition_layer_size(s.child_index(cpos), k)
            if got != n:
                fails.append(f"partition {cpos} layer {k} size {got} != {n}")
    if "stats" in kat:
        st = s.stats()
        got = [int(st.active_isogrid_partitions), int(st.active_delta_partitions)]
        if got != kat["stats"]:
            fails.append(f"stats {got} != {kat['stats']}")
    if "get" in kat:
        for pos, v in kat["get"]:
            if float(s.get([pos])[0]) != v:
                fails.append(f"get {pos}")
    if "children" in kat and s.num_partitions != kat["children"]:
        fails.append("children count")
    return fails
