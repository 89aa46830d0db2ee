"""Writes grid_kats.json: the known answers of the reference's grid tests for the numeric operators
a speed functor uses (curvature, interpolation, the gradients, divergence) and for the index
convention.

The reference keeps them in src/tests/test_Grid.cpp as inline grid literals, set() calls and
CHECK lines.  This script reads those numbers out of the file by line number (text only; nothing
of the reference is compiled or run) and pairs each with the operator and position it checks,
written here in this project's own encoding.  Run it where the reference is checked out:

    python tests/golden/make_grid_kats.py /path/to/Felt > tests/golden/grid_kats.json

Encoding (replayed by tests/grid_kats.py):
    {"D", "size", "grid": row-major values, last axis fastest (the literal's own order),
     "checks": [{"op", "pos" (integer) or "rpos" (real), "dx", "expect" (number or vector),
                 "exact": the reference compares with ==, else with Catch's Approx}]}
A grid of size s has offset -s/2 (s = 3: -1 as in the reference's 3x3 and 3x3x3 grids, s = 4:
-2).  The reference's static interpolation examples (a vector of 2^D corner values reduced at
a position in the unit cell) are embedded as the corners (0..1)^D of a 4^D grid.
"""
import json
import re
import sys

NUM = r"-?\d+(?:\.\d+)?(?:e-?\d+)?"


def block_at(lines, line_no):
    """Numbers of the brace block that opens on 1-based line line_no."""
    text = []
    i = line_no - 1
    assert "{" in lines[i], lines[i]
    depth = 0
    started = False
    while True:
        for ch in lines[i]:
            if ch == "{":
                depth += 1
                started = True
            elif ch == "}":
                depth -= 1
        text.append(lines[i])
        if started and depth == 0:
            break
        i += 1
    body = "\n".join(text)
    body = body[body.index("{") + 1: body.rindex("}")]
    body = re.sub(r"//.*", "", body)
    return [float(x) for x in re.findall(NUM, body)]


def rhs_at(lines, line_no, token):
    """Expected value of the CHECK on line line_no, which must mention `token`:
    (value or vector, exact)."""
    s = lines[line_no - 1]
    assert "CHECK" in s and token in s, (line_no, s)
    rhs = s[s.index("==") + 2:]
    if "Vec" in rhs:
        inner = rhs[rhs.index("(") + 1: rhs.index(")")]
        return [float(x) for x in re.findall(NUM, inner)], True
    if "Approx(" in rhs:
        inner = rhs[rhs.index("Approx(") + 7:]
        return float(re.findall(NUM, inner)[0]), False
    return float(re.findall(NUM, rhs)[0]), True


def set_at(lines, line_no):
    """(position, value) of a grid.set(VecNi(...), v) call."""
    s = lines[line_no - 1]
    assert "grid.set(" in s, (line_no, s)
    vec = s[s.index("Vec") :]
    pos = [int(float(x)) for x in re.findall(NUM, vec[vec.index("(") + 1: vec.index(")")])]
    val = float(re.findall(NUM, vec[vec.index(")") + 1:])[0])
    return pos, val


def decl_has(lines, line_no, token):
    s = lines[line_no - 1]
    assert token in s, (line_no, s)


def lin(pos, size, offset):
    idx = 0
    for p, s, o in zip(pos, size, offset):
        idx = idx * s + (p - o)
    return idx


def main(root):
    path = root.rstrip("/") + "/src/tests/test_Grid.cpp"
    lines = open(path).read().split("\n")
    kats = []

    def add(name, cite, D, size, grid, checks):
        kats.append({"name": name, "cite": f"test_Grid.cpp:{cite}", "D": D, "size": size,
                     "grid": grid, "checks": checks})

    def chk(op, line, token, pos=None, rpos=None, dx=1.0, decl=None):
        if decl is not None:
            decl_has(lines, decl, op + "(")
        expect, exact = rhs_at(lines, line, token)
        c = {"op": op, "dx": dx, "expect": expect, "exact": exact, "line": line}
        if pos is not None:
            c["pos"] = pos
        else:
            c["rpos"] = rpos
        return c

    def grid3x3(D, sets=()):
        size = [3] * D
        g = [0.0] * (3 ** D)
        for p, v in sets:
            g[lin(p, size, [-1] * D)] = v
        return g

    # -- 3x3 grid, offset (-1,-1), background 0 (test_Grid.cpp:1993-2222) --
    sets = [set_at(lines, n) for n in (2018, 2019, 2020, 2021)]
    add("interp_2d_grid", "2016-2045", 2, [3, 3], grid3x3(2, sets), [
        chk("get", 2031, "val1", rpos=[0.0, 0.0]),
        chk("get", 2032, "val2", rpos=[-0.5, -0.5]),
        chk("get", 2033, "val3", rpos=[0.5, 0.5]),
        chk("get", 2043, "val", rpos=[0.5, 0.5]),
    ])
    add("grad_2d_flat", "2048-2077", 2, [3, 3], grid3x3(2), [
        chk("gradF", 2055, "grad", pos=[0, 0], decl=2051),
        chk("gradB", 2065, "grad", pos=[0, 0], decl=2061),
        chk("gradC", 2075, "grad", pos=[0, 0], decl=2071),
    ])
    add("grad_2d_peak", "2079-2111", 2, [3, 3], grid3x3(2, [set_at(lines, 2081)]), [
        chk("gradF", 2089, "grad", pos=[0, 0], decl=2085),
        chk("gradB", 2099, "grad", pos=[0, 0], decl=2095),
        chk("gradC", 2109, "grad", pos=[0, 0], decl=2105),
    ])
    # the literal's orientation: get(-1, 0) == 2 puts x outermost, the last axis fastest
    orient = []
    for n in range(2126, 2135):
        s = lines[n - 1]
        pos = [int(float(x)) for x in re.findall(NUM, s[s.index("Vec2i(") + 6: s.index(")")])]
        orient.append({"op": "get", "pos": pos, "dx": 1.0,
                       "expect": float(re.findall(NUM, s[s.index("==") + 2:])[0]),
                       "exact": True, "line": n})
    add("literal_order_2d", "2115-2136", 2, [3, 3], block_at(lines, 2117), orient)
    for name, lit, chk_line in (("gradE_2d_pos_flow", 2140, 2149), ("gradE_2d_neg_flow", 2155, 2164),
                                ("gradE_2d_pos_div", 2170, 2178), ("gradE_2d_neg_div", 2184, 2192)):
        add(name, f"{lit - 2}-{chk_line + 1}", 2, [3, 3], block_at(lines, lit),
            [chk("gradE", chk_line, "gradE", pos=[0, 0])])
    add("curv_2d_max", "2196-2208", 2, [3, 3], block_at(lines, 2198),
        [chk("curv", 2206, "curv", pos=[0, 0])])
    add("curv_2d_corner", "2210-2222", 2, [3, 3], block_at(lines, 2212),
        [chk("curv", 2220, "curv", pos=[0, 0])])

    # -- 3x3x3 grid, offset (-1,-1,-1), background 0 (test_Grid.cpp:2226-2415) --
    add("grad_3d", "2234-2315", 3, [3, 3, 3], block_at(lines, 2236), [
        chk("gradF", 2253, "grad", pos=[0, 0, 0], decl=2252),
        chk("gradB", 2260, "grad", pos=[0, 0, 0], decl=2258),
        chk("gradC", 2267, "grad", pos=[0, 0, 0], decl=2265),
        chk("grad", 2274, "grad", pos=[0, 0, 0], decl=2272),
        chk("grad", 2281, "grad", pos=[0, -1, 0], decl=2279),
        chk("grad", 2288, "grad", pos=[1, 0, 1], decl=2286),
        chk("gradF", 2298, "grad", pos=[0, 0, 0], dx=2.0, decl=2297),
        chk("gradB", 2305, "grad", pos=[0, 0, 0], dx=2.0, decl=2303),
        chk("gradC", 2312, "grad", pos=[0, 0, 0], dx=2.0, decl=2310),
    ])
    add("divergence_3d", "2317-2352", 3, [3, 3, 3], block_at(lines, 2319), [
        chk("divergence", 2335, "divergence", pos=[0, 0, 0]),
        chk("divergence", 2344, "divergence", pos=[0, 0, 0], dx=2.0),
        chk("curv", 2350, "curv", pos=[0, 0, 0]),
    ])
    add("curv_3d_corner", "2354-2374", 3, [3, 3, 3], block_at(lines, 2356),
        [chk("curv", 2372, "curv", pos=[0, 0, 0])])
    add("curv_3d_sharp_corner", "2376-2396", 3, [3, 3, 3], block_at(lines, 2378),
        [chk("curv", 2394, "curv", pos=[0, 0, 0])])

    # -- static interpolation examples (test_Grid.cpp:1873-1990), embedded in a 4^D grid --
    def embed(D, corners):
        size = [4] * D
        g = [0.0] * (4 ** D)
        for i, v in enumerate(corners):
            p = [(i >> a) & 1 for a in range(D)]
            g[lin(p, size, [-2] * D)] = v
        return size, g

    vec1 = block_at(lines, 1876)  # {1.0f, 0}
    assert len(vec1) == 2
    size, g = embed(2, [vec1[0], vec1[1], vec1[0], vec1[1]])
    add("interp_1d_vector", "1873-1892", 2, size, g,
        [chk("get", 1888, "input[0]", rpos=[0.3, 0.0])])
    c2 = [0.0] * 4
    for n in range(1898, 1902):
        s = lines[n - 1]
        c2[int(re.findall(r"input\[(\d)", s)[0])] = float(re.findall(NUM, s.split("=")[1])[0])
    size, g = embed(2, c2)
    add("interp_2d_vector", "1894-1928", 2, size, g,
        [chk("get", 1922, "input[0]", rpos=[0.8, 0.5])])
    c3 = [0.0] * 8
    for n in range(1942, 1950):
        s = lines[n - 1]
        c3[int(re.findall(r"input\[(\d)", s)[0])] = float(re.findall(NUM, s.split("=")[1])[0])
    size, g = embed(3, c3)
    add("interp_3d_vector", "1930-1990", 3, size, g,
        [chk("get", 1983, "input[0", rpos=[0.5, 0.75, 0.5])])

    # -- index convention (test_Grid.cpp:72-89) --
    s = lines[80]
    assert "613" in s and "Vec3i(1, 0, -1)" in s, s
    index = {"name": "index_613", "cite": "test_Grid.cpp:72-89", "size": [7, 11, 13],
             "offset": [-3, -3, -3], "pos": [1, 0, -1],
             "index": int(re.findall(r"== (\d+)", s)[0])}

    # -- neighbour order (test_Grid.cpp:2398-2415) --
    nb = []
    for n in range(2406, 2412):
        nb += [[int(x) for x in re.findall(NUM, m)] for m in re.findall(r"Vec3i\(([^)]*)\)", lines[n - 1])]
    neighs = {"name": "neighs_order", "cite": "test_Grid.cpp:2398-2415", "pos": [-67, 54, 3],
              "order": nb}
    assert len(nb) == 6, nb

    json.dump({"source": "feltech/Felt src/tests/test_Grid.cpp", "kats": kats, "index": index,
               "neighs": neighs}, sys.stdout, indent=1)
    sys.stdout.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
