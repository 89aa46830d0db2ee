"""Writes surface_kats.json: the known answers held by the reference's own unit tests.

The reference keeps its expected isogrids as inline literals in src/tests/test_Surface.cpp and
its expected counts, sets and vertices in src/tests/test_Surface.cpp and test_Poly.cpp.  This
script reads the numeric literals out of test_Surface.cpp by line number (text only; nothing of
the reference is compiled or run) and pairs each with the operations of its scenario, written
here in this project's own encoding.  Run it where the reference is checked out:

    python tests/golden/make_reference_kats.py /path/to/Felt > tests/golden/surface_kats.json

Operation encoding (replayed by tests/kat_replay.py):
    ["seed", pos]                      Surface::seed
    ["update", v]                      update(fn) with fn returning the constant v
    ["update_points", [[pos, v],...]]  update(fn) with fn returning v at pos and 0 elsewhere
    ["local", [[pos, v], ...]]         update_start(); delta(pos, v)...; update_end()
    ["box", lo, hi, v]                 update(lo, hi, fn) with fn returning v
    ["save_load"]                      Surface::save, then continue on Surface::load of it
"""
import json
import re
import sys


def grid_at(lines, line_no):
    """Numbers of the brace block that opens on 1-based line line_no."""
    text = []
    i = line_no - 1
    assert "{" in lines[i], lines[i]
    depth = 0
    started = False
    while True:
        s = lines[i]
        for ch in s:
            if ch == "{":
                depth += 1
                started = True
            elif ch == "}":
                depth -= 1
        text.append(s)
        if started and depth == 0:
            break
        i += 1
    body = "".join(text)
    body = body[body.index("{") + 1: body.rindex("}")]
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?", body)]


def main(root):
    path = root.rstrip("/") + "/src/tests/test_Surface.cpp"
    lines = open(path).read().split("\n")
    g = lambda n: grid_at(lines, n)
    cite = lambda a, b: f"test_Surface.cpp:{a}-{b}"
    seed9 = [["seed", [0, 0]]]
    kats = []

    def add(name, cit, D, L, size, part, ops, grid=None, sizes=None, **extra):
        k = {"name": name, "cite": cit, "D": D, "L": L, "size": size, "part": part, "ops": ops}
        if grid is not None:
            k["grid"] = grid
        if sizes is not None:
            k["layer_sizes"] = sizes
        k.update(extra)
        kats.append(k)

    S9 = dict(D=2, L=2, size=[9, 9], part=[3, 3])
    add("seed", cite(49, 103), ops=seed9, grid=g(79), sizes=[0, 0, 1, 4, 8], **S9)
    add("expand_1", cite(105, 160), ops=seed9 + [["update", -1.0]], grid=g(108),
        sizes=[0, 1, 4, 8, 12], zero_layer=[[-1, 0], [1, 0], [0, 1], [0, -1]], **S9)
    add("expand_2", cite(185, 212), ops=seed9 + [["update", -1.0]] * 2, grid=g(192), **S9)
    add("expand_consume", cite(214, 253), ops=seed9 + [["update", -1.0]] * 11, grid=g(226),
        sizes=[0, 0, 0, 0, 0], **S9)
    add("expand_contract", cite(256, 290), ops=seed9 + [["update", -1.0], ["update", 1.0]],
        grid=g(267), sizes=[0, 0, 1, 4, 8], **S9)
    add("expand_contract_collapse", cite(292, 350),
        ops=seed9 + [["update", -1.0], ["update", 1.0], ["update", 1.0]], grid=g(324),
        sizes=[0, 0, 0, 0, 0], zero_layer=[], **S9)
    add("expand_0.6", cite(353, 388), ops=seed9 + [["update", -0.6]], grid=g(364),
        sizes=[0, 1, 4, 8, 12], **S9)
    add("expand_0.6_contract_0.6", cite(390, 425),
        ops=seed9 + [["update", -0.6], ["update", 0.6]], grid=g(402),
        sizes=[0, 0, 1, 4, 8], **S9)
    add("expand_0.6_contract_0.6_twice", cite(427, 461),
        ops=seed9 + [["update", -0.6], ["update", 0.6], ["update", 0.6]], grid=g(437),
        sizes=[0, 0, 0, 0, 0], **S9)
    add("save_load_update", cite(464, 509),
        ops=seed9 + [["update", -0.6], ["save_load"], ["update", 0.6]], grid=g(489),
        sizes=[0, 0, 1, 4, 8], **S9)
    add("deactivate_centre", cite(519, 580), D=2, L=2, size=[21, 21], part=[2, 2],
        ops=seed9 + [["update", -1.0]] * 5, grid=g(540),
        active=[[[0, 1], True], [[0, -1], True], [[1, 0], True], [[-1, 0], True],
                [[0, 0], False]])
    S9p2 = dict(D=2, L=2, size=[9, 9], part=[2, 2])
    add("local_collapse", cite(598, 633), ops=seed9 + [["local", [[[0, 0], 1.0]]]],
        grid=g(608), sizes=[0, 0, 0, 0, 0], **S9p2)
    add("local_expand", cite(635, 666), ops=seed9 + [["local", [[[0, 0], -1.0]]]],
        grid=g(643), sizes=[0, 1, 4, 8, 12], **S9p2)
    add("local_affected", cite(668, 759),
        ops=seed9 + [["local", [[[0, 0], -1.0]]], ["local", [[[0, 1], 0.3], [[1, 0], 0.3]]]],
        affected={"-2": [], "-1": [[0, 0]],
                  "1": [[-1, 1], [1, 1], [0, 2], [2, 0], [1, -1]],
                  "2": [[-2, 1], [2, 1], [-1, 2], [1, 2], [0, 3], [3, 0], [1, -2], [2, -1]]},
        **S9p2)
    add("box_visits_partial", cite(761, 784),
        ops=seed9 + [["local", [[[0, 0], -1.0]]]], box_visits=[[1, 0], [3, 3], [[1, 0]]],
        **S9p2)
    add("box_visits_all", cite(786, 808),
        ops=seed9 + [["local", [[[0, 0], -1.0]]]],
        box_visits=[[-100, -100], [100, 100], [[1, 0], [-1, 0], [0, 1], [0, -1]]], **S9p2)
    add("box_expand_0.6", cite(810, 878), ops=seed9 + [["box", [-1, -1], [1, 1], -0.6]],
        grid=g(816), sizes=[0, 1, 4, 8, 12], **S9p2)
    S16 = dict(D=2, L=2, size=[16, 9], part=[3, 3])
    two = [["seed", [-4, 0]], ["seed", [4, 0]], ["update", -1.0]]
    add("two_seeds", cite(888, 935), ops=two, grid=g(911), sizes=[0, 2, 8, 16, 24],
        partition_layer=[[[0, 0], 2, 3], [[1, 0], 2, 3]], **S16)
    add("two_seeds_towards", cite(937, 977),
        ops=two + [["local", [[[-3, 0], -1.0], [[3, 0], -1.0]]]], grid=g(951),
        sizes=[0, 4, 12, 20, 27], partition_layer=[[[0, 0], 2, 3], [[1, 0], 2, 2]], **S16)
    cyc = []
    for _ in range(10):
        cyc.append(["local", [[[-3, 0], -1.0], [[3, 0], -1.0]]])
        cyc.append(["local", [[[-3, 1], 1.0], [[-2, 0], 1.0], [[-3, -1], 1.0],
                              [[3, 1], 1.0], [[2, 0], 1.0], [[3, -1], 1.0]]])
    add("two_seeds_cycle", cite(979, 1028), ops=two + cyc, grid=g(1005),
        sizes=[0, 2, 8, 16, 24], partition_layer=[[[0, 0], 2, 3], [[1, 0], 2, 3]], **S16)
    S11 = dict(D=3, L=3, size=[11, 11, 11], part=[3, 3, 3])
    base3 = [["seed", [0, 0, 0]], ["update", -1.0]]
    add("layers_3d", cite(1031, 1053), ops=base3, sizes=[0, 0, 1, 6, 18, 38, 66], **S11)
    add("layers_3d_local", cite(1055, 1080),
        ops=base3 + [["local", [[[0, 1, 0], -1.0]]],
                     ["local", [[[0, 2, 0], 1.0], [[1, 1, 0], 1.0], [[-1, 1, 0], 1.0],
                                [[0, 1, 1], 1.0], [[0, 1, -1], 1.0]]]],
        sizes=[0, 0, 1, 6, 18, 38, 66], **S11)
    S12 = dict(D=2, L=3, size=[12, 12], part=[2, 2])
    pre = [["seed", [-2, -2]], ["seed", [2, 2]],
           ["local", [[[-2, -2], -1.0]]], ["local", [[[-1, -2], -1.0]]],
           ["local", [[[0, -2], -1.0]]]]
    add("local_vs_global_local", cite(1155, 1214),
        ops=pre + [["local", [[[1, -2], -1.0], [[2, 2], 1.0]]], ["local", [[[1, -1], -1.0]]]],
        grid=g(1188), sizes=[0, 0, 5, 11, 15, 19, 23], **S12)
    add("local_vs_global_global", cite(1216, 1272),
        ops=pre + [["update_points", [[[1, -2], -1.0], [[2, 2], 1.0]]],
                   ["update_points", [[[1, -1], -1.0]]]],
        grid=g(1245), sizes=[0, 0, 5, 11, 15, 19, 23], **S12)
    S7 = dict(D=2, L=2, size=[7, 7], part=[3, 3])
    add("stats_initial", cite(1543, 1557), ops=[], stats=[0, 0], **S7)
    add("stats_seeded", cite(1559, 1573), ops=seed9, stats=[4, 0], **S7)
    add("stats_expanded", cite(1575, 1590), ops=seed9 + [["update", -1.0]], stats=[6, 4], **S7)
    add("initial", cite(32, 48), ops=[], get=[[[0, 0], 3.0]],
        children=9, **S7)
    json.dump({"source": "feltech/Felt src/tests/test_Surface.cpp", "kats": kats}, sys.stdout,
              indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
