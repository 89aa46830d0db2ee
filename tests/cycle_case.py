"""A band whose round loop never ends under the reference's algorithm (Surface.hpp:739-762,
988-1125, 1364-1413).

16^3 grid, L = 3, everything -4 (inside, out of band) except a city-block ball about the
centre c = (0, 0, 0): c at -1, its face neighbours at -2, distance-2 voxels at -3, and c's
neighbour n0 = (-1, 0, 0) at +4 (outside, out of band).  There is no zero layer.  distance()
takes the neighbour nearest the zero curve on either side (its `dist_neigh * dir < dist`), so:

round 1 (band, layers -1, -2, -3 in turn):  c sees n0 = +4 and becomes 3; the -2 voxels then see
  c = 3 and become 2; the -3 voxels see a +2 neighbour and become 1, except the five next to n0,
  which see +4 and become 3;
round 2 (the changed voxels, layers 1, 2, 3): the +1 voxels see the -4 background and become -3,
  the +2 voxels then see a -3 and become -2, c and the +3 voxels see -2 or -4 and become -1 or
  -3: the state of round 1's start.

So rounds repeat in pairs and the loop never ends (the reference's debug build throws at the
first jump of more than one layer, Surface.hpp:1044-1059)."""
import numpy as np

SIZE, PART, L = (16, 16, 16), (8, 8, 8), 3
OFF = np.array([8, 8, 8])
N0 = (-1, 0, 0)


def initial():
    phi = np.full(SIZE, -4.0, np.float32)
    for x in range(-3, 4):
        for y in range(-3, 4):
            for z in range(-3, 4):
                r = abs(x) + abs(y) + abs(z)
                if r <= 2:
                    phi[tuple(OFF + [x, y, z])] = -1.0 - r
    phi[tuple(OFF + list(N0))] = 4.0
    return phi


def flipped():
    """The state after round 1 (and after every odd round)."""
    phi = initial()
    out = phi.copy()
    for x in range(-3, 4):
        for y in range(-3, 4):
            for z in range(-3, 4):
                r = abs(x) + abs(y) + abs(z)
                p = (x, y, z)
                if r > 2 or p == N0:
                    continue
                if r == 0:
                    v = 3.0
                elif r == 1:
                    v = 2.0
                else:
                    next_to_n0 = abs(x - N0[0]) + abs(y - N0[1]) + abs(z - N0[2]) == 1
                    v = 3.0 if next_to_n0 else 1.0
                out[tuple(OFF + list(p))] = v
    return out
