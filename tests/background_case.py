"""A partition emptied by one flush with untracks out of both sides (the background it is left
with, Surface.hpp:1183-1244; Tracked.hpp untrack).

2D, 16x16 grid (offset -8) in 4x4 partitions, L = 3.  phi = +4 for x >= 4, and for 2 <= x < 4
inside the partition P = [0,4) x [0,4); -4 elsewhere (so every other partition is uniform), except
two band voxels in P: (0,1) at -3 and (3,2) at +3.  Every
neighbour of (0,1) is -4, so distance() gives it max(-3, -4) - 1 = -4: it leaves layer -3 for -4;
every neighbour of (3,2) is +4, so it gets min(3, 4) + 1 = 4 and leaves layer 3 for 4.  The flush
runs the from-layers in ascending order (its outer loop, Surface.hpp:1187), so (3,2)'s untrack comes
last and empties P: P is deactivated with background Distance(4) = 4, and every voxel of P then
reads 4.  With only (0,1) in the band, P is left with -4."""
import numpy as np

SIZE, PART, L = (16, 16), (4, 4), 3
OFF = np.array([8, 8])
P_LO, P_HI = (0, 0), (4, 4)


def field(with_plus=True):
    x = np.broadcast_to(np.arange(SIZE[0])[:, None] - OFF[0], SIZE)
    y = np.broadcast_to(np.arange(SIZE[1])[None, :] - OFF[1], SIZE)
    in_p = (x >= P_LO[0]) & (x < P_HI[0]) & (y >= P_LO[1]) & (y < P_HI[1])
    plus = (x >= 4) | (in_p & (x >= 2))
    phi = np.where(plus, 4.0, -4.0).astype(np.float32)
    phi[tuple(OFF + [0, 1])] = -3.0
    if with_plus:
        phi[tuple(OFF + [3, 2])] = 3.0
    return phi


def expected(with_plus=True):
    phi = field(with_plus)
    out = phi.copy()
    bg = 4.0 if with_plus else -4.0
    out[OFF[0] + P_LO[0]:OFF[0] + P_HI[0], OFF[1] + P_LO[1]:OFF[1] + P_HI[1]] = bg
    return out
