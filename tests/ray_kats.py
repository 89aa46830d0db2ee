"""The reference's raycasting scenarios (test_Surface.cpp:1278-1535) over any Surface built on
the felt ABI, and rays for parity runs."""
import numpy as np

from felt_amd._binding import RAY_MISS, speed_constant

F32 = np.float32


def approx_vec(got, want, eps=float(np.finfo(np.float32).eps) * 100, scale=1.0) -> bool:
    """ApproxVec: |a - b| < eps * (scale + max(|a|, |b|)) per component (tests/Utils.hpp:92-98)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    return bool(np.all(np.abs(got - want) < eps * (scale + np.maximum(np.abs(got),
                                                                        np.abs(want)))))


def is_miss(hit) -> bool:
    return bool(np.all(np.asarray(hit) == F32(RAY_MISS)))


def sphere3(make):
    """32^3 grid, 5^3 partitions, 3 layers; seed at the origin and three updates of -1."""
    s = make(3, 3, (32, 32, 32), (5, 5, 5))
    s.seed((0, 0, 0))
    for _ in range(3):
        s.update(speed_constant(-1.0))
    return s


def circle2(make):
    """16^2 grid, 3^2 partitions, 2 layers; seed at the origin and three updates of -1."""
    s = make(2, 2, (16, 16), (3, 3))
    s.seed((0, 0))
    for _ in range(3):
        s.update(speed_constant(-1.0))
    return s


def unit(v):
    v = np.asarray(v, dtype=F32)
    return (v / np.sqrt(np.sum(v * v, dtype=F32))).astype(F32)


def angle_axis(angle: float, axis) -> np.ndarray:
    """Eigen::AngleAxisf(angle, axis).matrix() in fp32 (Rodrigues' formula)."""
    k = np.asarray(axis, dtype=F32)
    c, s = F32(np.cos(F32(angle))), F32(np.sin(F32(angle)))
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]], dtype=F32)
    return (c * np.eye(3, dtype=F32) + s * K + (F32(1) - c) * np.outer(k, k)).astype(F32)


def rotation_rays(axis):
    """Origins R*(0,0,-10) and directions R*(0,0,1) for angles rot_mult*pi, rot_mult stepping
    0, 0.1, ... below 2 in fp32, as the reference's rotation sweep does."""
    org, dirs = [], []
    r = F32(0.0)
    while r < F32(2.0):
        R = angle_axis(r * F32(np.pi), axis)
        org.append(R @ np.array([0, 0, -10], dtype=F32))
        dirs.append(unit(R @ np.array([0, 0, 1], dtype=F32)))
        r = F32(r + F32(0.1))
    return np.array(org, dtype=F32), np.array(dirs, dtype=F32)


def check_ray_kats(make):
    """Every raycasting check of the reference, as a list of failure strings."""
    fails = []
    s = sphere3(make)
    cases = [((-35, 0, 0), (1, 0, 0), (-3, 0, 0)),   # from outside the grid
             ((-6, 0, 0), (1, 0, 0), (-3, 0, 0)),    # from inside the grid
             ((-3, 0, 0), (1, 0, 0), (-3, 0, 0))]    # from the zero layer
    for o, d, want in cases:
        h = s.ray([o], [d])[0]
        if not approx_vec(h, want):
            fails.append(f"ray {o} -> {h}, want {want}")
    h = s.ray([(0, 0, 0)], [(1, 0, 0)])[0]               # from inside the surface: a miss
    if not is_miss(h):
        fails.append(f"ray from the centre hit {h}")
    h = s.ray([(-10, -10, 0)], [unit((1, 1, 0))])[0]
    if not approx_vec(h, (-1.5, -1.5, 0)):
        fails.append(f"diagonal ray -> {h}")
    h = s.ray([(10, 10, 10)], [unit((-1, -1, -1))])[0]
    if is_miss(h):
        fails.append("ray from (10,10,10) missed")
    for axis in ((0, 1, 0), unit((1, 1, 1)), unit((0, 1, 1))):
        o, d = rotation_rays(axis)
        hits = s.ray(o, d)
        for i in range(len(o)):
            if is_miss(hits[i]):
                fails.append(f"rotation about {tuple(axis)}: ray {i} from {o[i]} missed")
    s.update(speed_constant(-0.3))                     # expanded slightly
    h = s.ray([(-10, 0, 0)], [(1, 0, 0)])[0]
    if not approx_vec(h, (-3.3, 0, 0)):
        fails.append(f"after -0.3: ray -> {h}, want (-3.3, 0, 0)")

    s2 = circle2(make)
    h = s2.ray([(-2.4, -10)], [(0, 1)])[0]
    if not approx_vec(h, (-2.21609, -0.78391), eps=0.1):
        fails.append(f"2D ray -> {h}")
    return fails


def random_rays(rng, D, n, size, inside_frac=0.5):
    """Unit-direction rays from around the grid: some start inside, the rest up to one grid
    size outside."""
    size = np.asarray(size, dtype=np.float64)
    half = size / 2
    o = rng.uniform(-half - size, half + size, size=(n, D))
    k = int(n * inside_frac)
    o[:k] = rng.uniform(-half, half, size=(k, D))
    d = rng.normal(size=(n, D))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # aim half of them at the centre so that most hit
    aim = rng.random(n) < 0.5
    c = -o[aim] + rng.normal(scale=2.0, size=(int(aim.sum()), D))
    d[aim] = c / np.linalg.norm(c, axis=1, keepdims=True)
    o = o.astype(F32)
    d = d.astype(F32)
    d = (d / np.sqrt(np.sum(d * d, axis=1, keepdims=True, dtype=F32))).astype(F32)
    return o, d
