"""Synthetic surface workloads (seeded, host-generated) shared by tests and the benchmark.

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014) with floats drawn as
(u >> 40) * 2^-24, so the same seed gives the same inputs on every machine.  Initial fields are
computed in float64 and stored as float32: phi = clamp(distance-to-surface, -(L+1), L+1).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 40) * (1.0 / 16777216.0)
        return lo + (hi - lo) * u

    def randint(self, lo: int, hi: int) -> int:
        """Integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


def grid_coords(size: Sequence[int]):
    """World coordinate arrays (offset -size/2, C division) with indexing 'ij'."""
    axes = [np.arange(s, dtype=np.float64) - (s // 2) for s in size]
    return np.meshgrid(*axes, indexing="ij")


def spheres_phi(size: Sequence[int], L: int, centres, radii) -> np.ndarray:
    """min over spheres of (|x - c| - r), clamped to +-(L+1), in float64 then float32.  Each
    sphere is evaluated only inside its bounding box grown by L+1: outside it |x - c| - r > L+1,
    which the clamp maps to L+1 anyway."""
    D = len(size)
    lim = float(L + 1)
    phi = np.full(tuple(size), lim)
    off = [s // 2 for s in size]
    for c, r in zip(centres, radii):
        sl, ax = [], []
        for a in range(D):
            lo = max(0, int(math.floor(c[a] - r - lim)) + off[a])
            hi = min(size[a], int(math.ceil(c[a] + r + lim)) + off[a] + 1)
            if lo >= hi:
                break
            sl.append(slice(lo, hi))
            ax.append(np.arange(lo, hi, dtype=np.float64) - off[a])
        if len(sl) < D:
            continue
        X = np.meshgrid(*ax, indexing="ij")
        d2 = sum((X[a] - c[a]) ** 2 for a in range(D))
        sub = phi[tuple(sl)]
        np.minimum(sub, np.sqrt(d2) - r, out=sub)
    return np.clip(phi, -lim, lim).astype(np.float32)


def _real_sph_harm(l: int, m: int, theta: np.ndarray, ph: np.ndarray) -> np.ndarray:
    """Real spherical harmonic (unnormalised associated Legendre times cos/sin)."""
    x = np.cos(theta)
    am = abs(m)
    # associated Legendre P_l^|m|(x) by recurrence
    pmm = np.ones_like(x)
    if am > 0:
        somx2 = np.sqrt(np.maximum(0.0, 1 - x * x))
        fact = 1.0
        for _ in range(am):
            pmm = -pmm * fact * somx2
            fact += 2.0
    if l == am:
        plm = pmm
    else:
        pmmp1 = x * (2 * am + 1) * pmm
        if l == am + 1:
            plm = pmmp1
        else:
            pll = pmmp1
            for ll in range(am + 2, l + 1):
                pll = ((2 * ll - 1) * x * pmmp1 - (ll + am - 1) * pmm) / (ll - am)
                pmm, pmmp1 = pmmp1, pll
            plm = pll
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    if m > 0:
        return math.sqrt(2) * norm * plm * np.cos(am * ph)
    if m < 0:
        return math.sqrt(2) * norm * plm * np.sin(am * ph)
    return norm * plm


def harmonic_sphere_phi(size: Sequence[int], L: int, radius: float, seed: int,
                        n_terms: int = 8, lmax: int = 6, amp: float = 6.0) -> np.ndarray:
    """Sphere of `radius` at the centre, radially perturbed by a sum of real spherical
    harmonics (degree <= lmax) with amplitudes U[0, amp), normalised so the largest single
    term reaches its amplitude."""
    rng = SplitMix64(seed)
    terms = []
    for _ in range(n_terms):
        l = rng.randint(0, lmax)
        m = rng.randint(-l, l)
        terms.append((l, m, rng.uniform(0.0, amp)))
    peaks = []
    for l, m, _ in terms:
        peaks.append(max(1e-12, float(np.abs(_real_sph_harm(
            l, m, np.linspace(0, math.pi, 181)[:, None],
            np.linspace(-math.pi, math.pi, 361)[None, :])).max())))
    reach = sum(a for _, _, a in terms) + L + 2
    out = np.empty(tuple(size), dtype=np.float32)
    ys = np.arange(size[1], dtype=np.float64) - (size[1] // 2)
    zs = np.arange(size[2], dtype=np.float64) - (size[2] // 2)
    Y2, Z2 = np.meshgrid(ys, zs, indexing="ij")
    slab = 16
    for x0 in range(0, size[0], slab):
        xs = np.arange(x0, min(size[0], x0 + slab), dtype=np.float64) - (size[0] // 2)
        X = xs[:, None, None] + 0.0 * Y2[None]
        Yb = np.broadcast_to(Y2[None], X.shape)
        Zb = np.broadcast_to(Z2[None], X.shape)
        r = np.sqrt(X ** 2 + Yb ** 2 + Zb ** 2)
        phi = r - radius
        # the perturbation moves the surface by at most sum(amplitudes): outside that shell
        # the clamped value is already decided
        near = np.abs(phi) <= reach
        if near.any():
            rn, xn, yn, zn = r[near], X[near], Yb[near], Zb[near]
            theta = np.arccos(np.clip(zn / np.maximum(rn, 1e-12), -1.0, 1.0))
            ph = np.arctan2(yn, xn)
            R = np.full_like(rn, radius)
            for (l, m, a), pk in zip(terms, peaks):
                R += a * _real_sph_harm(l, m, theta, ph) / pk
            phi[near] = rn - R
        out[x0:x0 + len(xs)] = np.clip(phi, -(L + 1), L + 1)
    return out


def speed_of(index: int):
    """The three per-surface speeds of the mixed workloads: constant -0.4, 0.15 * curvature,
    -0.3 + 0.1 * curvature."""
    from ._binding import speed_const_curv, speed_constant, speed_curvature
    return [speed_constant(-0.4), speed_curvature(0.15), speed_const_curv(-0.3, 0.1)][index]


@dataclass
class SurfaceSpec:
    D: int
    L: int
    size: Tuple[int, ...]
    part: Tuple[int, ...]
    centres: List[Tuple[float, ...]] = field(default_factory=list)
    radii: List[float] = field(default_factory=list)
    speed: int = 0  # speed_of(speed)

    def phi(self) -> np.ndarray:
        return spheres_phi(self.size, self.L, self.centres, self.radii)


def mixed_surfaces(n: int, seed: int, edge_lo: int, edge_hi: int, D: int = 3, L: int = 3,
                   part_small: int = 8, part_big: int = 16, small_below: int = 96,
                   fixed_part: int | None = None) -> List[SurfaceSpec]:
    """Independent surfaces of uneven size: edge log-uniform in [edge_lo, edge_hi] rounded to
    a multiple of the partition size, 1-8 spheres of radius U[0.05, 0.3] * edge."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        e = math.exp(rng.uniform(math.log(edge_lo), math.log(edge_hi)))
        p = fixed_part if fixed_part else (part_small if e < small_below else part_big)
        edge = max(p, int(round(e / p)) * p)
        ns = rng.randint(1, 8)
        cs, rs = [], []
        for _ in range(ns):
            rs.append(rng.uniform(0.05, 0.3) * edge)
            cs.append(tuple(rng.uniform(-edge / 2, edge / 2) for _ in range(D)))
        out.append(SurfaceSpec(D, L, (edge,) * D, (p,) * D, cs, rs, rng.randint(0, 2)))
    return out


def served_requests(specs, window_s: float, rate: float, seed: int):
    """Poisson request arrivals at `rate` per second over `window_s` seconds: (arrival time,
    surface, box lo, box hi, speed index, steps) per request.  The surface is uniform over specs,
    the box side U[0.05, 1] of the grid edge per axis placed uniformly inside the grid, the speed
    uniform over speed_of's three, the step count U{1..50}."""
    rng = SplitMix64(seed)
    out, t = [], 0.0
    while True:
        t += -math.log(1.0 - rng.uniform()) / rate
        if t >= window_s:
            break
        i = rng.randint(0, len(specs) - 1)
        lo, hi = [], []
        for s in specs[i].size:
            side = max(1, int(rng.uniform(0.05, 1.0) * s))
            a = rng.randint(-(s // 2), s - s // 2 - side)
            lo.append(a)
            hi.append(a + side - 1)
        out.append((t, i, lo, hi, rng.randint(0, 2), rng.randint(1, 50)))
    return out

