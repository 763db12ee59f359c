"""Deterministic test inputs, regenerated bit-for-bit from the reference tests' LCGs.

All generators use the 64-bit LCG state = state*6364136223846793005 +
1442695040888963407 (mod 2^64), started from seed|1, exactly as the reference tests
do; every draw is exact integer or exact-double arithmetic.
"""
from __future__ import annotations

import numpy as np

_MUL = 6364136223846793005
_ADD = 1442695040888963407
_MASK = (1 << 64) - 1


class Lcg:
    def __init__(self, seed: int):
        self.state = (seed | 1) & _MASK

    def next_u64(self) -> int:
        self.state = (self.state * _MUL + _ADD) & _MASK
        return self.state

    def next_f64(self) -> float:
        """(state >> 11) / 2^53, in [0, 1)."""
        return float(self.next_u64() >> 11) / float(1 << 53)


def cluster(seed: int, n: int, with_vel: bool = False):
    """solvers/tests/lbvh.rs:27-44 (and gpu/tests/gpu_lbvh.rs:38-55). With
    with_vel, the velocity-carrying variant of gpu/tests/gpu_resident.rs:33-50."""
    r = Lcg(seed)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    mass = np.zeros(n)
    for i in range(n):
        x = r.next_f64() - 0.5
        y = r.next_f64() - 0.5
        z = r.next_f64() - 0.5
        pos[i] = (x * 3.0, y * 3.0, z * 3.0)
        if with_vel:
            vx = r.next_f64() - 0.5
            vy = r.next_f64() - 0.5
            vz = r.next_f64() - 0.5
            vel[i] = (vx * 0.1, vy * 0.1, vz * 0.1)
        mass[i] = 0.1 + 0.9 * r.next_f64()
    return pos, vel, mass


def cloud(seed: int, n: int, r: float, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """gpu/tests/lbvh_morton.rs:25-36: center + (u-0.5)*(2r) per axis."""
    g = Lcg(seed)
    c = np.asarray(center, dtype=np.float64)
    out = np.zeros((n, 3))
    two_r = 2.0 * r
    for i in range(n):
        x = g.next_f64() - 0.5
        y = g.next_f64() - 0.5
        z = g.next_f64() - 0.5
        out[i] = (c[0] + x * two_r, c[1] + y * two_r, c[2] + z * two_r)
    return out


def random_codes(seed: int, n: int) -> np.ndarray:
    """gpu/tests/lbvh_sort.rs:41-51: ((state >> 33) as u32) & (2^30 - 1)."""
    g = Lcg(seed)
    return np.array([(g.next_u64() >> 33) & ((1 << 30) - 1) for _ in range(n)], dtype=np.uint32)


def masses(seed: int, n: int) -> np.ndarray:
    """gpu/tests/lbvh_tree.rs:47-57: 0.5 + u."""
    g = Lcg(seed)
    return np.array([0.5 + g.next_f64() for _ in range(n)])


def cube(seed: int, n: int, r: float):
    """gpu/tests/gpu_lbvh.rs:231-247."""
    g = Lcg(seed)
    pos = np.zeros((n, 3))
    mass = np.zeros(n)
    two_r = 2.0 * r
    for i in range(n):
        x = g.next_f64() - 0.5
        y = g.next_f64() - 0.5
        z = g.next_f64() - 0.5
        pos[i] = (x * two_r, y * two_r, z * two_r)
        mass[i] = 0.1 + 0.9 * g.next_f64()
    return pos, mass


def cluster_with_coincident_knot():
    """solvers/tests/lbvh.rs:249-261: cluster(3, 60) plus 8 coincident bodies."""
    pos, _, mass = cluster(3, 60)
    knot = np.tile(np.array([[0.25, -0.1, 0.4]]), (8, 1))
    km = np.array([0.5 + 0.05 * k for k in range(8)])
    return np.vstack([pos, knot]), np.concatenate([mass, km])


def plummer_numpy(n: int, seed: int = 0, a: float = 1.0, m: float = 1.0):
    """Plummer sphere (G = 1) by the closed-form inversion of M(<r) (ic/src/plummer.rs:104-121),
    vectorised with numpy's generator: a statistical sample, not the reference's stream."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, n)
    r = a / np.sqrt(x ** (-2.0 / 3.0) - 1.0)
    ct = rng.uniform(-1.0, 1.0, n)
    st = np.sqrt(np.maximum(1.0 - ct * ct, 0.0))
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    pos = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0.0, 1.0, todo.size)
        hh = 0.1 * rng.uniform(0.0, 1.0, todo.size)
        ok = hh < qq * qq * (1.0 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0 * m) * (r * r + a * a) ** -0.25
    ct = rng.uniform(-1.0, 1.0, n)
    st = np.sqrt(np.maximum(1.0 - ct * ct, 0.0))
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    vel = (q * vesc)[:, None] * np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    pos -= pos.mean(axis=0)
    vel -= vel.mean(axis=0)
    return pos, vel, np.full(n, m / n)


def rms_rel_err(approx, exact) -> float:
    """solvers/tests/lbvh.rs:62-72"""
    rms = max(np.sqrt(np.mean(np.sum(exact * exact, axis=1))), 1e-300)
    return float(np.sqrt(np.mean(np.sum((approx - exact) ** 2, axis=1))) / rms)


def worst_rel_err(approx, exact) -> float:
    """solvers/tests/lbvh.rs:75-82"""
    rms = max(np.sqrt(np.mean(np.sum(exact * exact, axis=1))), 1e-300)
    return float(np.max(np.linalg.norm(approx - exact, axis=1)) / rms) if len(exact) else 0.0


def rel_err_quantiles(approx, exact, qs=(0.5, 0.99)):
    """Per-particle e_i = |a_i - a_i^ref| / |a_i^ref|, returned at the quantiles qs."""
    num = np.linalg.norm(approx - exact, axis=1)
    den = np.maximum(np.linalg.norm(exact, axis=1), 1e-300)
    e = num / den
    return [float(np.quantile(e, q)) for q in qs]
