"""CPU tests of the workload generators (galaxy_amd/csrc/glx_ic.cpp, host code)."""
from __future__ import annotations

import math

import numpy as np

from galaxy_amd import ic
from oracle import oracle as O


def test_plummer_truncated_recentered_f32():
    pos, vel, mass = ic.plummer(16384, 42, rmax=10.0)
    assert pos.shape == (16384, 3) and abs(mass.sum() - 1.0) < 1e-12
    r = np.linalg.norm(pos, axis=1)
    assert r.max() < 10.0 + 0.1  # truncated at r = 10 (up to the recentring shift)
    assert np.abs(pos.mean(axis=0)).max() < 1e-6 and np.abs(vel.mean(axis=0)).max() < 1e-6
    assert np.array_equal(pos, pos.astype(np.float32).astype(np.float64))
    # half-mass radius of a Plummer sphere truncated at 10a: M(<r)/M(<10) = 1/2
    xh = 0.5 * 1000 / 101 ** 1.5
    rh = 1.0 / math.sqrt(xh ** (-2.0 / 3.0) - 1.0)
    assert abs(np.median(r) - rh) / rh < 0.03
    # virial ratio 2T/|W| ~ 1 (truncation perturbs it slightly)
    ke = 0.5 * (mass * (vel ** 2).sum(1)).sum()
    w = O.potential_serial(pos[:4096], mass[:4096] * 4, 1.0, 1e-3)
    assert 0.8 < 2 * ke / abs(w) < 1.25
    # deterministic in the seed
    p2, _, _ = ic.plummer(16384, 42, rmax=10.0)
    assert np.array_equal(pos, p2)


def test_collision_masses_split_and_orbit():
    n = 1 << 16
    pos, vel, mass, info = ic.collision(n, 5)
    m1, m2 = mass[: n // 2].sum(), mass[n // 2:].sum()
    assert abs(m1 - 6.33) < 1e-9 and abs(m2 - 6.33) < 1e-9
    # zero total momentum and COM at the origin
    assert np.abs((pos * mass[:, None]).sum(0)).max() / mass.sum() < 1e-3
    assert np.abs((vel * mass[:, None]).sum(0)).max() < 1e-6 * n
    c1 = (pos[: n // 2] * mass[: n // 2, None]).sum(0) / m1
    c2 = (pos[n // 2:] * mass[n // 2:, None]).sum(0) / m2
    assert abs(np.linalg.norm(c2 - c1) - 80.0) < 0.5
    # mass-ratio variant scales the second galaxy
    _, _, mass3, _ = ic.collision(n, 5, mass_ratio=3.0)
    assert abs(mass3[n // 2:].sum() - 3 * 6.33) < 1e-9


def test_63bit_codes_resolve_collision_cells():
    """On the collision input the root cube is ~6000 wide (Hernquist tails truncated at
    300a), so 30-bit codes (cells ~6 units) collapse the disks into few cells; 63-bit
    codes (cells ~3e-3) separate almost every particle."""
    pos, _, _, _ = ic.collision(1 << 17, 5)
    u30 = np.unique(O.reference_morton(pos)[2]).size
    u63 = np.unique(O.reference_morton63(pos)[2]).size
    assert u63 > 0.99 * pos.shape[0] and u30 < 0.3 * pos.shape[0]
