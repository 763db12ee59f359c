"""CPU tests of the per-particle-stream workload generators.

The oracle (oracle/ic_oracle.c) is pinned by what the mathematics fixes, following the
reference's own sampler tests (ic/tests/sampling.rs, ic/tests/hernquist_sampling.rs,
ic/tests/disk_sampling.rs, ic/tests/collision.rs): realised mass profiles against the
closed-form enclosed-mass fractions (4-sigma binomial bands), half-mass radii, virial
ratios against the analytic potential energies, the Hernquist DF integrating to the
closed-form density, disk kinematics against the circular speed of the closed-form
enclosed masses, and the Kepler elements and spin orientations of the placed galaxies.
The host entry point glx_ic_generate is then checked against the oracle.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from galaxy_amd import ic
from galaxy_amd import lbvh as L
from oracle import oracle as O


def _band(frac, expected, n):
    return abs(frac - expected) < 4.0 * math.sqrt(expected * (1 - expected) / n) + 1e-3


def _radii(pos):
    return np.sort(np.linalg.norm(pos, axis=1))


# ---- oracle pins -------------------------------------------------------------------

def test_plummer_stream_profile_half_mass_and_virial():
    """ic/tests/sampling.rs:88-160 on the per-particle-stream Plummer sampler."""
    n = 40000
    pos, vel, mass = O.ic_sphere(0, n, 0x5EED, rmax=0.0)
    r = _radii(pos)
    for rr in (0.5, 1.0, 1.305, 2.0, 4.0):
        frac = np.searchsorted(r, rr, side="right") / n
        assert _band(frac, rr ** 3 / (rr * rr + 1.0) ** 1.5, n), rr
    rh = 1.0 / math.sqrt(2.0 ** (2.0 / 3.0) - 1.0)
    assert abs(r[n // 2] - rh) < 0.03 * rh
    ke = 0.5 * (mass * (vel ** 2).sum(1)).sum()
    assert abs(2 * ke / (3 * math.pi / 32) - 1.0) < 0.05
    assert abs(mass.sum() - 1.0) < 1e-12
    assert np.abs((mass[:, None] * pos).sum(0)).max() < 1e-6


def test_plummer_stream_speed_distribution():
    """q = v / v_esc(r) follows g(q) ~ q^2 (1 - q^2)^3.5 (ic/src/plummer.rs:111-122)."""
    n = 40000
    pos, vel, _ = O.ic_sphere(0, n, 3, rmax=0.0)
    r = np.linalg.norm(pos, axis=1)
    q = np.sort(np.linalg.norm(vel, axis=1) / (math.sqrt(2.0) * (r * r + 1.0) ** -0.25))
    grid = np.linspace(0, 1, 20001)
    pdf = grid ** 2 * (1 - grid ** 2) ** 3.5
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    ks = np.max(np.abs(np.interp(q, grid, cdf) - np.arange(1, n + 1) / n))
    assert ks < 2.0 / math.sqrt(n)


def test_plummer_stream_truncation():
    n = 20000
    pos, _, _ = O.ic_sphere(0, n, 9, rmax=10.0)
    r = _radii(pos)
    mtr = 1000.0 / 101.0 ** 1.5
    for rr in (0.5, 1.0, 2.0, 5.0):
        assert _band(np.searchsorted(r, rr, side="right") / n,
                     rr ** 3 / (rr * rr + 1.0) ** 1.5 / mtr, n), rr
    assert r[-1] < 10.0 + 0.05


def test_hernquist_df_integrates_to_density():
    """ic/tests/hernquist_sampling.rs:101-123: 4 pi int f(E) sqrt(2 (psi - E)) dE = rho(r)."""
    for r in (0.5, 1.0, 2.0, 4.0):
        psi = 1.0 / (r + 1.0)
        k = 20000
        de = psi / k
        acc = sum(O.ic_hernquist_df(1.0, 1.0, 1.0, (j + 0.5) * de) * math.sqrt(2 * (psi - (j + 0.5) * de))
                  for j in range(k))
        rho = 1.0 / (2 * math.pi * r * (r + 1.0) ** 3)
        assert abs(4 * math.pi * acc * de - rho) < 3e-3 * rho, r
    assert O.ic_hernquist_df(1.0, 1.0, 1.0, 0.0) == 0.0
    assert O.ic_hernquist_df(1.0, 1.0, 1.0, -0.3) == 0.0
    assert O.ic_hernquist_df(1.0, 1.0, 1.0, 1.0) == 0.0


def test_hernquist_stream_profile_and_virial():
    """ic/tests/hernquist_sampling.rs:141-207 (truncated at 300a, X uniform on [0, X_max))."""
    n = 20000
    pos, vel, mass = O.ic_sphere(1, n, 0xBEEF, a=1.0)
    r = _radii(pos)
    xmax = (300 / 301) ** 2
    for rr in (0.5, 1.0, 2.0, 5.0, 20.0):
        assert _band(np.searchsorted(r, rr, side="right") / n, (rr / (rr + 1)) ** 2 / xmax, n), rr
    rh = 1.0 / (math.sqrt(0.5 * xmax) ** -1 - 1.0)
    assert abs(r[n // 2] - rh) < 0.03 * rh
    ke = 0.5 * (mass * (vel ** 2).sum(1)).sum()
    assert abs(2 * ke / (1.0 / 6.0) - 1.0) < 0.05
    assert r[-1] < 300.0 + 1.0


def _vcirc(r, scale=1.0):
    def hq(m, a):
        return m * r * r / (r + a) ** 2
    x = np.minimum(r, 12.0) / 3.0
    disk = scale * (1 - (1 + x) * np.exp(-x)) / (1 - 5 * math.exp(-4))
    return np.sqrt((hq(0.33 * scale, 0.5) + hq(5.0 * scale, 10.0) + disk) / r)


@pytest.mark.parametrize("scale", [1.0, 3.0])
def test_galaxy_components(scale):
    """Disk (ic/tests/disk_sampling.rs: exponential surface density cut at 4 R_d, sech^2
    height, v_phi about v_c with 10% dispersion), bulge and halo profiles, masses."""
    n = 32768
    pos, vel, mass = O.ic_galaxy_body(n, 17, scale)
    nd, nb = int(math.floor(0.25 * n + 0.5)), int(math.floor(0.10 * n + 0.5))
    d, b, h = slice(0, nd), slice(nd, nd + nb), slice(nd + nb, n)
    assert abs(mass[d].sum() - scale) < 1e-12 and abs(mass[b].sum() - 0.33 * scale) < 1e-12
    assert abs(mass[h].sum() - 5.0 * scale) < 1e-12
    R = np.hypot(pos[d, 0], pos[d, 1])
    Rs = np.sort(R)
    norm = 1 - 5 * math.exp(-4)
    for rr in (1.0, 3.0, 6.0, 9.0):
        x = rr / 3.0
        assert _band(np.searchsorted(Rs, rr, side="right") / nd, (1 - (1 + x) * math.exp(-x)) / norm, nd)
    assert Rs[-1] <= 12.0
    t = np.sort(np.tanh(pos[d, 2] / 0.3))  # sech^2 layer: tanh(z / h) uniform on (-1, 1)
    for u in (-0.5, 0.0, 0.5):
        assert _band(np.searchsorted(t, u) / nd, (u + 1) / 2, nd)
    vc = _vcirc(R, scale)
    phi_hat = np.stack([-pos[d, 1] / R, pos[d, 0] / R], 1)
    vphi = (vel[d, :2] * phi_hat).sum(1)
    res = (vphi - vc) / vc
    assert abs(res.mean()) < 0.005 and abs(res.std() - 0.1) < 0.005
    assert abs((vel[d, 2] / vc).std() - 0.1) < 0.005
    for sl, a, m in ((b, 0.5, 0.33), (h, 10.0, 5.0)):
        r = _radii(pos[sl])
        k = r.size
        xmax = (300 / 301) ** 2
        for rr in (0.5 * a, a, 3 * a):
            assert _band(np.searchsorted(r, rr, side="right") / k, (rr / (rr + a)) ** 2 / xmax, k)


def test_collision_orbit_and_orientation():
    """ic/tests/collision.rs:65-153 and ic/tests/orientation.rs: parabolic relative orbit
    (zero energy, h^2 = 2 mu r_p, approaching), zero-COM/zero-momentum frame, and each disk's
    spin along (sin i sin w, -sin i cos w, cos i)."""
    n = 1 << 15
    for ratio in (1.0, 3.0):
        pos, vel, mass, info = O.ic_collision(n, 5, mass_ratio=ratio)
        h = n // 2
        m1, m2 = mass[:h].sum(), mass[h:].sum()
        c1 = (mass[:h, None] * pos[:h]).sum(0) / m1
        c2 = (mass[h:, None] * pos[h:]).sum(0) / m2
        v1 = (mass[:h, None] * vel[:h]).sum(0) / m1
        v2 = (mass[h:, None] * vel[h:]).sum(0) / m2
        r, v = c2 - c1, v2 - v1
        mu = m1 + m2
        assert abs(np.linalg.norm(r) - 80.0) < 1e-4
        assert abs(0.5 * v @ v - mu / np.linalg.norm(r)) < 1e-5 * mu / 80.0
        assert abs(np.linalg.norm(np.cross(r, v)) ** 2 - 2 * mu * 8.0) < 1e-4 * 2 * mu * 8.0
        assert r @ v < 0
        assert np.linalg.norm(m1 * c1 + m2 * c2) < 1e-4 * mu
        assert np.linalg.norm(m1 * v1 + m2 * v2) < 1e-5 * mu
        nd = int(math.floor(0.25 * h + 0.5))
        for k, (lo, c, vv) in enumerate(((0, c1, v1), (h, c2, v2))):
            dp, dv = pos[lo:lo + nd] - c, vel[lo:lo + nd] - vv
            spin = (mass[lo:lo + nd, None] * np.cross(dp, dv)).sum(0)
            i, w = info[2 * k], info[2 * k + 1]
            want = np.array([math.sin(i) * math.sin(w), -math.sin(i) * math.cos(w), math.cos(i)])
            assert math.acos(min(1.0, spin @ want / np.linalg.norm(spin))) < 0.05


# ---- host entry point against the oracle ----------------------------------------

def _agree(a, b, frac=0.9999):
    """Bit-equal except where the recentring sums round differently: at most 1 f32 ulp."""
    ulp = np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(a - b) / ulp) <= 1.0
    assert np.mean(a == b) >= frac


@pytest.mark.parametrize("kind", ["plummer", "plummer_cut", "hernquist", "collision", "collision3"])
def test_host_generate_matches_oracle(kind):
    n = 30001
    if kind.startswith("plummer"):
        rmax = 10.0 if kind == "plummer_cut" else 0.0
        sp, ref = ic.spec_plummer(n, 21, rmax=rmax), O.ic_sphere(0, n, 21, rmax=rmax)
    elif kind == "hernquist":
        sp, ref = ic.spec_hernquist(n, 21, a=2.0, m=3.0), O.ic_sphere(1, n, 21, m=3.0, a=2.0)
    else:
        ratio = 3.0 if kind == "collision3" else 1.0
        sp = ic.spec_collision(n, 21, mass_ratio=ratio)
        ref = O.ic_collision(n, 21, mass_ratio=ratio)[:3]
    pos, vel, mass = ic.generate(sp)
    _agree(pos, ref[0])
    _agree(vel, ref[1])
    assert np.array_equal(mass, ref[2])
    assert np.array_equal(pos, pos.astype(np.float32).astype(np.float64))


def test_collision_entry_point_is_the_spec_generator():
    pos, vel, mass, info = ic.collision(4096, 8, mass_ratio=2.0)
    p2, v2, m2 = ic.generate(ic.spec_collision(4096, 8, mass_ratio=2.0))
    assert np.array_equal(pos, p2) and np.array_equal(vel, v2) and np.array_equal(mass, m2)
    ref = O.ic_collision(4096, 8, mass_ratio=2.0)[3]
    assert np.array_equal([info["incl"][0], info["arg"][0], info["incl"][1], info["arg"][1]], ref)


def test_generate_deterministic_and_small_cases():
    a = ic.generate(ic.spec_collision(3000, 4))
    b = ic.generate(ic.spec_collision(3000, 4))
    c = ic.generate(ic.spec_collision(3000, 5))
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    for n in (1, 2, 3, 255, 256, 257):
        p, v, m = ic.generate(ic.spec_plummer(n, 1))
        assert np.isfinite(p).all() and np.isfinite(v).all() and abs(m.sum() - 1.0) < 1e-12
    for n in (2, 3, 5, 513):
        p, v, m = ic.generate(ic.spec_collision(n, 1))
        assert np.isfinite(p).all() and np.isfinite(v).all() and (m > 0).all()
        if n >= 40:  # every galaxy has disk, bulge and halo particles
            assert abs(m.sum() - 12.66) < 1e-12 * n


def test_invalid_specs_rejected():
    lib = ic._bind()
    buf = np.zeros(30)
    bad = [
        ic.spec_plummer(0, 1),
        ic.spec_plummer(10, 1, a=-1.0),
        ic.spec_plummer(10, 1, m=float("nan")),
        ic.spec_hernquist(10, 1, g=0.0),
        ic.spec_collision(1, 1),
        ic.spec_collision(10, 1, rp=90.0),
        ic.spec_collision(10, 1, mass_ratio=0.0),
        ic.spec_plummer((1 << 29) + 1, 1),
    ]
    s = ic.spec_plummer(10, 1)
    s.kind = 7
    bad.append(s)
    for sp in bad:
        p = buf.ctypes.data_as(C.c_void_p)
        assert lib.glx_ic_generate(C.addressof(sp), p, p, p, None) == L.GLX_EINVAL
        assert lib.glx_last_error()
        assert lib.glx_ic_generate_device(C.addressof(sp), p, p, p) == L.GLX_EINVAL
    assert lib.glx_ic_generate(None, None, None, None, None) == L.GLX_EINVAL
