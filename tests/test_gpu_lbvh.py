"""End-to-end GPU forces (the ForceSolver path) against the oracle.

Ports the reference's GPU gates (gpu/tests/gpu_lbvh.rs) and adds two comparisons: the
same-theta reference CPU path (oracle) and an FP64 direct sum,
each as a median and 99th-percentile per-particle relative error with the tolerances
written below. Because Morton codes, sort and topology are bit-exact, the same-theta
difference is f32 arithmetic plus rare opening-test flips near the acceptance boundary.
"""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from oracle import oracle as O
from tests import inputs as I

pytestmark = pytest.mark.gpu
G, EPS = 1.0, 0.05


def say(*a):
    print("[case]", *a, flush=True)


def test_theta_to_zero_matches_direct_sum():
    """gpu/tests/gpu_lbvh.rs:100-117 (rms < 3e-4, worst < 5e-2); ours: rms < 2e-5."""
    s = glx.Lbvh(G, EPS, 1e-6)
    worst_rms = 0.0
    for seed in range(64):
        pos, _, mass = I.cluster(seed, 128)
        exact = O.direct_sum(pos, mass, G, EPS)
        got = s.accelerations(pos, mass)
        rms, worst = I.rms_rel_err(got, exact), I.worst_rel_err(got, exact)
        worst_rms = max(worst_rms, rms)
        assert rms < 3e-4 and worst < 5e-2, (seed, rms, worst)
        assert rms < 2e-5, (seed, rms)
    say("theta0 max rms", worst_rms)


def test_finite_theta_bounded_and_grows():
    """gpu/tests/gpu_lbvh.rs:124-147"""
    lo_s, hi_s = glx.Lbvh(G, EPS, 0.3), glx.Lbvh(G, EPS, 0.6)
    for seed in range(32):
        pos, _, mass = I.cluster(seed, 128)
        exact = O.direct_sum(pos, mass, G, EPS)
        rl = I.rms_rel_err(lo_s.accelerations(pos, mass), exact)
        rh = I.rms_rel_err(hi_s.accelerations(pos, mass), exact)
        assert rl < 0.012 and rh < 0.05 and rh > rl, (seed, rl, rh)


def test_tracks_oracle_same_theta():
    """gpu/tests/gpu_lbvh.rs:154-168 (rms < 5e-2). Ours: per-particle relative error vs
    the same-theta oracle, median < 1e-5 and p99 < 1e-3."""
    s = glx.Lbvh(G, EPS, 0.5)
    meds, p99s = [], []
    for seed in range(16):
        pos, _, mass = I.cluster(seed, 256)
        cpu, _ = O.lbvh_accelerations(pos, mass, G, EPS, 0.5)
        got = s.accelerations(pos, mass)
        assert I.rms_rel_err(got, cpu) < 5e-2
        med, p99 = I.rel_err_quantiles(got, cpu)
        meds.append(med)
        p99s.append(p99)
    say("same-theta median", max(meds), "p99", max(p99s))
    assert max(meds) < 1e-5 and max(p99s) < 1e-3


def test_bit_deterministic_same_device():
    """gpu/tests/gpu_lbvh.rs:175-185"""
    s = glx.Lbvh(G, EPS, 0.5)
    pos, _, mass = I.cluster(7, 300)
    assert np.array_equal(s.accelerations(pos, mass), s.accelerations(pos, mass))


def test_bit_deterministic_multi_tile_trees():
    """Forces repeat bit for bit on systems whose trees span many bottom-up tiles (the
    order in which a node's children finish varies between runs), in the reference and
    the fast build."""
    from galaxy_amd import ic

    pos, _, mass = ic.plummer(20000, 7)
    fast = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2, far_theta=0.55,
                mono_ratio=0.8)
    for kw in ({}, fast):
        s = glx.Lbvh(1.0, 0.01, 0.5, **kw)
        a = s.accelerations(pos, mass)
        for _ in range(3):
            assert np.array_equal(s.accelerations(pos, mass), a), kw


def test_momentum_flux_at_theta_zero():
    """gpu/tests/gpu_lbvh.rs:191-207"""
    s = glx.Lbvh(G, EPS, 1e-6)
    for seed in range(16):
        pos, _, mass = I.cluster(seed, 256)
        a = s.accelerations(pos, mass)
        net = (a * mass[:, None]).sum(axis=0)
        scale = (np.linalg.norm(a, axis=1) * mass).sum()
        assert np.linalg.norm(net) / scale < 1e-4


def test_empty_and_single():
    """gpu/tests/gpu_lbvh.rs:212-226"""
    s = glx.Lbvh(G, EPS, 0.5)
    assert s.accelerations(np.zeros((0, 3)), np.zeros(0)).shape == (0, 3)
    a = s.accelerations(np.array([[1.0, 2.0, 3.0]]), np.ones(1))
    assert np.array_equal(a, np.zeros((1, 3)))


def test_coincident_knot_theta_zero():
    pos, mass = I.cluster_with_coincident_knot()
    exact = O.direct_sum(pos, mass, G, EPS)
    got = glx.Lbvh(G, EPS, 1e-6).accelerations(pos, mass)
    assert I.rms_rel_err(got, exact) < 2e-5
    a1 = glx.Lbvh(G, EPS, 0.5).accelerations(pos, mass)
    cpu, _ = O.lbvh_accelerations(pos, mass, G, EPS, 0.5)
    assert I.rms_rel_err(a1, cpu) < 1e-4


def test_large_cube_topology_matches_and_theta0_exact():
    """gpu/tests/gpu_lbvh.rs:259-308 asked for an f32 topology straddle; with bit-exact
    FP64 Morton codes there is none: the GPU topology equals the oracle's on every seed."""
    for seed in range(4):
        pos, mass = I.cube(seed, 20000, 3.0)
        codes = glx.morton(pos).codes
        _, _, want = O.reference_morton(pos)
        assert np.array_equal(codes, want)
        order, sc = glx.sort(codes)
        assert np.array_equal(order, O.reference_sort(want))
        l, r, p = glx.karras(sc)
        lw, rw, pw = O.reference_karras(want[order])
        assert np.array_equal(l, lw) and np.array_equal(r, rw) and np.array_equal(p, pw)
    pos, mass = I.cube(0, 20000, 3.0)
    tg = np.arange(0, 20000, 13, dtype=np.uint32)
    exact = O.direct_sum_targets(pos, mass, G, EPS, tg, nthreads=8)
    got = glx.Lbvh(G, EPS, 1e-6).accelerations(pos, mass)[tg]
    assert I.rms_rel_err(got, exact) < 3e-4


def test_plummer_errors_vs_oracle_and_fp64_direct():
    """C0-shaped input (Plummer, N=16384, eps=0.01, theta=0.5): stated median/p99 of the
    per-particle relative error vs the same-theta oracle and vs the FP64 direct sum."""
    pos, _, mass = I.plummer_numpy(16384, seed=1)
    eps, theta = 0.01, 0.5
    got = glx.Lbvh(1.0, eps, theta).accelerations(pos, mass)
    cpu, _ = O.lbvh_accelerations(pos, mass, 1.0, eps, theta, nthreads=8)
    tg = np.arange(16384, dtype=np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, eps, tg)
    sub = tg[::64]
    d64_cpu = O.direct_sum_targets(pos, mass, 1.0, eps, sub, nthreads=8)
    assert np.allclose(d64[sub], d64_cpu, rtol=1e-12, atol=1e-12)
    m_cpu, p_cpu = I.rel_err_quantiles(got, cpu)
    m_d, p_d = I.rel_err_quantiles(got, d64)
    m_ref, p_ref = I.rel_err_quantiles(cpu, d64)
    say("plummer16k vs oracle median", m_cpu, "p99", p_cpu)
    say("plummer16k vs fp64 median", m_d, "p99", p_d, "| oracle vs fp64", m_ref, p_ref)
    assert m_cpu < 1e-5 and p_cpu < 1e-3
    assert m_d < 1.2 * m_ref + 1e-5 and p_d < 1.2 * p_ref + 1e-4


def test_kahan_precision_mode_not_worse():
    pos, _, mass = I.plummer_numpy(16384, seed=2)
    tg = np.arange(0, 16384, 4, dtype=np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, 0.01, tg)
    a32 = glx.Lbvh(1.0, 0.01, 1e-6).accelerations(pos, mass)[tg]
    ak = glx.Lbvh(1.0, 0.01, 1e-6, precision=glx.PREC_KAHAN).accelerations(pos, mass)[tg]
    e32 = I.rel_err_quantiles(a32, d64)
    ek = I.rel_err_quantiles(ak, d64)
    say("theta0 f32 median/p99", e32, "kahan", ek)
    assert ek[0] <= e32[0] * 1.05 + 1e-9 and ek[1] <= e32[1] * 1.05 + 1e-9
    assert ek[0] < 2e-6


def test_interaction_count_matches_oracle():
    pos, _, mass = I.plummer_numpy(16384, seed=3)
    s = glx.Lbvh(1.0, 0.01, 0.5, count=True)
    s.accelerations(pos, mass)
    st = s.stats()
    _, cnt = O.lbvh_accelerations(pos, mass, 1.0, 0.01, 0.5, nthreads=8)
    say("interactions gpu", st["interactions"], "oracle", cnt, "visits", st["node_visits"])
    assert abs(st["interactions"] - cnt) / cnt < 1e-3


def test_potential_energy_fp64():
    pos, _, mass = I.cluster(2, 2000)
    u = glx.Lbvh(G, EPS, 0.5).potential_energy(pos, mass)
    want = O.potential_serial(pos, mass, G, EPS)
    assert abs(u - want) <= 1e-12 * abs(want)
