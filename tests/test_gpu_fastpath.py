"""GPU tests of the throughput modes: 63-bit Morton keys, quadrupole moments and the
wave-cooperative group walk.

- 63-bit codes, sort and topology are bit-exact with the oracle's 63-bit restatement
  (oracle/lbvh_oracle.c orc_expand21/reference_morton63/sort63/karras63, itself pinned in
  tests/test_oracle.py against an independent numpy restatement, a mask-based bit spread,
  numpy's stable argsort, a brute-force radix tree, and the exact reductions
  code63 >> 33 == code30 and karras63(code30 << 33) == karras30(code30)).
- Quadrupole moments (no reference counterpart: the reference is monopole-only,
  solvers/src/barnes_hut.rs:33) are checked against brute-force sums over each node's
  leaves, and through the force error they buy against an FP64 direct sum.
- The group walk accepts a node only when the reference acceptance test holds for every
  point of the group's box, so at theta -> 0 it is the direct sum, and at finite theta
  its error vs FP64 must not exceed the reference per-target walk's.
Stated tolerances are in each test.
"""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic
from oracle import oracle as O
from tests import inputs as I

pytestmark = pytest.mark.gpu
G, EPS = 1.0, 0.05
FAST = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2)
GROUP_MONO = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=0)


def say(*a):
    print("[case]", *a, flush=True)


# ------------------------------------------------------------------ 63-bit build stages


def assert_63_exact(pos, label):
    say("stages63", label, pos.shape)
    g = glx.morton(pos, bits=63)
    b, lanes, codes = O.reference_morton63(pos)
    assert np.array_equal(g.codes, codes), f"{label}: codes"
    assert np.array_equal(g.lanes, lanes), f"{label}: lanes"
    assert g.scale == b["scale"] and np.array_equal(g.bmin, b["bmin"])
    order, sc = glx.sort(g.codes)
    want = O.reference_sort63(codes)
    assert np.array_equal(order, want), f"{label}: order"
    assert np.array_equal(sc, codes[want])
    l, r, p = glx.karras(sc)
    lw, rw, pw = O.reference_karras63(codes[want])
    assert np.array_equal(l, lw) and np.array_equal(r, rw) and np.array_equal(p, pw), label


def test_63bit_stages_bit_exact():
    for seed in (1, 7):
        assert_63_exact(I.cloud(seed, 8000, 3.0), f"cloud{seed}")
    assert_63_exact(I.cloud(3, 8000, 3.0, (1.0e6, 1.0e6, 1.0e6)), "cloud_1e6")
    assert_63_exact(np.tile(np.array([[1.0, 1.0, 1.0]]), (300, 1)), "coincident")
    pos, _, _ = I.plummer_numpy(16384, seed=4)
    assert_63_exact(pos, "plummer16k")
    pos, _, _, _ = ic.collision(1 << 17, 3)
    assert_63_exact(pos, "collision128k")


def test_63bit_stages_bit_exact_million():
    rng = np.random.default_rng(8)
    assert_63_exact(rng.standard_normal((1 << 20, 3)) * 5.0, "gauss_1M")


# ------------------------------------------------------------------ quadrupole moments


def test_quadrupole_moments_match_bruteforce():
    pos = I.cloud(21, 1500, 3.0)
    mass = I.masses(22, 1500)
    _, _, codes = O.reference_morton(pos)
    order = O.reference_sort(codes)
    sc, sp, sm = codes[order], pos[order], mass[order]
    g = glx.aggregate(sc, sp, sm, quad=True)
    l, r, _ = O.reference_karras(sc)
    from oracle import numpy_ref as R

    sub = R.subtree_leaves(l, r, sc.size)
    leaves = (sp - g["origin"]).astype(np.float32).astype(np.float64)
    m32 = sm.astype(np.float32).astype(np.float64)
    worst = 0.0
    for u, idx in sub.items():
        if u < sc.size:
            assert np.all(g["quad"][u] == 0.0)
            continue
        x, m = leaves[idx], m32[idx]
        com = (x * m[:, None]).sum(0) / m.sum()
        d = x - com
        d2 = (d * d).sum(1)
        q = np.array([
            (m * (3 * d[:, 0] ** 2 - d2)).sum(), (m * (3 * d[:, 1] ** 2 - d2)).sum(),
            (m * (3 * d[:, 2] ** 2 - d2)).sum(), (m * 3 * d[:, 0] * d[:, 1]).sum(),
            (m * 3 * d[:, 0] * d[:, 2]).sum(), (m * 3 * d[:, 1] * d[:, 2]).sum(),
        ])
        scale = (m * d2).sum() + 1e-30
        err = np.abs(g["quad"][u].astype(np.float64) - q).max() / scale
        worst = max(worst, err)
    say("quadrupole worst rel err", worst)
    assert worst < 1e-4


# ------------------------------------------------------------------ group walk


def test_group_walk_theta_zero_is_direct_sum():
    for kw in (GROUP_MONO, FAST, dict(walk=glx.WALK_GROUP)):
        s = glx.Lbvh(G, EPS, 1e-6, **kw)
        for seed in range(8):
            pos, _, mass = I.cluster(seed, 300)
            exact = O.direct_sum(pos, mass, G, EPS)
            got = s.accelerations(pos, mass)
            assert I.rms_rel_err(got, exact) < 2e-5, (kw, seed)


def test_group_walk_momentum_and_determinism():
    s = glx.Lbvh(G, EPS, 1e-6, **FAST)
    pos, _, mass = I.cluster(4, 2000)
    a = s.accelerations(pos, mass)
    net = (a * mass[:, None]).sum(axis=0)
    assert np.linalg.norm(net) / (np.linalg.norm(a, axis=1) * mass).sum() < 1e-4
    s5 = glx.Lbvh(G, EPS, 0.5, **FAST)
    assert np.array_equal(s5.accelerations(pos, mass), s5.accelerations(pos, mass))


def test_group_walk_empty_single_coincident():
    s = glx.Lbvh(G, EPS, 0.5, **FAST)
    assert s.accelerations(np.zeros((0, 3)), np.zeros(0)).shape == (0, 3)
    assert np.array_equal(s.accelerations(np.array([[1.0, 2.0, 3.0]]), np.ones(1)), np.zeros((1, 3)))
    pos, mass = I.cluster_with_coincident_knot()
    exact = O.direct_sum(pos, mass, G, EPS)
    got = glx.Lbvh(G, EPS, 1e-6, **FAST).accelerations(pos, mass)
    assert I.rms_rel_err(got, exact) < 2e-5


def _errs(a, ref):
    return I.rel_err_quantiles(a, ref)


@pytest.mark.parametrize("leaf", [2, 8, 32])
def test_leaf_size_theta_zero_and_small_systems_are_direct_sum(leaf):
    """Buckets are summed body by body, so theta -> 0 stays the direct sum, and a system
    of at most leaf_size bodies is the direct sum at any theta."""
    for kw in (FAST, GROUP_MONO):
        s = glx.Lbvh(G, EPS, 1e-6, leaf_size=leaf, **kw)
        for seed in range(4):
            pos, _, mass = I.cluster(seed, 300)
            assert I.rms_rel_err(s.accelerations(pos, mass), O.direct_sum(pos, mass, G, EPS)) < 2e-5
        small = glx.Lbvh(G, EPS, 0.9, leaf_size=leaf, **kw)
        pos, _, mass = I.cluster(7, leaf)
        assert I.rms_rel_err(small.accelerations(pos, mass), O.direct_sum(pos, mass, G, EPS)) < 2e-6


def test_leaf_size_error_and_work():
    """Plummer N=16384, theta=0.5: larger buckets never raise the error (a bucket is
    exact where leaf size 1 might accept a descendant) and cut node tests."""
    pos, _, mass = I.plummer_numpy(16384, seed=9)
    eps = 0.01
    tg = np.arange(16384, dtype=np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, eps, tg)
    prev = None
    for leaf in (1, 4, 16):
        s = glx.Lbvh(1.0, eps, 0.5, count=True, leaf_size=leaf, **FAST)
        e = _errs(s.accelerations(pos, mass), d64)
        st = s.stats()
        say(f"leaf {leaf}: err {e} stats {st}")
        assert st["fallbacks"] == 0
        if prev is not None:
            assert e[0] <= prev[0][0] * 1.01 and e[1] <= prev[0][1] * 1.01
            assert st["node_tests"] < prev[1]["node_tests"]
        prev = (e, st)


def test_group_walk_error_not_worse_than_reference_walk():
    """Plummer N=16384, eps=0.01: per-particle relative error vs FP64 direct sum.
    Group monopole <= reference per-target walk (same theta); quadrupole at least 3x
    lower median than group monopole; quadrupole median <= 1e-3 at theta 0.5 and 0.6."""
    pos, _, mass = I.plummer_numpy(16384, seed=9)
    eps = 0.01
    tg = np.arange(16384, dtype=np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, eps, tg)
    for theta in (0.5, 0.6):
        ref = glx.Lbvh(1.0, eps, theta).accelerations(pos, mass)
        gm = glx.Lbvh(1.0, eps, theta, **GROUP_MONO).accelerations(pos, mass)
        sq = glx.Lbvh(1.0, eps, theta, count=True, **FAST)
        gq = sq.accelerations(pos, mass)
        e_ref, e_gm, e_gq = _errs(ref, d64), _errs(gm, d64), _errs(gq, d64)
        say(f"theta {theta}: ref {e_ref} group-mono {e_gm} group-quad {e_gq} stats {sq.stats()}")
        assert e_gm[0] <= e_ref[0] and e_gm[1] <= e_ref[1] * 1.05
        assert e_gq[0] * 3.0 <= e_gm[0]
        assert e_gq[0] <= 1e-3
        assert sq.stats()["fallbacks"] == 0


def test_fast_mode_collision_error():
    """Two-galaxy collision input (2^18): FAST mode at theta = 0.6, eps = 0.05; median
    relative error vs FP64 direct sum on 4096 seeded targets <= 1e-3."""
    pos, _, mass, _ = ic.collision(1 << 18, 12)
    rng = np.random.default_rng(2)
    tg = np.sort(rng.choice(pos.shape[0], 4096, replace=False)).astype(np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, 0.05, tg)
    s = glx.Lbvh(1.0, 0.05, 0.6, count=True, **FAST)
    a = s.accelerations(pos, mass)[tg]
    med, p99 = _errs(a, d64)
    say("collision 2^18 fast median", med, "p99", p99, "stats", s.stats())
    assert med <= 1e-3 and p99 <= 2e-2
    assert s.stats()["fallbacks"] == 0


def test_fast_resident_roundtrip_determinism_and_exact_kdk():
    """FAST resident stepping: bit-identical to upload/step/snapshot round trips and
    run to run; the velocity change over 20 steps matches the oracle's exact (direct
    sum) KDK to 5e-3 of its size."""
    pos, vel, mass = I.cluster(3, 2048, with_vel=True)
    outs = []
    for _ in range(2):
        r = glx.GpuResidentLeapfrog(G, EPS, 0.5, **FAST)
        r.upload(pos, vel, mass)
        r.step_many(1e-3, 20)
        outs.append(r.snapshot())
    assert np.array_equal(outs[0].pos, outs[1].pos) and np.array_equal(outs[0].vel, outs[1].vel)
    rt = glx.GpuResidentLeapfrog(G, EPS, 0.5, **FAST)
    p, v = pos, vel
    for _ in range(20):
        rt.upload(p, v, mass)
        rt.step(1e-3)
        st = rt.snapshot()
        p, v = st.pos, st.vel
    assert np.array_equal(p, outs[0].pos) and np.array_equal(v, outs[0].vel)
    pe, ve, _ = O.kdk_run(pos, vel, mass, G, EPS, 0.5, 1e-3, 20, solver="direct")
    dv = np.linalg.norm(outs[0].vel - ve) / np.linalg.norm(ve - vel)
    say("fast resident dv vs exact KDK (relative to dv_total)", dv)
    assert dv < 5e-3


FAR = dict(FAST, far_theta=0.4)


def test_far_field_theta_zero_is_direct_sum():
    s = glx.Lbvh(G, EPS, 1e-6, **FAR)
    for seed in range(4):
        pos, _, mass = I.cluster(seed, 300)
        assert I.rms_rel_err(s.accelerations(pos, mass), O.direct_sum(pos, mass, G, EPS)) < 2e-5


def test_far_field_error_and_work():
    """Plummer N=16384, theta=0.5 against the FP64 direct sum: the third-order local
    expansion adds almost nothing to the walk's error at far_theta 0.4 (median 1.41e-4
    against 1.41e-4 without; the second-order expansion gave 1.44e-4 in a CPU emulation)
    and little at 0.5 (1.54e-4; second order 2.42e-4), while taking accepted nodes out of
    the per-body lists."""
    pos, _, mass = I.plummer_numpy(16384, seed=9)
    eps = 0.01
    d64 = glx.direct_sum_f64(pos, mass, 1.0, eps, np.arange(16384, dtype=np.uint32))
    base = glx.Lbvh(1.0, eps, 0.5, count=True, **FAST)
    e0 = _errs(base.accelerations(pos, mass), d64)
    st0 = base.stats()
    for ft, med_max, p99_max in ((0.4, 1.5e-4, 1.3e-3), (0.5, 2.0e-4, 1.6e-3)):
        s = glx.Lbvh(1.0, eps, 0.5, count=True, **dict(FAST, far_theta=ft))
        a = s.accelerations(pos, mass)
        e = _errs(a, d64)
        st = s.stats()
        say(f"far_theta {ft}: err {e} (none: {e0}) stats {st}")
        assert e[0] <= med_max and e[1] <= p99_max
        assert st["far_nodes"] > 0 and st["quad_terms"] < st0["quad_terms"]
        assert st["fallbacks"] == 0
        assert np.array_equal(a, s.accelerations(pos, mass))  # deterministic


def test_far_field_is_third_order():
    """64 bodies of negligible mass in a unit cube feel a close pair of unit masses at
    distance ~ sqrt(3) R: the pair's node is a far node of the cube's group, so the walk's
    force is the local expansion of its field about the group centre. Its error against the
    FP64 direct sum falls as (r_group / distance)^4 for a third-order expansion: 21x when R
    doubles from 4 to 8 in an f64 emulation of the same expansion (7.5e-5 -> 3.6e-6), where
    a second-order one falls 9.7x (8.6e-4 -> 8.8e-5)."""
    rng = np.random.default_rng(5)
    tg = rng.uniform(0.0, 1.0, (64, 3))
    errs = []
    for R in (4.0, 8.0):
        pos = np.concatenate([tg, [[R, R, R], [R + 1e-4, R, R]]])
        mass = np.concatenate([np.full(64, 1e-12), [1.0, 1.0]])
        s = glx.Lbvh(1.0, 0.01, 0.5, count=True, **dict(FAST, far_theta=0.9))
        a = s.accelerations(pos, mass)[:64]
        assert s.stats()["far_nodes"] > 0
        d64 = glx.direct_sum_f64(pos, mass, 1.0, 0.01, np.arange(64, dtype=np.uint32))
        e = np.median(np.linalg.norm(a - d64, axis=1) / np.linalg.norm(d64, axis=1))
        say(f"far field at R {R}: median {e:.3e}")
        errs.append(e)
    assert errs[0] < 1.5e-4
    assert errs[0] / errs[1] > 14.0


@pytest.mark.parametrize("mode", ["monopole", "quadrupole"])
def test_group_walk_above_theta_one_tests_disjoint_boxes(mode):
    """Up to theta = 1 acceptance implies that a node's box and the group's box are
    disjoint (tests/test_walk_acceptance.py), so the group walk reads no box records there;
    above it the walk also requires disjoint boxes, from the box record. At theta = 1.2, on
    a Plummer sphere with 200 coincident pairs, the group walk's error vs FP64 stays below
    the reference per-target walk's at the same theta, and every force is finite."""
    pos, _, mass = ic.plummer(20000, 31)
    pos = np.concatenate([pos, pos[:200]])
    mass = np.concatenate([mass, mass[:200]])
    tg = np.arange(0, pos.shape[0], 37, dtype=np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, 0.01, tg)

    def err(solver):
        a = solver.accelerations(pos, mass)
        assert np.isfinite(a).all()
        e = np.linalg.norm(a[tg] - d64, axis=1) / np.linalg.norm(d64, axis=1)
        return float(np.median(e)), float(np.quantile(e, 0.99))

    ref = err(glx.Lbvh(1.0, 0.01, 1.2))
    grp = err(glx.Lbvh(1.0, 0.01, 1.2, **(GROUP_MONO if mode == "monopole" else FAST)))
    say(f"theta 1.2 {mode}: group {grp}, reference walk {ref}")
    assert grp[0] <= ref[0] and grp[1] <= ref[1]


@pytest.mark.parametrize("multipole", [0, 2])
def test_stack_overflow_falls_back_to_the_per_target_walk(monkeypatch, multipole):
    """A group whose LDS stack would overflow is walked per target instead (the reference
    criterion and monopoles, the lane walk of the reference build). GLX_WALK_STACK_CAP=1
    sends every group there: with 30-bit keys (the reference's tree) the forces then equal
    the reference build's bit for bit, in either node-record layout, the far field and the
    monopole band taking no part, and every group is counted as a fallback."""
    pos, _, mass = ic.plummer(20000, 41)
    ref = glx.Lbvh(1.0, 0.01, 0.5, morton_bits=30, walk=glx.WALK_REFERENCE,
                   multipole=0).accelerations(pos, mass)
    monkeypatch.setenv("GLX_WALK_STACK_CAP", "1")
    knobs = dict(far_theta=0.6, mono_ratio=0.9) if multipole == 2 else {}
    s = glx.Lbvh(1.0, 0.01, 0.5, count=True, morton_bits=30, walk=glx.WALK_GROUP,
                 multipole=multipole, **knobs)
    a = s.accelerations(pos, mass)
    st = s.stats()
    say("stack cap 1, multipole", multipole, ": fallbacks", st["fallbacks"], "far nodes",
        st["far_nodes"])
    assert st["fallbacks"] >= 20000 // 64 and st["far_nodes"] == 0
    assert np.array_equal(a, ref)


def test_monopole_band_limits():
    """mono_ratio -> 0 keeps every quadrupole (bit-equal to the band off); mono_ratio -> 1
    turns (nearly) every accepted node into a monopole, i.e. the group walk without
    quadrupoles; theta -> 0 accepts nothing, so the direct sum."""
    pos, _, mass = I.plummer_numpy(8192, seed=4)
    off = glx.Lbvh(1.0, 0.01, 0.5, **FAST).accelerations(pos, mass)
    tiny = glx.Lbvh(1.0, 0.01, 0.5, **dict(FAST, mono_ratio=1e-6)).accelerations(pos, mass)
    assert np.array_equal(off, tiny)
    mono_walk = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=0)
    m0 = glx.Lbvh(1.0, 0.01, 0.5, **mono_walk).accelerations(pos, mass)
    m1 = glx.Lbvh(1.0, 0.01, 0.5, **dict(FAST, mono_ratio=0.999999)).accelerations(pos, mass)
    d = np.linalg.norm(m1 - m0, axis=1) / np.linalg.norm(m0, axis=1)
    say(f"mono_ratio ~1 vs monopole group walk: median {np.median(d):.3g} max {d.max():.3g}")
    assert np.median(d) < 1e-5
    s = glx.Lbvh(G, EPS, 1e-6, **dict(FAST, mono_ratio=0.9))
    for seed in range(3):
        p, _, m = I.cluster(seed, 300)
        assert I.rms_rel_err(s.accelerations(p, m), O.direct_sum(p, m, G, EPS)) < 2e-5


def test_monopole_band_error_and_work():
    """Plummer N=16384, theta=0.5 against the FP64 direct sum: the band converts accepted
    quadrupoles into monopoles and the error grows with mono_ratio, staying below the
    monopole-only group walk's."""
    pos, _, mass = I.plummer_numpy(16384, seed=9)
    eps = 0.01
    d64 = glx.direct_sum_f64(pos, mass, 1.0, eps, np.arange(16384, dtype=np.uint32))
    mono_walk = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=0)
    e_mono = _errs(glx.Lbvh(1.0, eps, 0.5, **mono_walk).accelerations(pos, mass), d64)
    prev, prev_q = None, None
    for mr, med_max in ((0.0, 2e-4), (0.7, 5e-4), (0.8, 7e-4), (0.9, 9e-4)):
        s = glx.Lbvh(1.0, eps, 0.5, count=True, **dict(FAST, mono_ratio=mr))
        e = _errs(s.accelerations(pos, mass), d64)
        st = s.stats()
        say(f"mono_ratio {mr}: err {e} stats {st} (monopole walk {e_mono})")
        assert e[0] <= med_max and e[0] < e_mono[0]
        if prev is not None:
            assert e[0] >= prev * 0.98 and st["quad_terms"] < prev_q
        prev, prev_q = e[0], st["quad_terms"]


def test_far_field_collision_error():
    pos, _, mass, _ = ic.collision(1 << 18, 12)
    rng = np.random.default_rng(2)
    tg = np.sort(rng.choice(1 << 18, 4096, replace=False)).astype(np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, 0.05, tg)
    a = glx.Lbvh(1.0, 0.05, 0.6, **dict(FAST, far_theta=0.5)).accelerations(pos, mass)
    e = _errs(a[tg], d64)
    say(f"collision 2^18 far_theta 0.5: {e}")
    assert e[0] <= 5e-4


# ------------------------------------------------------------------ tuning knobs


def test_radix_width_leaves_forces_and_stepping_bit_identical():
    """The sort's digit width changes only the pass count (and, for an odd count, the
    buffer the sort starts in): forces and resident steps are bit-identical for 8, 9
    and 10 bits, in the reference build (30-bit keys: 4, 4, 3 passes) and the tuned fast
    build (63-bit keys: 8, 7, 7 passes)."""
    pos, vel, mass = I.plummer_numpy(20000, seed=11)
    tuned = dict(FAST, far_theta=0.55, mono_ratio=0.8)
    for kw in ({}, tuned):
        base = glx.Lbvh(1.0, 0.01, 0.5, **kw).accelerations(pos, mass)
        for bits in (9, 10):
            say("radix width", bits, kw)
            got = glx.Lbvh(1.0, 0.01, 0.5, radix_bits=bits, **kw).accelerations(pos, mass)
            assert np.array_equal(got, base), (bits, kw)
    runs = []
    for bits in (8, 9, 10):
        sim = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5, radix_bits=bits, **tuned)
        sim.upload(pos, vel, mass)
        sim.step_many(1e-3, 4)
        runs.append(sim.snapshot())
        sim.close()
    for r in runs[1:]:
        assert np.array_equal(r.pos, runs[0].pos) and np.array_equal(r.vel, runs[0].vel)


def test_walk_block_shape_group_size_and_lanes():
    """Walk block shape: group_size bounds the walk groups (subtrees of at most that many
    bodies), group_lanes the lanes per body in groups of at most 32. Every shape is the
    direct sum at theta -> 0. At theta = 0.5 (Plummer N=16384 against the FP64 direct sum)
    smaller groups test more nodes and sum fewer terms per body, accepting nodes nearer to
    their bodies: the median error grows (1.4e-4 at 64 bodies, 2.2e-4 at 16) but every
    shape stays within 5e-4, well below the per-target reference walk's. With one group
    size the lane count changes only the summation order."""
    for gs in (8, 16, 32, 48):
        s = glx.Lbvh(G, EPS, 1e-6, group_size=gs, **FAST)
        for seed in range(3):
            pos, _, mass = I.cluster(seed, 300)
            assert I.rms_rel_err(s.accelerations(pos, mass), O.direct_sum(pos, mass, G, EPS)) < 2e-5
    pos, _, mass = I.plummer_numpy(16384, seed=9)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, 0.01, np.arange(16384, dtype=np.uint32))
    rows = {}
    for gs in (64, 32, 16):
        for lanes in (1, 2, 4):
            s = glx.Lbvh(1.0, 0.01, 0.5, count=True, group_size=gs, group_lanes=lanes, **FAST)
            a = s.accelerations(pos, mass)
            st = s.stats()
            rows[gs, lanes] = (a, _errs(a, d64), st)
            say(f"group_size {gs} lanes {lanes}: err {rows[gs, lanes][1]} stats {st}")
            assert rows[gs, lanes][1][0] <= 5e-4
        for lanes in (1, 2):
            assert I.rms_rel_err(rows[gs, lanes][0], rows[gs, 4][0]) < 1e-5, (gs, lanes)
    t64, t16 = rows[64, 4][2], rows[16, 4][2]
    assert t16["rounds"] > t64["rounds"]
    assert t16["point_terms"] + t16["quad_terms"] < t64["point_terms"] + t64["quad_terms"]
