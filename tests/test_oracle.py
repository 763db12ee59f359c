"""CPU tests pinning the oracle (oracle/lbvh_oracle.c) before it is trusted as the
checker: known answers from the reference's own tests, an independent numpy
restatement, brute-force constructions and the mathematical invariants the reference
gates on. No GPU needed."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

from oracle import numpy_ref as R
from oracle import oracle as O
from tests import inputs as I

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))
G, EPS = 1.0, 0.05


# ----------------------------------------------------------------- Morton primitives


def test_expand10_known_answers():
    for v, want in KA["expand10"]["cases"]:
        assert O.expand10(v) == want


def test_morton3_known_answers():
    for (x, y, z), want in KA["morton3"]["cases"]:
        assert O.morton3(x, y, z) == want


def test_expand10_matches_bit_loop_for_every_lane_value():
    for v in range(1024):
        assert O.expand10(v) == R.expand10_naive(v)


def test_morton3_monotone_along_each_axis():
    for axis in range(3):
        prev = -1
        for v in range(1024):
            lanes = [0, 0, 0]
            lanes[axis] = v
            c = O.morton3(*lanes)
            assert c > prev
            prev = c


# ----------------------------------------------------------------- reference_morton


@pytest.mark.parametrize("seed", [1, 7, 42, 1000])
def test_morton_matches_numpy_restatement(seed):
    pos = I.cloud(seed, 8000, 3.0)
    b, lanes, codes = O.reference_morton(pos)
    b2, lanes2, codes2 = R.morton(pos)
    assert b["size"] == b2["size"] and b["scale"] == b2["scale"]
    assert np.array_equal(b["bmin"], b2["bmin"])
    assert np.array_equal(lanes, lanes2)
    assert np.array_equal(codes, codes2)


def test_morton_large_coordinates_match_numpy_restatement():
    pos = I.cloud(3, 8000, 3.0, (1.0e6, 1.0e6, 1.0e6))
    _, lanes, codes = O.reference_morton(pos)
    _, lanes2, codes2 = R.morton(pos)
    assert np.array_equal(lanes, lanes2) and np.array_equal(codes, codes2)


def test_morton_codes_are_interleave_of_in_range_lanes():
    pos = I.cloud(9, 5000, 3.0)
    _, lanes, codes = O.reference_morton(pos)
    assert lanes.max() < 1024
    assert codes.max() < (1 << 30)
    for l, c in zip(lanes[:500], codes[:500]):
        assert c == O.morton3(*[int(v) for v in l])


def test_morton_collinear_known_bbox():
    ka = KA["bbox_collinear"]
    pos = np.array([[i * 0.1, 2.0, -1.0] for i in range(ka["n"])])
    b, lanes, _ = O.reference_morton(pos)
    # lo/hi of the reference cube: bmin + size/2 is the centre, extents from lanes
    center = b["bmin"] + 0.5 * b["size"]
    assert np.float32(center[1]) == np.float32(ka["bbox_min_f32"][1])
    assert np.float32(center[2]) == np.float32(ka["bbox_min_f32"][2])
    half = 0.5 * b["size"]
    assert np.float32(center[0] - half / (1 + 1e-9)) == np.float32(ka["bbox_min_f32"][0])
    assert np.float32(center[0] + half / (1 + 1e-9)) == np.float32(ka["bbox_max_f32"][0])
    assert lanes[0, 0] == 0 and lanes[-1, 0] == 1023
    assert np.all(np.diff(lanes[:, 0].astype(np.int64)) >= 0)
    # degenerate axes: every particle shares one lane next to the cube centre (the f64
    # rounding of p - bmin decides 511 vs 512; the numpy restatement must agree)
    for a in (1, 2):
        assert np.all(lanes[:, a] == lanes[0, a]) and lanes[0, a] in (511, 512)
    assert np.array_equal(lanes, R.morton(pos)[1])


def test_morton_single_and_coincident():
    _, lanes, codes = O.reference_morton(np.array(KA["single_particle"]["pos"]))
    assert lanes.shape == (1, 3) and np.all(lanes == 512)
    ka = KA["coincident_share_code"]
    pos = np.tile(np.array(ka["pos"]), (ka["n"], 1))
    _, lanes, codes = O.reference_morton(pos)
    assert np.all(codes == codes[0]) and np.all(lanes == lanes[0])


def test_morton_rejects_empty():
    with pytest.raises(ValueError):
        O.reference_morton(np.zeros((0, 3)))


# ----------------------------------------------------------------- reference_sort


def _stable(codes):
    return np.argsort(codes, kind="stable").astype(np.uint32)


@pytest.mark.parametrize(
    "codes",
    [
        I.random_codes(2, 10000),
        I.random_codes(99, 10000),
        np.arange(5000, dtype=np.uint32) % 4,
        (0x00ABCD00 | ((np.arange(4096, dtype=np.uint64) * 37) & 0xFF)).astype(np.uint32),
        ((np.arange(4096, dtype=np.uint32) % 64) << 24).astype(np.uint32),
        I.random_codes(0xBEEF, 65536),
    ],
    ids=["random2", "random99", "ties4", "lowbyte", "highbyte", "large"],
)
def test_sort_matches_stable_argsort(codes):
    assert np.array_equal(O.reference_sort(codes), _stable(codes))


def test_sort_known_answers():
    n = KA["sort_identity"]["n"]
    assert np.array_equal(O.reference_sort(np.arange(n, dtype=np.uint32)), np.arange(n))
    n = KA["sort_reversed"]["n"]
    assert np.array_equal(O.reference_sort(np.arange(n, dtype=np.uint32)[::-1]), np.arange(n)[::-1])
    ka = KA["sort_all_equal"]
    assert np.array_equal(O.reference_sort(np.full(ka["n"], ka["code"], np.uint32)), np.arange(ka["n"]))
    ka = KA["sort_single"]
    assert list(O.reference_sort(np.array(ka["codes"], np.uint32))) == ka["order"]


# ----------------------------------------------------------------- reference_karras


def _check_karras_vs_bruteforce(sc):
    l1, r1, p1 = O.reference_karras(sc)
    l2, r2, p2 = R.radix_tree_bruteforce(sc)
    assert np.array_equal(l1, l2) and np.array_equal(r1, r2) and np.array_equal(p1, p2)
    _check_structural(l1, r1, p1, len(sc))


def _check_structural(left, right, parent, n):
    total = 2 * n - 1
    root = 0 if n == 1 else n
    cc = np.zeros(total, np.int64)
    for i in range(n - 1):
        me = n + i
        l, r = int(left[i]), int(right[i])
        assert l < total and r < total and l != r
        assert parent[l] == me and parent[r] == me
        cc[l] += 1
        cc[r] += 1
    for u in range(total):
        if u == root:
            assert cc[u] == 0 and parent[u] == O.NO_PARENT
        else:
            assert cc[u] == 1


@pytest.mark.parametrize("n", [2, 3, 5, 17, 128, 257, 2000])
def test_karras_matches_bruteforce_on_morton_codes(n):
    for seed in (1, 7):
        pos = I.cloud(seed, n, 3.0)
        _, _, codes = O.reference_morton(pos)
        _check_karras_vs_bruteforce(np.sort(codes))


def test_karras_all_equal_heavy_duplicates_chain():
    _check_karras_vs_bruteforce(np.full(300, 42, np.uint32))
    _check_karras_vs_bruteforce((np.arange(2000, dtype=np.uint32) // 100).astype(np.uint32))
    _check_karras_vs_bruteforce(np.arange(1500, dtype=np.uint32))


def test_karras_known_answers():
    l, r, p = O.reference_karras(np.array([5], np.uint32))
    assert list(p) == KA["tree_single"]["parent"] and l.size == 0
    ka = KA["tree_two"]
    pos = np.array(ka["pos"])
    _, _, codes = O.reference_morton(pos)
    order = O.reference_sort(codes)
    l, r, p = O.reference_karras(codes[order])
    assert list(p) == ka["parent"]


# ----------------------------------------------------------------- 63-bit keys
# The 63-bit variant (2^21 cells per axis) is the reference's deferred refinement
# (solvers/src/lbvh.rs:29-31) over the cube convention of :105-129. Each function is
# pinned three ways: an independent numpy restatement, a different construction (a
# magic-mask bit spread, numpy's stable argsort, the brute-force top-down radix tree),
# and an exact reduction to the already pinned 30-bit path.


def _expand21_masks(v: int) -> int:
    """Bit spread by shift-and-mask doubling (a different construction from the oracle's
    per-bit loop)."""
    x = v & 0x1FFFFF
    x = (x | (x << 32)) & 0x001F00000000FFFF
    x = (x | (x << 16)) & 0x001F0000FF0000FF
    x = (x | (x << 8)) & 0x100F00F00F00F00F
    x = (x | (x << 4)) & 0x10C30C30C30C30C3
    x = (x | (x << 2)) & 0x1249249249249249
    return x


def test_expand21_known_answers_and_mask_construction():
    assert O.expand21(0) == 0
    assert O.expand21(1) == 1
    assert O.expand21(2) == 8
    assert O.expand21(0x1FFFFF) == 0x1249249249249249
    assert O.expand21(1 << 20) == 1 << 60
    assert O.expand21(0x3FFFFF) == O.expand21(0x1FFFFF)  # bits above 21 are dropped
    rng = np.random.default_rng(21)
    sample = np.concatenate([np.arange(4096), rng.integers(0, 1 << 21, 20000),
                             (1 << 21) - 1 - np.arange(4096)])
    for v in sample:
        v = int(v)
        assert O.expand21(v) == _expand21_masks(v), v


@pytest.mark.parametrize("seed", [1, 7, 42, 1000])
def test_morton63_matches_numpy_restatement(seed):
    pos = I.cloud(seed, 8000, 3.0)
    b, lanes, codes = O.reference_morton63(pos)
    b2, lanes2, codes2 = R.morton63(pos)
    assert b["size"] == b2["size"] and b["scale"] == b2["scale"]
    assert np.array_equal(b["bmin"], b2["bmin"])
    assert np.array_equal(lanes, lanes2)
    assert np.array_equal(codes, codes2)
    assert lanes.max() < (1 << 21) and codes.max() < (1 << 63)


def test_morton63_large_coordinates_match_numpy_restatement():
    pos = I.cloud(3, 8000, 3.0, (1.0e6, 1.0e6, 1.0e6))
    _, lanes, codes = O.reference_morton63(pos)
    _, lanes2, codes2 = R.morton63(pos)
    assert np.array_equal(lanes, lanes2) and np.array_equal(codes, codes2)


@pytest.mark.parametrize("seed", [1, 7, 42])
def test_morton63_refines_the_30bit_codes(seed):
    """2^21 / size is 2^11 · (1024 / size) exactly (a power-of-two scaling commutes with
    rounding), so every 21-bit lane shifted right by 11 is the 30-bit lane, both clamps
    agree, and the 63-bit code shifted right by 33 is the reference's 30-bit code."""
    for pos in (I.cloud(seed, 6000, 3.0), I.cloud(seed, 3000, 3.0, (1.0e6, -2.0e5, 7.0)),
                I.plummer_numpy(4000, seed=seed)[0]):
        b30, lanes30, codes30 = O.reference_morton(pos)
        b63, lanes63, codes63 = O.reference_morton63(pos)
        assert np.array_equal(b30["bmin"], b63["bmin"]) and b30["size"] == b63["size"]
        assert b63["scale"] == 2048.0 * b30["scale"]
        assert np.array_equal(lanes63 >> 11, lanes30)
        assert np.array_equal((codes63 >> np.uint64(33)).astype(np.uint32), codes30)
        for l, c in zip(lanes63[:300], codes63[:300]):
            want = 0
            for a in range(3):
                want |= _expand21_masks(int(l[a])) << a
            assert int(c) == want


def test_morton63_collinear_single_and_coincident():
    n = KA["bbox_collinear"]["n"]
    pos = np.array([[i * 0.1, 2.0, -1.0] for i in range(n)])
    _, lanes, _ = O.reference_morton63(pos)
    assert lanes[0, 0] == 0 and lanes[-1, 0] == (1 << 21) - 1
    assert np.all(np.diff(lanes[:, 0].astype(np.int64)) > 0)
    for a in (1, 2):
        assert np.all(lanes[:, a] == lanes[0, a])
        assert lanes[0, a] >> 11 in (511, 512)
    # a lone particle sits at the centre of a 2e-12 cube: f64 rounding of p - bmin at
    # |p| ~ 1 moves it by many of the 2^21 fine cells, but not out of the coarse cell 512
    _, lanes, _ = O.reference_morton63(np.array(KA["single_particle"]["pos"]))
    assert np.all(lanes >> 11 == 512)
    ka = KA["coincident_share_code"]
    _, lanes, codes = O.reference_morton63(np.tile(np.array(ka["pos"]), (ka["n"], 1)))
    assert np.all(codes == codes[0]) and np.all(lanes == lanes[0])
    with pytest.raises(ValueError):
        O.reference_morton63(np.zeros((0, 3)))


def _codes63(seed, n):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << 63, n, dtype=np.uint64)


@pytest.mark.parametrize(
    "codes",
    [
        _codes63(2, 10000),
        _codes63(3, 65536),
        (np.arange(5000, dtype=np.uint64) % 4) << np.uint64(40),
        # keys equal in the low 32 bits and different above them: a sort that drops the
        # high word fails here
        (np.arange(4096, dtype=np.uint64)[::-1] << np.uint64(33)) | np.uint64(0x5A5A5A5A),
        # keys equal above bit 32 and different below
        np.uint64(0x7000000000000000) | (_codes63(4, 4096) & np.uint64(0xFFFFFFFF)),
        np.full(3000, (1 << 63) - 1, np.uint64),
    ],
    ids=["random", "large", "ties4_high", "high_word_only", "low_word_only", "all_equal"],
)
def test_sort63_matches_stable_argsort(codes):
    assert np.array_equal(O.reference_sort63(codes),
                          np.argsort(codes, kind="stable").astype(np.uint32))


def test_sort63_known_answers():
    assert np.array_equal(O.reference_sort63(np.arange(2000, dtype=np.uint64)), np.arange(2000))
    assert np.array_equal(O.reference_sort63(np.arange(2000, dtype=np.uint64)[::-1]),
                          np.arange(2000)[::-1])
    assert list(O.reference_sort63(np.array([9], np.uint64))) == [0]
    assert O.reference_sort63(np.zeros(0, np.uint64)).size == 0


def _check_karras63_vs_bruteforce(sc):
    l1, r1, p1 = O.reference_karras63(sc)
    l2, r2, p2 = R.radix_tree_bruteforce63(sc)
    assert np.array_equal(l1, l2) and np.array_equal(r1, r2) and np.array_equal(p1, p2)
    _check_structural(l1, r1, p1, len(sc))


@pytest.mark.parametrize("n", [2, 3, 5, 17, 128, 257, 2000])
def test_karras63_matches_bruteforce_on_morton_codes(n):
    for seed in (1, 7):
        _, _, codes = O.reference_morton63(I.cloud(seed, n, 3.0))
        _check_karras63_vs_bruteforce(np.sort(codes))


def test_karras63_all_equal_duplicates_chain_and_word_boundaries():
    _check_karras63_vs_bruteforce(np.full(300, 42, np.uint64))
    _check_karras63_vs_bruteforce(np.full(200, (1 << 63) - 1, np.uint64))
    _check_karras63_vs_bruteforce((np.arange(2000, dtype=np.uint64) // 100) << np.uint64(35))
    _check_karras63_vs_bruteforce(np.arange(1500, dtype=np.uint64))
    # differences only above bit 32, only below it, and straddling it
    _check_karras63_vs_bruteforce(np.sort(_codes63(5, 1000) & np.uint64(0x7FFFFFFF00000000)))
    _check_karras63_vs_bruteforce(np.sort(_codes63(6, 1000) & np.uint64(0x00000000FFFFFFFF)))
    _check_karras63_vs_bruteforce(np.sort(_codes63(7, 1000) & np.uint64(0x0000000FFFF00000)))


def test_karras63_known_answers():
    l, r, p = O.reference_karras63(np.array([5], np.uint64))
    assert list(p) == KA["tree_single"]["parent"] and l.size == 0
    ka = KA["tree_two"]
    _, _, codes = O.reference_morton63(np.array(ka["pos"]))
    order = O.reference_sort63(codes)
    _, _, p = O.reference_karras63(codes[order])
    assert list(p) == ka["parent"]


@pytest.mark.parametrize("seed", [1, 7, 42])
def test_karras63_of_shifted_30bit_codes_is_the_30bit_tree(seed):
    """For codes c << 33, δ63 maps the 30-bit δ order-preservingly (distinct codes:
    δ30 − 1; ties: δ30 + 32; out of range: −1), so the Karras search takes the same
    branches and the topology equals reference_karras on c."""
    for n in (1000, 20000):
        _, _, codes = O.reference_morton(I.cloud(seed, n, 3.0))
        sc = np.sort(codes)
        l30, r30, p30 = O.reference_karras(sc)
        l63, r63, p63 = O.reference_karras63(sc.astype(np.uint64) << np.uint64(33))
        assert np.array_equal(l30, l63) and np.array_equal(r30, r63)
        assert np.array_equal(p30, p63)


# ----------------------------------------------------------------- reference_aggregate


def test_aggregate_matches_bruteforce_subtrees():
    pos = I.cloud(3, 600, 3.0)
    mass = I.masses(3 ^ 0xABCD, 600)
    _, _, codes = O.reference_morton(pos)
    order = O.reference_sort(codes)
    sp, sm = pos[order], mass[order]
    l, r, _ = O.reference_karras(codes[order])
    agg = O.reference_aggregate(l, r, sp, sm)
    sub = R.subtree_leaves(l, r, 600)
    for u, leaves in sub.items():
        ps, ms = sp[leaves], sm[leaves]
        assert np.array_equal(agg["aabb_min"][u], ps.min(axis=0))
        assert np.array_equal(agg["aabb_max"][u], ps.max(axis=0))
        assert abs(agg["mass"][u] - ms.sum()) <= 1e-12 * ms.sum()
        com = (ps * ms[:, None]).sum(axis=0) / ms.sum()
        assert np.all(np.abs(agg["com"][u] - com) <= 1e-12 * 3.0)


def test_aggregate_massless_uses_midpoint():
    pos = I.cloud(5, 64, 2.0)
    mass = np.zeros(64)
    _, _, codes = O.reference_morton(pos)
    order = O.reference_sort(codes)
    sp = pos[order]
    l, r, _ = O.reference_karras(codes[order])
    agg = O.reference_aggregate(l, r, sp, np.zeros(64))
    root = 64
    assert agg["mass"][root] == 0.0
    # every internal node's com lies within its AABB (midpoint fallback)
    for u in range(64, 127):
        assert np.all(agg["com"][u] >= agg["aabb_min"][u] - 1e-12)
        assert np.all(agg["com"][u] <= agg["aabb_max"][u] + 1e-12)


# ----------------------------------------------------------------- reference_flatten


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 128, 257])
def test_flatten_structure_wellformed(n):
    """Port of solvers/tests/lbvh.rs:164-245."""
    for seed in range(4):
        pos, _, mass = I.cluster(seed, n)
        nodes, lb = O.build_flat(pos, mass)
        assert nodes.size == 2 * n - 1
        leaves = nodes["body_count"] > 0
        assert leaves.sum() == n and (~leaves).sum() == n - 1
        assert np.all(nodes["body_count"][leaves] == 1)
        assert sorted(lb[nodes["body_start"][leaves]].tolist()) == list(range(n))
        end = nodes.size
        k = np.arange(end)
        assert np.all(nodes["next"] > k) and np.all(nodes["next"] <= end)
        assert nodes["next"][0] == end
        for q in np.nonzero(~leaves)[0]:
            left = q + 1
            right = nodes["next"][left]
            assert right < nodes["next"][q]
            assert nodes["next"][right] == nodes["next"][q]
            for c in (left, right):
                pmin = nodes["center"][q] - nodes["half"][q]
                pmax = nodes["center"][q] + nodes["half"][q]
                cmin = nodes["center"][c] - nodes["half"][c]
                cmax = nodes["center"][c] + nodes["half"][c]
                assert np.all(cmin >= pmin - 1e-9) and np.all(cmax <= pmax + 1e-9)


def test_flatten_known_answers():
    ka = KA["flatten_single"]
    nodes, lb = O.build_flat(np.array([[2.0, -1.0, 0.5]]), np.array([1.25]))
    assert list(nodes["next"]) == ka["next"] and list(nodes["body_count"]) == ka["body_count"]
    assert list(nodes["delta"]) == ka["delta"] and list(lb) == [0]
    ka = KA["flatten_two"]
    nodes, _ = O.build_flat(np.array(KA["tree_two"]["pos"]), np.array(KA["tree_two"]["mass"]))
    assert list(nodes["next"]) == ka["next"] and list(nodes["body_count"]) == ka["body_count"]


# ----------------------------------------------------------------- walk / solver

GATES = KA["force_gates"]


def test_lbvh_theta_to_zero_reproduces_direct_sum():
    """solvers/tests/lbvh.rs:89-102"""
    for seed in range(100):
        pos, _, mass = I.cluster(seed, 120)
        exact = O.direct_sum(pos, mass, G, EPS)
        got, _ = O.lbvh_accelerations(pos, mass, G, EPS, 1e-6)
        assert I.worst_rel_err(got, exact) < GATES["cpu_theta0_worst"]


def test_lbvh_finite_theta_bounded_and_grows():
    """solvers/tests/lbvh.rs:108-135"""
    for seed in range(100):
        pos, _, mass = I.cluster(seed, 120)
        exact = O.direct_sum(pos, mass, G, EPS)
        lo, _ = O.lbvh_accelerations(pos, mass, G, EPS, 0.3)
        hi, _ = O.lbvh_accelerations(pos, mass, G, EPS, 0.6)
        rl, rh = I.rms_rel_err(lo, exact), I.rms_rel_err(hi, exact)
        assert rl < GATES["cpu_theta03_rms"] and rh < GATES["cpu_theta06_rms"]
        assert I.worst_rel_err(lo, exact) < GATES["cpu_theta03_worst"]
        assert I.worst_rel_err(hi, exact) < GATES["cpu_theta06_worst"]
        assert rh > rl


def test_lbvh_momentum_flux_at_theta_zero():
    """solvers/tests/lbvh.rs:141-157"""
    for seed in range(32):
        pos, _, mass = I.cluster(seed, 200)
        a, _ = O.lbvh_accelerations(pos, mass, G, EPS, 1e-6)
        net = (a * mass[:, None]).sum(axis=0)
        scale = (np.linalg.norm(a, axis=1) * mass).sum()
        assert np.linalg.norm(net) / max(scale, 1e-300) < 1e-9


def test_lbvh_coincident_knot():
    """solvers/tests/lbvh.rs:268-292"""
    pos, mass = I.cluster_with_coincident_knot()
    a1, _ = O.lbvh_accelerations(pos, mass, G, EPS, 0.5)
    a2, _ = O.lbvh_accelerations(pos, mass, G, EPS, 0.5)
    assert np.array_equal(a1, a2)
    exact = O.direct_sum(pos, mass, G, EPS)
    got, _ = O.lbvh_accelerations(pos, mass, G, EPS, 1e-6)
    assert I.worst_rel_err(got, exact) < 1e-9


def test_lbvh_empty_and_single():
    a, c = O.lbvh_accelerations(np.zeros((0, 3)), np.zeros(0), G, EPS, 0.5)
    assert a.shape == (0, 3) and c == 0
    a, _ = O.lbvh_accelerations(np.array(KA["lone_particle"]["pos"]), np.ones(1), G, EPS, 0.5)
    assert np.array_equal(a, np.array(KA["lone_particle"]["acc"]))


def test_lbvh_targets_subset_and_threads_agree():
    pos, _, mass = I.cluster(4, 300)
    full, cnt = O.lbvh_accelerations(pos, mass, G, EPS, 0.5)
    tg = np.array([0, 17, 299, 150], np.uint32)
    sub, _ = O.lbvh_accelerations(pos, mass, G, EPS, 0.5, targets=tg, nthreads=4)
    assert np.array_equal(sub, full[tg])
    par, cnt2 = O.lbvh_accelerations(pos, mass, G, EPS, 0.5, nthreads=4)
    assert np.array_equal(par, full) and cnt2 == cnt


# ----------------------------------------------------------------- direct sum, energy


def test_diagnostics_two_body_hand_values():
    ka = KA["diagnostics_two_body"]
    d = O.diagnostics(np.array(ka["pos"]), np.array(ka["vel"]), np.array(ka["mass"]))
    assert abs(d["ke"] - ka["kinetic"]) < 1e-12
    assert np.allclose(d["p"], ka["momentum"], atol=1e-12)
    assert np.allclose(d["l"], ka["angular_momentum"], atol=1e-12)
    assert np.allclose(d["com"], ka["center_of_mass"], atol=1e-12)


def test_direct_sum_two_body_hand_values():
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    mass = np.array([1.0, 3.0])
    a = O.direct_sum(pos, mass, 1.0, 0.0)
    assert np.allclose(a, [[0.75, 0, 0], [-0.25, 0, 0]], rtol=0, atol=1e-15)
    u = O.potential_serial(pos, mass, 1.0, 0.0)
    assert abs(u + 1.5) < 1e-15
    u = O.potential_serial(pos, mass, 2.0, 1.5)  # sqrt(4 + 2.25) = 2.5
    assert abs(u + 2.0 * 3.0 / 2.5) < 1e-15


def test_direct_sum_targets_matches_pairwise():
    pos, _, mass = I.cluster(8, 150)
    full = O.direct_sum(pos, mass, G, EPS)
    tg = np.arange(0, 150, 7, dtype=np.uint32)
    sub = O.direct_sum_targets(pos, mass, G, EPS, tg, nthreads=2)
    assert np.allclose(sub, full[tg], rtol=1e-12, atol=1e-12)


def test_potential_plummer_closed_form():
    """ic/src/plummer.rs:75-78: W = -3 pi G M^2 / (32 a) for a Plummer sphere."""
    pos, _, mass = I.plummer_numpy(8192, seed=5)
    u = O.potential_serial(pos, mass, 1.0, 1e-3)
    w = -3.0 * math.pi / 32.0
    assert abs(u - w) / abs(w) < 0.03


def test_lbvh_potential_energy_is_direct():
    pos, _, mass = I.cluster(2, 100)
    u = O.potential_serial(pos, mass, G, EPS)
    brute = 0.0
    for i in range(100):
        for j in range(i + 1, 100):
            brute -= mass[i] * mass[j] / math.sqrt(((pos[j] - pos[i]) ** 2).sum() + EPS * EPS)
    assert abs(u - brute) < 1e-12 * abs(brute)


# ----------------------------------------------------------------- leapfrog


def test_kdk_two_body_energy_and_reversibility():
    # circular binary, G = 1, m = 1 each at separation 1: v = sqrt(G M_tot / r)/2 each
    pos = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0]])
    v = math.sqrt(2.0) / 2.0
    vel = np.array([[0.0, -v, 0.0], [0.0, v, 0.0]])
    mass = np.ones(2)
    e0 = O.diagnostics(pos, vel, mass)["ke"] + O.potential_serial(pos, mass, 1.0, 0.0)
    p1, v1, _ = O.kdk_run(pos, vel, mass, 1.0, 0.0, 0.5, 1e-3, 2000, solver="direct")
    e1 = O.diagnostics(p1, v1, mass)["ke"] + O.potential_serial(p1, mass, 1.0, 0.0)
    assert abs(e1 - e0) / abs(e0) < 1e-6
    # circular orbit: separation stays 1
    assert abs(np.linalg.norm(p1[1] - p1[0]) - 1.0) < 1e-5
    # time reversal: flip velocities and step back
    p2, v2, _ = O.kdk_run(p1, -v1, mass, 1.0, 0.0, 0.5, 1e-3, 2000, solver="direct")
    assert np.allclose(p2, pos, atol=1e-10) and np.allclose(-v2, vel, atol=1e-10)


def test_kdk_lbvh_theta0_equals_direct_kdk():
    pos, vel, mass = I.cluster(6, 64, with_vel=True)
    pa, va, _ = O.kdk_run(pos, vel, mass, G, EPS, 1e-6, 1e-3, 5, solver="lbvh")
    pb, vb, _ = O.kdk_run(pos, vel, mass, G, EPS, 1e-6, 1e-3, 5, solver="direct")
    assert np.allclose(pa, pb, rtol=0, atol=1e-13) and np.allclose(va, vb, rtol=0, atol=1e-11)


# ----------------------------------------------------------------- committed fixture


def test_oracle_matches_committed_vectors():
    d = np.load(os.path.join(HERE, "golden", "oracle_vectors.npz"))
    for seed in (1, 7, 42):
        pos = d[f"cloud{seed}_pos"]
        assert np.array_equal(pos, I.cloud(seed, 2000, 3.0))
        b, _, codes = O.reference_morton(pos)
        assert np.array_equal(codes, d[f"cloud{seed}_codes"])
        assert np.array_equal(np.concatenate([b["bmin"], [b["size"], b["scale"]]]),
                              d[f"cloud{seed}_bounds"])
        order = O.reference_sort(codes)
        assert np.array_equal(order, d[f"cloud{seed}_order"])
        l, r, p = O.reference_karras(codes[order])
        assert np.array_equal(l, d[f"cloud{seed}_left"])
        assert np.array_equal(r, d[f"cloud{seed}_right"])
        assert np.array_equal(p, d[f"cloud{seed}_parent"])
    assert np.array_equal(O.reference_sort(d["random2_codes"]), d["random2_order"])
    for seed in (0, 1):
        a, c = O.lbvh_accelerations(d[f"cluster{seed}_pos"], d[f"cluster{seed}_mass"], 1.0, 0.05, 0.5)
        assert np.array_equal(a, d[f"cluster{seed}_lbvh_theta05"])
        assert c == int(d[f"cluster{seed}_lbvh_theta05_count"][0])
