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
