"""GPU stage parity: the HIP kernels against the oracle, stage by stage.

Morton codes/lanes/bounds, the sort permutation and the Karras topology are pure
integer (or exact FP64) work and must be bit-exact with the oracle — stricter than the
reference's own GPU gates, which allow +-1 Morton lane (gpu/tests/lbvh_morton.rs:104-136).
The aggregate's AABB is exact (min/max), so it is bit-exact too; centre of mass and mass
are f32 folds and are held to the reference's tolerances (gpu/tests/lbvh_tree.rs:237-265)
and to tighter ones of our own.
"""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic
from oracle import oracle as O
from tests import inputs as I

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


# ------------------------------------------------------------------------- Morton


def assert_morton_exact(pos, label):
    say("morton", label, pos.shape)
    g = glx.morton(pos)
    b, lanes, codes = O.reference_morton(pos)
    assert np.array_equal(g.codes, codes), f"{label}: codes differ at {np.nonzero(g.codes != codes)[0][:5]}"
    assert np.array_equal(g.lanes, lanes), f"{label}: lanes differ"
    assert np.array_equal(g.bmin, b["bmin"]) and g.size == b["size"] and g.scale == b["scale"]


@pytest.mark.parametrize("seed", [1, 7, 42, 1000])
def test_morton_bit_exact_clouds(seed):
    assert_morton_exact(I.cloud(seed, 8000, 3.0), f"cloud{seed}")


def test_morton_bit_exact_large_coordinates():
    # the reference's f32 stage drifts ~6 lanes here (gpu/tests/lbvh_morton.rs:143-165)
    assert_morton_exact(I.cloud(3, 8000, 3.0, (1.0e6, 1.0e6, 1.0e6)), "cloud_1e6")


def test_morton_collinear_single_coincident():
    assert_morton_exact(np.array([[i * 0.1, 2.0, -1.0] for i in range(500)]), "collinear")
    assert_morton_exact(np.array([[2.0, -3.0, 4.0]]), "single")
    assert_morton_exact(np.tile(np.array([[1.0, 1.0, 1.0]]), (64, 1)), "coincident")


def test_morton_bit_exact_plummer_and_large_n():
    pos, _, _ = I.plummer_numpy(16384, seed=11)
    assert_morton_exact(pos, "plummer16k")
    rng = np.random.default_rng(5)
    assert_morton_exact(rng.standard_normal((1 << 20, 3)) * 7.0, "gauss_1M")


def test_morton_is_deterministic():
    pos = I.cloud(0xD00, 6000, 3.0)
    a, b = glx.morton(pos), glx.morton(pos)
    assert np.array_equal(a.codes, b.codes) and np.array_equal(a.lanes, b.lanes)


def test_morton_rejects_empty():
    with pytest.raises(glx.GlxError):
        glx.morton(np.zeros((0, 3)))


# ------------------------------------------------------------------------- sort


def assert_sort_exact(codes, label):
    say("sort", label, codes.size)
    order, sc = glx.sort(codes)
    want = O.reference_sort(codes)
    assert np.array_equal(order, want), f"{label}: order differs"
    assert np.array_equal(sc, codes[want])


@pytest.mark.parametrize("seed", [1, 7, 42, 1000])
def test_sort_bit_exact_morton_codes(seed):
    _, _, codes = O.reference_morton(I.cloud(seed, 8000, 3.0))
    assert_sort_exact(codes, f"morton{seed}")


def test_sort_bit_exact_reference_cases():
    for seed in (2, 99, 12345):
        assert_sort_exact(I.random_codes(seed, 10000), f"random{seed}")
    assert_sort_exact(np.arange(5000, dtype=np.uint32) % 4, "ties4")
    assert_sort_exact((0x00ABCD00 | ((np.arange(4096, dtype=np.uint64) * 37) & 0xFF)).astype(np.uint32), "lowbyte")
    assert_sort_exact(((np.arange(4096, dtype=np.uint32) % 64) << 24).astype(np.uint32), "highbyte")
    assert_sort_exact(np.arange(3000, dtype=np.uint32)[::-1].copy(), "reversed")
    assert_sort_exact(I.random_codes(0xBEEF, 65536), "large65536")
    order, sc = glx.sort(np.arange(3000, dtype=np.uint32))
    assert np.array_equal(order, np.arange(3000))
    order, _ = glx.sort(np.full(2048, 42, np.uint32))
    assert np.array_equal(order, np.arange(2048))
    order, sc = glx.sort(np.array([7], np.uint32))
    assert list(order) == [0] and list(sc) == [7]
    order, sc = glx.sort(np.zeros(0, np.uint32))
    assert order.size == 0


def test_sort_bit_exact_million_and_ragged():
    """Sizes around the sort tile (2048 keys) and above 256 tiles, where the digit scan runs
    over chunks of 32 tiles (2^20: 16 whole chunks; 301 tiles: a last chunk of 13; 257
    tiles: a last chunk of one tile; 256 tiles: scanned whole)."""
    rng = np.random.default_rng(9)
    for n in (1 << 20, 2048 * 300 + 5, 2048 * 256 + 1, 2048 * 256, 2048 * 3 + 17, 2047, 2049, 255):
        codes = rng.integers(0, 1 << 30, n, dtype=np.uint32)
        assert_sort_exact(codes, f"uniform{n}")
    codes = (rng.integers(0, 50, 1 << 18, dtype=np.uint32) << 20).astype(np.uint32)
    assert_sort_exact(codes, "clustered_ties")


@pytest.mark.parametrize("bits", [8, 9, 10])
def test_sort_radix_widths_bit_exact(bits):
    """Every digit width gives the reference (code, index) order: a stable LSD sort's
    result does not depend on the width. 30-bit and 63-bit codes, ragged tiles, ties;
    10-bit digits over 30-bit codes (3 passes) and 9- or 10-bit digits over 63-bit codes
    (7 passes) take an odd pass count, which starts in the second buffer."""
    rng = np.random.default_rng(40 + bits)
    top = 1 << (30 if bits == 10 else 32)
    for n in (1 << 20, 2048 * 3 + 17, 2047, 255, 1):
        codes = rng.integers(0, top, n, dtype=np.uint64).astype(np.uint32)
        say("sort bits", bits, "u32", n)
        order, sc = glx.sort(codes, radix_bits=bits)
        want = O.reference_sort(codes)
        assert np.array_equal(order, want) and np.array_equal(sc, codes[want]), (bits, n)
    codes = (rng.integers(0, 50, 1 << 18, dtype=np.uint32) << 20).astype(np.uint32)
    order, _ = glx.sort(codes, radix_bits=bits)
    assert np.array_equal(order, O.reference_sort(codes)), "clustered ties"
    top63 = 1 << 63
    for n in (1 << 20, 2048 * 5 + 3, 100):
        codes = rng.integers(0, top63, n, dtype=np.uint64)
        codes[: n // 4] = codes[n // 2]  # ties
        say("sort bits", bits, "u64", n)
        order, sc = glx.sort(codes, radix_bits=bits)
        want = O.reference_sort63(codes)
        assert np.array_equal(order, want) and np.array_equal(sc, codes[want]), (bits, n)


def test_sort_is_deterministic():
    _, _, codes = O.reference_morton(I.cloud(0x50, 6000, 3.0))
    a, b = glx.sort(codes), glx.sort(codes)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ------------------------------------------------------------------------- Karras


def assert_topology_exact(sc, label):
    say("karras", label, sc.size)
    l, r, p = glx.karras(sc)
    lw, rw, pw = O.reference_karras(sc)
    assert np.array_equal(l, lw), f"{label}: left differs"
    assert np.array_equal(r, rw), f"{label}: right differs"
    assert np.array_equal(p, pw), f"{label}: parent differs"


def test_karras_bit_exact_reference_cases():
    for seed in (1, 7, 42, 1000):
        _, _, codes = O.reference_morton(I.cloud(seed, 6000, 3.0))
        assert_topology_exact(np.sort(codes), f"morton{seed}")
    assert_topology_exact(np.full(4096, 42, np.uint32), "all_equal")
    assert_topology_exact((np.arange(5000, dtype=np.uint32) // 500).astype(np.uint32), "dup500")
    assert_topology_exact(np.arange(4096, dtype=np.uint32), "chain")
    _, _, codes = O.reference_morton(I.cloud(0xBEEF, 65536, 5.0))
    assert_topology_exact(np.sort(codes), "large65536")
    l, r, p = glx.karras(np.array([9], np.uint32))
    assert l.size == 0 and list(p) == [glx.NO_PARENT]
    _, _, codes = O.reference_morton(np.array([[-1.0, 0, 0], [1.0, 0, 0]]))
    l, r, p = glx.karras(np.sort(codes))
    assert list(p) == [2, 2, glx.NO_PARENT]


def test_karras_bit_exact_million():
    rng = np.random.default_rng(3)
    pos = rng.standard_normal((1 << 20, 3))
    _, _, codes = O.reference_morton(pos)
    assert_topology_exact(np.sort(codes), "gauss_1M")


# ------------------------------------------------------------------------- aggregate


def assert_aggregate(sc, sp, sm, r, label):
    say("aggregate", label, sc.size)
    n = sc.size
    g = glx.aggregate(sc, sp, sm)
    l, rr, _ = O.reference_karras(sc)
    o = g["origin"]
    # the oracle folds the same leaves the device sees: origin-relative, f32-narrowed
    leaves = (sp - o).astype(np.float32).astype(np.float64)
    ms = sm.astype(np.float32).astype(np.float64)
    agg = O.reference_aggregate(l, rr, leaves, ms)
    assert np.array_equal(g["aabb_min"].astype(np.float64), agg["aabb_min"]), f"{label}: aabb_min"
    assert np.array_equal(g["aabb_max"].astype(np.float64), agg["aabb_max"]), f"{label}: aabb_max"
    com_err = np.abs(g["com"].astype(np.float64) - agg["com"]).max()
    mass_err = (np.abs(g["mass"] - agg["mass"]) / np.maximum(agg["mass"], 1.0)).max()
    say(label, "com_err", com_err, "mass_err", mass_err)
    assert com_err <= 1e-3 * r + 1e-5  # gpu/tests/lbvh_tree.rs:237
    assert mass_err <= 1e-4            # gpu/tests/lbvh_tree.rs:263
    assert com_err <= 2e-5 * r + 1e-6  # f32 fold depth bound of this implementation
    assert mass_err <= 2e-6
    # child AABB within parent, root bounds all
    for i in range(n - 1):
        for c in (l[i], rr[i]):
            assert np.all(g["aabb_min"][c] >= g["aabb_min"][n + i])
            assert np.all(g["aabb_max"][c] <= g["aabb_max"][n + i])
    return g


def _sorted_inputs(pos, mass):
    _, _, codes = O.reference_morton(pos)
    order = O.reference_sort(codes)
    return codes[order], pos[order], mass[order]


def test_aggregate_reference_cases():
    for seed in (3, 17, 271):
        pos = I.cloud(seed, 4000, 3.0)
        mass = I.masses(seed ^ 0xABCD, 4000)
        assert_aggregate(*_sorted_inputs(pos, mass), 3.0, f"cloud{seed}")
    pos = np.tile(np.array([[1.5, -2.0, 0.25]]), (512, 1))
    assert_aggregate(*_sorted_inputs(pos, I.masses(99, 512)), 2.0, "coincident")
    n = 2048
    assert_aggregate(np.arange(n, dtype=np.uint32), I.cloud(0xC0FFEE, n, 4.0), I.masses(0xBEE, n), 4.0,
                     "chain")
    # one key per bit (0, 1, 2, 4, ..., 2^29), each 33 times: a caterpillar of 30 levels
    # over balanced tie subtrees, about 36 levels inside one tile
    cat = np.repeat(np.array([0] + [1 << k for k in range(30)], np.uint32), 33)
    m = cat.size
    assert_aggregate(cat, I.cloud(0xCA7, m, 2.0), I.masses(0xCA8, m), 2.0, "caterpillar")
    g = glx.aggregate(np.array([5], np.uint32), np.array([[2.0, -1.0, 0.5]]), np.array([1.25]))
    assert g["mass"][0] == np.float32(1.25)
    assert np.array_equal(g["aabb_min"][0] + g["origin"].astype(np.float32),
                          np.array([2.0, -1.0, 0.5], np.float32))


# ------------------------------------------------------------------------- flatten


def assert_flatten(sc, sp, sm, order, label):
    """Pre-order layout against the oracle's reference_flatten over the same tree (the
    oracle's aggregate taken on the f32 origin-relative leaves, as in assert_aggregate):
    next / body_start / body_count and leaf_bodies exact; geometry to f32 rounding."""
    say("flatten", label, sc.size)
    g = glx.flatten(sc, sp, sm)
    n = sc.size
    if n == 1:
        assert list(g["next"]) == [1] and list(g["body_start"]) == [0]
        assert list(g["body_count"]) == [1]
        return
    l, rr, _ = O.reference_karras(sc)
    leaves = (sp - g["origin"]).astype(np.float32).astype(np.float64)
    agg = O.reference_aggregate(l, rr, leaves, sm.astype(np.float32).astype(np.float64))
    nodes, lb = O.reference_flatten(l, rr, agg, order)
    for k in ("next", "body_start", "body_count"):
        assert np.array_equal(g[k], nodes[k]), f"{label}: {k}"
    leaf_slots = g["body_count"] == 1
    assert np.array_equal(order[g["body_start"][leaf_slots]], lb), f"{label}: leaf_bodies"
    ext = float(np.abs(leaves).max()) + 1e-30
    for k in ("center", "half"):
        err = np.abs(g[k].astype(np.float64) - nodes[k]).max()
        assert err <= 2e-7 * ext, (label, k, err)
    com_err = np.abs(g["com"].astype(np.float64) - nodes["com"]).max()
    assert com_err <= 2e-5 * ext + 1e-6, (label, com_err)
    d_err = np.abs(g["delta"].astype(np.float64) - nodes["delta"]).max()
    assert d_err <= 4e-5 * ext + 2e-6, (label, d_err)
    m_err = (np.abs(g["mass"] - nodes["mass"]) / np.maximum(nodes["mass"], 1.0)).max()
    assert m_err <= 2e-6, (label, m_err)


def _flatten_case(pos, mass, label):
    _, _, codes = O.reference_morton(pos)
    order = O.reference_sort(codes)
    assert_flatten(codes[order], pos[order], mass[order], order.astype(np.uint32), label)


def test_flatten_matches_reference_layout():
    """reference_flatten (solvers/src/lbvh.rs:545-614), which the walks here do not need,
    exported for parity: clouds, coincident bodies (index tie-breaks), the N = 1 and N = 2
    known answers of gpu/tests/lbvh_flatten.rs:355-388, a 64k Plummer sphere and a 2^18
    collision whose tree is deep."""
    for seed in (3, 17):
        _flatten_case(I.cloud(seed, 4000, 3.0), I.masses(seed ^ 0xABCD, 4000), f"cloud{seed}")
    _flatten_case(np.tile(np.array([[1.5, -2.0, 0.25]]), (512, 1)), I.masses(99, 512),
                  "coincident")
    _flatten_case(np.array([[2.0, -1.0, 0.5]]), np.array([1.25]), "single")
    g = glx.flatten(np.array([1, 6], np.uint32), np.array([[-1.0, 0, 0], [1.0, 0, 0]]),
                    np.array([1.0, 2.0]))
    assert list(g["next"]) == [3, 2, 3] and list(g["body_count"]) == [0, 1, 1]
    assert list(g["body_start"]) == [0, 0, 1]
    pos, _, mass = I.plummer_numpy(65536, seed=21)
    _flatten_case(pos, mass, "plummer64k")
    pos, _, mass, _ = ic.collision(1 << 18, 5)
    _flatten_case(pos, mass, "collision256k")


def test_aggregate_deterministic_on_a_deep_cross_tile_tree():
    """A clustered input (2^18 two-galaxy collision, 30-bit keys) gives the top bottom-up
    pass a deep tree of nodes that straddle tiles, where which child arrives last varies
    from run to run: every node's aggregate must not depend on it (bit-identical over
    repeated builds, with and without quadrupoles)."""
    pos, _, mass, _ = ic.collision(1 << 18, 3)
    sc, sp, sm = _sorted_inputs(pos, mass)
    for quad in (False, True):
        say("aggregate determinism", "collision_256k", quad)
        a = glx.aggregate(sc, sp, sm, quad=quad)
        for _ in range(3):
            b = glx.aggregate(sc, sp, sm, quad=quad)
            for k in ("aabb_min", "aabb_max", "com", "mass") + (("quad",) if quad else ()):
                assert np.array_equal(a[k], b[k]), f"{k} not deterministic (quad={quad})"


def test_aggregate_super_tile_edges():
    """Systems above 2^18 bodies, for which the middle bottom-up pass runs over super-tiles
    of 32768 leaves, with a last super-tile of a single leaf and one ending inside a tile:
    AABBs bit-exact against the oracle, the rest to the aggregate tolerances, and the
    quadrupole build bit-identical over repeated builds."""
    for n in (262145, 262144 + 3 * 32768 + 517):
        pos = I.cloud(n, n, 3.0)
        args = _sorted_inputs(pos, I.masses(n ^ 0x55, n))
        assert_aggregate(*args, 3.0, f"cloud{n}")
        a = glx.aggregate(*args, quad=True)
        b = glx.aggregate(*args, quad=True)
        for k in ("aabb_min", "aabb_max", "com", "mass", "quad"):
            assert np.array_equal(a[k], b[k]), f"{k} not deterministic (n={n})"


def test_aggregate_large_and_deterministic():
    rng = np.random.default_rng(12)
    pos = rng.standard_normal((1 << 18, 3)) * 4.0
    mass = rng.uniform(0.5, 1.5, 1 << 18)
    sc, sp, sm = _sorted_inputs(pos, mass)
    say("aggregate", "gauss_256k")
    a = glx.aggregate(sc, sp, sm)
    b = glx.aggregate(sc, sp, sm)
    for k in ("aabb_min", "aabb_max", "com", "mass"):
        assert np.array_equal(a[k], b[k]), f"{k} not deterministic"
    n = sc.size
    l, rr, _ = O.reference_karras(sc)
    leaves = (sp - a["origin"]).astype(np.float32).astype(np.float64)
    agg = O.reference_aggregate(l, rr, leaves, sm.astype(np.float32).astype(np.float64))
    assert np.array_equal(a["aabb_min"].astype(np.float64), agg["aabb_min"])
    assert np.array_equal(a["aabb_max"].astype(np.float64), agg["aabb_max"])
    root = n
    assert abs(a["mass"][root] - agg["mass"][root]) / agg["mass"][root] < 1e-5
