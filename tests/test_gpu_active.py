"""GPU tests of the active-subset and cached-tree walks (ForceSolver::accelerations_active,
rebuild_gravity_cache, gravity_active_cached; core/src/traits.rs:43-91).

* accelerations_active builds the same tree and walks only the groups that hold an
  active body, so the active rows equal the full pass bit for bit, with fewer terms.
* gravity_active_cached on unmoved bodies is the full pass; after a drift it walks the
  kept tree at the new positions: near-field bodies current, accepted nodes stale — the
  error against an FP64 direct sum stays at the fresh walk's level for small drifts.
"""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic
from galaxy_amd.lbvh import GLX_EINVAL, GlxError
from tests import inputs as I

pytestmark = pytest.mark.gpu
MODES = {
    "reference": dict(morton_bits=30, walk=glx.WALK_REFERENCE, multipole=0),
    "fast": dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2),
}


@pytest.mark.parametrize("mode", ["reference", "fast"])
def test_active_rows_equal_the_full_pass(mode):
    pos, _, mass = ic.plummer(20000, 7)
    s = glx.Lbvh(1.0, 0.01, 0.5, count=True, **MODES[mode])
    full = s.accelerations(pos, mass)
    terms_full = s.stats()["interactions"]
    rng = np.random.default_rng(1)
    for k in (1, 37, 2000):
        act = np.sort(rng.choice(20000, k, replace=False))
        a = s.accelerations_active(pos, mass, act)
        assert np.array_equal(a[act], full[act]), k
        assert s.stats()["interactions"] < terms_full
    assert np.array_equal(s.accelerations_active(pos, mass, None), full)


@pytest.mark.parametrize("mode", ["reference", "fast"])
def test_cached_walk_on_unmoved_bodies_is_the_full_pass(mode):
    pos, _, mass = ic.plummer(8000, 8)
    s = glx.Lbvh(1.0, 0.01, 0.5, **MODES[mode])
    full = s.accelerations(pos, mass)
    s.rebuild_gravity_cache(pos, mass)
    act = np.arange(0, 8000, 3)
    assert np.array_equal(s.gravity_active_cached(pos, mass, act)[act], full[act])
    assert np.array_equal(s.gravity_active_cached(pos, mass), full)


def test_cached_walk_after_a_drift_tracks_the_fresh_walk():
    pos, vel, mass = ic.plummer(16384, 9)
    s = glx.Lbvh(1.0, 0.01, 0.5, **MODES["fast"])
    s.rebuild_gravity_cache(pos, mass)
    moved = pos + vel * 2e-3  # a few fine ticks of a block step
    act = np.sort(np.random.default_rng(4).choice(16384, 1024, replace=False))
    cached = s.gravity_active_cached(moved, mass, act)[act]
    fresh = glx.Lbvh(1.0, 0.01, 0.5, **MODES["fast"]).accelerations(moved, mass)[act]
    exact = glx.direct_sum_f64(moved, mass, 1.0, 0.01, act.astype(np.uint32))
    e_cached = I.rel_err_quantiles(cached, exact)
    e_fresh = I.rel_err_quantiles(fresh, exact)
    print("[case] cached", e_cached, "fresh", e_fresh, flush=True)
    assert e_cached[0] < 2.0 * e_fresh[0] + 1e-4 and e_cached[0] < 1e-3
    assert I.rel_err_quantiles(cached, fresh)[0] < 1e-3


def test_errors():
    pos, _, mass = ic.plummer(500, 3)
    s = glx.Lbvh(1.0, 0.01, 0.5)
    with pytest.raises(GlxError) as e:
        s.gravity_active_cached(pos, mass, [0, 1])
    assert e.value.code == GLX_EINVAL  # nothing kept yet
    s.rebuild_gravity_cache(pos, mass)
    with pytest.raises(GlxError) as e:
        s.gravity_active_cached(pos[:400], mass[:400], [0])
    assert e.value.code == GLX_EINVAL  # a different n
    with pytest.raises(GlxError) as e:
        s.accelerations_active(pos, mass, [0, 500])
    assert e.value.code == GLX_EINVAL  # index out of range
    assert np.array_equal(s.accelerations_active(pos[:1], mass[:1], [0]), np.zeros((1, 3)))


@pytest.mark.parametrize("mode", ["reference", "fast"])
def test_potential_energy_leaves_the_cached_tree_alone(mode):
    """potential_energy takes &self in the reference (core/src/traits.rs:95): a call on
    other bodies between rebuild_gravity_cache and gravity_active_cached, with more
    bodies than the solver's storage holds, must not disturb the kept tree."""
    pos, _, mass = ic.plummer(6000, 11)
    big, _, bmass = ic.plummer(30000, 12)
    act = np.arange(0, 6000, 5)
    ref = glx.Lbvh(1.0, 0.01, 0.5, **MODES[mode])
    ref.rebuild_gravity_cache(pos, mass)
    want = ref.gravity_active_cached(pos, mass, act)[act]
    s = glx.Lbvh(1.0, 0.01, 0.5, **MODES[mode])
    s.rebuild_gravity_cache(pos, mass)
    u_big = s.potential_energy(big, bmass)
    assert np.array_equal(s.gravity_active_cached(pos, mass, act)[act], want)
    u_small = s.potential_energy(pos[:3000], mass[:3000])
    assert np.array_equal(s.gravity_active_cached(pos, mass, act)[act], want)
    assert u_big < 0 and u_small < 0
    assert u_big == s.potential_energy(big, bmass)
