"""The group walk skips the disjoint-box test for theta <= 1 (galaxy_amd/csrc/glx_kernels.hip,
k_walk_group): every point of a node's box lies within sqrt(3) s / 2 + delta of its com,
and an accepted com is at least s / theta + delta from the group box, so acceptance
implies disjoint boxes whenever 1 / theta > sqrt(3) / 2. This checks the claim on the f32
arithmetic the kernels use (the acceptance radius of store_node, the group-box distance
of the walk) with group boxes placed at 0.95-1.05 acceptance radii from the com in random
directions and coms at box corners, and shows that the bound matters: above
theta = 2 / sqrt(3) accepted nodes do overlap the group box.
"""
from __future__ import annotations

import numpy as np

f = np.float32


def _cases(rng, n, theta):
    scale = f(rng.choice([1e-3, 1.0, 1e3]))
    c = (rng.normal(size=(n, 3)) * scale).astype(f)
    h = ((rng.random((n, 3)) ** 3) * f(rng.choice([1e-4, 1e-2, 1.0]))).astype(f)
    mn, mx = (c - h).astype(f), (c + h).astype(f)
    com = (mn + rng.random((n, 3)).astype(f) * (mx - mn)).astype(f)
    corner = rng.random(n) < 0.5  # coms at a box corner: the farthest case
    com[corner] = np.where(rng.random((int(corner.sum()), 3)) < 0.5, mn[corner], mx[corner])
    # acceptance radius as the bottom-up pass stores it (f32, no contraction)
    cen = ((mn + mx) * f(0.5)).astype(f)
    half = ((mx - mn) * f(0.5)).astype(f)
    size = (f(2) * half.max(axis=1)).astype(f)
    dd = (com - cen).astype(f)
    delta = np.sqrt((dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]).astype(f)
    rcrit = (size * f(f(1) / f(theta)) + delta).astype(f)
    # group boxes with their nearest point at about rcrit from the com
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u = u.astype(f)
    gh = (rng.random((n, 3)) * size[:, None] * f(rng.choice([0.1, 1.0, 10.0]))).astype(f)
    dist = (rcrit * rng.uniform(0.95, 1.05, n).astype(f)).astype(f)
    gc = (com + u * (dist[:, None] + gh * np.abs(u))).astype(f)
    g0, g1 = (gc - gh).astype(f), (gc + gh).astype(f)
    # the walk's test: squared distance from the group box to the com >= rcrit^2
    gcen = (f(0.5) * (g0 + g1)).astype(f)
    ghalf = (f(0.5) * (g1 - g0)).astype(f)
    e = np.maximum(np.abs((com - gcen).astype(f)) - ghalf, f(0)).astype(f)
    dmin2 = ((e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]).astype(f)
    accept = (dmin2 >= (rcrit * rcrit).astype(f)) & (size > 0)
    apart = ((mx < g0) | (mn > g1)).any(axis=1)
    return accept, apart


def test_acceptance_implies_disjoint_boxes_up_to_theta_one():
    rng = np.random.default_rng(7)
    for theta in (0.3, 0.5, 0.56, 0.6, 0.7, 0.85, 1.0):
        accepted = 0
        for _ in range(4):
            accept, apart = _cases(rng, 100_000, theta)
            assert not (accept & ~apart).any(), theta
            accepted += int(accept.sum())
        assert accepted > 2_000, (theta, accepted)  # the placement does reach acceptance


def test_above_two_over_sqrt3_accepted_nodes_can_overlap():
    rng = np.random.default_rng(8)
    overlap = 0
    for _ in range(8):
        accept, apart = _cases(rng, 100_000, 1.6)
        overlap += int((accept & ~apart).sum())
    assert overlap > 0
