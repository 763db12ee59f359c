"""CPU pin of the pre-order placement the GPU flatten uses (no device work): for every
node u of a Karras tree, slot(u) = 2 lo(u) + (left edges on the root path of u) and
next = slot + 1 (leaf) or slot + 2 (hi - lo + 1) - 1 (internal node) reproduce the
oracle's serial reference_flatten (solvers/src/lbvh.rs:545-614) exactly."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests import inputs as I


def closed_form(left, right, parent, n):
    lo = np.zeros(max(n - 1, 0), np.int64)
    hi = np.zeros(max(n - 1, 0), np.int64)
    stack, post = [n], []
    while stack:  # iterative post-order for the leaf ranges
        u = stack.pop()
        post.append(u)
        if u >= n:
            stack += [int(left[u - n]), int(right[u - n])]
    for u in reversed(post):
        if u >= n:
            a, b = int(left[u - n]), int(right[u - n])
            lo[u - n] = a if a < n else lo[a - n]
            hi[u - n] = b if b < n else hi[b - n]
    slot = np.zeros(2 * n - 1, np.int64)
    for u in range(2 * n - 1):
        first = u if u < n else lo[u - n]
        turns, v, p = 0, u, int(parent[u])
        while p != 0xFFFFFFFF:  # no parent (the root)
            turns += int(left[p - n]) == v
            v, p = p, int(parent[p])
        slot[u] = 2 * first + turns
    nxt = np.zeros(2 * n - 1, np.int64)
    for u in range(2 * n - 1):
        nxt[slot[u]] = slot[u] + 1 if u < n else slot[u] + 2 * (hi[u - n] - lo[u - n] + 1) - 1
    return slot, nxt


@pytest.mark.parametrize("case", ["cloud", "coincident", "pair", "chain"])
def test_closed_form_slots_equal_the_serial_flatten(case):
    if case == "cloud":
        pos = I.cloud(3, 3000, 3.0)
    elif case == "coincident":
        pos = np.tile(np.array([[1.5, -2.0, 0.25]]), (300, 1))
    elif case == "pair":
        pos = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    else:
        pos = np.stack([np.linspace(0.0, 1.0, 700)] * 3, axis=1)
    n = len(pos)
    _, _, codes = O.reference_morton(pos)
    order = O.reference_sort(codes)
    left, right, parent = O.reference_karras(codes[order])
    agg = O.reference_aggregate(left, right, pos[order], np.ones(n))
    nodes, _ = O.reference_flatten(left, right, agg, order.astype(np.uint32))
    slot, nxt = closed_form(left, right, parent, n)
    assert sorted(slot.tolist()) == list(range(2 * n - 1))
    assert np.array_equal(nxt, nodes["next"].astype(np.int64))
    # leaves sit in sorted order along the pre-order
    assert np.array_equal(np.sort(slot[:n]), slot[:n])
