"""Independent numpy restatement of the reference's integer/bbox stages.

TEST INFRASTRUCTURE ONLY. Written separately from oracle/lbvh_oracle.c so that the
two can pin each other: a transcription slip in one shows up as a disagreement.

- morton(): solvers/src/lbvh.rs:105-129 in numpy float64 (IEEE, no FMA).
- stable_order(): solvers/src/lbvh.rs:143-147 as numpy's stable argsort.
- radix_tree_bruteforce(): the binary radix tree over the augmented 64-bit keys
  (code << 32 | sorted position), built top-down by splitting each range at its
  highest differing bit. This is a different construction from Karras' parallel
  range search (solvers/src/lbvh.rs:195-240) but defines the same tree, with Karras
  node numbering: the root is internal 0, a left child internal node is numbered by
  its range's last leaf and a right child internal node by its range's first leaf.
- expand10_naive(): the bit-spreading definition (bit b of v -> bit 3b).
"""
from __future__ import annotations

import numpy as np

NO_PARENT = 0xFFFFFFFF


def expand10_naive(v: int) -> int:
    out = 0
    for b in range(10):
        if (v >> b) & 1:
            out |= 1 << (3 * b)
    return out


def morton(pos: np.ndarray):
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    lo = np.fmin.reduce(pos, axis=0)
    hi = np.fmax.reduce(pos, axis=0)
    center = (lo + hi) * 0.5
    ext = hi - lo
    maxe = max(ext[0], ext[1], ext[2])
    half = np.float64(max(np.float64(0.5) * maxe, np.float64(1e-12))) * (np.float64(1.0) + np.float64(1e-9))
    bmin = center - half
    size = np.float64(2.0) * half
    scale = np.float64(1024.0) / size
    v = (pos - bmin) * scale
    f = np.fmax(np.floor(v), 0.0)
    lanes = np.where(f >= 1023.0, 1023.0, f).astype(np.uint32)
    lanes = np.minimum(lanes, 1023)
    table = np.array([expand10_naive(i) for i in range(1024)], dtype=np.uint32)
    codes = table[lanes[:, 0]] | (table[lanes[:, 1]] << 1) | (table[lanes[:, 2]] << 2)
    return {"bmin": bmin, "size": size, "scale": scale}, lanes, codes.astype(np.uint32)


def morton63(pos: np.ndarray):
    """63-bit variant: 2^21 cells per axis, same cube convention."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    lo = np.fmin.reduce(pos, axis=0)
    hi = np.fmax.reduce(pos, axis=0)
    center = (lo + hi) * 0.5
    ext = hi - lo
    maxe = max(ext[0], ext[1], ext[2])
    half = np.float64(max(np.float64(0.5) * maxe, np.float64(1e-12))) * (np.float64(1.0) + np.float64(1e-9))
    bmin = center - half
    size = np.float64(2.0) * half
    scale = np.float64(2.0 ** 21) / size
    v = (pos - bmin) * scale
    f = np.fmax(np.floor(v), 0.0)
    top = float((1 << 21) - 1)
    lanes = np.where(f >= top, top, f).astype(np.uint64)
    codes = np.zeros(lanes.shape[0], np.uint64)
    for b in range(21):
        for a in range(3):
            codes |= ((lanes[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    return {"bmin": bmin, "size": size, "scale": scale}, lanes.astype(np.uint32), codes


def radix_tree_bruteforce63(sorted_codes):
    """Brute-force radix tree over 96-bit augmented keys (code63 << 32 | position)."""
    sc = [int(c) for c in np.asarray(sorted_codes, dtype=np.uint64)]
    n = len(sc)
    keys = [(c << 32) | k for k, c in enumerate(sc)]
    left = np.zeros(max(n - 1, 0), np.uint32)
    right = np.zeros(max(n - 1, 0), np.uint32)
    parent = np.full(2 * n - 1, NO_PARENT, np.uint32)
    if n == 1:
        return left, right, parent

    def clz96(x):
        return 96 - x.bit_length()

    stack = [(0, n - 1, 0)]
    while stack:
        lo, hi, idx = stack.pop()
        d_node = clz96(keys[lo] ^ keys[hi])
        g = lo
        for k in range(lo, hi):
            if clz96(keys[lo] ^ keys[k]) > d_node:
                g = k
        if g == lo:
            lu = g
        else:
            lu = n + g
            stack.append((lo, g, g))
        if g + 1 == hi:
            ru = g + 1
        else:
            ru = n + g + 1
            stack.append((g + 1, hi, g + 1))
        left[idx] = lu
        right[idx] = ru
        parent[lu] = n + idx
        parent[ru] = n + idx
    return left, right, parent


def stable_order(codes: np.ndarray) -> np.ndarray:
    return np.argsort(np.asarray(codes, dtype=np.uint32), kind="stable").astype(np.uint32)


def _clz64(x: int) -> int:
    return 64 - x.bit_length()


def radix_tree_bruteforce(sorted_codes):
    sc = [int(c) for c in np.asarray(sorted_codes, dtype=np.uint32)]
    n = len(sc)
    keys = [(c << 32) | k for k, c in enumerate(sc)]
    left = np.zeros(max(n - 1, 0), np.uint32)
    right = np.zeros(max(n - 1, 0), np.uint32)
    parent = np.full(2 * n - 1, NO_PARENT, np.uint32)
    if n == 1:
        return left, right, parent
    # (lo, hi, karras index of the node covering [lo, hi])
    stack = [(0, n - 1, 0)]
    while stack:
        lo, hi, idx = stack.pop()
        d_node = _clz64(keys[lo] ^ keys[hi])
        g = lo
        for k in range(lo, hi):
            if _clz64(keys[lo] ^ keys[k]) > d_node:
                g = k
        # children: [lo, g] and [g+1, hi]
        if g == lo:
            lu = g
        else:
            lu = n + g  # a left child internal node is numbered by its last leaf
            stack.append((lo, g, g))
        if g + 1 == hi:
            ru = g + 1
        else:
            ru = n + g + 1  # a right child internal node is numbered by its first leaf
            stack.append((g + 1, hi, g + 1))
        left[idx] = lu
        right[idx] = ru
        parent[lu] = n + idx
        parent[ru] = n + idx
    return left, right, parent


def subtree_leaves(left, right, n):
    """Leaf set (sorted positions) under every unified node, by brute force."""
    out = {}
    for k in range(n):
        out[k] = [k]

    def rec(u):
        if u in out:
            return out[u]
        i = u - n
        out[u] = rec(int(left[i])) + rec(int(right[i]))
        return out[u]

    if n > 1:
        import sys

        old = sys.getrecursionlimit()
        sys.setrecursionlimit(max(old, 4 * n + 100))
        try:
            rec(n)
        finally:
            sys.setrecursionlimit(old)
    return out
