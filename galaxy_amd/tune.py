"""Autotuner for the group walk's leaf size and far-field ratio on a given workload.

* leaf size: a node of at most `leaf_size` bodies that fails the opening test is summed
  body by body instead of opened further — larger buckets trade quadrupole terms and node
  tests for cheaper point-mass terms; they never raise the error.
* far_theta: an accepted node well separated from the whole group (node radius + group
  radius <= far_theta x distance to the group centre) enters the group's local expansion
  instead of one term per body — fewer operations, a larger error as far_theta grows.

The best pair depends on the particle distribution and theta, so the tuner times pool
steps of a sample of the workload for each candidate pair and keeps the fastest one whose
median force error on seeded targets (against an FP64 direct sum) stays within the bound.

The walk's block shape (one wave per group of at most 64 bodies) and the sort's radix width
(8 bits) are fixed by the kernels' LDS layout and are not tuned here.
"""
from __future__ import annotations

import time

import numpy as np

from . import lbvh
from .pool import Pool

CANDIDATES = (1, 2, 4, 8, 16)
LEAF_SIZES = (1, 2, 4)
FAR_THETAS = (0.0, 0.4, 0.5)


def _pool_rate(systems, leaf, mode, softening, theta, dt, steps, far=0.0):
    need = sum(Pool.job_bytes(m.size) for _, _, m in systems)
    pool = Pool(need + (1 << 20), max_jobs=len(systems), leaf_size=leaf, far_theta=far, **mode)
    ids = [pool.admit(p, v, m, softening=softening, theta=theta, dt=dt) for p, v, m in systems]
    pool.step()  # captures the graph
    t0 = time.perf_counter()
    dev = [pool.step() for _ in range(steps)]
    wall = time.perf_counter() - t0
    n = sum(m.size for _, _, m in systems)
    return pool, ids, n * steps / wall, float(np.median(dev))


def tune_leaf_size(systems, *, softening, theta, dt=1e-3, mode=None, steps=3,
                   err_bound=1e-3, per_system=8, candidates=None, seed=5, leaf_sizes=None,
                   far_thetas=None):
    """Time `steps` pool steps of `systems` (list of (pos, vel, mass)) for each candidate
    (leaf size, far_theta); return {"leaf_size", "far_theta", "table"}, where each table row
    has the rate (particle-forces/s), the median device step time and the force error.
    `candidates` alone tunes the leaf size with the far field off."""
    mode = mode or dict(morton_bits=63, walk=lbvh.WALK_GROUP, multipole=2)
    if candidates is not None:
        pairs = [(leaf, 0.0) for leaf in candidates]
    else:
        pairs = [(leaf, far) for leaf in (leaf_sizes or LEAF_SIZES)
                 for far in (far_thetas or FAR_THETAS)]
    rng = np.random.default_rng(seed)
    picks = [np.sort(rng.choice(m.size, min(m.size, per_system), replace=False))
             for _, _, m in systems]
    table = []
    for leaf, far in pairs:
        pool, ids, rate, ms = _pool_rate(systems, leaf, mode, softening, theta, dt, steps, far)
        errs = []
        for j, (_, _, m), tg in zip(ids, systems, picks):
            pos, _ = pool.snapshot(j)
            acc = pool.snapshot_accel(j)
            ref = lbvh.direct_sum_f64(pos, m, 1.0, softening, tg.astype(np.uint32))
            errs.append(np.linalg.norm(acc[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1))
        pool.close()
        e = np.concatenate(errs)
        table.append({"leaf_size": leaf, "far_theta": far, "rate": rate, "device_ms": ms,
                      "median_err": float(np.median(e)), "p99_err": float(np.quantile(e, 0.99))})
    ok = [r for r in table if r["median_err"] <= err_bound] or table
    best = max(ok, key=lambda r: r["rate"])
    return {"leaf_size": best["leaf_size"], "far_theta": best["far_theta"], "table": table}
