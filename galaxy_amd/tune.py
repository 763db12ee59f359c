"""Autotuner for the group walk's knobs on a given workload.

* leaf size: a node of at most `leaf_size` bodies that fails the opening test is summed
  body by body instead of opened further — larger buckets trade quadrupole terms and node
  tests for cheaper point-mass terms; they never raise the error.
* far_theta: an accepted node well separated from the whole group (node radius + group
  radius <= far_theta x distance to the group centre) enters the group's local expansion
  instead of one term per body — fewer operations, a larger error as far_theta grows.
* mono_ratio: an accepted node that also passes the opening test at theta x mono_ratio is
  summed as a monopole alone (14 operations per body instead of 33) — the quadrupole is
  kept for the nearer accepted nodes, where it matters.

The best combination depends on the particle distribution and theta, so the tuner times
pool steps of a sample of the workload for each candidate and keeps the fastest whose
force error on seeded targets (against an FP64 direct sum) stays within both bounds: the
median within `err_bound` x `margin` (the margin absorbs the difference between the
sample and the whole workload) and the 99th percentile within `p99_bound`. The default
p99 bound, 1e-2, is below the reference's own per-target monopole walk at theta = 0.5
(p99 1.4e-2 on a 16k Plummer sphere).

The walk's block shape (one wave per group of at most 64 bodies, 2-4 lanes per body in
groups of at most 32) and the sort's radix width (8 bits) are fixed by the kernels' LDS
layout and are not tuned here.
"""
from __future__ import annotations

import time

import numpy as np

from . import lbvh
from .pool import Pool

CANDIDATES = (1, 2, 4, 8, 16)
LEAF_SIZES = (1, 2)
FAR_THETAS = (0.0, 0.5, 0.55, 0.6)
MONO_RATIOS = (0.0, 0.7, 0.8, 0.9)


def _pool_rate(systems, leaf, mode, softening, theta, dt, steps, far=0.0, mono=0.0):
    need = sum(Pool.job_bytes(m.size) for _, _, m in systems)
    pool = Pool(need + (1 << 20), max_jobs=len(systems), leaf_size=leaf, far_theta=far,
                mono_ratio=mono, **mode)
    ids = [pool.admit(p, v, m, softening=softening, theta=theta, dt=dt) for p, v, m in systems]
    pool.step()  # captures the graph
    t0 = time.perf_counter()
    dev = [pool.step() for _ in range(steps)]
    wall = time.perf_counter() - t0
    n = sum(m.size for _, _, m in systems)
    return pool, ids, n * steps / wall, float(np.median(dev))


def tune_leaf_size(systems, *, softening, theta, dt=1e-3, mode=None, steps=3,
                   err_bound=1e-3, p99_bound=1e-2, margin=0.8, per_system=32, candidates=None,
                   seed=5, leaf_sizes=None, far_thetas=None, mono_ratios=None):
    """Time `steps` pool steps of `systems` (list of (pos, vel, mass)) for each candidate
    (leaf size, far_theta, mono_ratio); return {"leaf_size", "far_theta", "mono_ratio",
    "table"}, where each table row has the rate (particle-forces/s), the median device step
    time and the median and p99 force errors. `candidates` alone tunes the leaf size with
    the far field and the monopole band off."""
    mode = mode or dict(morton_bits=63, walk=lbvh.WALK_GROUP, multipole=2)
    if candidates is not None:
        combos = [(leaf, 0.0, 0.0) for leaf in candidates]
    else:
        combos = [(leaf, far, mono) for leaf in (leaf_sizes or LEAF_SIZES)
                  for far in (far_thetas if far_thetas is not None else FAR_THETAS)
                  for mono in (mono_ratios if mono_ratios is not None else MONO_RATIOS)]
    rng = np.random.default_rng(seed)
    picks = [np.sort(rng.choice(m.size, min(m.size, per_system), replace=False))
             for _, _, m in systems]
    table = []
    for leaf, far, mono in combos:
        pool, ids, rate, ms = _pool_rate(systems, leaf, mode, softening, theta, dt, steps, far,
                                         mono)
        errs = []
        for j, (_, _, m), tg in zip(ids, systems, picks):
            pos, _ = pool.snapshot(j)
            acc = pool.snapshot_accel(j)
            ref = lbvh.direct_sum_f64(pos, m, 1.0, softening, tg.astype(np.uint32))
            errs.append(np.linalg.norm(acc[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1))
        pool.close()
        e = np.concatenate(errs)
        table.append({"leaf_size": leaf, "far_theta": far, "mono_ratio": mono, "rate": rate,
                      "device_ms": ms, "median_err": float(np.median(e)),
                      "p99_err": float(np.quantile(e, 0.99))})
    ok = [r for r in table
          if r["median_err"] <= err_bound * margin and r["p99_err"] <= p99_bound] or table
    best = max(ok, key=lambda r: r["rate"])
    return {"leaf_size": best["leaf_size"], "far_theta": best["far_theta"],
            "mono_ratio": best["mono_ratio"], "table": table}
