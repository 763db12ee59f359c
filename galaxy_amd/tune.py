"""Autotuner for the batched path's knobs on a given workload.

Accuracy knobs of the group walk (each candidate is error-checked):

* leaf size: a node of at most `leaf_size` bodies that fails the opening test is summed
  body by body instead of opened further — larger buckets trade quadrupole terms and node
  tests for cheaper point-mass terms; they never raise the error.
* far_theta: an accepted node well separated from the whole group (node radius + group
  radius <= far_theta x distance to the group centre) enters the group's third-order
  local expansion instead of one term per body — fewer operations, a larger error as
  far_theta grows.
* mono_ratio: an accepted node that also passes the opening test at theta x mono_ratio is
  summed as a monopole alone (14 operations per body instead of 33) — the quadrupole is
  kept for the nearer accepted nodes, where it matters.

Walk block shape (error-checked too: the groups' boxes decide which nodes are accepted):

* group_size: bodies per walk group, one wave per group (the subtrees of at most this
  many bodies) — smaller groups accept nodes nearer to their bodies and sum fewer terms
  per body, but test more nodes per body and leave lanes idle.
* group_lanes: most lanes per body in groups of at most 32 bodies (the lanes of a body
  take every lanes-th list entry and are folded before the store).

Morton key width (30 bits, the reference's codes, or 63 bits, 2^21 cells per axis): with
30-bit keys the codes, the sorted order and the tree are bit-exact with the reference's
reference_morton / reference_sort / reference_karras and the sort moves half the bytes;
63-bit keys resolve clustered inputs whose 30-bit cells hold many bodies (the two-galaxy
collision's disks). Each width is timed by the whole step and error-checked; among the
widths within the error bounds the 30-bit build is kept unless the 63-bit one is faster
by more than `width_margin`.

Sort radix width (8, 9 or 10 bits per LSD pass): fewer passes against more bins per
tile. The sorted order, hence every result, does not depend on it, so it is timed only.

The best combination depends on the particle distribution and theta, so the tuner times
pool steps of a sample of the workload for each candidate (device time of the step graph,
which separates candidates a few per cent apart that wall-clock time of a few steps does
not) and keeps the fastest whose force error on seeded targets (against an FP64 direct
sum) stays within both bounds: the median within `err_bound` x `margin` (the margin
absorbs the difference between the sample and the whole workload) and the 99th
percentile within `p99_bound`. The default
p99 bound, 1e-2, is below the reference's own per-target monopole walk at theta = 0.5
(p99 1.4e-2 on a 16k Plummer sphere). The search runs in stages (accuracy knobs, then key
width, block shape and radix width, each at the knobs chosen before it) rather than over
the full product. `TuneCache` keeps results in a JSON file keyed by the
workload's size bucket and theta.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import lbvh
from .pool import Pool

CANDIDATES = (1, 2, 4, 8, 16)
LEAF_SIZES = (1, 2)
FAR_THETAS = (0.0, 0.55, 0.6, 0.65, 0.7)
MONO_RATIOS = (0.0, 0.7, 0.8, 0.9)
GROUP_SHAPES = ((64, 4), (64, 1), (48, 4), (32, 4))
RADIX_WIDTHS = (8, 9, 10)
KEY_WIDTHS = (30, 63)


def _pool(systems, knobs, mode, softening, theta, dt):
    need = sum(Pool.job_bytes(m.size) for _, _, m in systems)
    pool = Pool(need + (1 << 20), max_jobs=len(systems), **mode, **knobs)
    ids = [pool.admit(p, v, m, softening=softening, theta=theta, dt=dt) for p, v, m in systems]
    return pool, ids


def _pool_rate(systems, leaf, mode, softening, theta, dt, steps, far=0.0, mono=0.0, **shape):
    """Pool of `systems` with the given knobs; particle-forces/s over the median device
    time of `steps` graph-launched steps, and that time (ms)."""
    knobs = dict(leaf_size=leaf, far_theta=far, mono_ratio=mono, **shape)
    pool, ids = _pool(systems, knobs, mode, softening, theta, dt)
    pool.step()  # captures the graph
    ms = float(np.median([pool.step() for _ in range(steps)]))
    n = sum(m.size for _, _, m in systems)
    return pool, ids, n / (ms * 1e-3), ms


def _errors(pool, ids, systems, picks, softening):
    errs = []
    for j, (_, _, m), tg in zip(ids, systems, picks):
        pos, _ = pool.snapshot(j)
        acc = pool.snapshot_accel(j)
        ref = lbvh.direct_sum_f64(pos, m, 1.0, softening, tg.astype(np.uint32))
        errs.append(np.linalg.norm(acc[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1))
    e = np.concatenate(errs)
    return float(np.median(e)), float(np.quantile(e, 0.99))


def _picks(systems, per_system, seed):
    rng = np.random.default_rng(seed)
    return [np.sort(rng.choice(m.size, min(m.size, per_system), replace=False))
            for _, _, m in systems]


def _best(table, err_bound, margin, p99_bound):
    """The fastest row within both error bounds (of all rows if none is)."""
    ok = [r for r in table
          if r["median_err"] <= err_bound * margin and r["p99_err"] <= p99_bound] or table
    return max(ok, key=lambda r: r["rate"])


def tune_leaf_size(systems, *, softening, theta, dt=1e-3, mode=None, steps=3,
                   err_bound=1e-3, p99_bound=1e-2, margin=0.8, per_system=32, candidates=None,
                   seed=5, leaf_sizes=None, far_thetas=None, mono_ratios=None, shape=None):
    """Time `steps` pool steps of `systems` (list of (pos, vel, mass)) for each candidate
    (leaf size, far_theta, mono_ratio); return {"leaf_size", "far_theta", "mono_ratio",
    "table"}, where each table row has the rate (particle-forces/s over the median device
    step time), that time and the median and p99 force errors. `candidates` alone tunes the leaf size with
    the far field and the monopole band off. `shape` (group_size, group_lanes,
    radix_bits as keywords) is passed to every candidate's pool."""
    mode = mode or dict(morton_bits=63, walk=lbvh.WALK_GROUP, multipole=2)
    shape = shape or {}
    if candidates is not None:
        combos = [(leaf, 0.0, 0.0) for leaf in candidates]
    else:
        combos = [(leaf, far, mono) for leaf in (leaf_sizes or LEAF_SIZES)
                  for far in (far_thetas if far_thetas is not None else FAR_THETAS)
                  for mono in (mono_ratios if mono_ratios is not None else MONO_RATIOS)]
    picks = _picks(systems, per_system, seed)
    table = []
    for leaf, far, mono in combos:
        pool, ids, rate, ms = _pool_rate(systems, leaf, mode, softening, theta, dt, steps, far,
                                         mono, **shape)
        med, p99 = _errors(pool, ids, systems, picks, softening)
        pool.close()
        table.append({"leaf_size": leaf, "far_theta": far, "mono_ratio": mono, "rate": rate,
                      "device_ms": ms, "median_err": med, "p99_err": p99})
    best = _best(table, err_bound, margin, p99_bound)
    return {"leaf_size": best["leaf_size"], "far_theta": best["far_theta"],
            "mono_ratio": best["mono_ratio"], "table": table}


def _phase_ms(systems, knobs, mode, softening, theta, dt, steps, phase):
    """Median over `steps` timed pool steps of one phase's device time (ms)."""
    pool, ids = _pool(systems, knobs, mode, softening, theta, dt)
    pool.step()
    ms = float(np.median([pool.step_timed()[phase] for _ in range(steps)]))
    return pool, ids, ms


def tune_radix(systems, knobs, *, softening, theta, dt=1e-3, mode=None, steps=3,
               radix_widths=None):
    """The sort's digit width for `systems` at the given knobs: each width (default 8, 9
    and 10 bits) timed by the sort phase of a pool of the systems. The sorted order does
    not depend on it, so there is no error check; the key width in `mode` sets the number
    of passes (30-bit keys: 4, 4, 3; 63-bit keys: 8, 7, 7). Returns (width, table)."""
    mode = dict(mode or dict(morton_bits=63, walk=lbvh.WALK_GROUP, multipole=2))
    table = []
    for bits in (radix_widths or RADIX_WIDTHS):
        pool, _, ms = _phase_ms(systems, dict(knobs, radix_bits=bits), mode, softening, theta,
                                dt, steps, "sort")
        pool.close()
        table.append({"radix_bits": bits, "sort_ms": ms})
    return min(table, key=lambda row: row["sort_ms"])["radix_bits"], table


def tune(systems, *, softening, theta, dt=1e-3, mode=None, steps=3, err_bound=1e-3,
         p99_bound=1e-2, margin=0.8, per_system=32, seed=5, leaf_sizes=None, far_thetas=None,
         mono_ratios=None, group_shapes=None, radix_widths=None, radix_systems=None,
         key_widths=None, width_margin=0.02, shape_margin=0.01):
    """All knobs, in four stages on the sample `systems`:
    1. leaf size x far_theta x mono_ratio at the default block shape and radix width
       (`tune_leaf_size`), with 63-bit keys;
    2. Morton key width (`key_widths`, default 30 and 63) at the chosen accuracy knobs:
       each width is timed by the median device time of whole pool steps and
       error-checked; the 30-bit (reference-exact) build is kept unless the 63-bit one
       is more than `width_margin` faster;
    3. block shape (group_size, group_lanes) at the chosen knobs and width: each shape is
       timed by its walk phase and error-checked; the first shape listed (the default,
       64 bodies with up to 4 lanes each) is kept unless another within the error bounds
       is more than `shape_margin` faster (shapes a per cent apart on the sample swap
       places from run to run);
    4. radix width at the chosen knobs, key width and shape, timed by the sort phase on
       `radix_systems` (default `systems`; the whole workload where it is at hand: the
       sort's cost per key changes with the number of tiles in flight, and a sample
       fills the GPU less than the workload).
    Returns {"leaf_size", "far_theta", "mono_ratio", "morton_bits", "group_size",
    "group_lanes", "radix_bits", "table", "width_table", "shape_table", "radix_table"}."""
    mode = dict(mode or dict(morton_bits=63, walk=lbvh.WALK_GROUP, multipole=2))
    r = tune_leaf_size(systems, softening=softening, theta=theta, dt=dt, mode=mode,
                       steps=steps, err_bound=err_bound, p99_bound=p99_bound, margin=margin,
                       per_system=per_system, seed=seed, leaf_sizes=leaf_sizes,
                       far_thetas=far_thetas, mono_ratios=mono_ratios)
    knobs = dict(leaf_size=r["leaf_size"], far_theta=r["far_theta"],
                 mono_ratio=r["mono_ratio"])
    n = sum(m.size for _, _, m in systems)
    picks = _picks(systems, per_system, seed)
    width_table = []
    for bits in (key_widths or KEY_WIDTHS):
        pool, ids, rate, ms = _pool_rate(systems, knobs["leaf_size"], dict(mode, morton_bits=bits),
                                         softening, theta, dt, steps, knobs["far_theta"],
                                         knobs["mono_ratio"])
        med, p99 = _errors(pool, ids, systems, picks, softening)
        pool.close()
        width_table.append({"morton_bits": bits, "rate": rate, "device_ms": ms,
                            "median_err": med, "p99_err": p99})
    ok = [w for w in width_table
          if w["median_err"] <= err_bound * margin and w["p99_err"] <= p99_bound] or width_table
    fastest = max(ok, key=lambda w: w["rate"])
    exact = [w for w in ok if w["morton_bits"] == 30]
    width = (exact[0] if exact and exact[0]["rate"] * (1.0 + width_margin) >= fastest["rate"]
             else fastest)
    mode["morton_bits"] = width["morton_bits"]
    shape_table = []
    for gs, lanes in (group_shapes or GROUP_SHAPES):
        k = dict(knobs, group_size=gs, group_lanes=lanes)
        pool, ids, ms = _phase_ms(systems, k, mode, softening, theta, dt, steps, "walk")
        med, p99 = _errors(pool, ids, systems, picks, softening)
        pool.close()
        shape_table.append({"group_size": gs, "group_lanes": lanes, "walk_ms": ms,
                            "rate": n / (ms * 1e-3), "median_err": med, "p99_err": p99})
    ok = [t for t in shape_table
          if t["median_err"] <= err_bound * margin and t["p99_err"] <= p99_bound] or shape_table
    fastest = max(ok, key=lambda t: t["rate"])
    shape = ok[0] if ok[0]["rate"] * (1.0 + shape_margin) >= fastest["rate"] else fastest
    knobs.update(group_size=shape["group_size"], group_lanes=shape["group_lanes"])
    radix_bits, radix_table = tune_radix(radix_systems or systems, knobs, softening=softening,
                                         theta=theta, dt=dt, mode=mode, steps=steps,
                                         radix_widths=radix_widths)
    out = dict(knobs, morton_bits=mode["morton_bits"], radix_bits=radix_bits)
    out.update(table=r["table"], width_table=width_table, shape_table=shape_table,
               radix_table=radix_table)
    return out


KNOBS = ("leaf_size", "far_theta", "mono_ratio", "morton_bits", "group_size", "group_lanes",
         "radix_bits")


class TuneCache:
    """Tuning results in a JSON file, keyed by the workload's size bucket (log2 of the
    total and of the mean bodies per system), theta and build (walk and expansion; the
    key width is one of the tuned knobs). `lookup` returns the knobs
    of a stored entry or None; `store` records `tune`'s result and rewrites the file."""

    def __init__(self, path: str):
        self.path = path
        self.entries = {}
        if os.path.exists(path):
            with open(path) as f:
                self.entries = json.load(f)

    @staticmethod
    def key(sizes, theta: float, mode: dict | None = None) -> str:
        sizes = np.asarray(sizes, dtype=np.float64)
        total, mean = float(sizes.sum()), float(sizes.mean())
        mode = mode or {}
        build = f"w{mode.get('walk', lbvh.WALK_GROUP)}q{mode.get('multipole', 2)}"
        return (f"N2^{int(math.floor(math.log2(max(total, 1.0))))}"
                f"/n2^{int(math.floor(math.log2(max(mean, 1.0))))}/theta{theta:g}/{build}")

    def lookup(self, key: str):
        e = self.entries.get(key)
        # entries written before the key width was tuned used 63-bit keys
        return None if e is None else {k: e.get(k, 63) for k in KNOBS}

    def store(self, key: str, result: dict):
        self.entries[key] = {k: result[k] for k in KNOBS}
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.entries, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)
