#!/usr/bin/env python3
"""Benchmark of the LBVH gravity step on one MI355X.

Headline workload (the mixed-size ensemble the served-vs-plain metric is quoted on):
512 independent Plummer spheres (G = M = a = 1, truncated at r = 10), sizes
N_s = floor(exp(U(ln 1024, ln 65536))) from a seed (7,482,934 particles in all),
eps = 0.01, theta = 0.5, dt = 1e-3. Every system is resident in one pool and all are
advanced together by one hipGraph launch per KDK step (segment-aware kernels:
kick-drift + per-system bbox, 63-bit Morton encode, segmented radix sort, gather,
Karras build, bottom-up monopole + quadrupole aggregate, group walk + closing kick).

  value            aggregate particle-forces/s of the pool = sum(N_s) * K / wall(K steps)
  plain            the same systems with the same build, one system at a time through
                   the single-system resident path (glx_sys_*)
  p50/p99 step     per-step device time of the pool step (HIP events on its stream)
  steps/s          pool steps per second, against the step's roofline (algorithmic HBM
                   bytes of every phase at 8 TB/s + the walk's executed flops at the
                   FP32 peak); particle-forces/s is N times steps/s, so the fraction is
                   the same for both
  force error      |a - a_fp64| / |a_fp64| on seeded targets of every system, against an
                   FP64 direct sum at the same positions

Also measured: the two-galaxy collision, N = 2^20, eps = 0.05, theta = 0.6, dt = 5e-3,
through the plain single-system path, in the fast build and in the reference build
(30-bit keys, per-target walk, monopole: codes, order and topology bit-exact with the
reference CPU path).

Roofline: the walk (k_walk_group) dominates; it is FP32-VALU bound. `achieved` counts 20
flops per interaction of the reference algorithm on the same positions (counted by the
reference build); the executed figure counts 20 flops per point-mass term, 56 per
quadrupole term, 230 per far node and 140 per body of a local expansion (terms counted by
the kernel); both over the walk's average device time from HIP events on the pool's
stream.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-forces/sec at 1e-3 median force err, served vs plain; p50/p99 step"
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector with packed FP32; 78.6 without
FP32_NONPACKED_TFLOPS = 78.6
FLOPS_POINT = 20
FLOPS_QUAD = 56
# far field: one node folded into a group's third-order local expansion (about 140 VALU
# operations, FMAs counted twice) and the evaluation of the expansion at one body
FLOPS_M2L = 230
FLOPS_L2P = 140
# Counters of one pool-wide walk launch on this workload, per tuned configuration (leaf size,
# far_theta, mono_ratio, group size, lanes per body, key bits) — the two the tuner lands on,
# a per cent apart: HBM bytes from two rocprofv3 passes of `bench.py --profile --leaf-size 1
# --far-theta F --mono-ratio M --group-size 64 --group-lanes 4 --radix-bits 8
# --morton-bits 30` (--pmc FETCH_SIZE, then WRITE_SIZE), scaled by the same counter on a
# 1 GiB device copy (`bench.py --copy-calibration 1024`): FETCH_SIZE reads half the copied
# bytes (x2), WRITE_SIZE all of them (x1) — the tested node records come from beyond L2;
# and SQ_INSTS_VALU / SQ_BUSY_CU_CYCLES, VALU wave-instructions per CU-cycle: one per
# CU-cycle (four SIMDs each taking four cycles per wave64 instruction) is the issue limit,
# so the walk is VALU-issue bound at whatever clock the GPU runs.
WALK_COUNTERS = {
    (1, 0.6, 0.9, 64, 4, 30): {"traffic": 4.641e9 + 0.684e9, "valu_per_cu_cycle": 1.02},
    (1, 0.65, 0.8, 64, 4, 30): {"traffic": 4.951e9 + 0.679e9, "valu_per_cu_cycle": 1.02},
}
# algorithmic bytes per particle for the HBM-bound phases (per force evaluation)


def phase_bytes(morton_bits, radix_bits, quad, boxes=True):
    """Algorithmic HBM bytes per particle of each HBM-bound phase: kick-drift reads v and a
    (f32) and x (f64) and writes v and x; encode reads x and writes key + payload; each of
    the ceil(bits / radix_bits) LSD sort passes reads the keys for its histogram, then reads
    and writes key + payload; the gather reads the order, x and m (f64) and writes a 16-B
    leaf; Karras reads the codes and writes the 16-B topology record, a leaf parent and a
    skip entry; the bottom-up pass reads the leaf, its parent and the topology and writes
    the 32-B classifier record, the 32-B box record (boxes: only the builds that read it;
    group builds at theta <= 1 with leaf size 1 write it for a few handed-over roots only)
    and the 32-B quadrupole."""
    kb = 4 if morton_bits == 30 else 8
    passes = -(-(30 if morton_bits == 30 else 63) // (radix_bits or 8))
    return {"kick_drift_bbox": 84, "encode": 28 + kb, "sort": (3 * kb + 8) * passes,
            "gather": 52, "karras": 24 + kb,
            "aggregate": 68 + (32 if boxes else 0) + (32 if quad else 0)}


def box_records(mode, theta):
    """Whether the build writes every node's box record (the per-target walk, the group walk
    above theta = 1 or with buckets)."""
    return (mode.get("walk") != 1 or theta > 1.0 or (mode.get("leaf_size") or 1) > 1)


HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def step_roofline(bytes_tab, n, walk_flops, ms_per_step):
    """Steps/s against the step's roofline: every HBM-bound phase moving its algorithmic
    bytes at the HBM peak plus the walk's executed flops at the FP32 peak (the executed
    terms, not the reference algorithm's interactions, which on clustered inputs would put
    the roofline below the measured rate)."""
    hbm_bytes = float(sum(bytes_tab.values())) * n
    t = hbm_bytes / (HBM_PEAK_GBS * 1e9) + walk_flops / (FP32_PEAK_TFLOPS * 1e12)
    sps = 1e3 / ms_per_step
    return {"steps_per_s": sps, "roofline_steps_per_s": 1.0 / t, "frac": sps * t,
            "hbm_bytes": hbm_bytes, "walk_flops_executed": walk_flops,
            "roofline_ms": t * 1e3}


def ensemble_sizes(seed: int, count: int = 512, lo: int = 1024, hi: int = 65536):
    rng = np.random.default_rng(seed)
    n = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    return np.clip(n, lo, hi)


def quantiles(e):
    return {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)),
            "targets": int(np.size(e))}


def rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)


# ------------------------------------------------------------------ CPU baseline


def cpu_ensemble(systems, eps, theta, budget_s, seed=3):
    """The oracle (C restatement of solvers/src/lbvh.rs, f64) on whole systems of the
    ensemble drawn in seeded random order until `budget_s` is spent: full build + walk of
    every target per system; particle-forces/s = sum N / sum t. One thread (the
    reference's accelerations loop is serial), then all cores on a few of the same
    systems."""
    from oracle import oracle as O

    O.lib()
    order = np.random.default_rng(seed).permutation(len(systems))
    n_done, t_done, used = 0, 0.0, []
    for k in order:
        pos, _, mass = systems[k]
        t0 = time.perf_counter()
        O.lbvh_accelerations(pos, mass, 1.0, eps, theta)
        t_done += time.perf_counter() - t0
        n_done += mass.size
        used.append(int(k))
        if t_done > budget_s:
            break
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    n_mt, t_mt = 0, 0.0
    for k in used[: max(1, len(used) // 2)]:
        pos, _, mass = systems[k]
        t0 = time.perf_counter()
        O.lbvh_accelerations(pos, mass, 1.0, eps, theta, nthreads=cores)
        t_mt += time.perf_counter() - t0
        n_mt += mass.size
    return {
        "value": n_done / t_done,
        "unit": "particle-forces/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (C restatement of solvers/src/lbvh.rs, f64, -O3, 1 thread): full "
                   f"build + walk of every target for {len(used)} of the 512 systems drawn in "
                   f"seeded order ({n_done} particles, {t_done:.1f} s)"),
        "all_cores": {"value": n_mt / t_mt, "cores": cores,
                      "sample": f"first {max(1, len(used) // 2)} of the same systems"},
        "cpu_model": cpu_model(),
    }


def cpu_model() -> str:
    """Host CPU model name (/proc/cpuinfo), for the baseline's record."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_collision(pos, mass, eps, theta, gpu_acc, budget_s, seed=7):
    """The oracle on the 2^20 collision: one full build plus the walk of S seeded
    targets (1 thread), extrapolated to one full force evaluation; also the error of the
    reference-build GPU accelerations against it on those targets."""
    from oracle import oracle as O

    n = mass.size
    rng = np.random.default_rng(seed)
    O.lib()
    t0 = time.perf_counter()
    O.lbvh_accelerations(pos, mass, 1.0, eps, theta, targets=np.zeros(0, np.uint32))
    t_build = time.perf_counter() - t0
    probe = rng.choice(n, 256, replace=False).astype(np.uint32)
    t0 = time.perf_counter()
    O.lbvh_accelerations(pos, mass, 1.0, eps, theta, targets=probe)
    per_target = max((time.perf_counter() - t0 - t_build) / probe.size, 1e-7)
    s = int(np.clip((budget_s - 2 * t_build) / per_target, 256, n))
    tg = np.sort(rng.choice(n, s, replace=False)).astype(np.uint32)
    t0 = time.perf_counter()
    acc, cnt = O.lbvh_accelerations(pos, mass, 1.0, eps, theta, targets=tg)
    t_walk = time.perf_counter() - t0 - t_build
    t_full = t_build + n * t_walk / s
    return {
        "value": n / t_full, "unit": "particle-forces/s", "cores": 1, "kind": "port",
        "sample": (f"full build {t_build:.2f}s + walk of {s} seeded targets {t_walk:.2f}s, "
                   f"extrapolated to all {n} targets = {t_full:.1f}s per force evaluation"),
        "interactions_per_particle": cnt / s,
        "reference_build_gpu_vs_oracle": quantiles(rel_err(gpu_acc[tg], acc)),
    }


# ------------------------------------------------------------------ GPU runs


def reference_interactions(glx, Pool, pool, ids, systems, eps, theta, dt):
    """I_ref of the pool's current positions: the interactions (leaf bodies and accepted
    monopole nodes, per target) of the reference algorithm (per-target walk, monopole,
    leaf size 1, 30-bit keys, same theta; solvers/src/lbvh.rs:469-524), counted by the
    reference build on the GPU, which reproduces the reference's interaction lists."""
    need = sum(Pool.job_bytes(m.size) for _, _, m in systems)
    ref = Pool(need + (1 << 20), max_jobs=len(systems), morton_bits=30,
               walk=glx.WALK_REFERENCE, multipole=0)
    for j, (_, _, m) in zip(ids, systems):
        pos, vel = pool.snapshot(j)
        ref.admit(pos, vel, m, softening=eps, theta=theta, dt=dt)
    c = ref.count()
    ref.close()
    return int(c["point_terms"])


def batched(glx, Pool, systems, eps, theta, dt, steps, warmup, mode, err_per_system,
            with_ref=True):
    n_tot = int(sum(m.size for _, _, m in systems))
    need = sum(Pool.job_bytes(m.size) for _, _, m in systems)
    pool = Pool(need + (1 << 20), max_jobs=len(systems), **mode)
    t0 = time.perf_counter()
    ids = [pool.admit(p, v, m, softening=eps, theta=theta, dt=dt) for p, v, m in systems]
    t_admit = time.perf_counter() - t0
    assert all(j is not None for j in ids)
    for _ in range(max(1, warmup)):  # the first step captures the graph
        pool.step()
    cnt = pool.count()
    i_ref = (reference_interactions(glx, Pool, pool, ids, systems, eps, theta, dt)
             if with_ref else 0)
    dev = np.zeros(steps)
    t0 = time.perf_counter()
    for q in range(steps):
        dev[q] = pool.step()
    wall = time.perf_counter() - t0
    # per-phase device times (HIP events between phases on the pool's stream), after the
    # timed steps so that the clocks have settled as they had for them
    timed = [pool.step_timed() for _ in range(5)]
    phases = {k: float(np.mean([t[k] for t in timed])) for k in timed[0]}
    rng = np.random.default_rng(17)
    errs = []
    for j, (_, _, m) in zip(ids, systems):
        pos, _ = pool.snapshot(j)
        acc = pool.snapshot_accel(j)
        tg = np.sort(rng.choice(m.size, min(m.size, err_per_system), replace=False))
        errs.append(rel_err(acc[tg], glx.direct_sum_f64(pos, m, 1.0, eps, tg.astype(np.uint32))))
    walk_ms = phases["walk"]
    bytes_tab = phase_bytes(mode.get("morton_bits"), mode.get("radix_bits"),
                            mode.get("multipole") == 2, box_records(mode, theta))
    flops = (FLOPS_POINT * cnt["point_terms"] + FLOPS_QUAD * cnt["quad_terms"] +
             FLOPS_M2L * cnt["far_nodes"] + (FLOPS_L2P * n_tot if mode.get("far_theta") else 0))
    pool.close()
    return {
        "value": n_tot * steps / wall,
        "ms_per_step": wall * 1e3 / steps,
        "steps": step_roofline(bytes_tab, n_tot, flops, wall * 1e3 / steps),
        "device_ms": {"p50": float(np.percentile(dev, 50)), "p99": float(np.percentile(dev, 99)),
                      "mean": float(dev.mean())},
        "phase_ms_per_step": phases,
        "phase_hbm_gbs": {k: bytes_tab[k] * n_tot / (phases[k] * 1e-3) / 1e9
                          for k in bytes_tab},
        "walk": {"ms": walk_ms, "flops": flops, "tflops": flops / (walk_ms * 1e-3) / 1e12,
                 "i_ref_per_particle": i_ref / n_tot,
                 "flops_ref": FLOPS_POINT * i_ref,
                 "tflops_ref": FLOPS_POINT * i_ref / (walk_ms * 1e-3) / 1e12,
                 "point_terms_per_particle": cnt["point_terms"] / n_tot,
                 "quad_terms_per_particle": cnt["quad_terms"] / n_tot,
                 "node_tests_per_particle": cnt["node_tests"] / n_tot,
                 "far_nodes_per_particle": cnt["far_nodes"] / n_tot,
                 "round_fill": cnt["node_tests"] / max(1, 64 * cnt["rounds"]),
                 "fallbacks": cnt["fallbacks"], "max_stack": cnt["max_stack"]},
        "error": quantiles(np.concatenate(errs)),
        "admit_s": t_admit,
        "particles": n_tot,
    }


def plain(glx, systems, eps, theta, dt, steps, mode):
    n_f, t_f, t_all = 0, 0.0, 0.0
    t0a = time.perf_counter()
    for pos, vel, mass in systems:
        r = glx.GpuResidentLeapfrog(1.0, eps, theta, **mode)
        r.upload(pos, vel, mass)
        r.step_many(dt, 1)  # captures the graph
        t0 = time.perf_counter()
        r.step_many(dt, steps)
        t_f += time.perf_counter() - t0
        n_f += mass.size * steps
        r.close()
    t_all = time.perf_counter() - t0a
    return {"value": n_f / t_f, "ms_per_ensemble_step": t_f * 1e3 / steps,
            "steps_per_system": steps, "wall_incl_upload_s": t_all}


def single(glx, pos, vel, mass, eps, theta, dt, steps, warmup, mode, d64_targets):
    n = mass.size
    sysm = glx.GpuResidentLeapfrog(1.0, eps, theta, **mode)
    sysm.upload(pos, vel, mass)
    if warmup:
        sysm.step_many(dt, warmup)
    snap = sysm.snapshot()
    acc0 = sysm.snapshot_accel()
    c0 = sysm.count()
    t0 = time.perf_counter()
    phases = sysm.step_timed(dt, steps)
    wall = time.perf_counter() - t0
    g0 = time.perf_counter()
    sysm.step_many(dt, steps)
    graph_ms = (time.perf_counter() - g0) * 1e3 / steps
    d64 = glx.direct_sum_f64(snap.pos, mass, 1.0, eps, d64_targets)
    walk_ms = phases["walk"] / steps
    flops = (FLOPS_POINT * c0["point_terms"] + FLOPS_QUAD * c0["quad_terms"] +
             FLOPS_M2L * c0["far_nodes"] + (FLOPS_L2P * n if mode.get("far_theta") else 0))
    bytes_tab = phase_bytes(mode.get("morton_bits"), mode.get("radix_bits"),
                            mode.get("multipole") == 2, box_records(mode, theta))
    sysm.close()
    return {
        "snap": snap, "acc0": acc0,
        "value": n * steps / wall,
        "ms_per_step": wall * 1e3 / steps,
        "graph_ms_per_step": graph_ms,
        "steps": step_roofline(bytes_tab, n, flops, graph_ms),
        "phase_ms_per_step": {k: v / steps for k, v in phases.items()},
        "phase_hbm_gbs": {k: bytes_tab[k] * n / (phases[k] / steps * 1e-3) / 1e9
                          for k in bytes_tab},
        "walk": {"ms": walk_ms, "tflops": flops / (walk_ms * 1e-3) / 1e12, "flops": flops,
                 "point_terms_per_particle": c0["point_terms"] / n,
                 "quad_terms_per_particle": c0["quad_terms"] / n,
                 "node_tests_per_particle": c0["node_tests"] / n,
                 "fallbacks": c0["fallbacks"], "max_stack": c0["max_stack"]},
        "error": quantiles(rel_err(acc0[d64_targets], d64)),
    }


PRECISION_FORMS = {
    # FP32 walk in node/cube-relative coordinates with Kahan-compensated accumulation, FP64
    # drift and FP64 diagnostics (the precision form)
    "kahan": 1,
    # the same walk with plain f32 accumulation
    "f32": 0,
    # the reference's plain f32 path: absolute f32 coordinates, f32 drift
    "abs32": 2,
}


def _energy_run(glx, pos, vel, mass, eps, theta, dt, steps, prec, every, with_potential,
                potential0=None):
    """`steps` KDK steps of one precision form; FP64 diagnostics every `every` steps (the
    exact FP64 potential too when with_potential; potential0, when given, is the initial
    state's, already known). Returns the initial force array, the energy and momentum
    traces and the step time."""
    r = glx.GpuResidentLeapfrog(1.0, eps, theta, precision=prec, morton_bits=63,
                                walk=glx.WALK_GROUP, multipole=2)
    r.upload(pos, vel, mass)
    acc0 = r.snapshot_accel()
    d0 = r.diag(with_potential=with_potential and potential0 is None)
    if potential0 is not None:
        d0["potential"] = potential0
    trace = [(0, d0["kinetic"], d0["potential"], d0["momentum"])]
    t_steps, done = 0.0, 0
    while done < steps:
        k = min(every, steps - done)
        t0 = time.perf_counter()
        r.step_many(dt, k)  # returns when the steps are done
        t_steps += time.perf_counter() - t0
        done += k
        d = r.diag(with_potential=with_potential)
        trace.append((done, d["kinetic"], d["potential"], d["momentum"]))
    r.close()
    return acc0, trace, t_steps * 1e3 / max(steps, 1)


def _drifts(trace, mass, vel):
    """Energy drift max_t |E(t) - E(0)| / |E(0)| over the trace (E = kinetic + FP64
    potential) and momentum drift max_t |P(t) - P(0)| / sum m |v|."""
    e = np.array([k + u for _, k, u, _ in trace])
    e0 = e[0]
    p = np.array([m for _, _, _, m in trace])
    psc = float(np.sum(mass * np.linalg.norm(vel, axis=1)))
    return (float(np.max(np.abs(e - e0)) / abs(e0)),
            float(np.max(np.linalg.norm(p - p[0], axis=1)) / psc))


def precision(glx, ic, n, seed, steps, sample=65536, eps=0.02, theta=0.5, dt=2e-3,
              drift_n=1 << 18, drift_steps=1000, drift_every=10):
    """Precision stress: the 1:3 two-galaxy collision (eps 0.02, theta 0.5, dt 2e-3).

    At N = n: each form's force error against an FP64 direct sum on `sample` seeded
    bodies, and its energy drift after `steps` KDK steps (exact FP64 potential at both
    ends; an O(N^2) potential at N = 2^23 takes about 20 s, so the precision form and the
    all-f32 path are run; the plain-f32 form's force error is reported too).
    At N = drift_n: the energy drift as the maximum over steps, with the exact FP64
    potential every `drift_every` steps over `drift_steps` steps, for all three forms."""
    pos, vel, mass, _ = ic.collision(n, seed, mass_ratio=3.0)
    tg = np.sort(np.random.default_rng(5).choice(n, min(n, sample), replace=False))
    tg = tg.astype(np.uint32)
    t0 = time.perf_counter()
    d64 = glx.direct_sum_f64(pos, mass, 1.0, eps, tg)
    log(f"precision: fp64 direct sum on {tg.size} targets {time.perf_counter() - t0:.1f}s")
    res = {"n": n, "mass_ratio": 3.0, "eps": eps, "theta": theta, "dt": dt, "steps": steps,
           "sample": int(tg.size), "build": "63-bit Morton, group walk, quadrupole",
           "forms": {"kahan": "f32 walk relative to the root-cube centre, Kahan-compensated "
                              "accumulation, f64 drift, f64 diagnostics",
                     "f32": "f32 walk relative to the root-cube centre, plain accumulation, "
                            "f64 drift",
                     "abs32": "all-f32: absolute f32 coordinates, plain accumulation, f32 "
                              "drift"}}
    u0 = None  # FP64 potential of the initial state (the same for every form)
    for name, prec in PRECISION_FORMS.items():
        energy = name != "f32"
        t0 = time.perf_counter()
        acc0, trace, ms = _energy_run(glx, pos, vel, mass, eps, theta, dt,
                                      steps if energy else 0, prec, max(steps, 1), energy,
                                      potential0=u0 if energy else None)
        row = {"force_error_vs_fp64_direct": quantiles(rel_err(acc0[tg], d64))}
        if energy:
            u0 = trace[0][2]
            de, dp = _drifts(trace, mass, vel)
            row.update(energy_drift_end=de, momentum_drift=dp, ms_per_step=ms,
                       energy=[trace[0][1] + trace[0][2], trace[-1][1] + trace[-1][2]])
        res[name] = row
        log(f"precision {name} ({time.perf_counter() - t0:.0f}s): {row}")
    # the energy drift as a maximum over steps, at a size where the exact potential is cheap
    pos, vel, mass, _ = ic.collision(drift_n, seed + 1, mass_ratio=3.0)
    sub = {"n": drift_n, "steps": drift_steps, "every": drift_every}
    for name, prec in PRECISION_FORMS.items():
        _, trace, ms = _energy_run(glx, pos, vel, mass, eps, theta, dt, drift_steps, prec,
                                   drift_every, True)
        de, dp = _drifts(trace, mass, vel)
        sub[name] = {"energy_drift_max": de, "momentum_drift_max": dp, "ms_per_step": ms}
    res["drift_over_steps"] = sub
    log(f"precision drift over steps: {sub}")
    return res


def ic_timing(ic, n, seed, repeats=3):
    """Host generation (glx_ic_collision, up to 16 threads) against device generation
    (glx_ic_generate_device, which returns once the arrays are written) of the same
    n-body collision; device time is the median of `repeats` calls into buffers allocated
    beforehand, after one warm-up call."""
    t0 = time.perf_counter()
    ic.collision(n, seed)
    host_s = time.perf_counter() - t0
    sp = ic.spec_collision(n, seed)
    bufs = ic.generate_device(sp)
    lib = ic._bind()
    dev = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        rc = lib.glx_ic_generate_device(ic.C.addressof(sp), bufs[0].ptr, bufs[1].ptr, bufs[2].ptr)
        dev.append(time.perf_counter() - t0)
        if rc:
            raise RuntimeError(f"glx_ic_generate_device failed ({rc})")
    for b in bufs:
        b.free()
    d = float(np.median(dev))
    return {"n": n, "host_s": host_s, "host_threads": min(16, os.cpu_count() or 1),
            "device_ms": 1e3 * d, "host_over_device": host_s / d}


# the serving build: 63-bit keys (collision inputs), far-field ratio 0.6 and monopole band 0.8
# (on a 60-s stream: median error 5.7e-4 over all jobs' checked bodies, job-step p50 5.8 ms,
# against 4.8e-4 / 6.8 ms at 0.5 and 7.4e-4 / 4.5 ms at 0.6 with band 0.9)
SERVE_MODE = dict(morton_bits=63, walk=1, multipole=2, far_theta=0.6, mono_ratio=0.8)


def serve(glx, Pool, seconds, rate, seed, warmup_s, step_budget_ms, budget=200 * 10**9,
          device_ic=True, check_targets=64):
    """Serving stream (galaxy_amd/serve.py): two-galaxy collision jobs arrive as a Poisson
    process at `rate` per second (N log-uniform in [16,384, 2,097,152], theta uniform in
    {0.4, 0.5, 0.7}, steps uniform in [50, 500], eps = 0.05, dt = 5e-3); one pool with a
    device-memory budget admits them first come, first served as memory and the step-time
    budget allow, advances every resident job one KDK step per pool step and retires
    finished jobs. Initial conditions are drawn on the device at admission (inside the
    server's clock). Latency and throughput are reported over the window after
    `warmup_s` (the stream starts from an empty pool); each finished job's forces are
    checked against the FP64 direct sum on `check_targets` seeded bodies."""
    from galaxy_amd import ic
    from galaxy_amd.serve import Server, poisson_collisions

    pool = Pool(budget, max_jobs=4096, **SERVE_MODE)
    srv = Server(pool, step_budget_ms=step_budget_ms, direct_sum=glx.direct_sum_f64)
    arrivals = poisson_collisions(rate, seconds, seed, device_ic=device_ic,
                                  check_targets=check_targets,
                                  make_state=lambda n, s: ic.collision(n, s)[:3])
    srv.run(arrivals, seconds, progress=lambda v: log(
        f"serving {v.now():.0f}s: {len(v.results)} done, {len(v.active)} active, "
        f"{len(v.waiting)} waiting, {len(v.records)} steps"))
    st = srv.stats(since=warmup_s)
    st.update(seconds=seconds, warmup_s=warmup_s, rate_per_s=rate, budget_bytes=budget,
              step_budget_ms=step_budget_ms, build=dict(SERVE_MODE),
              initial_conditions="device, at admission" if device_ic else "host, at arrival",
              evictions_total=pool.status()["evictions"])
    pool.close()
    return st


def strip(d):
    return {k: v for k, v in d.items() if k not in ("snap", "acc0")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--collision-n", type=int, default=1 << 20)
    ap.add_argument("--collision-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-collision", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--precision-n", type=int, default=1 << 23,
                    help="N of the 1:3 collision precision stress (0: skip)")
    ap.add_argument("--precision-steps", type=int, default=200,
                    help="KDK steps of the precision stress at that N (the workload names 1000)")
    ap.add_argument("--drift-n", type=int, default=1 << 18,
                    help="N of the energy drift taken as a maximum over steps")
    ap.add_argument("--drift-steps", type=int, default=1000)
    ap.add_argument("--serve-seconds", type=float, default=90.0,
                    help="serving stream length in seconds (0: skip; the workload names 600)")
    ap.add_argument("--serve-warmup", type=float, default=30.0,
                    help="seconds of the stream before the reported window")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="job arrivals per second")
    ap.add_argument("--serve-budget-ms", type=float, default=25.0,
                    help="device-time budget of one serving step (0: none)")
    ap.add_argument("--leaf-size", type=int, default=0,
                    help="walk leaf size of the ensemble (0: autotune on 64 of its systems)")
    ap.add_argument("--far-theta", type=float, default=-1.0,
                    help="far-field ratio of the group walk (0: off; negative: autotune)")
    ap.add_argument("--mono-ratio", type=float, default=-1.0,
                    help="monopole band of the group walk (0: off; negative: autotune)")
    ap.add_argument("--group-size", type=int, default=0,
                    help="bodies per walk group, 8..64 (0: autotune)")
    ap.add_argument("--group-lanes", type=int, default=0,
                    help="most lanes per body in groups of <= 32 bodies, 1..4 (0: autotune)")
    ap.add_argument("--radix-bits", type=int, default=0,
                    help="sort digit width, 8, 9 or 10 (0: autotune)")
    ap.add_argument("--morton-bits", type=int, default=0, choices=(0, 30, 63),
                    help="Morton key width of the ensemble build, 30 or 63 (0: autotune)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving stream")
    ap.add_argument("--serve-host-ic", action="store_true",
                    help="serving jobs drawn on the host and uploaded (default: on the device)")
    ap.add_argument("--precision-only", action="store_true", help="only the precision stress")
    ap.add_argument("--profile", action="store_true",
                    help="pool steps only (for profilers): no plain, collision or CPU legs")
    ap.add_argument("--profile-collision", action="store_true",
                    help="collision fast build only (for profilers)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 3 times (HBM counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; runs on several GPUs are "
                                   "not part of it (independent systems would shard across "
                                   "GPUs with no communication)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration:
        import torch

        x = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4, "copies": 3}), flush=True)
        return
    if args.profile:
        args.no_cpu = args.no_plain = args.no_collision = True
        args.serve_seconds = 0.0
        args.precision_n = 0

    import galaxy_amd as glx
    from galaxy_amd import ic
    from galaxy_amd.pool import Pool

    fast_mode = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2)
    ref_mode = dict(morton_bits=30, walk=glx.WALK_REFERENCE, multipole=0)
    if args.serve_only or args.precision_only:
        out = {}
        if args.serve_only:
            out["serving"] = serve(glx, Pool, args.serve_seconds or 60.0, args.serve_rate,
                                   args.seed, args.serve_warmup, args.serve_budget_ms,
                                   device_ic=not args.serve_host_ic)
        if args.precision_only:
            out["precision_stress"] = precision(glx, ic, args.precision_n or (1 << 23), args.seed,
                                                args.precision_steps, drift_n=args.drift_n,
                                                drift_steps=args.drift_steps)
        out["device"] = glx.device_info()
        print(json.dumps(out), flush=True)
        return
    if args.profile_collision:
        fast_mode["far_theta"] = max(args.far_theta, 0.0)
        fast_mode["mono_ratio"] = max(args.mono_ratio, 0.0)
        fast_mode.update(group_size=args.group_size, group_lanes=args.group_lanes,
                         radix_bits=args.radix_bits)  # 0: the defaults (64, 4, 8)
        n = args.collision_n
        pos, vel, mass, _ = ic.collision(n, args.seed)
        tg = np.arange(min(n, 256), dtype=np.uint32)
        f = single(glx, pos, vel, mass, 0.05, 0.6, 5e-3, args.collision_steps, 1, fast_mode, tg)
        print(json.dumps({"collision_2p20_fast": strip(f)}), flush=True)
        return
    eps, theta, dt = 0.01, 0.5, 1e-3

    sizes = ensemble_sizes(args.seed, args.systems)
    t0 = time.perf_counter()
    systems = [ic.plummer(int(n), args.seed + 1 + k) for k, n in enumerate(sizes)]
    log(f"ensemble: {len(systems)} systems, {int(sizes.sum())} particles, "
        f"ICs {time.perf_counter() - t0:.1f}s")
    tuning = None
    leaf, far, mono = args.leaf_size, max(args.far_theta, 0.0), max(args.mono_ratio, 0.0)
    gsize, glanes, rbits = args.group_size or 64, args.group_lanes or 4, args.radix_bits or 8
    kbits = args.morton_bits or 30
    if not (args.leaf_size and args.far_theta >= 0 and args.mono_ratio >= 0 and
            args.group_size and args.group_lanes and args.radix_bits and args.morton_bits):
        from galaxy_amd.tune import GROUP_SHAPES, tune

        pick = np.random.default_rng(args.seed + 7).choice(len(systems), 64, replace=False)
        shapes = [(g, n) for g, n in GROUP_SHAPES
                  if args.group_size in (0, g) and args.group_lanes in (0, n)] or [(gsize, glanes)]
        t0 = time.perf_counter()
        tuning = tune([systems[k] for k in sorted(pick)], softening=eps, theta=theta, dt=dt,
                      leaf_sizes=(args.leaf_size,) if args.leaf_size else None,
                      far_thetas=(args.far_theta,) if args.far_theta >= 0 else None,
                      mono_ratios=(args.mono_ratio,) if args.mono_ratio >= 0 else None,
                      group_shapes=shapes,
                      radix_widths=(args.radix_bits,) if args.radix_bits else None,
                      key_widths=(args.morton_bits,) if args.morton_bits else None,
                      radix_systems=systems)
        leaf, far, mono = tuning["leaf_size"], tuning["far_theta"], tuning["mono_ratio"]
        gsize, glanes, rbits = tuning["group_size"], tuning["group_lanes"], tuning["radix_bits"]
        kbits = tuning["morton_bits"]
        log(f"tuned in {time.perf_counter() - t0:.1f}s: leaf size {leaf}, far_theta {far}, "
            f"mono_ratio {mono}, {kbits}-bit keys, group {gsize}x{glanes} lanes, radix {rbits} "
            f"bits: {tuning['table']} {tuning['width_table']} {tuning['shape_table']} "
            f"{tuning['radix_table']}")
    ens_mode = dict(fast_mode, morton_bits=kbits, leaf_size=leaf, far_theta=far,
                    mono_ratio=mono, group_size=gsize, group_lanes=glanes, radix_bits=rbits)
    fast_mode.update(leaf_size=leaf, far_theta=far, mono_ratio=mono, group_size=gsize,
                     group_lanes=glanes, radix_bits=rbits)
    knobs = (leaf, far, mono, gsize, glanes, kbits)
    b = batched(glx, Pool, systems, eps, theta, dt, args.steps, args.warmup, ens_mode, 8)
    log(f"batched: {b['value']:.4g} pf/s, {b['ms_per_step']:.2f} ms/step, err {b['error']}")
    out = {
        "metric": METRIC,
        "value": b["value"],
        "unit": "particle-forces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched ensemble of Plummer spheres, one pool, one graph "
                               "launch per KDK step",
                   "systems": len(systems), "particles": int(sizes.sum()),
                   "n_range": [1024, 65536], "theta": theta, "eps": eps, "dt": dt,
                   "seed": args.seed, "build": f"{kbits}-bit Morton, group walk, quadrupole",
                   # 30-bit keys: the codes, sorted order and tree of every system are those of
                   # the reference's reference_morton / reference_sort / reference_karras, bit
                   # for bit (the same kernels as the stage exports the GPU tests check
                   # against the oracle); 63-bit keys have no reference counterpart
                   "tree_bit_exact_with_reference": kbits == 30,
                   "morton_bits": kbits, "leaf_size": leaf, "far_theta": far, "mono_ratio": mono,
                   "group_size": gsize, "group_lanes": glanes, "radix_bits": rbits},
        "steps_per_s": b["steps"]["steps_per_s"],
        # steps/s (and so particle-forces/s) against the whole step's roofline
        "step_roofline": b["steps"],
        "p50_step_ms": b["device_ms"]["p50"],
        "p99_step_ms": b["device_ms"]["p99"],
        "force_error_vs_fp64_direct": b["error"],
        # the metric's accuracy condition, checked on this run's forces
        "median_force_error_within_1e-3": b["error"]["median"] <= 1e-3,
        "roofline": {
            "bound": "fp32-valu",
            "kernel": "k_walk_group",
            # algorithmic flops: 20 per interaction of the reference algorithm on the same
            # positions (I_ref), whatever the kernel executes to reach the same answer
            "achieved": b["walk"]["tflops_ref"],
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": b["walk"]["tflops_ref"] / FP32_PEAK_TFLOPS,
            "frac_vs_non_packed_peak": b["walk"]["tflops_ref"] / FP32_NONPACKED_TFLOPS,
            "i_ref_per_particle": b["walk"]["i_ref_per_particle"],
            # the terms the kernel executes (20 per point term, 56 per quadrupole term,
            # 230 per far node, 140 per body of a local expansion)
            "achieved_executed": b["walk"]["tflops"],
            "frac_executed": b["walk"]["tflops"] / FP32_PEAK_TFLOPS,
            "traffic": WALK_COUNTERS.get(knobs, {}).get("traffic"),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE (x2) + WRITE_SIZE (x1), 1 GiB copy scale,"
                              " on the same knobs (measured for the tuner's two usual picks;"
                              " null for others)",
            "flops_per_launch": b["walk"]["flops_ref"],
            "executed_flops_per_launch": b["walk"]["flops"],
            "walk_ms": b["walk"]["ms"],
            # VALU wave-instructions per CU busy cycle (rocprofv3 SQ_INSTS_VALU /
            # SQ_BUSY_CU_CYCLES, profiles/sq_counters_walk_ensemble512.md): at the issue
            # limit of one per cycle although its counted flops are a fraction of peak
            "valu_per_cu_cycle": WALK_COUNTERS.get(knobs, {}).get("valu_per_cu_cycle"),
        },
        "batched": strip(b),
        "tuning": tuning,
    }
    if not args.no_plain:
        p = plain(glx, systems, eps, theta, dt, args.plain_steps, ens_mode)
        log(f"plain: {p['value']:.4g} pf/s")
        out["plain"] = p
        out["served_vs_plain"] = b["value"] / p["value"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_ensemble(systems, eps, theta, args.cpu_budget)
        log(f"cpu: {out['cpu_baseline']['value']:.4g} pf/s")
    del systems
    if not args.no_collision:
        ce, ct, cdt = 0.05, 0.6, 5e-3
        n = args.collision_n
        pos, vel, mass, _ = ic.collision(n, args.seed)
        tg = np.sort(np.random.default_rng(11).choice(n, min(n, 4096), replace=False))
        tg = tg.astype(np.uint32)
        col_mode = dict(fast_mode)
        if not args.radix_bits:  # the digit width for this system's 63-bit keys
            from galaxy_amd.tune import tune_radix

            cknobs = {k: col_mode.pop(k) for k in ("leaf_size", "far_theta", "mono_ratio",
                                                    "group_size", "group_lanes", "radix_bits")
                      if k in col_mode}
            col_mode["radix_bits"], radix_table = tune_radix(
                [(pos, vel, mass)], cknobs, softening=ce, theta=ct, dt=cdt, mode=col_mode)
            col_mode.update({k: v for k, v in cknobs.items() if k != "radix_bits"})
            log(f"collision radix width {col_mode['radix_bits']} bits: {radix_table}")
        f = single(glx, pos, vel, mass, ce, ct, cdt, args.collision_steps, 2, col_mode, tg)
        log(f"collision fast: {f['value']:.4g} pf/s")
        r = single(glx, pos, vel, mass, ce, ct, cdt, 1, 1, ref_mode, tg)
        log(f"collision reference build: {r['value']:.4g} pf/s")
        # I_ref: the reference build's interactions per particle (one step later than the
        # fast build's count; the positions barely move in one step)
        i_ref = r["walk"]["point_terms_per_particle"]
        f["walk"]["i_ref_per_particle"] = i_ref
        f["walk"]["tflops_ref"] = FLOPS_POINT * i_ref * n / (f["walk"]["ms"] * 1e-3) / 1e12
        out["collision_2p20"] = {
            "workload": "two-galaxy collision, plain single-system resident path",
            "n": n, "theta": ct, "eps": ce, "dt": cdt,
            "build": {k: col_mode.get(k) for k in ("morton_bits", "leaf_size", "far_theta",
                                                   "mono_ratio", "group_size", "group_lanes",
                                                   "radix_bits")},
            "fast": strip(f),
            "reference_build": strip(r),
        }
        if not args.no_cpu:
            out["collision_2p20"]["cpu_baseline"] = cpu_collision(
                r["snap"].pos, mass, ce, ct, r["acc0"], args.cpu_budget)
        out["collision_2p20"]["initial_conditions"] = ic_timing(ic, n, args.seed)
        log(f"collision ICs: {out['collision_2p20']['initial_conditions']}")
    if args.serve_seconds > 0:
        out["serving"] = serve(glx, Pool, args.serve_seconds, args.serve_rate, args.seed,
                               args.serve_warmup, args.serve_budget_ms,
                               device_ic=not args.serve_host_ic)
        log(f"serving: {out['serving']}")
    if args.precision_n:
        out["precision_stress"] = precision(glx, ic, args.precision_n, args.seed,
                                            args.precision_steps, drift_n=args.drift_n,
                                            drift_steps=args.drift_steps)
    out["device"] = glx.device_info()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
