#!/usr/bin/env python3
"""Benchmark of the LBVH gravity step on one MI355X.

Workload (the configuration the headline metric is quoted on that fits one GPU):
two-galaxy collision, N = 2^20 f32-representable particles, eps = 0.05, theta = 0.6,
dt = 5e-3, G = 1, through the plain single-system resident path. One step = one KDK
step with a full rebuild: kick-drift+bbox, Morton encode, radix sort, gather, Karras
build, bottom-up aggregate, walk + closing kick.

Two builds of the step are measured on the same input:
  throughput  63-bit Morton keys, group walk, quadrupole moments (the headline `value`)
  reference   30-bit keys, per-target walk, monopole: the reference's algorithm, whose
              codes/sort/topology are bit-exact with the reference CPU path
Each reports its force error against an FP64 direct sum on 4096 seeded targets.

Prints ONE JSON line. `value` = particle-forces/s = N * K / (wall time of K steps), inputs
resident in HBM, bracketed by device synchronisation. The walk's roofline counts
20 flops per point-mass term and 56 per quadrupole term (terms counted by the kernel on
the same input) over the walk's average device time from HIP events on its stream.
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
FLOPS_POINT = 20
FLOPS_QUAD = 56
# algorithmic bytes per particle for the HBM-bound phases (per force evaluation)
PHASE_BYTES_30 = {"kick_drift_bbox": 72, "encode": 32, "sort": 68, "gather": 36, "karras": 20,
                  "aggregate": 168}
PHASE_BYTES_63 = {"kick_drift_bbox": 72, "encode": 36, "sort": 196, "gather": 36, "karras": 24,
                  "aggregate": 264}


def cpu_baseline(pos, mass, eps, theta, gpu_acc, budget_s=12.0, seed=7):
    """Time the oracle (C restatement of the reference's f64 LBVH, 1 thread) on a
    bounded sample: one full build plus the walk of S seeded targets, extrapolated to
    one full force evaluation. Also returns the reference-mode GPU error against it."""
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
    e = np.linalg.norm(gpu_acc[tg] - acc, axis=1) / np.maximum(np.linalg.norm(acc, axis=1), 1e-300)
    return {
        "value": n / t_full,
        "unit": "particle-forces/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (C restatement of solvers/src/lbvh.rs, -O3, 1 thread) on the same "
                   f"positions: full build {t_build:.2f}s + walk of {s} seeded targets "
                   f"{t_walk:.2f}s, extrapolated to all {n} targets = {t_full:.1f}s per force "
                   f"evaluation"),
        "steps_per_s": 1.0 / t_full,
        "interactions_per_particle": cnt / s,
        "reference_mode_gpu_vs_oracle": {"median": float(np.median(e)),
                                         "p99": float(np.quantile(e, 0.99)), "targets": s},
    }


def run_mode(glx, pos, vel, mass, eps, theta, dt, steps, warmup, mode, d64_targets):
    n = mass.size
    sysm = glx.GpuResidentLeapfrog(1.0, eps, theta, **mode)
    sysm.upload(pos, vel, mass)
    if warmup:
        sysm.step_many(dt, warmup)
    snap = sysm.snapshot()
    acc0 = sysm.snapshot_accel()
    c0 = sysm.count()
    sysm.snapshot()  # synchronises the stream
    t0 = time.perf_counter()
    phases = sysm.step_timed(dt, steps)
    wall = time.perf_counter() - t0
    c1 = sysm.count()
    g0 = time.perf_counter()
    sysm.step_many(dt, steps)
    graph_ms = (time.perf_counter() - g0) * 1e3 / steps
    d64 = glx.direct_sum_f64(snap.pos, mass, 1.0, eps, d64_targets)
    e = np.linalg.norm(acc0[d64_targets] - d64, axis=1) / np.linalg.norm(d64, axis=1)
    pt = 0.5 * (c0["point_terms"] + c1["point_terms"])
    qt = 0.5 * (c0["quad_terms"] + c1["quad_terms"])
    walk_ms = phases["walk"] / steps
    flops = FLOPS_POINT * pt + FLOPS_QUAD * qt
    achieved = flops / (walk_ms * 1e-3) / 1e12
    bytes_tab = PHASE_BYTES_63 if mode.get("morton_bits") == 63 else PHASE_BYTES_30
    if mode.get("multipole") != 2:
        bytes_tab = dict(bytes_tab, aggregate=168)
    return {
        "snap": snap, "acc0": acc0,
        "value": n * steps / wall,
        "ms_per_step": wall * 1e3 / steps,
        "graph_ms_per_step": graph_ms,
        "phase_ms_per_step": {k: v / steps for k, v in phases.items()},
        "phase_hbm_gbs": {k: bytes_tab[k] * n / (phases[k] / steps * 1e-3) / 1e9 for k in bytes_tab},
        "walk": {"ms": walk_ms, "tflops": achieved, "flops": flops,
                 "point_terms_per_particle": pt / n, "quad_terms_per_particle": qt / n,
                 "node_tests_per_particle": 0.5 * (c0["node_tests"] + c1["node_tests"]) / n,
                 "fallbacks": c0["fallbacks"] + c1["fallbacks"]},
        "error": {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)),
                  "targets": int(d64_targets.size)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-reference-mode", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--profile", action="store_true", help="short run for profilers")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part "
                                   "of it (independent systems shard across GPUs with no "
                                   "communication; see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    import galaxy_amd as glx
    from galaxy_amd import ic

    eps, theta, dt = 0.05, 0.6, 5e-3
    n = args.n
    pos, vel, mass, info = ic.collision(n, args.seed)
    rng = np.random.default_rng(11)
    tg = np.sort(rng.choice(n, min(n, 4096), replace=False)).astype(np.uint32)
    fast_mode = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2)
    ref_mode = dict(morton_bits=30, walk=glx.WALK_REFERENCE, multipole=0)

    fast = run_mode(glx, pos, vel, mass, eps, theta, dt, args.steps, args.warmup, fast_mode, tg)
    out = {
        "metric": METRIC,
        "value": fast["value"],
        "unit": "particle-forces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fast["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "two-galaxy collision, plain single-system resident path",
                   "n": n, "theta": theta, "eps": eps, "dt": dt, "seed": args.seed,
                   "build": "63-bit Morton, group walk, quadrupole", "leaf_size": 1},
        "roofline": {
            "bound": "fp32-valu",
            "kernel": "k_walk_group",
            "achieved": fast["walk"]["tflops"],
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": fast["walk"]["tflops"] / FP32_PEAK_TFLOPS,
            "frac_vs_non_packed_peak": fast["walk"]["tflops"] / 78.6,
            "traffic": None,
            "flops_per_launch": fast["walk"]["flops"],
            "walk_ms": fast["walk"]["ms"],
        },
        "force_error_vs_fp64_direct": fast["error"],
        "steps_per_s": 1e3 / fast["ms_per_step"],
        "graph_ms_per_step": fast["graph_ms_per_step"],
        "phase_ms_per_step": fast["phase_ms_per_step"],
        "phase_hbm_gbs": fast["phase_hbm_gbs"],
        "walk_stats": fast["walk"],
        "device": glx.device_info(),
    }
    ref = None
    if not args.no_reference_mode and not args.profile:
        ref = run_mode(glx, pos, vel, mass, eps, theta, dt, 1, 1, ref_mode, tg)
        out["reference_mode"] = {
            "build": "30-bit Morton, per-target walk, monopole (bit-exact codes/sort/topology)",
            "value": ref["value"], "ms_per_step": ref["ms_per_step"],
            "phase_ms_per_step": ref["phase_ms_per_step"], "walk_stats": ref["walk"],
            "force_error_vs_fp64_direct": ref["error"],
        }
    if not args.no_cpu and not args.profile:
        snap = (ref or fast)["snap"]
        acc_ref = ref["acc0"] if ref else fast["acc0"]
        out["cpu_baseline"] = cpu_baseline(snap.pos, mass, eps, theta, acc_ref, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
