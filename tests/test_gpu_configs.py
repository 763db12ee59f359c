"""GPU tests at the workloads' own sizes (the benchmark's configurations), stopping at the
first failure of each:

  C0  Plummer, N = 16,384 (ic.plummer, truncated at r = 10), eps 0.01, theta 0.5: one force
      evaluation and 10 KDK steps at dt = 1e-3 of the reference build against the oracle
      (core/src/integrator.rs:37-78 with solvers/src/lbvh.rs), and the fast build against
      the FP64 direct sum.
  C1  two-galaxy collision, N = 2^20, eps 0.05, theta 0.6: 30-bit codes, order and tree
      bit-exact with reference_morton / reference_sort / reference_karras; reference-build
      forces against the oracle on seeded targets; fast build against FP64.
  C2  the 512-system Plummer ensemble (N log-uniform in [1024, 65536]) in one pool with the
      benchmark's knobs: one step; sampled jobs bit-identical to the same job alone, their
      trees bit-exact with the oracle's; force error against FP64 on every system.
  C4  the 1:3 collision at N = 2^23, eps 0.02, theta 0.5: a few steps of the precision
      form; error against FP64 on a seeded subset, FP64 diagnostics finite, momentum kept.

The reference's own gates run at scale too (gpu/tests/gpu_lbvh.rs:259-308,
gpu/tests/lbvh_sort.rs:209)."""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic
from galaxy_amd.pool import Pool
from oracle import oracle as O
from tests import inputs as I

pytestmark = pytest.mark.gpu
REF = dict(morton_bits=30, walk=glx.WALK_REFERENCE, multipole=0)
FAST = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2)
ENSEMBLE = dict(morton_bits=30, walk=glx.WALK_GROUP, multipole=2, leaf_size=1, far_theta=0.6,
                mono_ratio=0.9, group_size=64, group_lanes=4, radix_bits=8)


def say(*a):
    print("[case]", *a, flush=True)


def rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


def test_c0_plummer_force_and_ten_kdk_steps():
    pos, vel, mass = ic.plummer(16384, 20261202)
    say("C0 plummer 16384")
    eps, theta, dt = 0.01, 0.5, 1e-3
    got = glx.Lbvh(1.0, eps, theta).accelerations(pos, mass)
    cpu, _ = O.lbvh_accelerations(pos, mass, 1.0, eps, theta, nthreads=8)
    med, p99 = I.rel_err_quantiles(got, cpu)
    say("C0 force vs oracle median", med, "p99", p99)
    assert med < 1e-5 and p99 < 1e-3
    pr, vr, ar = O.kdk_run(pos, vel, mass, 1.0, eps, theta, dt, 10, nthreads=8)
    r = glx.GpuResidentLeapfrog(1.0, eps, theta, **REF)
    r.upload(pos, vel, mass)
    r.step_many(dt, 10)
    s = r.snapshot()
    dp, dv = np.abs(s.pos - pr).max(), np.abs(s.vel - vr).max()
    say("C0 10 KDK steps vs oracle dp", dp, "dv", dv)
    assert dp < 1e-5 and dv < 1e-4
    assert I.rel_err_quantiles(r.snapshot_accel(), ar)[0] < 1e-5
    r.close()
    fast = glx.Lbvh(1.0, eps, theta, **FAST).accelerations(pos, mass)
    d64 = glx.direct_sum_f64(pos, mass, 1.0, eps, np.arange(16384, dtype=np.uint32))
    med, p99 = I.rel_err_quantiles(fast, d64)
    say("C0 fast build vs fp64 median", med, "p99", p99)
    assert med < 1e-3 and p99 < 1e-2


def test_c1_collision_2p20_bit_exact_build_and_forces():
    n = 1 << 20
    pos, vel, mass, _ = ic.collision(n, 20261202)
    eps, theta = 0.05, 0.6
    say("C1 morton/sort/karras 2^20")
    m = glx.morton(pos)
    bw, lw, cw = O.reference_morton(pos)
    assert np.array_equal(m.codes, cw) and np.array_equal(m.lanes, lw)
    order, sc = glx.sort(m.codes)
    ow = O.reference_sort(cw)
    assert np.array_equal(order, ow) and np.array_equal(sc, cw[ow])
    l, r, p = glx.karras(sc)
    lw_, rw_, pw_ = O.reference_karras(cw[ow])
    assert np.array_equal(l, lw_) and np.array_equal(r, rw_) and np.array_equal(p, pw_)
    say("C1 reference-build forces vs oracle")
    tg = np.sort(np.random.default_rng(3).choice(n, 2048, replace=False)).astype(np.uint32)
    got = glx.Lbvh(1.0, eps, theta, **REF).accelerations(pos, mass)
    cpu, _ = O.lbvh_accelerations(pos, mass, 1.0, eps, theta, targets=tg, nthreads=8)
    e = rel_err(got[tg], cpu)
    say("C1 reference build vs oracle median", np.median(e), "p99", np.quantile(e, 0.99))
    assert np.median(e) < 1e-5 and np.quantile(e, 0.99) < 1e-4
    say("C1 fast build vs fp64")
    fast = glx.Lbvh(1.0, eps, theta, **FAST).accelerations(pos, mass)
    tg = np.sort(np.random.default_rng(4).choice(n, 4096, replace=False)).astype(np.uint32)
    e = rel_err(fast[tg], glx.direct_sum_f64(pos, mass, 1.0, eps, tg))
    say("C1 fast build vs fp64 median", np.median(e), "p99", np.quantile(e, 0.99))
    assert np.median(e) < 1e-3 and np.quantile(e, 0.99) < 1e-2


def test_c2_ensemble_512_systems_one_pool_step():
    seed = 20261202
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.floor(np.exp(rng.uniform(np.log(1024), np.log(65536), 512))), 1024,
                    65536).astype(int)
    systems = [ic.plummer(int(n), seed + 1 + k) for k, n in enumerate(sizes)]
    say("C2 ensemble", len(systems), "systems", int(sizes.sum()), "particles")
    need = sum(Pool.job_bytes(m.size) for _, _, m in systems)
    pool = Pool(need + (1 << 20), max_jobs=512, **ENSEMBLE)
    ids = [pool.admit(p, v, m, softening=0.01, theta=0.5, dt=1e-3) for p, v, m in systems]
    assert all(j is not None for j in ids)
    pool.step()
    errs = []
    for k, (j, (_, _, m)) in enumerate(zip(ids, systems)):
        p, _ = pool.snapshot(j)
        a = pool.snapshot_accel(j)
        tg = np.sort(np.random.default_rng(k).choice(m.size, 8, replace=False)).astype(np.uint32)
        errs.append(rel_err(a[tg], glx.direct_sum_f64(p, m, 1.0, 0.01, tg)))
    e = np.concatenate(errs)
    say("C2 force error vs fp64 median", np.median(e), "p99", np.quantile(e, 0.99))
    assert np.median(e) <= 1e-3 and np.quantile(e, 0.99) < 1e-2
    for k in np.random.default_rng(9).choice(512, 16, replace=False):
        pos, vel, mass = systems[k]
        alone = Pool(Pool.job_bytes(mass.size) + (1 << 20), max_jobs=1, **ENSEMBLE)
        ja = alone.admit(pos, vel, mass, softening=0.01, theta=0.5, dt=1e-3)
        alone.step()
        pa, va = alone.snapshot(ja)
        pb, vb = pool.snapshot(ids[k])
        assert np.array_equal(pa, pb) and np.array_equal(va, vb), int(k)
        alone.close()
        # the job's 30-bit codes, order and tree at its current positions are the oracle's
        mo = glx.morton(pb)
        _, _, cw = O.reference_morton(pb)
        order, sc = glx.sort(mo.codes)
        ow = O.reference_sort(cw)
        assert np.array_equal(mo.codes, cw) and np.array_equal(order, ow)
        assert all(np.array_equal(x, y) for x, y in zip(glx.karras(sc), O.reference_karras(cw[ow])))
    pool.close()


def test_c4_collision_2p23_precision_form():
    n = 1 << 23
    pos, vel, mass, _ = ic.collision(n, 20261202, mass_ratio=3.0)
    eps, theta, dt = 0.02, 0.5, 2e-3
    say("C4 collision 2^23, 1:3")
    r = glx.GpuResidentLeapfrog(1.0, eps, theta, precision=glx.PREC_KAHAN, **FAST)
    r.upload(pos, vel, mass)
    tg = np.sort(np.random.default_rng(5).choice(n, 4096, replace=False)).astype(np.uint32)
    e = rel_err(r.snapshot_accel()[tg], glx.direct_sum_f64(pos, mass, 1.0, eps, tg))
    say("C4 force error vs fp64 median", np.median(e), "p99", np.quantile(e, 0.99))
    assert np.median(e) < 1e-3 and np.quantile(e, 0.99) < 1e-2
    d0 = r.diag(with_potential=False)
    r.step_many(dt, 5)
    d1 = r.diag(with_potential=False)
    s = r.snapshot()
    assert np.isfinite(s.pos).all() and np.isfinite(s.vel).all()
    assert all(np.isfinite(v).all() for v in d1.values())
    psc = float(np.sum(mass * np.linalg.norm(vel, axis=1)))
    dp = float(np.linalg.norm(d1["momentum"] - d0["momentum"])) / psc
    say("C4 momentum drift over 5 steps", dp)
    assert dp < 1e-5
    assert abs(d1["mass"] - d0["mass"]) <= 1e-12 * d0["mass"]
    r.close()
