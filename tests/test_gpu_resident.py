"""Resident KDK stepper on the GPU, porting gpu/tests/gpu_resident.rs's gates.

Positions are FP64 on the device (the reference carries a double-single f32 pair), so
several of the reference's gates tighten: the resident/round-trip comparison and the
force-free drift are bit-exact here.
"""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from oracle import oracle as O
from tests import inputs as I

pytestmark = pytest.mark.gpu
G, EPS, THETA, DT = 1.0, 0.05, 0.5, 1e-3


def say(*a):
    print("[case]", *a, flush=True)


def resident(theta=THETA, precision=glx.PREC_F32):
    return glx.GpuResidentLeapfrog(G, EPS, theta, precision)


def f32(v):
    return np.asarray(v, np.float64).astype(np.float32).astype(np.float64)


def test_resident_matches_roundtrip_bit_for_bit():
    """gpu/tests/gpu_resident.rs:113-140 (there: dp < 1e-12); FP64 positions make the
    upload/snapshot round trip lossless, so here both agree bit for bit."""
    pos, vel, mass = I.cluster(7, 512, with_vel=True)
    r = resident()
    r.upload(pos, vel, mass)
    r.step_many(DT, 20)
    out = r.snapshot()
    rt = resident()
    p, v = pos, vel
    for _ in range(20):
        rt.upload(p, v, mass)
        rt.step(DT)
        s = rt.snapshot()
        p, v = s.pos, s.vel
    assert np.array_equal(out.pos, p) and np.array_equal(out.vel, v)


def test_resident_tracks_host_leapfrog():
    """gpu/tests/gpu_resident.rs:148-185: vs the host f64 KDK (here: the oracle's KDK with
    the same-theta f64 LBVH), dp, dv < 1e-4 over 30 steps."""
    pos, vel, mass = I.cluster(3, 512, with_vel=True)
    pr, vr, _ = O.kdk_run(pos, vel, mass, G, EPS, THETA, DT, 30)
    r = resident()
    r.upload(pos, vel, mass)
    r.step_many(DT, 30)
    s = r.snapshot()
    dp, dv = np.abs(s.pos - pr).max(), np.abs(s.vel - vr).max()
    say("tracks host KDK dp", dp, "dv", dv)
    assert dp < 1e-4 and dv < 1e-4


def test_resident_conserves_momentum_theta0():
    """gpu/tests/gpu_resident.rs:196-222"""
    pos, vel, mass = I.cluster(11, 512, with_vel=True)
    p0 = (vel * mass[:, None]).sum(axis=0)
    r = resident(theta=1e-6)
    r.upload(pos, vel, mass)
    r.step_many(DT, 40)
    s = r.snapshot()
    p1 = (s.vel * mass[:, None]).sum(axis=0)
    scale = (np.linalg.norm(s.vel, axis=1) * mass).sum()
    assert np.linalg.norm(p1 - p0) / scale < 1e-4


def test_resident_energy_bounded():
    """gpu/tests/gpu_resident.rs:233-250"""
    pos, vel, mass = I.cluster(5, 256, with_vel=True)
    e0 = 0.5 * (mass * (vel ** 2).sum(axis=1)).sum() + O.potential_serial(pos, mass, G, EPS)
    r = resident()
    r.upload(pos, vel, mass)
    r.step_many(DT, 200)
    s = r.snapshot()
    e1 = 0.5 * (mass * (s.vel ** 2).sum(axis=1)).sum() + O.potential_serial(s.pos, mass, G, EPS)
    say("energy drift 200 steps", abs((e1 - e0) / e0))
    assert abs((e1 - e0) / e0) < 5e-2


def test_zero_steps_round_trips_input():
    """gpu/tests/gpu_resident.rs:263-283"""
    pos, vel, mass = I.cluster(2, 300, with_vel=True)
    r = resident()
    r.upload(pos, vel, mass)
    s = r.snapshot()
    assert np.array_equal(s.pos, pos)
    assert np.array_equal(s.vel, f32(vel))
    assert r.time() == 0.0 and len(r) == 300


def test_empty_and_single():
    """gpu/tests/gpu_resident.rs:288-316"""
    r0 = resident()
    r0.upload(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    r0.step_many(DT, 5)
    assert len(r0.snapshot()) == 0
    x0 = np.array([[1.0, -2.0, 0.5]])
    v0 = np.array([[0.05, 0.0, -0.02]])
    r1 = resident()
    r1.upload(x0, v0, np.ones(1))
    r1.step_many(DT, 10)
    s = r1.snapshot()
    x = x0.copy()
    for _ in range(10):
        x = x + f32(v0) * DT
    assert np.array_equal(s.pos, x)
    assert np.array_equal(s.vel, f32(v0))


def test_resident_is_deterministic():
    """gpu/tests/gpu_resident.rs:323-349"""
    pos, vel, mass = I.cluster(9, 400, with_vel=True)
    outs = []
    for _ in range(2):
        r = resident()
        r.upload(pos, vel, mass)
        r.step_many(DT, 25)
        outs.append(r.snapshot())
    assert np.array_equal(outs[0].pos, outs[1].pos) and np.array_equal(outs[0].vel, outs[1].vel)


def test_time_advances():
    """gpu/tests/gpu_resident.rs:354-366"""
    pos, vel, mass = I.cluster(1, 64, with_vel=True)
    r = resident()
    r.upload(pos, vel, mass)
    r.step_many(DT, 17)
    assert abs(r.time() - 17 * DT) < 1e-12


def test_fp64_drift_tracks_exact_sum():
    """gpu/tests/gpu_resident.rs:384-411 (double-single there, err < 1e-5). FP64 drift
    here: bit-exact with the sequential f64 sum x += fl32(v) * dt."""
    x0 = np.array([[1024.0, -1024.0, 512.0]])
    v0 = np.array([[1.0, -1.0, 0.5]])
    r = resident()
    r.upload(x0, v0, np.ones(1))
    r.step_many(DT, 10_000)
    s = r.snapshot()
    x = x0[0].copy()
    d = f32(v0)[0]
    for _ in range(10_000):
        x = x + d * DT
    assert np.array_equal(s.pos[0], x)


def test_diagnostics_fp64_match_oracle():
    pos, vel, mass = I.cluster(4, 3000, with_vel=True)
    r = resident()
    r.upload(pos, vel, mass)
    r.step_many(DT, 3)
    s = r.snapshot()
    d = r.diag(with_potential=True)
    want = O.diagnostics(s.pos, s.vel, mass)
    assert abs(d["kinetic"] - want["ke"]) <= 1e-12 * want["ke"]
    assert np.allclose(d["momentum"], want["p"], rtol=0, atol=1e-12)
    assert np.allclose(d["angular_momentum"], want["l"], rtol=0, atol=1e-12)
    assert np.allclose(d["center_of_mass"], want["com"], rtol=0, atol=1e-12)
    u = O.potential_serial(s.pos, mass, G, EPS)
    assert abs(d["potential"] - u) <= 1e-12 * abs(u)
    acc = r.snapshot_accel()
    cpu, _ = O.lbvh_accelerations(s.pos, mass, G, EPS, THETA)
    assert I.rms_rel_err(acc, cpu) < 1e-4


def test_kahan_resident_runs_and_tracks():
    pos, vel, mass = I.cluster(3, 512, with_vel=True)
    pr, vr, _ = O.kdk_run(pos, vel, mass, G, EPS, THETA, DT, 30)
    r = resident(precision=glx.PREC_KAHAN)
    r.upload(pos, vel, mass)
    r.step_many(DT, 30)
    s = r.snapshot()
    assert np.abs(s.pos - pr).max() < 1e-4 and np.abs(s.vel - vr).max() < 1e-4


def test_all_f32_path_drifts_in_f32():
    """The all-f32 form (GLX_PREC_ABS32) is the reference's plain f32 drift: bit-exact with
    the sequential f32 sum x = fl32(x + fl32(v dt)), which loses the increments below a
    coordinate's ulp (gpu/tests/gpu_resident.rs:378-411: ~3.5e-1 after 10^4 steps there),
    while the FP64 drift of the other forms tracks the exact sum."""
    x0 = np.array([[1024.0, -1024.0, 512.0]])
    v0 = np.array([[1.0, -1.0, 0.5]])
    r = resident(precision=glx.PREC_ABS32)
    r.upload(x0, v0, np.ones(1))
    r.step_many(DT, 10_000)
    s = r.snapshot()
    x = x0[0].astype(np.float32)
    d = (v0[0].astype(np.float32) * np.float32(DT)).astype(np.float32)
    for _ in range(10_000):
        x = (x + d).astype(np.float32)
    assert np.array_equal(s.pos[0], x.astype(np.float64))
    exact = x0[0] + f32(v0)[0] * DT * 10_000
    err = np.abs(s.pos[0] - exact).max()
    say("all-f32 drift error after 1e4 steps", err)
    assert err > 1e-2


def test_all_f32_walk_loses_resolution_away_from_the_origin():
    """A cluster centred 2000 units from the coordinate origin: the all-f32 walk (absolute
    f32 coordinates) is farther from the FP64 direct sum than the walk in coordinates
    relative to the root-cube centre, at theta -> 0 where the tree adds no error."""
    pos, vel, mass = I.cluster(4, 2000, with_vel=True)
    pos = pos + np.array([2000.0, -1500.0, 700.0])
    tg = np.arange(0, 2000, 3, dtype=np.uint32)
    d64 = glx.direct_sum_f64(pos, mass, G, EPS, tg)
    errs = {}
    for name, prec in (("rel", glx.PREC_F32), ("abs", glx.PREC_ABS32)):
        r = glx.GpuResidentLeapfrog(G, EPS, 1e-6, prec)
        r.upload(pos, vel, mass)
        errs[name] = I.rel_err_quantiles(r.snapshot_accel()[tg], d64)
        r.close()
    say("theta0 errors, centre-relative vs absolute f32", errs)
    assert errs["rel"][0] < 1e-5
    assert errs["abs"][0] > 10 * errs["rel"][0]
