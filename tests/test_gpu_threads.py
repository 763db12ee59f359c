"""Distinct handles used from distinct threads at the same time (include/galaxy_lbvh.h:
"a handle is used by one thread at a time; distinct handles may be used from distinct
threads — each owns its own HIP stream"; the reference shares its device across threads,
gpu/src/context.rs:15-18,49-53). ctypes releases the GIL around each library call, so the
calls overlap. Results equal those of the same calls made one after another, bit for
bit, and error messages stay with the thread that caused them."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic
from galaxy_amd import lbvh as L
from galaxy_amd.pool import Pool

pytestmark = pytest.mark.gpu
FAST = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2, far_theta=0.5)


def run_threads(fns):
    out = [None] * len(fns)
    errs = []

    def wrap(k, f):
        try:
            out[k] = f()
        except Exception as e:  # reported in the main thread
            errs.append(e)

    ts = [threading.Thread(target=wrap, args=(k, f)) for k, f in enumerate(fns)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errs, errs
    return out


def solver_work(pos, mass, mode):
    s = glx.Lbvh(1.0, 0.02, 0.5, **mode)
    res = [s.accelerations(pos, mass) for _ in range(4)]
    s.close()
    return res


def resident_work(pos, vel, mass):
    r = glx.GpuResidentLeapfrog(1.0, 0.05, 0.5, **FAST)
    r.upload(pos, vel, mass)
    r.step_many(1e-3, 20)
    s = r.snapshot()
    r.close()
    return s.pos, s.vel


def pool_work(systems):
    p = Pool(1 << 30, **FAST)
    ids = [p.admit(a, b, c, softening=0.01, theta=0.5, dt=1e-3) for a, b, c in systems]
    for _ in range(6):
        p.step()
    out = [p.snapshot(j) for j in ids]
    p.close()
    return out


def test_solvers_residents_and_pools_in_parallel_threads():
    inputs = [ic.collision(60000, 1)[:3], ic.plummer(40000, 2), ic.collision(30000, 3)[:3]]
    plummers = [[ic.plummer(n, 10 + k) for k, n in enumerate((3000, 7000, 1500))],
                [ic.plummer(n, 20 + k) for k, n in enumerate((5000, 2500))]]
    fns = [
        lambda: solver_work(inputs[0][0], inputs[0][2], FAST),
        lambda: solver_work(inputs[1][0], inputs[1][2], dict(morton_bits=30,
                                                             walk=glx.WALK_REFERENCE)),
        lambda: resident_work(*inputs[2]),
        lambda: resident_work(*inputs[1]),
        lambda: pool_work(plummers[0]),
        lambda: pool_work(plummers[1]),
    ]
    sequential = [f() for f in fns]
    for rep in range(2):
        print("[case] parallel threads, round", rep, flush=True)
        parallel = run_threads(fns)
        for a, b in zip(sequential, parallel):
            for x, y in zip(a, b):
                if isinstance(x, tuple):
                    assert all(np.array_equal(u, v) for u, v in zip(x, y))
                else:
                    assert np.array_equal(x, y)


def test_error_messages_stay_with_their_thread():
    lib = L.lib()
    seen = {}
    barrier = threading.Barrier(2)

    def bad():
        h = L.C.c_void_p()
        p = L._params(1.0, 0.01, -1.0)  # theta <= 0
        rc = lib.glx_solver_create(L.C.byref(p), L.C.byref(h))
        barrier.wait()
        seen["bad"] = (rc, lib.glx_last_error())
        barrier.wait()

    def good():
        h = L.C.c_void_p()
        p = L._params(1.0, 0.01, 0.5)
        barrier.wait()  # after the other thread's failure
        rc = lib.glx_solver_create(L.C.byref(p), L.C.byref(h))
        seen["good"] = (rc, lib.glx_last_error())
        lib.glx_solver_destroy(h)
        barrier.wait()

    run_threads([bad, good])
    assert seen["bad"][0] == L.GLX_EINVAL and b"theta" in seen["bad"][1]
    assert seen["good"] == (L.GLX_OK, b"")
