"""GPU tests of the served form: the request loop (galaxy_amd/serve.py) over a real pool
and the pool's step budget (glx_pool_set_step_budget).

A served job's result is the same alone and in the mix (every kernel works per system):
a job admitted mid-run, among jobs that arrive and leave around it, ends bit-identical to
the same job stepped alone. Jobs wait for memory in arrival order; the step budget pauses
the most recently admitted job and leaves the others' trajectories untouched, and a
resumed job continues from its kept state; jobs drawn on the device carry an FP64
force-error check."""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic
from galaxy_amd.pool import Pool
from galaxy_amd.serve import Request, Server

pytestmark = pytest.mark.gpu
GB = 1 << 30
MODE = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2, far_theta=0.5, mono_ratio=0.8)


def say(*a):
    print("[case]", *a, flush=True)


def job(tag, n, theta, steps, seed):
    pos, vel, mass, _ = ic.collision(n, seed)
    return Request(n, theta, steps, softening=0.05, dt=5e-3, state=(pos, vel, mass), tag=tag)


def alone(r: Request):
    p = Pool(GB, **MODE)
    pos, vel, mass = r.state
    j = p.admit(pos, vel, mass, softening=r.softening, theta=r.theta, dt=r.dt, steps=r.steps)
    for _ in range(r.steps):
        p.step()
    out = p.retire(j)
    p.close()
    return out


def test_served_job_equals_the_same_job_alone():
    say("served mix vs alone")
    srv = Server(Pool(GB, **MODE), keep_state=True)
    reqs = [job(0, 6000, 0.5, 12, 1), job(1, 3001, 0.7, 5, 2), job(2, 9000, 0.4, 9, 3),
            job(3, 2048, 0.5, 3, 4), job(4, 5000, 0.5, 7, 5)]
    srv.submit(reqs[0])
    srv.submit(reqs[1])
    srv.admit_ready()
    for _ in range(2):
        srv.step()
    srv.submit(reqs[2])  # arrives while 0 and 1 are on the GPU
    srv.admit_ready()
    for _ in range(3):
        srv.step()
    srv.submit(reqs[3])
    srv.submit(reqs[4])
    srv.drain()
    srv.pool.close()
    got = {r.tag: r for r in srv.results}
    assert sorted(got) == [0, 1, 2, 3, 4]
    for r in reqs:
        res = got[r.tag]
        assert res.steps_done == r.steps and not res.preempted
        pos, vel = alone(r)
        assert np.array_equal(res.state[0], pos) and np.array_equal(res.state[1], vel), r.tag


def test_jobs_wait_for_memory_in_arrival_order():
    sizes = [20000, 12000, 30000, 4000, 8000]
    budget = Pool.job_bytes(20000) + Pool.job_bytes(12000) + (1 << 20)
    pool = Pool(budget, **MODE)
    srv = Server(pool)
    for k, n in enumerate(sizes):
        srv.submit(job(k, n, 0.5, 4, 10 + k))
    admitted = []
    for _ in range(200):
        srv.admit_ready()
        admitted += [e.req.tag for e in srv.active.values() if e.req.tag not in admitted]
        if srv.step() is None and not srv.waiting:
            break
    pool.close()
    assert admitted == [0, 1, 2, 3, 4]
    assert srv.refused_memory >= 1
    assert sorted(r.tag for r in srv.results) == [0, 1, 2, 3, 4]
    assert all(r.steps_done == 4 and not r.preempted for r in srv.results)


def test_step_budget_pauses_the_newest_and_spares_the_rest():
    reqs = [job(k, n, 0.5, 0, 20 + k) for k, n in enumerate([4000, 2500, 6000])]
    a, b = Pool(GB, **MODE), Pool(GB, **MODE)
    ia = [a.admit(*r.state, softening=0.05, theta=0.5, dt=5e-3) for r in reqs]
    ib = [b.admit(*r.state, softening=0.05, theta=0.5, dt=5e-3) for r in reqs]
    a.set_step_budget(1e-6)  # every step overruns: one eviction per step, newest first
    for q in range(3):
        say("budget step", q)
        a.step()
        b.step()
        info = [a.job(j) for j in ia]
        assert [i["evicted"] for i in info] == [False] * (2 - q) + [True] * (q + 1)
    assert a.status()["evictions"] == 3 and a.status()["active"] == 0
    # job k was evicted after 3 - k steps; its state is that of the unbudgeted pool then
    for k in (2, 1, 0):
        assert a.job(ia[k])["steps_done"] == 3 - k
    pa, va = a.snapshot(ia[0])
    pb, vb = b.snapshot(ib[0])
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert a.step() == 0.0  # nothing left to step
    # a resumed job continues from its kept state: job 2 paused after one step, resumed
    # for two more, equals job 2 of the unbudgeted pool after three
    a.set_step_budget(0.0)
    a.resume(ia[2])
    a.step()
    a.step()
    assert a.job(ia[2])["steps_done"] == 3 and not a.job(ia[2])["evicted"]
    pa, va = a.snapshot(ia[2])
    pb, vb = b.snapshot(ib[2])
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    for j in ia:
        a.retire(j, copy=False)
    assert a.status()["bytes_used"] == 0
    a.close()
    b.close()


def test_device_drawn_jobs_with_fp64_error_check():
    srv = Server(Pool(4 * GB, **MODE), direct_sum=glx.direct_sum_f64)
    for k, (n, th) in enumerate([(50000, 0.5), (20000, 0.4), (30000, 0.7)]):
        srv.submit(Request(n, th, 6, softening=0.05, dt=5e-3,
                           spec=ic.spec_collision(n, 100 + k), tag=k, check_targets=256))
    srv.drain()
    srv.pool.close()
    st = srv.stats()
    say("device-drawn jobs", st["force_error_vs_fp64_direct"])
    assert st["jobs_finished"] == 3
    assert st["force_error_vs_fp64_direct"]["targets"] == 3 * 256
    assert st["force_error_vs_fp64_direct"]["median"] < 1e-3
    assert st["step_ms_wall"]["p50"] >= st["step_ms_device"]["p50"]
