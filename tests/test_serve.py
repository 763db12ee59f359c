"""CPU tests of the serving loop (galaxy_amd/serve.py) over a stand-in pool that keeps
the glx_pool contract on the host: admission by memory and slots, one step advances every
active job, jobs finish after their steps, and after a step over the step budget the most
recently admitted or resumed active job is evicted — paused until resumed
(include/galaxy_lbvh.h, glx_pool_set_step_budget, glx_pool_resume).
The stand-in's step takes `cost` ms per weighted particle, so admission control, eviction
and the reports can be checked without a GPU."""
from __future__ import annotations

import numpy as np
import pytest

from galaxy_amd.serve import Request, Server, THETA_REF


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakePool:
    def __init__(self, memory, clock, cost=1e-3, slots=64, per_job_cost=None, overhead=0.0):
        self.memory, self.clock, self.cost, self.slots = memory, clock, cost, slots
        self.overhead = overhead
        self.per_job_cost = per_job_cost or {}
        self.jobs, self.order, self.budget, self.next_id, self.used = {}, [], 0.0, 1, 0
        self.admitted = []
        self.activation = 0

    def set_step_budget(self, ms):
        self.budget = ms

    def _admit(self, n, theta, steps, tag):
        if len(self.jobs) >= self.slots or self.used + n > self.memory:
            return None
        jid = self.next_id
        self.next_id += 1
        self.used += n
        self.activation += 1
        self.jobs[jid] = dict(n=n, theta=theta, steps=steps, steps_done=0, done=False,
                              evicted=False, tag=tag, act=self.activation)
        self.admitted.append(tag)
        return jid

    def admit(self, pos, vel, mass, *, g, softening, theta, dt, steps):
        return self._admit(mass.size, theta, steps, int(mass[0]))

    def admit_ic(self, spec, *, softening, theta, dt, steps):
        return self._admit(spec["n"], theta, steps, spec["tag"])

    def step(self):
        act = [j for j in self.jobs.values() if not j["done"] and not j["evicted"]]
        ms = sum(self.per_job_cost.get(j["tag"], self.cost) * weight(j["n"], j["theta"])
                 for j in act) + (self.overhead if act else 0.0)
        self.clock.t += ms * 1e-3
        for j in act:
            j["steps_done"] += 1
            j["done"] = j["steps_done"] >= j["steps"]
        if self.budget > 0 and ms > self.budget:
            live = [k for k, j in self.jobs.items() if not j["done"] and not j["evicted"]]
            if live:
                self.jobs[max(live, key=lambda k: self.jobs[k]["act"])]["evicted"] = True
        return ms

    def resume(self, jid):
        j = self.jobs[jid]
        assert j["evicted"]
        self.activation += 1
        j["evicted"], j["act"] = False, self.activation

    def job(self, jid):
        j = self.jobs[jid]
        return {"id": jid, "n": j["n"], "steps_done": j["steps_done"], "steps": j["steps"],
                "done": j["done"], "evicted": j["evicted"]}

    def retire(self, jid, copy=False):
        self.used -= self.jobs.pop(jid)["n"]
        return None


def weight(n, theta):
    return n * (THETA_REF / theta) ** 3 * np.log2(max(n, 2)) / 16.0


def req(tag, n, theta=0.5, steps=5):
    mass = np.full(n, float(tag))  # the stand-in reads the tag from the first mass
    return Request(n, theta, steps, state=(np.zeros((n, 3)), np.zeros((n, 3)), mass), tag=tag)


def test_first_come_first_served_under_the_memory_budget():
    clk = Clock()
    pool = FakePool(memory=1000, clock=clk)
    srv = Server(pool, clock=clk)
    for tag, n in enumerate([600, 300, 500, 100, 200]):
        srv.submit(req(tag, n, steps=3))
    srv.drain()
    # job 2 (500) waits for memory and the jobs behind it wait with it
    assert pool.admitted == [0, 1, 2, 3, 4]
    assert sorted(r.tag for r in srv.results) == [0, 1, 2, 3, 4]
    assert all(r.steps_done == 3 and not r.preempted for r in srv.results)
    assert srv.refused_memory >= 1
    st = srv.stats()
    assert st["jobs_finished"] == 5 and st["jobs_waiting"] == 0 and st["jobs_active"] == 0


def test_admission_keeps_steps_within_the_budget():
    clk = Clock()
    pool = FakePool(memory=10**9, clock=clk, cost=1e-3)
    srv = Server(pool, step_budget_ms=10.0, fill=0.85, clock=clk)
    rng = np.random.default_rng(3)
    for tag in range(60):
        srv.submit(req(tag, int(rng.integers(200, 3000)), theta=float(rng.choice([0.4, 0.5, 0.7])),
                   steps=int(rng.integers(3, 20))))
    srv.drain()
    assert len(srv.results) == 60 and not any(r.preempted for r in srv.results)
    learned = [r for r in srv.records[2:]]  # the first steady step teaches the cost
    assert max(r.device_ms for r in learned) <= 10.0
    assert srv.cost == pytest.approx(1e-3)
    # the admitted load stays near the budget while the queue is long
    busy = [r.device_ms for r in srv.records[5:40]]
    assert np.median(busy) > 0.5 * 10.0


def test_an_evicted_job_is_paused_then_resumed_and_finishes():
    clk = Clock()
    # job 7 costs 20x its estimate: its first step overruns the budget, it is paused and
    # resumes (ahead of later arrivals) once the others have left room
    pool = FakePool(memory=10**9, clock=clk, cost=1e-3, per_job_cost={7: 2e-2})
    srv = Server(pool, step_budget_ms=6.0, clock=clk)
    for tag in range(4):
        srv.submit(req(tag, 1000, steps=8))
    srv.admit_ready()
    for _ in range(3):
        srv.step()
    srv.submit(req(7, 400, steps=8))
    srv.admit_ready()
    assert 7 in [e.req.tag for e in srv.active.values()]
    srv.step()
    assert [e.req.tag for e in srv.paused] == [7]
    srv.submit(req(8, 300, steps=2))  # arrives after 7 was paused: 7 goes first
    srv.drain()
    by_tag = {r.tag: r for r in srv.results}
    assert by_tag[7].preempted >= 1 and by_tag[7].steps_done == 8 and not by_tag[7].refused
    assert all(by_tag[t].steps_done == 8 and not by_tag[t].preempted for t in range(4))
    assert pool.admitted.index(8) > pool.admitted.index(7)
    st = srv.stats()
    assert st["jobs_preempted"] == 1 and st["jobs_finished"] == 6 and st["jobs_paused"] == 0


def test_a_job_over_the_budget_by_itself_is_retired_as_refused():
    clk = Clock()
    pool = FakePool(memory=10**9, clock=clk, cost=1e-3, per_job_cost={3: 1.0})
    srv = Server(pool, step_budget_ms=5.0, clock=clk)
    srv.submit(req(3, 500, steps=10))
    srv.drain()
    r = srv.results[0]
    assert r.refused and r.steps_done == 1 and r.preempted == 1


def test_a_job_over_the_budget_alone_is_refused_once_the_cost_is_known():
    clk = Clock()
    pool = FakePool(memory=10**9, clock=clk, cost=1e-3)
    srv = Server(pool, step_budget_ms=5.0, clock=clk)
    srv.submit(req(0, 1000, steps=4))
    srv.admit_ready()
    srv.step()
    srv.step()  # a steady step: cost learned
    assert srv.cost is not None
    assert not srv.submit(req(1, 20000))  # 20 ms estimated alone
    assert srv.submit(req(2, 2000))
    srv.drain()
    by_tag = {r.tag: r for r in srv.results}
    assert by_tag[1].refused and by_tag[1].admitted is None
    assert not by_tag[2].refused and by_tag[2].steps_done == 5
    assert srv.stats()["jobs_refused"] == 1


def test_run_submits_arrivals_at_their_times_and_reports_the_window():
    clk = Clock()
    pool = FakePool(memory=10**9, clock=clk, cost=1e-3)
    def sleep(dt):
        clk.t += dt

    srv = Server(pool, clock=clk, sleep=sleep)
    arrivals = [(0.0, req(0, 1000, steps=50)), (0.02, req(1, 500, steps=10)),
                (0.5, req(2, 800, steps=10))]
    srv.run(arrivals, seconds=1.0, idle_sleep=1e-3)
    # arrival times relative to the server's start are honoured: job 2 waits for t=0.5
    by_tag = {r.tag: r for r in srv.results}
    assert by_tag[2].submitted >= 0.5
    assert by_tag[1].submitted >= 0.02
    st = srv.stats(since=0.0)
    assert st["steps"] == len(srv.records) and st["particle_forces_per_s"] > 0


def test_cost_model_learns_the_fixed_step_cost():
    clk = Clock()
    pool = FakePool(memory=10**9, clock=clk, cost=1e-3, overhead=2.0)
    srv = Server(pool, step_budget_ms=12.0, fill=0.9, clock=clk)
    rng = np.random.default_rng(8)
    for tag in range(80):
        srv.submit(req(tag, int(rng.integers(100, 4000)), steps=int(rng.integers(2, 15))))
    srv.drain()
    assert srv.overhead == pytest.approx(2.0) and srv.cost == pytest.approx(1e-3)
    assert not any(r.preempted for r in srv.results)
    # a job is refused when overhead + cost x weight alone exceeds the budget
    assert srv.estimate_ms(weight(9000, 0.5)) == pytest.approx(2.0 + 1e-3 * weight(9000, 0.5))
    assert srv.submit(req(90, 5000)) and not srv.submit(req(91, 13000))


def test_an_unphysical_fit_falls_back_to_the_ratio():
    """Steps whose times do not grow with their weight (noise around a level) must not
    produce a near-zero slope that would admit everything."""
    clk = Clock()
    srv = Server(FakePool(memory=10**9, clock=clk), step_budget_ms=25.0, clock=clk)
    rng = np.random.default_rng(1)
    for k in range(64):
        w = float(rng.uniform(1e6, 3e6))
        srv._learn(w, 20.0 + float(rng.normal(0, 0.5)))
    assert srv.overhead == 0.0
    assert srv.cost == pytest.approx(20.0 / 2e6, rel=0.1)


def test_heavier_than_estimated_jobs_do_not_derail_the_model():
    """Every seventh job costs 4x its weight: its overruns are charged to it (it is paused
    and resumed), while the model keeps the typical cost and the steps stay near the
    budget."""
    clk = Clock()
    heavy = {t: 4e-3 for t in range(0, 200, 7)}
    pool = FakePool(memory=10**9, clock=clk, cost=1e-3, per_job_cost=heavy)
    srv = Server(pool, step_budget_ms=10.0, clock=clk)
    rng = np.random.default_rng(4)
    for tag in range(200):
        srv.submit(req(tag, int(rng.integers(200, 2500)), steps=int(rng.integers(3, 30))))
    srv.drain()
    assert len(srv.results) == 200 and not any(r.refused for r in srv.results)
    assert 0.5e-3 < srv.cost < 2e-3
    dev = np.array([r.device_ms for r in srv.records])
    assert np.mean(dev > 10.0) < 0.05 and dev.max() < 20.0
