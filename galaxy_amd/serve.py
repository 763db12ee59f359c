"""Served form: a request loop over a device-resident pool (continuous batching).

Jobs (simulation requests of any size) arrive while others are on the GPU. The server
keeps them in arrival order, admits them into one `Pool` as device memory, job slots
and the step-time budget allow, advances every resident job by one KDK step per pool
step (one graph launch for all of them), and retires each job when its steps are done.
There is no reference counterpart: the reference advances one system per
GpuResidentLeapfrog (gpu/src/gpu_resident.rs:1848) and bounds device work per submit
(MAX_BATCH, gpu/src/gpu_resident.rs:1271-1287); this loop bounds the device time of
each step instead.

Admission order is first come, first served: a job that does not fit waits at the head
of the queue and the jobs behind it wait too, so a large job is never overtaken
indefinitely. A job fits when the pool has a slot and the memory (`Pool.admit` returns
None otherwise) and, with a step budget, when the estimated step time with it stays
within `fill` x budget (before the first steady step has taught the cost, only into an
empty pool). The estimate is a fixed cost per step plus a cost per weighted particle,
fitted to the device time of the recent steps whose active set did not change, plus
the amount by which it underestimated the last step (the current regime): a job of
n particles at opening angle theta weighs n (0.5 / theta)^3 log2(n) / 16 (the interaction
count per particle of a tree walk grows about as theta^-3 log n).
The pool evicts (pauses) the most recently admitted or resumed job after a step that
overran the budget (what the estimate missed); the server charges the overrun to that job (an extra time
its resumption check adds), keeps its state on the device and resumes it, ahead of any new
job, as soon as the estimate with it fits — a paused job is slowed down, not lost. While
the last step took more than `fill` x budget, no job joins. A job whose own estimated step exceeds the budget is refused when
it is submitted, and a job evicted while it ran alone is retired and reported as refused
(it cannot meet the budget by itself).

Step latency is recorded two ways per step: the host wall time of `Pool.step` (which
includes the refresh after the active set changed: a(x0) of newly admitted jobs and the
graph capture) and the device time of the step graph. Percentiles are reported per pool
step and per job step (each pool step counted once for every job it advanced: the
latency the jobs see).
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

THETA_REF = 0.5


@dataclass
class Request:
    """One simulation job: `n` bodies at opening angle `theta`, advanced `steps` KDK
    steps of `dt`. The initial state is either host arrays (`state` = (pos, vel, mass))
    or a device generator spec (`spec`, `galaxy_amd.ic.Spec`, drawn at admission)."""
    n: int
    theta: float
    steps: int
    softening: float = 0.05
    dt: float = 5e-3
    g: float = 1.0
    state: tuple | None = None
    spec: object | None = None
    tag: object = None
    check_targets: int = 0  # FP64 force-error check on this many seeded bodies at the end

    def weight(self) -> float:
        return (float(self.n) * (THETA_REF / float(self.theta)) ** 3 *
                math.log2(max(self.n, 2)) / 16.0)


@dataclass
class Result:
    tag: object
    n: int
    steps_done: int
    preempted: int                    # times the step budget paused it
    refused: bool
    submitted: float
    admitted: float | None
    finished: float
    state: tuple | None = None        # (pos, vel) at the end, if kept
    force_error: np.ndarray | None = None  # |a - a_fp64| / |a_fp64| on the checked bodies


@dataclass
class StepRecord:
    t: float           # seconds since the server started
    wall_ms: float     # host time of Pool.step (refresh included)
    device_ms: float   # device time of the step graph
    jobs: int
    particles: int
    changed: bool      # the active set changed since the previous step
    estimate_ms: float  # the server's estimate of device_ms before the step (nan: none)


@dataclass
class _Entry:
    req: Request
    submitted: float
    admitted: float | None = None
    jid: int | None = None
    preempted: int = 0
    extra_ms: float = 0.0  # step time the model missed for this job (learned from overruns)

    def weight(self) -> float:
        return self.req.weight()


class Server:
    """Request loop over `pool` (a `galaxy_amd.pool.Pool`, or any object with its
    admit / admit_ic / step / job / resume / retire / snapshot / snapshot_accel /
    snapshot_mass methods).

    step_budget_ms: device-time budget of one step (0: none); fill: fraction of it the
    admission estimate may use; keep_state: keep each finished job's final (pos, vel);
    clock, sleep: time source in seconds and idle wait (for tests); direct_sum: the FP64
    direct sum (galaxy_amd.direct_sum_f64) for the force-error check of requests with
    check_targets."""

    def __init__(self, pool, *, step_budget_ms: float = 0.0, fill: float = 0.85,
                 keep_state: bool = False, clock=time.perf_counter, sleep=time.sleep,
                 direct_sum=None):
        self.pool = pool
        self.budget = float(step_budget_ms)
        self.fill = float(fill)
        self.keep_state = keep_state
        self.clock = clock
        self.sleep = sleep
        self.direct_sum = direct_sum
        if self.budget > 0:
            pool.set_step_budget(self.budget)
        self.t0 = clock()
        self.waiting: deque[_Entry] = deque()
        self.paused: list[_Entry] = []  # evicted by the budget, state on the device
        self.active: dict[int, _Entry] = {}
        self.results: list[Result] = []
        self.records: list[StepRecord] = []
        self.cost = None       # device ms per weighted particle (learned)
        self.overhead = 0.0    # device ms of a step independent of its load (learned)
        self._samples: deque = deque(maxlen=128)  # (weight, device ms) of steady steps
        self._last = None  # (weight, device ms) of the last step
        self._changed = True
        self.refused_memory = 0  # admissions the pool turned down for memory or slots

    # ------------------------------------------------------------------ bookkeeping

    def now(self) -> float:
        return self.clock() - self.t0

    def active_weight(self) -> float:
        return sum(e.weight() for e in self.active.values())

    def active_particles(self) -> int:
        return sum(e.req.n for e in self.active.values())

    def model_ms(self, weight: float) -> float | None:
        """The fitted model's device time of a step over jobs of total `weight`."""
        return None if self.cost is None else self.overhead + self.cost * weight

    def estimate_ms(self, weight: float, extra: float = 0.0) -> float | None:
        """Device time of a step over the active jobs plus more of total `weight` (the
        active jobs' included) and learned overruns `extra`: the model, plus the amount by
        which it underestimated the last step (the model is fitted over recent steps of
        any load; the last step is the current regime)."""
        if self.cost is None:
            return None
        est = self.model_ms(weight) + extra
        if self._last is not None:
            lw, ld = self._last
            est += max(ld - self.model_ms(lw), 0.0)
        return est

    def _learn(self, weight: float, ms: float) -> None:
        """Step cost model ms = overhead + cost x weight, fitted by least squares to the
        last 128 steady steps. The fit is kept only when it is physical — a fixed cost no
        larger than the cheapest step, a slope at least a quarter of the plain ratio of
        time to weight — and otherwise (weights too alike, noise) the model is that
        ratio alone."""
        self._samples.append((weight, ms))
        w = np.array([q[0] for q in self._samples])
        t = np.array([q[1] for q in self._samples])
        ratio = float(t.sum() / w.sum())
        if len(w) >= 8 and np.ptp(w) > 0.1 * w.mean():
            c = float(np.cov(w, t, bias=True)[0, 1] / np.var(w))
            a = float(t.mean() - c * w.mean())
            if c >= 0.25 * ratio and 0.0 <= a <= float(t.min()):
                self.cost, self.overhead = c, a
                return
        self.cost, self.overhead = ratio, 0.0

    # ------------------------------------------------------------------ the loop

    def submit(self, req: Request) -> bool:
        """Queue a job (arrival time = now). Returns False, and records a refused result,
        when its own estimated step exceeds the step budget."""
        t = self.now()
        est = self.model_ms(req.weight())  # the job alone
        if self.budget > 0 and est is not None and est > self.budget:
            self.results.append(Result(req.tag, req.n, 0, 0, True, t, None, t))
            return False
        self.waiting.append(_Entry(req, t))
        return True

    def _fits(self, e: _Entry) -> bool:
        if self.budget <= 0 or not self.active:
            return True
        # until a steady step has taught the cost, a job is admitted only into an empty pool
        # while the last step overran, no job joins; a paused job brings its learned
        # overrun along
        if self._last is not None and self._last[1] > self.fill * self.budget:
            return False
        est = self.estimate_ms(self.active_weight() + e.weight(), e.extra_ms)
        return est is not None and est <= self.fill * self.budget

    def admit_ready(self) -> int:
        """Resume paused jobs, then admit waiting jobs, in submission order while they
        fit; returns how many joined."""
        admitted = 0
        while self.paused:
            e = self.paused[0]
            if not self._fits(e):
                return admitted
            self.pool.resume(e.jid)
            self.paused.pop(0)
            self.active[e.jid] = e
            self._changed = True
            admitted += 1
        while self.waiting:
            e = self.waiting[0]
            if not self._fits(e):
                break
            r = e.req
            if r.spec is not None:
                jid = self.pool.admit_ic(r.spec, softening=r.softening, theta=r.theta, dt=r.dt,
                                         steps=r.steps)
            else:
                pos, vel, mass = r.state
                jid = self.pool.admit(pos, vel, mass, g=r.g, softening=r.softening,
                                      theta=r.theta, dt=r.dt, steps=r.steps)
            if jid is None:
                self.refused_memory += 1
                break
            self.waiting.popleft()
            e.jid, e.admitted = jid, self.now()
            self.active[jid] = e
            self._changed = True
            admitted += 1
        return admitted

    def step(self) -> StepRecord | None:
        """One pool step of every resident job; then retire the finished ones and set
        the evicted ones aside to resume. None when nothing is resident."""
        if not self.active:
            return None
        particles, weight = self.active_particles(), self.active_weight()
        est = self.estimate_ms(weight)
        t = self.now()
        w0 = self.clock()
        dev = self.pool.step()
        wall = (self.clock() - w0) * 1e3
        rec = StepRecord(t, wall, dev, len(self.active), particles, self._changed,
                         est if est is not None else float("nan"))
        self.records.append(rec)
        if not self._changed and weight > 0:  # steady steps teach the cost model
            self._learn(weight, dev)
        self._last = (weight, dev)
        self._changed = False
        alone = len(self.active) == 1
        for jid in list(self.active):
            info = self.pool.job(jid)
            if info["done"]:
                self._finish(jid, info)
            elif info["evicted"]:
                e = self.active.pop(jid)
                e.preempted += 1
                self._changed = True
                if est is not None and dev > est:
                    # the overrun is charged to the evicted job (the most recently joined,
                    # whose cost is the least known): it resumes only when it fits with it
                    e.extra_ms = max(e.extra_ms, dev - est)
                if alone:  # over the budget by itself: it cannot be served within it
                    self._finish(jid, info, entry=e, refused=True)
                else:
                    self.paused.append(e)
                    self.paused.sort(key=lambda q: q.submitted)
        return rec

    def _finish(self, jid: int, info: dict, entry: _Entry | None = None,
                refused: bool = False) -> None:
        e = entry if entry is not None else self.active.pop(jid)
        err = None
        if e.req.check_targets and self.direct_sum is not None and info["done"]:
            pos, _ = self.pool.snapshot(jid)
            acc = self.pool.snapshot_accel(jid)
            mass = e.req.state[2] if e.req.state is not None else self.pool.snapshot_mass(jid)
            rng = np.random.default_rng(int(jid))
            tg = np.sort(rng.choice(e.req.n, min(e.req.n, e.req.check_targets), replace=False))
            ref = self.direct_sum(pos, mass, e.req.g, e.req.softening, tg.astype(np.uint32))
            err = np.linalg.norm(acc[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        state = self.pool.retire(jid, copy=self.keep_state)
        self._changed = True
        self.results.append(Result(e.req.tag, e.req.n, int(info["steps_done"]), e.preempted,
                                   refused, e.submitted, e.admitted, self.now(),
                                   state if self.keep_state else None, err))

    def run(self, arrivals, seconds: float, idle_sleep: float = 1e-3, progress=None,
            progress_every: float = 30.0) -> None:
        """Serve for `seconds`: `arrivals` yields (arrival time in seconds since the
        server started, Request) in time order; each is submitted once its time has
        come, admitted when it fits, and stepped with the others. progress(server) is
        called every `progress_every` seconds."""
        it = iter(arrivals)
        nxt = next(it, None)
        tick = progress_every
        while self.now() < seconds:
            now = self.now()
            if progress is not None and now >= tick:
                progress(self)
                tick += progress_every
            while nxt is not None and nxt[0] <= now:
                self.submit(nxt[1])
                nxt = next(it, None)
            self.admit_ready()
            if self.step() is None:
                self.sleep(idle_sleep)

    def drain(self, max_steps: int = 1_000_000) -> None:
        """Admit and step until every submitted job has finished."""
        for _ in range(max_steps):
            self.admit_ready()
            if self.step() is None and not self.waiting and not self.paused:
                return
        raise RuntimeError("drain: jobs still resident")

    # ------------------------------------------------------------------ report

    def stats(self, since: float = 0.0) -> dict:
        """Latency and throughput over the steps taken at or after `since` seconds."""
        recs = [r for r in self.records if r.t >= since]
        wall = np.array([r.wall_ms for r in recs]) if recs else np.zeros(1)
        dev = np.array([r.device_ms for r in recs]) if recs else np.zeros(1)
        jobs = np.array([r.jobs for r in recs]) if recs else np.ones(1, int)
        wj = np.repeat(wall, jobs)  # one entry per job a step advanced
        span = max(self.now() - since, 1e-9)
        done = [r for r in self.results if not r.refused and r.finished >= since]
        delays = [r.admitted - r.submitted for r in self.results
                  if r.admitted is not None and r.admitted >= since]
        errs = [r.force_error for r in done if r.force_error is not None]
        out = {
            "window_s": span, "steps": len(recs),
            "step_ms_wall": {"p50": float(np.percentile(wall, 50)),
                             "p99": float(np.percentile(wall, 99)), "max": float(wall.max())},
            "step_ms_device": {"p50": float(np.percentile(dev, 50)),
                               "p99": float(np.percentile(dev, 99)), "max": float(dev.max())},
            "job_step_ms_wall": {"p50": float(np.percentile(wj, 50)),
                                 "p99": float(np.percentile(wj, 99))},
            "particle_forces_per_s": sum(r.particles for r in recs) / span,
            "particle_forces_per_s_stepping": sum(r.particles for r in recs) /
                                              max(sum(r.wall_ms for r in recs) * 1e-3, 1e-9),
            "jobs_finished": len(done),
            "jobs_preempted": sum(1 for r in done if r.preempted),
            "preemptions": sum(r.preempted for r in done) + sum(e.preempted for e in
                                                              self.paused + list(self.active.values())),
            "jobs_refused": sum(1 for r in self.results if r.refused),
            "jobs_waiting": len(self.waiting), "jobs_paused": len(self.paused),
            "jobs_active": len(self.active),
            "queue_delay_s": {"p50": float(np.percentile(delays, 50)) if delays else 0.0,
                              "p99": float(np.percentile(delays, 99)) if delays else 0.0},
        }
        if errs:
            e = np.concatenate(errs)
            out["force_error_vs_fp64_direct"] = {"median": float(np.median(e)),
                                                 "p99": float(np.quantile(e, 0.99)),
                                                 "targets": int(e.size), "jobs": len(errs)}
        return out


def poisson_collisions(rate: float, seconds: float, seed: int, *, nmin=16384, nmax=2097152,
                       thetas=(0.4, 0.5, 0.7), steps=(50, 500), softening=0.05, dt=5e-3,
                       device_ic=True, check_targets=0, make_state=None):
    """The serving stream's arrivals: two-galaxy collision jobs at Poisson times of
    `rate` per second, N log-uniform in [nmin, nmax], theta uniform in `thetas`, steps
    uniform in [steps[0], steps[1]]; yields (time, Request). device_ic: the job's state
    is drawn on the device at admission (galaxy_amd.ic.spec_collision); otherwise
    `make_state(n, seed)` returns host arrays (pos, vel, mass)."""
    from . import ic

    rng = np.random.default_rng(seed)
    t = 0.0
    k = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            return
        n = int(np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        theta = float(rng.choice(thetas))
        st = int(rng.integers(steps[0], steps[1] + 1))
        jseed = seed + 1000 + k
        if device_ic:
            req = Request(n, theta, st, softening, dt, spec=ic.spec_collision(n, jseed), tag=k,
                          check_targets=check_targets)
        else:
            pos, vel, mass = make_state(n, jseed)
            req = Request(n, theta, st, softening, dt, state=(pos, vel, mass), tag=k,
                          check_targets=check_targets)
        k += 1
        yield t, req
