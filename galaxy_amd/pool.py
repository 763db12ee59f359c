"""Batched / served forms: a device-resident pool of independent systems.

Python front end of the glx_pool_* C ABI (include/galaxy_lbvh.h). There is no reference
counterpart — the reference steps one system per GpuResidentLeapfrog
(gpu/src/gpu_resident.rs:1848) — so the interface follows the serving loop it exists
for: admit jobs while memory and slots allow, step every active job together, retire
finished jobs, repeat.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lbvh import (GLX_EBUDGET, GLX_ENOSLOT, PREC_F32, WALK_REFERENCE, Stats,
                   _check, _f64, _p, _params, lib)


class Job(C.Structure):
    _fields_ = [
        ("n", C.c_size_t),
        ("pos", C.c_void_p),
        ("vel", C.c_void_p),
        ("mass", C.c_void_p),
        ("g", C.c_double),
        ("softening", C.c_double),
        ("theta", C.c_double),
        ("dt", C.c_double),
        ("steps", C.c_uint64),
    ]


class JobInfo(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("n", C.c_size_t),
        ("steps_done", C.c_uint64),
        ("steps", C.c_uint64),
        ("time", C.c_double),
        ("done", C.c_int),
        ("evicted", C.c_int),
    ]


class PoolInfo(C.Structure):
    _fields_ = [
        ("jobs", C.c_uint32),
        ("active", C.c_uint32),
        ("particles", C.c_uint64),
        ("bytes_used", C.c_uint64),
        ("budget", C.c_uint64),
        ("steps", C.c_uint64),
        ("evictions", C.c_uint64),
        ("refresh_ms", C.c_double),
        ("step_budget_ms", C.c_double),
    ]


def _bind():
    L = lib()
    if getattr(L, "_pool_bound", False):
        return L
    P, S, D, I = C.c_void_p, C.c_size_t, C.c_double, C.c_int
    sig = {
        "glx_pool_create": (I, [P, C.c_uint64, C.c_uint32, C.POINTER(P)]),
        "glx_pool_destroy": (I, [P]),
        "glx_pool_admit": (I, [P, C.POINTER(Job), C.POINTER(C.c_uint64)]),
        "glx_pool_admit_ic": (I, [P, P, D, D, D, C.c_uint64, C.POINTER(C.c_uint64)]),
        "glx_pool_step": (I, [P, C.POINTER(D)]),
        "glx_pool_set_step_budget": (I, [P, D]),
        "glx_pool_resume": (I, [P, C.c_uint64]),
        "glx_pool_step_timed": (I, [P, P]),
        "glx_pool_count": (I, [P, C.POINTER(Stats)]),
        "glx_pool_job": (I, [P, C.c_uint64, C.POINTER(JobInfo)]),
        "glx_pool_snapshot": (I, [P, C.c_uint64, P, P]),
        "glx_pool_snapshot_accel": (I, [P, C.c_uint64, P]),
        "glx_pool_snapshot_mass": (I, [P, C.c_uint64, P]),
        "glx_pool_retire": (I, [P, C.c_uint64, P, P]),
        "glx_pool_status": (I, [P, C.POINTER(PoolInfo)]),
        "glx_pool_job_bytes": (C.c_uint64, [S]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._pool_bound = True
    return L


PHASES = ("kick_drift_bbox", "encode", "sort", "gather", "karras", "aggregate", "walk")


class Pool:
    """Resident multi-system stepper (batched when filled up front, served when jobs
    come and go between steps)."""

    def __init__(self, budget_bytes: int, max_jobs: int = 4096, morton_bits: int = 30,
                 walk: int = WALK_REFERENCE, multipole: int = 0, precision: int = PREC_F32,
                 leaf_size: int = 1, far_theta: float = 0.0, mono_ratio: float = 0.0,
                 radix_bits: int = 0, group_size: int = 0, group_lanes: int = 0):
        L = _bind()
        h = C.c_void_p()
        p = _params(1.0, 0.0, 1.0, precision, False, morton_bits, walk, multipole, leaf_size,
                    far_theta, mono_ratio, radix_bits, group_size, group_lanes)
        _check(L.glx_pool_create(C.byref(p), int(budget_bytes), int(max_jobs), C.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().glx_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def job_bytes(n: int) -> int:
        return int(_bind().glx_pool_job_bytes(int(n)))

    def admit(self, pos, vel, mass, *, g=1.0, softening=0.05, theta=0.5, dt=1e-3, steps=0):
        """Admit a job; returns its id, or None when the budget or the slots are full."""
        pos = _f64(pos, (-1, 3))
        vel = _f64(vel, (-1, 3))
        mass = _f64(mass)
        job = Job(mass.size, pos.ctypes.data, vel.ctypes.data, mass.ctypes.data, float(g),
                  float(softening), float(theta), float(dt), int(steps))
        jid = C.c_uint64(0)
        rc = _bind().glx_pool_admit(self._h, C.byref(job), C.byref(jid))
        if rc in (GLX_EBUDGET, GLX_ENOSLOT):
            return None
        _check(rc)
        self._n[jid.value] = mass.size
        return jid.value

    def admit_ic(self, spec, *, softening=0.05, theta=0.5, dt=1e-3, steps=0):
        """Admit a job whose state is drawn on the device by a generator
        (`galaxy_amd.ic.Spec`); returns its id, or None when the budget or the slots are
        full."""
        jid = C.c_uint64(0)
        rc = _bind().glx_pool_admit_ic(self._h, C.addressof(spec), float(softening),
                                       float(theta), float(dt), int(steps), C.byref(jid))
        if rc in (GLX_EBUDGET, GLX_ENOSLOT):
            return None
        _check(rc)
        self._n[jid.value] = int(spec.n)
        return jid.value

    def step(self) -> float:
        """One KDK step of every active job; returns the step's device time in ms."""
        ms = C.c_double(0.0)
        _check(_bind().glx_pool_step(self._h, C.byref(ms)))
        return ms.value

    def set_step_budget(self, ms: float) -> None:
        """Device-time budget of one step (0: none): after a step over it the most
        recently admitted or resumed active job is evicted (paused: not advanced, state
        kept until `resume` or `retire`)."""
        _check(_bind().glx_pool_set_step_budget(self._h, float(ms)))

    def resume(self, jid: int) -> None:
        """Advance an evicted job again from its kept state."""
        _check(_bind().glx_pool_resume(self._h, int(jid)))

    def step_timed(self) -> dict:
        ph = np.zeros(len(PHASES))
        _check(_bind().glx_pool_step_timed(self._h, _p(ph)))
        return dict(zip(PHASES, ph.tolist()))

    def count(self) -> dict:
        st = Stats()
        _check(_bind().glx_pool_count(self._h, C.byref(st)))
        return st.as_dict()

    def job(self, jid: int) -> dict:
        ji = JobInfo()
        _check(_bind().glx_pool_job(self._h, int(jid), C.byref(ji)))
        return {"id": ji.id, "n": ji.n, "steps_done": ji.steps_done, "steps": ji.steps,
                "time": ji.time, "done": bool(ji.done), "evicted": bool(ji.evicted)}

    def snapshot(self, jid: int):
        n = self._n[jid]
        pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
        _check(_bind().glx_pool_snapshot(self._h, int(jid), _p(pos), _p(vel)))
        return pos, vel

    def snapshot_accel(self, jid: int) -> np.ndarray:
        acc = np.zeros((self._n[jid], 3))
        _check(_bind().glx_pool_snapshot_accel(self._h, int(jid), _p(acc)))
        return acc

    def snapshot_mass(self, jid: int) -> np.ndarray:
        m = np.zeros(self._n[jid])
        _check(_bind().glx_pool_snapshot_mass(self._h, int(jid), _p(m)))
        return m

    def retire(self, jid: int, copy: bool = True):
        """Free the job; with copy, return its final (pos, vel) (else None)."""
        n = self._n.pop(jid)
        if not copy:
            _check(_bind().glx_pool_retire(self._h, int(jid), None, None))
            return None
        pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
        _check(_bind().glx_pool_retire(self._h, int(jid), _p(pos), _p(vel)))
        return pos, vel

    def status(self) -> dict:
        pi = PoolInfo()
        _check(_bind().glx_pool_status(self._h, C.byref(pi)))
        return {k: (float(getattr(pi, k)) if t is C.c_double else int(getattr(pi, k)))
                for k, t in PoolInfo._fields_}
