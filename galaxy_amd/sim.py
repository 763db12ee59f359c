"""Stepping driver and snapshot files (include/galaxy_sim.h).

run()            -> sim::run (sim/src/lib.rs:110-174) with a DirectorySink
                    (sim/src/lib.rs:80-106) over a GpuResidentLeapfrog
write_snapshot() -> io::write_file (io/src/snapshot.rs:130-194,319-328)
read_snapshot()  -> io::read_file (io/src/snapshot.rs:199-316,331-334)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from .lbvh import GlxError, _check, _f64, lib

GLX_EIO = 7
GLX_EFORMAT = 8


class SnapHeader(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("step", C.c_uint64),
        ("scale_factor", C.c_double),
        ("softening", C.c_double),
        ("n_particles", C.c_uint64),
        ("rng_seed", C.c_uint64),
        ("config_hash", C.c_uint64),
        ("units", C.c_char_p),
        ("code_version", C.c_char_p),
    ]


class SnapColumns(C.Structure):
    _fields_ = [
        ("id", C.c_void_p),
        ("progenitor", C.c_void_p),
        ("kind", C.c_void_p),
        ("u", C.c_void_p),
        ("formation_time", C.c_void_p),
    ]


class RunConfig(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("n_steps", C.c_uint64),
        ("snapshot_every", C.c_uint64),
        ("softening", C.c_double),
        ("rng_seed", C.c_uint64),
        ("config_hash", C.c_uint64),
        ("units", C.c_char_p),
        ("first_step", C.c_uint64),
    ]


class RunSummary(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("final_time", C.c_double),
                ("snapshots_emitted", C.c_uint64)]


def _bind():
    L = lib()
    if getattr(L, "_sim_bound", False):
        return L
    P, S, I = C.c_void_p, C.c_size_t, C.c_int
    sig = {
        "glx_snapshot_write": (I, [C.c_char_p, C.POINTER(SnapHeader), S, P, P, P,
                                   C.POINTER(SnapColumns)]),
        "glx_snapshot_read_header": (I, [C.c_char_p, C.POINTER(SnapHeader), C.c_char_p, S,
                                         C.c_char_p, S, C.POINTER(C.c_uint32)]),
        "glx_snapshot_read": (I, [C.c_char_p, S, P, P, P, P, P, P, P, P]),
        "glx_sys_run": (I, [P, C.POINTER(RunConfig), C.c_char_p, C.POINTER(SnapColumns),
                            C.POINTER(RunSummary)]),
        "glx_sys_set_time": (I, [P, C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._sim_bound = True
    return L


@dataclass
class Snapshot:
    """A snapshot in the reference's (Header, State) shape."""

    time: float
    step: int
    scale_factor: float
    softening: float
    rng_seed: int
    config_hash: int
    units: str
    code_version: str
    version: int
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    id: np.ndarray
    progenitor: np.ndarray
    kind: np.ndarray
    u: np.ndarray
    formation_time: np.ndarray = field(repr=False)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _columns(n, id=None, progenitor=None, kind=None, u=None, formation_time=None):
    keep = []

    def col(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        if a.shape != (n,):
            raise ValueError("column length must equal the particle count")
        keep.append(a)
        return _ptr(a)

    c = SnapColumns(col(id, np.uint64), col(progenitor, np.uint16), col(kind, np.uint8),
                    col(u, np.float64), col(formation_time, np.float64))
    return c, keep


def write_snapshot(path, pos, vel, mass, *, time=0.0, step=0, scale_factor=1.0, softening=0.0,
                   rng_seed=0, config_hash=0, units="", code_version="galaxy_amd 0.1.0",
                   **columns):
    pos = _f64(pos, (-1, 3))
    vel = _f64(vel, (-1, 3))
    mass = _f64(mass)
    n = mass.size
    h = SnapHeader(float(time), int(step), float(scale_factor), float(softening), n,
                   int(rng_seed), int(config_hash), units.encode(), code_version.encode())
    cols, _keep = _columns(n, **columns)
    _check(_bind().glx_snapshot_write(os.fsencode(path), C.byref(h), n, _ptr(pos), _ptr(vel),
                                      _ptr(mass), C.byref(cols)))


def read_snapshot(path) -> Snapshot:
    L = _bind()
    h = SnapHeader()
    ub = C.create_string_buffer(65537)
    cb = C.create_string_buffer(65537)
    ver = C.c_uint32(0)
    p = os.fsencode(path)
    _check(L.glx_snapshot_read_header(p, C.byref(h), ub, 65537, cb, 65537, C.byref(ver)))
    n = int(h.n_particles)
    pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
    mass, u, ft = np.zeros(n), np.zeros(n), np.zeros(n)
    ids = np.zeros(n, np.uint64)
    prog = np.zeros(n, np.uint16)
    kind = np.zeros(n, np.uint8)
    _check(L.glx_snapshot_read(p, n, _ptr(pos), _ptr(vel), _ptr(mass), _ptr(ids), _ptr(prog),
                               _ptr(kind), _ptr(u), _ptr(ft)))
    return Snapshot(h.time, int(h.step), h.scale_factor, h.softening, int(h.rng_seed),
                    int(h.config_hash), ub.value.decode(), cb.value.decode(), int(ver.value),
                    pos, vel, mass, ids, prog, kind, u, ft)


def run(system, directory, *, dt, n_steps, snapshot_every, softening, rng_seed=0,
        config_hash=0, units="", first_step=0, **columns) -> dict:
    """sim::run over a GpuResidentLeapfrog `system` (already uploaded). Returns the
    RunSummary as a dict."""
    L = _bind()
    n = len(system)
    cfg = RunConfig(float(dt), int(n_steps), int(snapshot_every), float(softening),
                    int(rng_seed), int(config_hash), units.encode(), int(first_step))
    cols, _keep = _columns(n, **columns)
    out = RunSummary()
    _check(L.glx_sys_run(system._h, C.byref(cfg), os.fsencode(directory), C.byref(cols),
                         C.byref(out)))
    return {"steps": int(out.steps), "final_time": out.final_time,
            "snapshots_emitted": int(out.snapshots_emitted)}


def set_time(system, time: float):
    _check(_bind().glx_sys_set_time(system._h, float(time)))


__all__ = ["GlxError", "Snapshot", "read_snapshot", "write_snapshot", "run", "set_time",
           "GLX_EIO", "GLX_EFORMAT"]
