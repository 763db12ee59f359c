"""Benchmark-workload initial conditions (libgalaxy_lbvh.so, include/galaxy_ic.h).

plummer()   -> Plummer::sample (ic/src/plummer.rs:94-139), truncated at rmax
collision() -> two disk+bulge+halo galaxies on a parabolic orbit
              (ic/src/disk.rs, ic/src/hernquist.rs, ic/src/encounter.rs, ic/src/orient.rs)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lbvh import lib as _lib


def _bind():
    L = _lib()
    if not getattr(L, "_ic_bound", False):
        P, S = C.c_void_p, C.c_size_t
        L.glx_ic_plummer.restype = C.c_int
        L.glx_ic_plummer.argtypes = [S, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double,
                                     P, P, P]
        L.glx_ic_collision.restype = C.c_int
        L.glx_ic_collision.argtypes = [S, C.c_uint64, C.c_double, C.c_double, C.c_double, P, P, P, P]
        L._ic_bound = True
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def plummer(n: int, seed: int, g: float = 1.0, m: float = 1.0, a: float = 1.0, rmax: float = 10.0):
    """Plummer sphere, f32-representable positions/velocities (stored as f64)."""
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    mass = np.zeros(n)
    rc = _bind().glx_ic_plummer(n, seed & 0xFFFFFFFFFFFFFFFF, g, m, a, rmax, _p(pos), _p(vel),
                                _p(mass))
    if rc:
        raise ValueError("invalid Plummer parameters")
    return pos, vel, mass


def collision(n: int, seed: int, mass_ratio: float = 1.0, rp: float = 8.0, r0: float = 80.0):
    """Two-galaxy collision (n/2 particles per galaxy). Returns pos, vel, mass, info."""
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    mass = np.zeros(n)
    info = np.zeros(4)
    rc = _bind().glx_ic_collision(n, seed & 0xFFFFFFFFFFFFFFFF, mass_ratio, rp, r0, _p(pos), _p(vel),
                                  _p(mass), _p(info))
    if rc:
        raise ValueError("invalid collision parameters")
    return pos, vel, mass, {"incl": (info[0], info[2]), "arg": (info[1], info[3])}
