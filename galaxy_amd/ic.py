"""Benchmark-workload initial conditions (libgalaxy_lbvh.so, include/galaxy_ic.h).

plummer()          -> Plummer::sample (ic/src/plummer.rs:94-139), truncated at rmax, in the
                      reference's sequential draw order (host)
collision()        -> two disk+bulge+halo galaxies on a parabolic orbit
                      (ic/src/disk.rs, ic/src/hernquist.rs, ic/src/encounter.rs,
                      ic/src/orient.rs) (host)
generate()         -> any generator of a Spec on the host (one stream per particle)
generate_device()  -> the same generator on the GPU, into device buffers

A Spec also feeds GpuResidentLeapfrog.generate and Pool.admit_ic, which draw the state
straight into the resident arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lbvh import _check
from .lbvh import lib as _lib

PLUMMER, HERNQUIST, COLLISION = 0, 1, 2


class Spec(C.Structure):
    """glx_ic_spec (include/galaxy_ic.h)."""
    _fields_ = [
        ("kind", C.c_int),
        ("n", C.c_size_t),
        ("seed", C.c_uint64),
        ("g", C.c_double),
        ("m", C.c_double),
        ("a", C.c_double),
        ("rmax", C.c_double),
        ("mass_ratio", C.c_double),
        ("rp", C.c_double),
        ("r0", C.c_double),
    ]


def spec_plummer(n: int, seed: int, g=1.0, m=1.0, a=1.0, rmax=10.0) -> Spec:
    return Spec(PLUMMER, int(n), seed & 0xFFFFFFFFFFFFFFFF, g, m, a, rmax, 0.0, 0.0, 0.0)


def spec_hernquist(n: int, seed: int, g=1.0, m=1.0, a=1.0) -> Spec:
    return Spec(HERNQUIST, int(n), seed & 0xFFFFFFFFFFFFFFFF, g, m, a, 0.0, 0.0, 0.0, 0.0)


def spec_collision(n: int, seed: int, mass_ratio=1.0, rp=8.0, r0=80.0) -> Spec:
    return Spec(COLLISION, int(n), seed & 0xFFFFFFFFFFFFFFFF, 1.0, 0.0, 0.0, 0.0, mass_ratio,
                rp, r0)


def _bind():
    L = _lib()
    if not getattr(L, "_ic_bound", False):
        P, S = C.c_void_p, C.c_size_t
        L.glx_ic_plummer.restype = C.c_int
        L.glx_ic_plummer.argtypes = [S, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double,
                                     P, P, P]
        L.glx_ic_collision.restype = C.c_int
        L.glx_ic_collision.argtypes = [S, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                       P, P, P, P]
        L.glx_ic_generate.restype = C.c_int
        L.glx_ic_generate.argtypes = [P, P, P, P, P]
        L.glx_ic_generate_device.restype = C.c_int
        L.glx_ic_generate_device.argtypes = [P, P, P, P]
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
    rc = _bind().glx_ic_collision(n, seed & 0xFFFFFFFFFFFFFFFF, mass_ratio, rp, r0, _p(pos),
                                  _p(vel), _p(mass), _p(info))
    if rc:
        raise ValueError("invalid collision parameters")
    return pos, vel, mass, {"incl": (info[0], info[2]), "arg": (info[1], info[3])}


def generate(spec: Spec):
    """Host generation of `spec`: (pos, vel, mass) as f64 arrays."""
    n = int(spec.n)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    mass = np.zeros(n)
    _check(_bind().glx_ic_generate(C.addressof(spec), _p(pos), _p(vel), _p(mass), None))
    return pos, vel, mass


class DeviceArray:
    """A float64 device buffer of `shape` owned by the library's allocator
    (glx_device_malloc); `ptr` is the device address, `numpy()` copies it to the host."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        p = C.c_void_p()
        _check(_lib().glx_device_malloc(self.nbytes, C.byref(p)))
        self.ptr = p.value or 0

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape)
        _check(_lib().glx_device_to_host(_p(out), self.ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", 0):
            _lib().glx_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_device(spec: Spec):
    """Device generation of `spec` into new device buffers (pos (n,3), vel (n,3), mass (n,),
    float64): the arrays are drawn in HBM and never exist on the host unless copied out
    with `.numpy()`."""
    n = int(spec.n)
    _bind()
    out = (DeviceArray((n, 3)), DeviceArray((n, 3)), DeviceArray((n,)))
    _check(_lib().glx_ic_generate_device(C.addressof(spec), out[0].ptr, out[1].ptr, out[2].ptr))
    return out
