"""GPU test of glx_direct_sum_f64_noalloc, called through the C ABI with ctypes: its
output equals glx_direct_sum_f64's bit for bit (same kernels, same chunk plan) on two
calls in a row, the second smaller than the first, and sizes beyond its static device
array are refused before any device work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from galaxy_amd import lbvh as L

pytestmark = pytest.mark.gpu

P, S, D = C.c_void_p, C.c_size_t, C.c_double


def _lib():
    lib = C.CDLL(L.LIB_PATH)
    for name in ("glx_direct_sum_f64", "glx_direct_sum_f64_noalloc"):
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = [S, P, P, D, D, P, S, P]
    lib.glx_last_error.restype = C.c_char_p
    return lib


def _call(lib, name, pos, mass, eps, tg):
    out = np.full((tg.size, 3), np.nan)
    rc = getattr(lib, name)(mass.size, pos.ctypes.data, mass.ctypes.data, 1.0, eps,
                            tg.ctypes.data, tg.size, out.ctypes.data)
    assert rc == 0, (name, rc, lib.glx_last_error())
    return out


def _cloud(seed, n):
    rng = np.random.default_rng(seed)
    pos = np.ascontiguousarray(rng.normal(size=(n, 3)) * rng.uniform(0.5, 3.0, (n, 1)))
    mass = np.ascontiguousarray(rng.uniform(0.5, 1.5, n) / n)
    return pos, mass


@pytest.mark.parametrize("case", [((0, 200_000, 5000), (1, 30_000, 700)),
                                  ((2, 65_536, 65_536), (3, 1_000, 1_000))])
def test_noalloc_equals_allocating_entry_point(case):
    lib = _lib()
    for seed, n, k in case:  # two calls in a row, the second smaller
        print(f"[no_alloc] seed {seed} n {n} targets {k}", flush=True)
        pos, mass = _cloud(seed, n)
        tg = np.ascontiguousarray(
            np.sort(np.random.default_rng(seed + 9).choice(n, k, replace=False)).astype(np.uint32))
        want = _call(lib, "glx_direct_sum_f64", pos, mass, 0.01, tg)
        got = _call(lib, "glx_direct_sum_f64_noalloc", pos, mass, 0.01, tg)
        assert np.isfinite(want).all()
        assert np.array_equal(got, want)


def test_noalloc_refuses_sizes_beyond_its_array():
    lib = _lib()
    n = 2_000_000  # 32 n bytes of staging exceed the 48 MiB staging area
    pos = np.zeros((n, 3))
    mass = np.ones(n)
    tg = np.zeros(4, dtype=np.uint32)
    out = np.zeros((4, 3))
    rc = lib.glx_direct_sum_f64_noalloc(n, pos.ctypes.data, mass.ctypes.data, 1.0, 0.01,
                                        tg.ctypes.data, tg.size, out.ctypes.data)
    assert rc == L.GLX_EINVAL
    assert b"staging" in lib.glx_last_error()
    # the array still serves a call that fits
    pos, mass = _cloud(4, 3000)
    tg = np.arange(3000, dtype=np.uint32)
    assert np.array_equal(_call(lib, "glx_direct_sum_f64_noalloc", pos, mass, 0.02, tg),
                          _call(lib, "glx_direct_sum_f64", pos, mass, 0.02, tg))
