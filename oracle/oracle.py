"""ctypes front end of the CPU oracle (oracle/lbvh_oracle.c, oracle/ic_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the timed CPU
baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module. Each wrapper names the reference function (file:line) its C
counterpart restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbvh_oracle.so")
_lib = None

NO_PARENT = 0xFFFFFFFF

# orc_node_t layout (solvers/src/lbvh.rs:394-414 LbvhNode)
NODE_DTYPE = np.dtype(
    [
        ("center", "<f8", 3),
        ("half", "<f8", 3),
        ("com", "<f8", 3),
        ("mass", "<f8"),
        ("delta", "<f8"),
        ("next", "<u4"),
        ("body_start", "<u4"),
        ("body_count", "<u4"),
        ("pad", "<u4"),
    ]
)


def build() -> str:
    """Compile the oracle with its committed Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        S = C.c_size_t
        D = C.c_double
        L.orc_expand10.restype = C.c_uint32
        L.orc_expand10.argtypes = [C.c_uint32]
        L.orc_morton3.restype = C.c_uint32
        L.orc_morton3.argtypes = [C.c_uint32] * 3
        L.orc_reference_morton.argtypes = [S, P, P, P, P]
        L.orc_reference_sort.argtypes = [S, P, P]
        L.orc_reference_karras.argtypes = [S, P, P, P, P]
        L.orc_reference_aggregate.argtypes = [S, P, P, P, P, P, P, P, P]
        L.orc_reference_flatten.argtypes = [S, P, P, P, P, P, P, P, P, P]
        L.orc_node_size.restype = S
        L.orc_lbvh_accelerations.argtypes = [S, P, P, D, D, D, P, S, P, P, C.c_int]
        L.orc_lbvh_build_flat.argtypes = [S, P, P, P, P]
        L.orc_direct_sum.argtypes = [S, P, P, D, D, P]
        L.orc_direct_sum.restype = None
        L.orc_direct_sum_targets.argtypes = [S, P, P, D, D, P, S, P, C.c_int]
        L.orc_direct_sum_targets.restype = None
        L.orc_potential_serial.argtypes = [S, P, P, D, D]
        L.orc_potential_serial.restype = D
        L.orc_diagnostics.argtypes = [S, P, P, P, P]
        L.orc_diagnostics.restype = None
        L.orc_kdk_run.argtypes = [S, P, P, P, D, D, D, D, S, C.c_int, P, C.c_int]
        L.orc_expand21.restype = C.c_uint64
        L.orc_expand21.argtypes = [C.c_uint64]
        L.orc_reference_morton63.argtypes = [S, P, P, P, P]
        L.orc_reference_sort63.argtypes = [S, P, P]
        L.orc_reference_karras63.argtypes = [S, P, P, P, P]
        L.orc_ic_hernquist_df.argtypes = [D, D, D, D]
        L.orc_ic_hernquist_df.restype = D
        L.orc_ic_galaxy_body.argtypes = [S, C.c_uint64, D, P, P, P]
        L.orc_ic_galaxy_body.restype = None
        L.orc_ic_sphere.argtypes = [C.c_int, S, C.c_uint64, D, D, D, D, P, P, P]
        L.orc_ic_sphere.restype = None
        L.orc_ic_collision.argtypes = [S, C.c_uint64, D, D, D, P, P, P, P]
        L.orc_ic_collision.restype = C.c_int
        assert L.orc_node_size() == NODE_DTYPE.itemsize
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def expand10(v: int) -> int:
    """solvers/src/lbvh.rs:37-44"""
    return int(lib().orc_expand10(v & 0xFFFFFFFF))


def morton3(x: int, y: int, z: int) -> int:
    """solvers/src/lbvh.rs:48-50"""
    return int(lib().orc_morton3(x, y, z))


def reference_morton(pos):
    """solvers/src/lbvh.rs:105-129 -> (bounds{bmin,size,scale}, lanes[n,3], codes[n])."""
    pos = _f64(pos).reshape(-1, 3)
    n = pos.shape[0]
    if n == 0:
        raise ValueError("reference_morton requires a non-empty system")
    bounds = np.zeros(5)
    lanes = np.zeros((n, 3), np.uint32)
    codes = np.zeros(n, np.uint32)
    lib().orc_reference_morton(n, _p(pos), _p(bounds), _p(lanes), _p(codes))
    return {"bmin": bounds[:3].copy(), "size": bounds[3], "scale": bounds[4]}, lanes, codes


def expand21(v: int) -> int:
    """The 21-bit lane spread of the 63-bit keys (bit b of v -> bit 3b)."""
    return int(lib().orc_expand21(v & 0xFFFFFFFFFFFFFFFF))


def reference_morton63(pos):
    """63-bit variant of solvers/src/lbvh.rs:105-129 (2^21 cells per axis)."""
    pos = _f64(pos).reshape(-1, 3)
    n = pos.shape[0]
    if n == 0:
        raise ValueError("reference_morton63 requires a non-empty system")
    bounds = np.zeros(5)
    lanes = np.zeros((n, 3), np.uint32)
    codes = np.zeros(n, np.uint64)
    lib().orc_reference_morton63(n, _p(pos), _p(bounds), _p(lanes), _p(codes))
    return {"bmin": bounds[:3].copy(), "size": bounds[3], "scale": bounds[4]}, lanes, codes


def reference_sort63(codes) -> np.ndarray:
    codes = np.ascontiguousarray(codes, dtype=np.uint64)
    order = np.zeros(codes.size, np.uint32)
    lib().orc_reference_sort63(codes.size, _p(codes), _p(order))
    return order


def reference_karras63(sorted_codes):
    sc = np.ascontiguousarray(sorted_codes, dtype=np.uint64)
    n = sc.size
    if n == 0:
        raise ValueError("reference_karras63 requires a non-empty system")
    left = np.zeros(max(n - 1, 0), np.uint32)
    right = np.zeros(max(n - 1, 0), np.uint32)
    parent = np.zeros(2 * n - 1, np.uint32)
    lib().orc_reference_karras63(n, _p(sc), _p(left), _p(right), _p(parent))
    return left, right, parent


def reference_sort(codes) -> np.ndarray:
    """solvers/src/lbvh.rs:143-147"""
    codes = _u32(codes)
    order = np.zeros(codes.size, np.uint32)
    lib().orc_reference_sort(codes.size, _p(codes), _p(order))
    return order


def reference_karras(sorted_codes):
    """solvers/src/lbvh.rs:276-306 -> (left[n-1], right[n-1], parent[2n-1])."""
    sc = _u32(sorted_codes)
    n = sc.size
    if n == 0:
        raise ValueError("reference_karras requires a non-empty system")
    left = np.zeros(max(n - 1, 0), np.uint32)
    right = np.zeros(max(n - 1, 0), np.uint32)
    parent = np.zeros(2 * n - 1, np.uint32)
    lib().orc_reference_karras(n, _p(sc), _p(left), _p(right), _p(parent))
    return left, right, parent


def reference_aggregate(left, right, sorted_pos, sorted_mass):
    """solvers/src/lbvh.rs:329-379 -> dict of per-unified-node aabb_min/max, com, mass."""
    sp = _f64(sorted_pos).reshape(-1, 3)
    sm = _f64(sorted_mass)
    n = sm.size
    t = 2 * n - 1
    out = {
        "aabb_min": np.zeros((t, 3)),
        "aabb_max": np.zeros((t, 3)),
        "com": np.zeros((t, 3)),
        "mass": np.zeros(t),
    }
    lib().orc_reference_aggregate(
        n, _p(_u32(left)), _p(_u32(right)), _p(sp), _p(sm),
        _p(out["aabb_min"]), _p(out["aabb_max"]), _p(out["com"]), _p(out["mass"]),
    )
    return out


def reference_flatten(left, right, agg, order):
    """solvers/src/lbvh.rs:545-614 -> (nodes structured array, leaf_bodies)."""
    order = _u32(order)
    n = order.size
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    lb = np.zeros(n, np.uint32)
    lib().orc_reference_flatten(
        n, _p(_u32(left)), _p(_u32(right)), _p(agg["aabb_min"]), _p(agg["aabb_max"]),
        _p(agg["com"]), _p(agg["mass"]), _p(order), _p(nodes), _p(lb),
    )
    return nodes, lb


def build_flat(pos, mass):
    """solvers/src/lbvh.rs:440-463 (LbvhFlat::build)."""
    pos = _f64(pos).reshape(-1, 3)
    mass = _f64(mass)
    n = mass.size
    nodes = np.zeros(max(2 * n - 1, 0), NODE_DTYPE)
    lb = np.zeros(n, np.uint32)
    if n:
        rc = lib().orc_lbvh_build_flat(n, _p(pos), _p(mass), _p(nodes), _p(lb))
        assert rc == 0
    return nodes, lb


def lbvh_accelerations(pos, mass, g, softening, theta, targets=None, nthreads=1):
    """solvers/src/lbvh.rs:662-676 (Lbvh::accelerations). Returns (acc[m,3], interactions)."""
    pos = _f64(pos).reshape(-1, 3)
    mass = _f64(mass)
    n = mass.size
    tg = None if targets is None else _u32(targets)
    m = n if tg is None else tg.size
    acc = np.zeros((m, 3))
    cnt = C.c_uint64(0)
    rc = lib().orc_lbvh_accelerations(
        n, _p(pos), _p(mass), g, softening, theta,
        None if tg is None else _p(tg), 0 if tg is None else tg.size,
        _p(acc), C.byref(cnt), int(nthreads),
    )
    if rc:
        raise MemoryError(f"oracle build failed ({rc})")
    return acc, int(cnt.value)


def direct_sum(pos, mass, g, softening):
    """solvers/src/direct_sum.rs:25-45"""
    pos = _f64(pos).reshape(-1, 3)
    mass = _f64(mass)
    acc = np.zeros_like(pos)
    lib().orc_direct_sum(mass.size, _p(pos), _p(mass), g, softening, _p(acc))
    return acc


def direct_sum_targets(pos, mass, g, softening, targets, nthreads=1):
    """Direct sum onto a target subset (same pair kernel as direct_sum)."""
    pos = _f64(pos).reshape(-1, 3)
    mass = _f64(mass)
    tg = _u32(targets)
    acc = np.zeros((tg.size, 3))
    lib().orc_direct_sum_targets(
        mass.size, _p(pos), _p(mass), g, softening, _p(tg), tg.size, _p(acc), int(nthreads)
    )
    return acc


def potential_serial(pos, mass, g, softening) -> float:
    """solvers/src/potential.rs:30-41"""
    pos = _f64(pos).reshape(-1, 3)
    mass = _f64(mass)
    return float(lib().orc_potential_serial(mass.size, _p(pos), _p(mass), g, softening))


def diagnostics(pos, vel, mass):
    """core/src/diagnostics.rs:8-68 -> dict(ke, p, l, com)."""
    pos = _f64(pos).reshape(-1, 3)
    vel = _f64(vel).reshape(-1, 3)
    mass = _f64(mass)
    out = np.zeros(10)
    lib().orc_diagnostics(mass.size, _p(pos), _p(vel), _p(mass), _p(out))
    return {"ke": out[0], "p": out[1:4].copy(), "l": out[4:7].copy(), "com": out[7:10].copy()}


def kdk_run(pos, vel, mass, g, softening, theta, dt, nsteps, solver="lbvh", nthreads=1):
    """core/src/integrator.rs:37-78 from a fresh integrator -> (pos, vel, acc)."""
    pos = _f64(pos).reshape(-1, 3).copy()
    vel = _f64(vel).reshape(-1, 3).copy()
    mass = _f64(mass)
    acc = np.zeros_like(pos)
    rc = lib().orc_kdk_run(
        mass.size, _p(pos), _p(vel), _p(mass), g, softening, theta, dt, int(nsteps),
        0 if solver == "lbvh" else 1, _p(acc), int(nthreads),
    )
    if rc:
        raise MemoryError(f"oracle build failed ({rc})")
    return pos, vel, acc


# --- workload generators with one stream per particle (oracle/ic_oracle.c) ---

def ic_hernquist_df(g, m, a, energy) -> float:
    """ic/src/hernquist.rs:78-96"""
    return float(lib().orc_ic_hernquist_df(g, m, a, energy))


def ic_sphere(kind, n, seed, g=1.0, m=1.0, a=1.0, rmax=0.0):
    """kind 0: Plummer (ic/src/plummer.rs:94-139, truncated at rmax > 0), kind 1:
    Hernquist (ic/src/hernquist.rs:124-176); recentred -> (pos, vel, mass)."""
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    lib().orc_ic_sphere(int(kind), n, seed & 0xFFFFFFFFFFFFFFFF, g, m, a, rmax, _p(pos),
                        _p(vel), _p(mass))
    return pos, vel, mass


def ic_galaxy_body(n, seed, scale=1.0):
    """One collision galaxy in its body frame before recentring (ic/src/disk.rs:333-371,
    ic/src/hernquist.rs:139-176) -> (pos, vel, mass)."""
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    lib().orc_ic_galaxy_body(n, seed & 0xFFFFFFFFFFFFFFFF, scale, _p(pos), _p(vel), _p(mass))
    return pos, vel, mass


def ic_collision(n, seed, mass_ratio=1.0, rp=8.0, r0=80.0):
    """Two galaxies on a parabolic orbit (ic/src/encounter.rs:45-87, ic/src/orient.rs:63-71)
    -> (pos, vel, mass, info {incl1, arg1, incl2, arg2})."""
    pos, vel, mass, info = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(4)
    rc = lib().orc_ic_collision(n, seed & 0xFFFFFFFFFFFFFFFF, mass_ratio, rp, r0, _p(pos),
                                _p(vel), _p(mass), _p(info))
    if rc:
        raise ValueError("invalid collision parameters")
    return pos, vel, mass, info
