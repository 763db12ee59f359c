"""Python front end of libgalaxy_lbvh.so (the C ABI in include/galaxy_lbvh.h).

Mirrors the reference's interfaces for the LBVH gravity path:
  Lbvh                  -> ForceSolver impl Lbvh / GpuLbvhFused
                           (solvers/src/lbvh.rs:643-683; gpu/src/gpu_lbvh_fused.rs:50-134)
  GpuResidentLeapfrog   -> gpu/src/gpu_resident.rs:1246-1968 (upload/step/step_many/snapshot)
  morton/sort/karras/aggregate -> the public stage oracles of solvers/src/lib.rs:15-18
There is no CPU fallback: every call runs the HIP library, and a missing library or
GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgalaxy_lbvh.so")

GLX_OK, GLX_EINVAL, GLX_ENOMEM, GLX_EDEVICE, GLX_EBUDGET, GLX_ENOSLOT, GLX_ENOTFOUND = range(7)
PREC_F32, PREC_KAHAN, PREC_ABS32 = 0, 1, 2
FLAG_COUNT = 1
WALK_REFERENCE, WALK_GROUP = 0, 1
NO_PARENT = 0xFFFFFFFF


class GlxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("g", C.c_double),
        ("softening", C.c_double),
        ("theta", C.c_double),
        ("leaf_size", C.c_uint32),
        ("multipole", C.c_uint32),
        ("precision", C.c_uint32),
        ("flags", C.c_uint32),
        ("morton_bits", C.c_uint32),
        ("walk", C.c_uint32),
        ("far_theta", C.c_double),
        ("mono_ratio", C.c_double),
        ("radix_bits", C.c_uint32),
        ("group_size", C.c_uint32),
        ("group_lanes", C.c_uint32),
    ]


class Diag(C.Structure):
    _fields_ = [
        ("kinetic", C.c_double),
        ("potential", C.c_double),
        ("momentum", C.c_double * 3),
        ("angular_momentum", C.c_double * 3),
        ("center_of_mass", C.c_double * 3),
        ("mass", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("point_terms", C.c_uint64),
        ("quad_terms", C.c_uint64),
        ("node_tests", C.c_uint64),
        ("fallbacks", C.c_uint64),
        ("max_stack", C.c_uint64),
        ("far_nodes", C.c_uint64),
        ("rounds", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        d = {k: int(getattr(self, k)) for k, _ in self._fields_}
        d["interactions"] = d["point_terms"] + d["quad_terms"]
        d["node_visits"] = d["node_tests"]
        return d


_lib = None


def lib():
    """Load the HIP library; raise if it is not built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m galaxy_amd.build` (needs hipcc)"
        )
    L = C.CDLL(LIB_PATH)
    P, S, D, I = C.c_void_p, C.c_size_t, C.c_double, C.c_int
    sig = {
        "glx_last_error": (C.c_char_p, []),
        "glx_device_info": (I, [C.c_char_p, S, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
        "glx_device_malloc": (I, [S, C.POINTER(P)]),
        "glx_device_free": (I, [P]),
        "glx_device_to_host": (I, [P, P, S]),
        "glx_solver_create": (I, [C.POINTER(Params), C.POINTER(P)]),
        "glx_solver_destroy": (I, [P]),
        "glx_solver_accelerations": (I, [P, S, P, P, P]),
        "glx_solver_accelerations_active": (I, [P, S, P, P, P, S, P]),
        "glx_solver_rebuild_cache": (I, [P, S, P, P]),
        "glx_solver_gravity_active_cached": (I, [P, S, P, P, P, S, P]),
        "glx_solver_potential_energy": (I, [P, S, P, P, C.POINTER(D)]),
        "glx_solver_stats": (I, [P, C.POINTER(Stats)]),
        "glx_direct_sum_f64": (I, [S, P, P, D, D, P, S, P]),
        "glx_direct_sum_f64_noalloc": (I, [S, P, P, D, D, P, S, P]),
        "glx_stage_morton": (I, [S, P, P, P, P]),
        "glx_stage_morton63": (I, [S, P, P, P, P]),
        "glx_stage_sort": (I, [S, P, P, P]),
        "glx_stage_sort63": (I, [S, P, P, P]),
        "glx_stage_sort_bits": (I, [S, P, P, P, C.c_uint32]),
        "glx_stage_sort63_bits": (I, [S, P, P, P, C.c_uint32]),
        "glx_stage_karras": (I, [S, P, P, P, P]),
        "glx_stage_karras63": (I, [S, P, P, P, P]),
        "glx_stage_aggregate": (I, [S, P, P, P, P, P, P, P, P, P]),
        "glx_stage_flatten": (I, [S, P, P, P, P, P, P, P, P, P, P, P, P]),
        "glx_sys_create": (I, [C.POINTER(Params), C.POINTER(P)]),
        "glx_sys_destroy": (I, [P]),
        "glx_sys_upload": (I, [P, S, P, P, P]),
        "glx_sys_generate": (I, [P, P]),
        "glx_sys_snapshot_mass": (I, [P, P]),
        "glx_sys_step": (I, [P, D, C.c_uint64]),
        "glx_sys_snapshot": (I, [P, P, P]),
        "glx_sys_step_timed": (I, [P, D, C.c_uint64, P]),
        "glx_sys_count": (I, [P, C.POINTER(Stats)]),
        "glx_sys_snapshot_accel": (I, [P, P]),
        "glx_sys_time": (I, [P, C.POINTER(D)]),
        "glx_sys_len": (I, [P, C.POINTER(S)]),
        "glx_sys_diag": (I, [P, I, C.POINTER(Diag)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != GLX_OK:
        raise GlxError(rc, lib().glx_last_error().decode())


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def device_info():
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    hbm = C.c_uint64(0)
    _check(lib().glx_device_info(name, 256, C.byref(cus), C.byref(hbm)))
    return {"name": name.value.decode(), "compute_units": cus.value, "hbm_bytes": hbm.value}


def _params(g, softening, theta, precision=PREC_F32, count=False, morton_bits=30,
            walk=WALK_REFERENCE, multipole=0, leaf_size=1, far_theta=0.0,
            mono_ratio=0.0, radix_bits=0, group_size=0, group_lanes=0) -> Params:
    return Params(float(g), float(softening), float(theta), int(leaf_size), int(multipole),
                  int(precision), FLAG_COUNT if count else 0, int(morton_bits), int(walk),
                  float(far_theta), float(mono_ratio), int(radix_bits), int(group_size),
                  int(group_lanes))


@dataclass
class State:
    """Particle state in the reference's SoA f64 layout (core/src/state.rs:29-62)."""

    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    time: float = 0.0

    def __len__(self):
        return int(self.mass.size)


class Lbvh:
    """ForceSolver over the GPU LBVH (Lbvh::new(g, softening, theta),
    solvers/src/lbvh.rs:653). accelerations() mirrors ForceSolver::accelerations."""

    def __init__(self, g: float, softening: float, theta: float, precision: int = PREC_F32,
                 count: bool = False, morton_bits: int = 30, walk: int = WALK_REFERENCE,
                 multipole: int = 0, leaf_size: int = 1, far_theta: float = 0.0,
                 mono_ratio: float = 0.0, radix_bits: int = 0, group_size: int = 0,
                 group_lanes: int = 0):
        self.g, self.softening, self.theta = float(g), float(softening), float(theta)
        h = C.c_void_p()
        p = _params(g, softening, theta, precision, count, morton_bits, walk, multipole,
                    leaf_size, far_theta, mono_ratio, radix_bits, group_size, group_lanes)
        _check(lib().glx_solver_create(C.byref(p), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().glx_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def accelerations(self, pos, mass, acc: np.ndarray | None = None) -> np.ndarray:
        pos = _f64(pos, (-1, 3))
        mass = _f64(mass)
        n = mass.size
        if pos.shape[0] != n:
            raise ValueError("pos and mass lengths differ")
        if acc is None:
            acc = np.zeros((n, 3))
        if acc.shape != (n, 3) or acc.dtype != np.float64 or not acc.flags.c_contiguous:
            raise ValueError("acc length must match particle count")
        _check(lib().glx_solver_accelerations(self._h, n, _p(pos), _p(mass), _p(acc)))
        return acc

    def _active_call(self, fn, pos, mass, active):
        pos = _f64(pos, (-1, 3))
        mass = _f64(mass)
        n = mass.size
        if pos.shape[0] != n:
            raise ValueError("pos and mass lengths differ")
        act = None if active is None else np.ascontiguousarray(active, dtype=np.uintp)
        acc = np.zeros((n, 3))
        _check(fn(self._h, n, _p(pos), _p(mass), None if act is None else _p(act),
                  0 if act is None else act.size, _p(acc)))
        return acc

    def accelerations_active(self, pos, mass, active) -> np.ndarray:
        """ForceSolver::accelerations_active (core/src/traits.rs:43-54): rows in `active`
        equal the full pass; the other rows are unspecified."""
        return self._active_call(lib().glx_solver_accelerations_active, pos, mass, active)

    def rebuild_gravity_cache(self, pos, mass):
        """ForceSolver::rebuild_gravity_cache (core/src/traits.rs:73-78)."""
        pos = _f64(pos, (-1, 3))
        mass = _f64(mass)
        _check(lib().glx_solver_rebuild_cache(self._h, mass.size, _p(pos), _p(mass)))

    def gravity_active_cached(self, pos, mass, active=None) -> np.ndarray:
        """ForceSolver::gravity_active_cached (core/src/traits.rs:80-91): the kept tree
        walked at the current positions for the `active` rows."""
        return self._active_call(lib().glx_solver_gravity_active_cached, pos, mass, active)

    def potential_energy(self, pos, mass) -> float:
        pos = _f64(pos, (-1, 3))
        mass = _f64(mass)
        out = C.c_double(0.0)
        _check(lib().glx_solver_potential_energy(self._h, mass.size, _p(pos), _p(mass),
                                                 C.byref(out)))
        return out.value

    def stats(self) -> dict:
        st = Stats()
        _check(lib().glx_solver_stats(self._h, C.byref(st)))
        return st.as_dict()


def direct_sum_f64(pos, mass, g: float, softening: float, targets,
                   noalloc: bool = False) -> np.ndarray:
    """FP64 direct sum onto targets (solvers/src/direct_sum.rs:25-45 pair kernel).
    noalloc: through glx_direct_sum_f64_noalloc (no device allocation; same bits)."""
    pos = _f64(pos, (-1, 3))
    mass = _f64(mass)
    tg = _u32(targets)
    out = np.zeros((tg.size, 3))
    f = lib().glx_direct_sum_f64_noalloc if noalloc else lib().glx_direct_sum_f64
    _check(f(mass.size, _p(pos), _p(mass), float(g), float(softening), _p(tg), tg.size,
             _p(out)))
    return out


# ------------------------------------------------------------------ stage exports


@dataclass
class MortonOut:
    bmin: np.ndarray
    size: float
    scale: float
    lanes: np.ndarray
    codes: np.ndarray


def morton(pos, bits: int = 30) -> MortonOut:
    """GPU counterpart of reference_morton (solvers/src/lbvh.rs:105-129); bits=63 for
    the 2^21-cells-per-axis refinement."""
    pos = _f64(pos, (-1, 3))
    n = pos.shape[0]
    b = np.zeros(5)
    lanes = np.zeros((n, 3), np.uint32)
    if bits == 63:
        codes = np.zeros(n, np.uint64)
        _check(lib().glx_stage_morton63(n, _p(pos), _p(b), _p(lanes), _p(codes)))
    else:
        codes = np.zeros(n, np.uint32)
        _check(lib().glx_stage_morton(n, _p(pos), _p(b), _p(lanes), _p(codes)))
    return MortonOut(b[:3].copy(), float(b[3]), float(b[4]), lanes, codes)


def sort(codes, radix_bits: int = 0):
    """GPU counterpart of reference_sort (solvers/src/lbvh.rs:143-147); uint64 codes
    select the 63-bit key sort. radix_bits (8, 9 or 10; 0 = 8) sets the digit width."""
    if np.asarray(codes).dtype == np.uint64:
        c = np.ascontiguousarray(codes, dtype=np.uint64)
        order = np.zeros(c.size, np.uint32)
        sc = np.zeros(c.size, np.uint64)
        if radix_bits:
            _check(lib().glx_stage_sort63_bits(c.size, _p(c), _p(order), _p(sc), radix_bits))
        else:
            _check(lib().glx_stage_sort63(c.size, _p(c), _p(order), _p(sc)))
        return order, sc
    codes = _u32(codes)
    order = np.zeros(codes.size, np.uint32)
    sc = np.zeros(codes.size, np.uint32)
    if radix_bits:
        _check(lib().glx_stage_sort_bits(codes.size, _p(codes), _p(order), _p(sc), radix_bits))
    else:
        _check(lib().glx_stage_sort(codes.size, _p(codes), _p(order), _p(sc)))
    return order, sc


def karras(sorted_codes):
    """GPU counterpart of reference_karras (solvers/src/lbvh.rs:276-306); uint64 codes
    select the 63-bit build."""
    wide = np.asarray(sorted_codes).dtype == np.uint64
    sc = np.ascontiguousarray(sorted_codes, dtype=np.uint64) if wide else _u32(sorted_codes)
    n = sc.size
    left = np.zeros(max(n - 1, 0), np.uint32)
    right = np.zeros(max(n - 1, 0), np.uint32)
    parent = np.zeros(max(2 * n - 1, 0), np.uint32)
    f = lib().glx_stage_karras63 if wide else lib().glx_stage_karras
    _check(f(n, _p(sc), _p(left), _p(right), _p(parent)))
    return left, right, parent


def aggregate(sorted_codes, sorted_pos, sorted_mass, quad: bool = False):
    """GPU counterpart of reference_aggregate (solvers/src/lbvh.rs:329-379): per unified
    node AABB/com/mass in f32 relative to `origin` (f64); with quad, also the traceless
    quadrupole {xx, yy, zz, xy, xz, yz} about each node's centre of mass."""
    sc = _u32(sorted_codes)
    sp = _f64(sorted_pos, (-1, 3))
    sm = _f64(sorted_mass)
    n = sc.size
    t = 2 * n - 1
    origin = np.zeros(3)
    out = {
        "aabb_min": np.zeros((t, 3), np.float32),
        "aabb_max": np.zeros((t, 3), np.float32),
        "com": np.zeros((t, 3), np.float32),
        "mass": np.zeros(t, np.float32),
    }
    q = np.zeros((t, 6), np.float32) if quad else None
    _check(lib().glx_stage_aggregate(n, _p(sc), _p(sp), _p(sm), _p(origin),
                                     _p(out["aabb_min"]), _p(out["aabb_max"]), _p(out["com"]),
                                     _p(out["mass"]), None if q is None else _p(q)))
    out["origin"] = origin
    if quad:
        out["quad"] = q
    return out


def flatten(sorted_codes, sorted_pos, sorted_mass):
    """GPU counterpart of reference_flatten (solvers/src/lbvh.rs:545-614) over the tree of
    the sorted inputs: per pre-order slot next / body_start / body_count and f32 geometry
    (center, half, com relative to `origin`; mass; delta). leaf_bodies = order[body_start]
    at the leaf slots."""
    sc = _u32(sorted_codes)
    sp = _f64(sorted_pos, (-1, 3))
    sm = _f64(sorted_mass)
    n = sc.size
    t = 2 * n - 1
    origin = np.zeros(3)
    out = {k: np.zeros(t, np.uint32) for k in ("next", "body_start", "body_count")}
    out.update({k: np.zeros((t, 3), np.float32) for k in ("center", "half", "com")})
    out.update({k: np.zeros(t, np.float32) for k in ("mass", "delta")})
    _check(lib().glx_stage_flatten(n, _p(sc), _p(sp), _p(sm), _p(origin), _p(out["next"]),
                                   _p(out["body_start"]), _p(out["body_count"]),
                                   _p(out["center"]), _p(out["half"]), _p(out["com"]),
                                   _p(out["mass"]), _p(out["delta"])))
    out["origin"] = origin
    return out


# ------------------------------------------------------------------ resident stepper


class GpuResidentLeapfrog:
    """Resident KDK stepping (gpu/src/gpu_resident.rs:1246-1968): upload -> step* ->
    snapshot. Positions FP64 on the device, velocities and forces FP32."""

    def __init__(self, g: float, softening: float, theta: float, precision: int = PREC_F32,
                 morton_bits: int = 30, walk: int = WALK_REFERENCE, multipole: int = 0,
                 leaf_size: int = 1, far_theta: float = 0.0, mono_ratio: float = 0.0,
                 radix_bits: int = 0, group_size: int = 0, group_lanes: int = 0):
        h = C.c_void_p()
        p = _params(g, softening, theta, precision, False, morton_bits, walk, multipole,
                    leaf_size, far_theta, mono_ratio, radix_bits, group_size, group_lanes)
        _check(lib().glx_sys_create(C.byref(p), C.byref(h)))
        self._h = h
        self._mass = np.zeros(0)

    def close(self):
        if getattr(self, "_h", None):
            lib().glx_sys_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, pos, vel, mass):
        pos = _f64(pos, (-1, 3))
        vel = _f64(vel, (-1, 3))
        mass = _f64(mass)
        n = mass.size
        if pos.shape[0] != n or vel.shape[0] != n:
            raise ValueError("pos/vel/mass lengths differ")
        self._mass = mass.copy()
        _check(lib().glx_sys_upload(self._h, n, _p(pos), _p(vel), _p(mass)))

    def generate(self, spec):
        """upload(sample(...)) with the state drawn on the device (`galaxy_amd.ic.Spec`,
        e.g. `ic.spec_collision(n, seed)`): no host arrays, no upload."""
        _check(lib().glx_sys_generate(self._h, C.addressof(spec)))
        self._mass = np.zeros(len(self))
        _check(lib().glx_sys_snapshot_mass(self._h, _p(self._mass)))

    def step(self, dt: float):
        _check(lib().glx_sys_step(self._h, float(dt), 1))

    def step_many(self, dt: float, steps: int):
        _check(lib().glx_sys_step(self._h, float(dt), int(steps)))

    PHASES = ("kick_drift_bbox", "encode", "sort", "gather", "karras", "aggregate", "walk")

    def step_timed(self, dt: float, steps: int) -> dict:
        """Direct launches with HIP events between phases; returns device ms per phase
        summed over the steps."""
        ph = np.zeros(len(self.PHASES))
        _check(lib().glx_sys_step_timed(self._h, float(dt), int(steps), _p(ph)))
        return dict(zip(self.PHASES, ph.tolist()))

    def count(self) -> dict:
        st = Stats()
        _check(lib().glx_sys_count(self._h, C.byref(st)))
        return st.as_dict()

    def snapshot(self) -> State:
        n = len(self)
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        _check(lib().glx_sys_snapshot(self._h, _p(pos), _p(vel)))
        return State(pos, vel, self._mass.copy(), self.time())

    def snapshot_accel(self) -> np.ndarray:
        acc = np.zeros((len(self), 3))
        _check(lib().glx_sys_snapshot_accel(self._h, _p(acc)))
        return acc

    def time(self) -> float:
        t = C.c_double(0.0)
        _check(lib().glx_sys_time(self._h, C.byref(t)))
        return t.value

    def __len__(self) -> int:
        n = C.c_size_t(0)
        _check(lib().glx_sys_len(self._h, C.byref(n)))
        return n.value

    def diag(self, with_potential: bool = False) -> dict:
        d = Diag()
        _check(lib().glx_sys_diag(self._h, 1 if with_potential else 0, C.byref(d)))
        return {
            "kinetic": d.kinetic,
            "potential": d.potential,
            "momentum": np.array(d.momentum[:]),
            "angular_momentum": np.array(d.angular_momentum[:]),
            "center_of_mass": np.array(d.center_of_mass[:]),
            "mass": d.mass,
        }
