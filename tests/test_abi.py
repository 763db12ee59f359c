"""CPU tests of the C ABI boundary: the library loads, exports every symbol the
public header declares, validates arguments before touching a device, and reports a
missing GPU as an error (never a silent fallback). No compute calls."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

import galaxy_amd
from galaxy_amd import lbvh as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if h.endswith(".h"):
            text = re.sub(r"/\*.*?\*/", "", open(os.path.join(INCLUDE, h)).read(), flags=re.S)
            text = re.sub(r"//[^\n]*", "", text)
            names |= set(re.findall(r"\b(glx_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("glx_solver_create", "glx_solver_accelerations", "glx_stage_morton",
                 "glx_stage_sort", "glx_stage_karras", "glx_stage_aggregate", "glx_sys_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    from galaxy_amd import ic, pool, sim
    pool._bind()
    ic._bind()
    sim._bind()
    for n in declared_functions():
        assert getattr(L.lib(), n).argtypes is not None or n == "glx_last_error", n


def test_pool_arguments_rejected_before_device():
    from galaxy_amd import pool
    lib = pool._bind()
    h = C.c_void_p()
    p = L.Params(1.0, 0.0, 1.0, 1, 0, 0, 0)
    assert lib.glx_pool_create(C.byref(p), 1 << 30, 0, C.byref(h)) == L.GLX_EINVAL
    assert lib.glx_pool_create(None, 1 << 30, 4, C.byref(h)) == L.GLX_EINVAL
    bad = L.Params(1.0, 0.0, 1.0, 1, 0, 0, 0, 31, 0)  # morton_bits
    assert lib.glx_pool_create(C.byref(bad), 1 << 30, 4, C.byref(h)) == L.GLX_EINVAL
    assert lib.glx_pool_step(None, None) == L.GLX_EINVAL
    assert lib.glx_pool_job_bytes(0) == 0
    assert lib.glx_pool_job_bytes(1000) > 1000 * 64
    assert lib.glx_pool_job_bytes(2000) > lib.glx_pool_job_bytes(1000)


def test_invalid_params_rejected_before_device():
    lib = L.lib()
    h = C.c_void_p()
    for bad in (
        L.Params(1.0, 0.05, 0.0, 1, 0, 0, 0),        # theta = 0
        L.Params(1.0, -1.0, 0.5, 1, 0, 0, 0),        # softening < 0
        L.Params(float("nan"), 0.05, 0.5, 1, 0, 0, 0),
        L.Params(1.0, 0.05, 0.5, 4, 0, 0, 0),        # leaf_size > 1 needs the group walk
        L.Params(1.0, 0.05, 0.5, 33, 0, 0, 0, 63, 1),  # leaf_size above 32
        L.Params(1.0, 0.05, 0.5, 1, 1, 0, 0),        # multipole
        L.Params(1.0, 0.05, 0.5, 1, 0, 7, 0),        # precision
        L.Params(1.0, 0.05, 0.5, 1, 0, 0, 0, 63, 1, 0.4),  # far field needs quadrupoles
        L.Params(1.0, 0.05, 0.5, 1, 2, 0, 0, 63, 0, 0.4),  # ... and the group walk
        L.Params(1.0, 0.05, 0.5, 1, 2, 0, 0, 63, 1, 1.0),  # far_theta >= 1
        L.Params(1.0, 0.05, 0.5, 1, 2, 0, 0, 63, 1, float("nan")),
        L.Params(1.0, 0.05, 0.5, 1, 2, 0, 0, 63, 1, 0.0, 1.0),  # mono_ratio >= 1
        L.Params(1.0, 0.05, 0.5, 1, 0, 0, 0, 63, 1, 0.0, 0.5),  # ... needs quadrupoles
        L.Params(1.0, 0.05, 0.5, 1, 2, 0, 0, 63, 1, 0.0, -0.1),
        L.Params(1.0, 0.05, 0.5, 1, 0, 0, 0, 30, 0, 0.0, 0.0, 7),   # radix_bits below 8
        L.Params(1.0, 0.05, 0.5, 1, 0, 0, 0, 30, 0, 0.0, 0.0, 11),  # ... above 10
        L.Params(1.0, 0.05, 0.5, 1, 0, 0, 0, 63, 1, 0.0, 0.0, 0, 4),   # group_size below 8
        L.Params(1.0, 0.05, 0.5, 1, 0, 0, 0, 63, 1, 0.0, 0.0, 0, 65),  # ... above 64
        L.Params(1.0, 0.05, 0.5, 1, 0, 0, 0, 63, 1, 0.0, 0.0, 0, 0, 5),  # group_lanes above 4
    ):
        assert lib.glx_solver_create(C.byref(bad), C.byref(h)) == L.GLX_EINVAL
        assert lib.glx_last_error()
    assert lib.glx_solver_create(None, C.byref(h)) == L.GLX_EINVAL


def test_sort_width_arguments_rejected_before_device():
    """A digit width outside 8..10, or a code beyond the bits its passes cover (30 for
    10-bit digits over 32-bit codes, 63 for 9-bit digits over 64-bit codes), is an
    argument error, found before any device work."""
    lib = L.lib()
    o32 = np.zeros(4, np.uint32)
    o64 = np.zeros(4, np.uint64)
    ok32 = np.array([1, 2, 3, 4], np.uint32)
    for bits in (7, 11):
        assert lib.glx_stage_sort_bits(4, L._p(ok32), L._p(o32), None, bits) == L.GLX_EINVAL
    wide32 = np.array([1, 2, 1 << 30, 4], np.uint32)
    assert lib.glx_stage_sort_bits(4, L._p(wide32), L._p(o32), None, 10) == L.GLX_EINVAL
    wide64 = np.array([1, 2, 1 << 63, 4], np.uint64)
    assert lib.glx_stage_sort63_bits(4, L._p(wide64), L._p(o32), L._p(o64), 9) == L.GLX_EINVAL
    assert "passes" in lib.glx_last_error().decode()


def test_missing_gpu_is_an_error_not_a_fallback():
    name = C.create_string_buffer(64)
    if L.lib().glx_device_info(name, 64, None, None) == L.GLX_OK:
        pytest.skip("a GPU is present")
    with pytest.raises(L.GlxError) as e:
        galaxy_amd.Lbvh(1.0, 0.05, 0.5)
    assert e.value.code == L.GLX_EDEVICE


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(ImportError):
        L.lib()


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Every ctypes mirror of a header struct has the C compiler's size and field offsets
    (field names as in the header)."""
    import subprocess

    from galaxy_amd import ic, pool, sim

    pairs = [(L.Params, "glx_params"), (L.Diag, "glx_diag"), (L.Stats, "glx_stats"),
             (pool.Job, "glx_job"), (pool.JobInfo, "glx_job_info"),
             (pool.PoolInfo, "glx_pool_info"), (ic.Spec, "glx_ic_spec"),
             (sim.SnapHeader, "glx_snap_header"), (sim.SnapColumns, "glx_snap_columns"),
             (sim.RunConfig, "glx_run_config"), (sim.RunSummary, "glx_run_summary")]
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "galaxy_lbvh.h"',
             '#include "galaxy_ic.h"', '#include "galaxy_sim.h"', "int main(void) {"]
    expect = []
    for cls, cname in pairs:
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(cls))
        for f in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {f[0]}));')
            expect.append(getattr(cls, f[0]).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == expect
