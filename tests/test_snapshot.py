"""CPU tests of the GLXYSNAP snapshot stream (host code of libgalaxy_lbvh.so).

The expected bytes come from an independent restatement of the layout in
io/src/snapshot.rs:1-37,130-194 written with `struct`; the cases follow
io/tests/roundtrip.rs (round trips, version 4 on write, bad magic, unsupported versions,
truncation, v1/v2/v3 streams with defaults, invalid species, empty state, count taken
from the data)."""
from __future__ import annotations

import struct

import numpy as np
import pytest

from galaxy_amd import sim
from galaxy_amd.lbvh import GLX_EINVAL, GlxError


def sample(n=7, seed=3):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * 10
    vel = rng.normal(size=(n, 3))
    mass = rng.uniform(0.1, 2.0, n)
    cols = dict(id=np.arange(100, 100 + n, dtype=np.uint64) * 7,
                progenitor=np.arange(n, dtype=np.uint16) % 3,
                kind=(np.arange(n) % 2).astype(np.uint8),
                u=rng.uniform(0, 5, n),
                formation_time=np.where(np.arange(n) % 3 == 0, -np.inf, rng.uniform(0, 1, n)))
    return pos, vel, mass, cols


HDR = dict(time=1.25, step=42, scale_factor=1.0, softening=0.05, rng_seed=99,
           config_hash=0xDEADBEEF12345678, units="nbody-G1", code_version="test 1")


def expected_bytes(pos, vel, mass, cols, version=4, n_header=None, **h):
    n = mass.size
    b = b"GLXYSNAP" + struct.pack("<I", version)
    b += struct.pack("<dQddQQQ", h["time"], h["step"], h["scale_factor"], h["softening"],
                     n if n_header is None else n_header, h["rng_seed"], h["config_hash"])
    for s in (h["units"], h["code_version"]):
        e = s.encode()
        b += struct.pack("<I", len(e)) + e
    b += b"".join(struct.pack("<3d", *p) for p in pos)
    b += b"".join(struct.pack("<3d", *v) for v in vel)
    b += b"".join(struct.pack("<f", m) for m in mass)
    b += b"".join(struct.pack("<Q", int(i)) for i in cols["id"])
    b += b"".join(struct.pack("<H", int(p)) for p in cols["progenitor"])
    if version >= 2:
        b += bytes(int(k) for k in cols["kind"])
    if version >= 3:
        b += b"".join(struct.pack("<d", u) for u in cols["u"])
    if version >= 4:
        b += b"".join(struct.pack("<d", t) for t in cols["formation_time"])
    return b


def test_writer_emits_the_reference_layout_at_version_4(tmp_path):
    pos, vel, mass, cols = sample()
    p = tmp_path / "a.snap"
    sim.write_snapshot(p, pos, vel, mass, **HDR, **cols)
    raw = p.read_bytes()
    assert raw[:8] == b"GLXYSNAP" and struct.unpack("<I", raw[8:12])[0] == 4
    assert raw == expected_bytes(pos, vel, mass, cols, **HDR)


def test_round_trip_recovers_header_and_columns(tmp_path):
    pos, vel, mass, cols = sample()
    p = tmp_path / "a.snap"
    sim.write_snapshot(p, pos, vel, mass, **HDR, **cols)
    s = sim.read_snapshot(p)
    for k, v in HDR.items():
        assert getattr(s, k) == v, k
    assert s.version == 4
    assert np.array_equal(s.pos, pos) and np.array_equal(s.vel, vel)
    assert np.array_equal(s.mass, mass.astype(np.float32).astype(np.float64))  # the lossy field
    for k in ("id", "progenitor", "kind", "u", "formation_time"):
        assert np.array_equal(getattr(s, k), cols[k]), k


def test_default_columns(tmp_path):
    pos, vel, mass, _ = sample(5)
    p = tmp_path / "d.snap"
    sim.write_snapshot(p, pos, vel, mass, **HDR)
    s = sim.read_snapshot(p)
    assert np.array_equal(s.id, np.arange(5))
    assert not s.progenitor.any() and not s.kind.any() and not s.u.any()
    assert np.all(np.isneginf(s.formation_time))


def test_empty_state_round_trips(tmp_path):
    p = tmp_path / "e.snap"
    sim.write_snapshot(p, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), **HDR)
    s = sim.read_snapshot(p)
    assert s.pos.shape == (0, 3) and s.step == 42


def _code(fn):
    with pytest.raises(GlxError) as e:
        fn()
    return e.value.code


def test_bad_magic_unsupported_version_and_garbage(tmp_path):
    pos, vel, mass, cols = sample()
    good = expected_bytes(pos, vel, mass, cols, **HDR)
    p = tmp_path / "x.snap"
    p.write_bytes(b"NOTASNAP" + good[8:])
    assert _code(lambda: sim.read_snapshot(p)) == sim.GLX_EFORMAT
    for v in (0, 5, 99):
        p.write_bytes(good[:8] + struct.pack("<I", v) + good[12:])
        assert _code(lambda: sim.read_snapshot(p)) == sim.GLX_EFORMAT
    for junk in (b"", b"GLXY", bytes(range(40))):
        p.write_bytes(junk)
        assert _code(lambda: sim.read_snapshot(p)) == sim.GLX_EFORMAT
    assert _code(lambda: sim.read_snapshot(tmp_path / "missing.snap")) == sim.GLX_EIO


def test_truncated_stream_is_an_error_not_a_crash(tmp_path):
    pos, vel, mass, cols = sample()
    good = expected_bytes(pos, vel, mass, cols, **HDR)
    p = tmp_path / "t.snap"
    for cut in (12, 30, 60, 100, len(good) // 2, len(good) - 1):
        p.write_bytes(good[:cut])
        assert _code(lambda: sim.read_snapshot(p)) == sim.GLX_EFORMAT, cut


def test_count_comes_from_the_data(tmp_path):
    import ctypes as C
    import os

    pos, vel, mass, cols = sample(4)
    p = tmp_path / "c.snap"
    h = sim.SnapHeader(1.0, 3, 1.0, 0.05, 999, 1, 2, b"u", b"v")  # header claims 999
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    assert sim._bind().glx_snapshot_write(os.fsencode(p), C.byref(h), 4, ptr(pos), ptr(vel),
                                          ptr(mass), None) == 0
    raw = p.read_bytes()
    assert struct.unpack("<Q", raw[12 + 32:12 + 40])[0] == 4
    assert sim.read_snapshot(p).pos.shape == (4, 3)


@pytest.mark.parametrize("version", [1, 2, 3])
def test_older_versions_read_with_defaults(tmp_path, version):
    pos, vel, mass, cols = sample()
    p = tmp_path / f"v{version}.snap"
    p.write_bytes(expected_bytes(pos, vel, mass, cols, version=version, **HDR))
    s = sim.read_snapshot(p)
    assert s.version == version and np.array_equal(s.pos, pos)
    if version == 1:
        assert not s.kind.any()
    else:
        assert np.array_equal(s.kind, cols["kind"])
    if version < 3:
        assert not s.u.any()
    else:
        assert np.array_equal(s.u, cols["u"])
    assert np.all(np.isneginf(s.formation_time))


def test_invalid_species_and_long_strings(tmp_path):
    pos, vel, mass, cols = sample()
    raw = bytearray(expected_bytes(pos, vel, mass, cols, **HDR))
    n = mass.size
    kind_at = len(raw) - n * 16 - n  # kind column precedes u and formation_time
    raw[kind_at] = 7
    p = tmp_path / "k.snap"
    p.write_bytes(bytes(raw))
    assert _code(lambda: sim.read_snapshot(p)) == sim.GLX_EFORMAT
    h = dict(HDR, units="x" * 70000)
    assert _code(lambda: sim.write_snapshot(tmp_path / "l.snap", pos, vel, mass, **h)) == GLX_EINVAL
    bad = dict(cols, kind=np.full(mass.size, 2, np.uint8))
    assert _code(lambda: sim.write_snapshot(tmp_path / "m.snap", pos, vel, mass, **HDR,
                                            **bad)) == GLX_EINVAL


def _with_units(pos, vel, mass, cols, units: bytes):
    raw = expected_bytes(pos, vel, mass, cols, **HDR)
    at = 12 + 56  # magic + version, then the 56-byte fixed header
    old = len(HDR["units"].encode())
    return raw[:at] + struct.pack("<I", len(units)) + units + raw[at + 4 + old:]


@pytest.mark.parametrize("units", [b"\xff", b"ab\xc0\xaf", b"\xed\xa0\x80", b"\xf4\x90\x80\x80",
                                   b"x\xe2\x82", b"\x80abc"],
                         ids=["ff", "overlong", "surrogate", "above_10ffff", "cut", "lone_cont"])
def test_header_strings_must_be_utf8(tmp_path, units):
    """io/src/snapshot.rs read_str decodes with String::from_utf8: a stream whose header
    strings are not UTF-8 is refused, and the writer (which the reference feeds &str) will
    not produce one."""
    import ctypes as C
    import os

    pos, vel, mass, cols = sample()
    p = tmp_path / "u.snap"
    p.write_bytes(_with_units(pos, vel, mass, cols, units))
    assert _code(lambda: sim.read_snapshot(p)) == sim.GLX_EFORMAT
    h = sim.SnapHeader(1.0, 3, 1.0, 0.05, 0, 1, 2, units, b"v")
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    q = tmp_path / "w.snap"
    assert sim._bind().glx_snapshot_write(os.fsencode(q), C.byref(h), mass.size, ptr(pos),
                                          ptr(vel), ptr(mass), None) == GLX_EINVAL


def test_multibyte_utf8_header_strings_round_trip(tmp_path):
    pos, vel, mass, cols = sample()
    units = "Msun·kpc·Gyr — ∑ 𝔊"
    p = tmp_path / "m.snap"
    p.write_bytes(_with_units(pos, vel, mass, cols, units.encode()))
    assert sim.read_snapshot(p).units == units
