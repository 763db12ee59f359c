"""GPU tests of the device workload generators (glx_ic_generate_device, glx_pool_admit_ic,
glx_sys_generate).

The device draws each particle from the same per-particle stream as the host and the
oracle (oracle/ic_oracle.c), and sums the recentring moments in the same fixed order as
the host entry point, so device and host agree bit for bit wherever the device math
library rounds its transcendental functions as the host's does; after rounding to f32
values nearly every coordinate is equal, and the rest differ by an f32 ulp or two. A
rejection test that flips on such a last-bit difference would redraw one particle
entirely: at most a handful of such particles are allowed.
"""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic
from galaxy_amd.pool import Pool
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GB = 1 << 30


def say(*a):
    print("[case]", *a, flush=True)


def close(dev, ref, equal_frac=0.999, ulps=4):
    """dev/ref (n, 3): at least equal_frac of the rows bit-equal, every row but a handful
    within `ulps` f32 ulps."""
    dev, ref = np.asarray(dev), np.asarray(ref)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    far = (np.abs(dev - ref) > ulps * ulp).any(axis=1)
    eq = (dev == ref).all(axis=1)
    assert eq.mean() >= equal_frac, eq.mean()
    assert far.sum() <= max(2, 1e-5 * len(ref)), far.sum()


SPECS = {
    "plummer": lambda n: ic.spec_plummer(n, 31, rmax=10.0),
    "plummer_untruncated": lambda n: ic.spec_plummer(n, 32, rmax=0.0),
    "hernquist": lambda n: ic.spec_hernquist(n, 33, m=2.0, a=0.7),
    "collision": lambda n: ic.spec_collision(n, 34),
    "collision_ratio3": lambda n: ic.spec_collision(n, 35, mass_ratio=3.0, rp=6.0, r0=60.0),
}


@pytest.mark.parametrize("kind", list(SPECS))
@pytest.mark.parametrize("n", [1, 3, 257, 100003])
def test_device_matches_host(kind, n):
    if kind.startswith("collision") and n == 1:
        pytest.skip("a collision needs two particles")
    say("device vs host", kind, n)
    sp = SPECS[kind](n)
    p, v, m = ic.generate_device(sp)
    hp, hv, hm = ic.generate(sp)
    close(p.numpy(), hp)
    close(v.numpy(), hv)
    assert np.array_equal(m.numpy(), hm)


@pytest.mark.parametrize("kind", ["plummer_untruncated", "hernquist", "collision"])
def test_device_matches_oracle(kind):
    n = 65536
    say("device vs oracle", kind, n)
    sp = SPECS[kind](n)
    p, v, m = (t.numpy() for t in ic.generate_device(sp))
    if kind == "collision":
        ref = O.ic_collision(n, sp.seed)[:3]
    elif kind == "hernquist":
        ref = O.ic_sphere(1, n, sp.seed, m=2.0, a=0.7)
    else:
        ref = O.ic_sphere(0, n, sp.seed, rmax=0.0)
    close(p, ref[0], equal_frac=0.99)
    close(v, ref[1], equal_frac=0.99)
    assert np.array_equal(m, ref[2])


def test_device_large_collision_properties():
    """2^21 bodies (the serving workload's upper size): zero COM and momentum, masses,
    pericentre geometry, f32-representable values."""
    n = 1 << 21
    say("device collision", n)
    p, v, m = (t.numpy() for t in ic.generate_device(ic.spec_collision(n, 77)))
    assert np.isfinite(p).all() and np.isfinite(v).all()
    assert abs(m.sum() - 12.66) < 1e-9
    assert np.linalg.norm((m[:, None] * p).sum(0)) < 1e-4 * 12.66
    assert np.linalg.norm((m[:, None] * v).sum(0)) < 1e-5 * 12.66
    h = n // 2
    c1 = (m[:h, None] * p[:h]).sum(0) / m[:h].sum()
    c2 = (m[h:, None] * p[h:]).sum(0) / m[h:].sum()
    assert abs(np.linalg.norm(c2 - c1) - 80.0) < 1e-3
    assert np.array_equal(p, p.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("mode", ["reference", "fast"])
def test_sys_generate_equals_upload_of_the_same_state(mode):
    """glx_sys_generate == glx_sys_upload of the state glx_ic_generate_device writes:
    same masses, bit-identical trajectory."""
    kw = dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2) if mode == "fast" else {}
    sp = ic.spec_collision(20000, 5)
    say("sys generate", mode)
    a = glx.GpuResidentLeapfrog(1.0, 0.05, 0.6, **kw)
    a.generate(sp)
    p, v, m = (t.numpy() for t in ic.generate_device(sp))
    b = glx.GpuResidentLeapfrog(1.0, 0.05, 0.6, **kw)
    b.upload(p, v, m)
    sa, sb = a.snapshot(), b.snapshot()
    assert np.array_equal(sa.mass, m) and np.array_equal(sa.pos, sb.pos)
    assert np.array_equal(sa.vel, sb.vel)
    assert np.array_equal(a.snapshot_accel(), b.snapshot_accel())
    a.step_many(1e-3, 4)
    b.step_many(1e-3, 4)
    sa, sb = a.snapshot(), b.snapshot()
    assert np.array_equal(sa.pos, sb.pos) and np.array_equal(sa.vel, sb.vel)
    a.close()
    b.close()


def test_pool_admit_ic_equals_admit_of_the_same_state():
    """Jobs drawn on the device step bit-identically to jobs uploaded from the same state,
    in one pool with uploaded jobs beside them; budget and slot refusals as admit."""
    specs = [ic.spec_plummer(3000, 1), ic.spec_collision(5001, 2), ic.spec_hernquist(700, 3)]
    say("pool admit_ic")
    pool = Pool(GB, max_jobs=8, morton_bits=63, walk=glx.WALK_GROUP, multipole=2)
    gen, up = [], []
    for sp in specs:
        gen.append(pool.admit_ic(sp, softening=0.05, theta=0.5, dt=1e-3))
        p, v, m = (t.numpy() for t in ic.generate_device(sp))
        g = 1.0 if sp.kind == ic.COLLISION else sp.g
        up.append(pool.admit(p, v, m, g=g, softening=0.05, theta=0.5, dt=1e-3))
    for _ in range(3):
        pool.step()
    for a, b in zip(gen, up):
        pa, va = pool.snapshot(a)
        pb, vb = pool.snapshot(b)
        assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    pool.close()
    small = Pool(Pool.job_bytes(1000), max_jobs=1, morton_bits=63, walk=glx.WALK_GROUP,
                 multipole=2)
    assert small.admit_ic(ic.spec_plummer(2000, 1)) is None  # over budget
    assert small.admit_ic(ic.spec_plummer(1000, 1)) is not None
    assert small.admit_ic(ic.spec_plummer(10, 1)) is None  # no slot
    small.close()


def test_invalid_specs_rejected_on_device_paths():
    say("invalid specs")
    pool = Pool(GB, max_jobs=2)
    with pytest.raises(glx.GlxError):
        pool.admit_ic(ic.spec_collision(1, 1))
    s = glx.GpuResidentLeapfrog(1.0, 0.05, 0.5)
    with pytest.raises(glx.GlxError):
        s.generate(ic.spec_plummer(10, 1, a=0.0))
    pool.close()
    s.close()
