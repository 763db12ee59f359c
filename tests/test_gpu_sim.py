"""GPU tests of the stepping driver glx_sys_run (sim::run, sim/src/lib.rs:110-174).

Cadence as in the reference (the IC, every snapshot_every steps, always the final step;
sim/tests cover the same rules), snapshot contents identical to the stepper's own state
after the same number of steps, header times, config validation, and a resume from a
snapshot that continues bit for bit (masses are stored as f32, and the walk reads f32
masses, so the reloaded system builds the same tree and forces)."""
from __future__ import annotations

import numpy as np
import pytest

import galaxy_amd as glx
from galaxy_amd import ic, sim
from galaxy_amd.lbvh import GLX_EINVAL, GlxError

pytestmark = pytest.mark.gpu
MODES = {
    "reference": dict(morton_bits=30, walk=glx.WALK_REFERENCE, multipole=0),
    "fast": dict(morton_bits=63, walk=glx.WALK_GROUP, multipole=2),
}


def files(d):
    return sorted(p.name for p in d.iterdir())


@pytest.mark.parametrize("n_steps,every,want", [
    (10, 3, [0, 3, 6, 9, 10]),
    (9, 3, [0, 3, 6, 9]),
    (0, 5, [0]),
    (4, 10, [0, 4]),
    (5, 1, [0, 1, 2, 3, 4, 5]),
])
def test_cadence(tmp_path, n_steps, every, want):
    pos, vel, mass = ic.plummer(500, 1)
    r = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5, **MODES["fast"])
    r.upload(pos, vel, mass)
    out = sim.run(r, tmp_path, dt=1e-3, n_steps=n_steps, snapshot_every=every, softening=0.01)
    assert files(tmp_path) == [f"snapshot_{s:08d}.snap" for s in want]
    assert out["snapshots_emitted"] == len(want) and out["steps"] == n_steps
    assert out["final_time"] == pytest.approx(n_steps * 1e-3, abs=1e-15)


@pytest.mark.parametrize("mode", ["reference", "fast"])
def test_snapshots_equal_the_stepped_state(tmp_path, mode):
    pos, vel, mass = ic.plummer(3000, 2)
    a = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5, **MODES[mode])
    a.upload(pos, vel, mass)
    sim.run(a, tmp_path, dt=1e-3, n_steps=7, snapshot_every=3, softening=0.01, rng_seed=5,
            config_hash=11, units="nbody-G1")
    b = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5, **MODES[mode])
    b.upload(pos, vel, mass)
    prev, t = 0, 0.0
    for step in (0, 3, 6, 7):
        k = step - prev
        if k:
            b.step_many(1e-3, k)
            for _ in range(k):
                t += 1e-3
        prev = step
        st = b.snapshot()
        s = sim.read_snapshot(tmp_path / f"snapshot_{step:08d}.snap")
        assert s.step == step and s.time == t
        assert s.softening == 0.01 and s.rng_seed == 5 and s.config_hash == 11
        assert s.units == "nbody-G1" and s.version == 4
        assert np.array_equal(s.pos, st.pos) and np.array_equal(s.vel, st.vel)
        assert np.array_equal(s.mass, mass.astype(np.float32).astype(np.float64))
        assert np.array_equal(s.id, np.arange(3000)) and np.all(np.isneginf(s.formation_time))


def test_resume_from_a_snapshot_is_bit_exact(tmp_path):
    pos, vel, mass = ic.plummer(4000, 3)
    mass = mass.astype(np.float32).astype(np.float64)
    full = tmp_path / "full"
    part = tmp_path / "part"
    full.mkdir()
    part.mkdir()
    a = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5, **MODES["fast"])
    a.upload(pos, vel, mass)
    sim.run(a, full, dt=2e-3, n_steps=10, snapshot_every=4, softening=0.01)
    s4 = sim.read_snapshot(full / "snapshot_00000004.snap")
    b = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5, **MODES["fast"])
    b.upload(s4.pos, s4.vel, s4.mass)
    sim.set_time(b, s4.time)
    sim.run(b, part, dt=2e-3, n_steps=6, snapshot_every=4, softening=0.01, first_step=4)
    assert files(part) == ["snapshot_00000004.snap", "snapshot_00000008.snap",
                           "snapshot_00000010.snap"]
    for step in (8, 10):
        x = sim.read_snapshot(full / f"snapshot_{step:08d}.snap")
        y = sim.read_snapshot(part / f"snapshot_{step:08d}.snap")
        assert np.array_equal(x.pos, y.pos) and np.array_equal(x.vel, y.vel), step
        assert x.time == y.time


def test_columns_pass_through(tmp_path):
    pos, vel, mass = ic.plummer(200, 4)
    r = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5)
    r.upload(pos, vel, mass)
    ids = np.arange(200, dtype=np.uint64) * 3 + 1
    prog = (np.arange(200) % 2).astype(np.uint16)
    sim.run(r, tmp_path, dt=1e-3, n_steps=2, snapshot_every=2, softening=0.01, id=ids,
            progenitor=prog)
    s = sim.read_snapshot(tmp_path / "snapshot_00000002.snap")
    assert np.array_equal(s.id, ids) and np.array_equal(s.progenitor, prog)


def test_config_validation(tmp_path):
    pos, vel, mass = ic.plummer(100, 5)
    r = glx.GpuResidentLeapfrog(1.0, 0.01, 0.5)
    r.upload(pos, vel, mass)
    for kw in (dict(dt=1e-3, snapshot_every=0), dict(dt=0.0, snapshot_every=1),
               dict(dt=-1e-3, snapshot_every=1), dict(dt=float("nan"), snapshot_every=1)):
        with pytest.raises(GlxError) as e:
            sim.run(r, tmp_path, n_steps=3, softening=0.01, **kw)
        assert e.value.code == GLX_EINVAL
    with pytest.raises(GlxError) as e:
        sim.run(r, tmp_path / "missing_dir", dt=1e-3, n_steps=1, snapshot_every=1,
                softening=0.01)
    assert e.value.code == sim.GLX_EIO
