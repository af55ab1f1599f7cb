"""GSD writer and restart through the harness.

A frame holds the state rounded to the schema's float32 fields and the
step.  A simulation created from it continues exactly as one created from a
snapshot of those rounded values at that step: the DPD random forces are
keyed on (seed, timestep, tags), so the restored step is what makes the
continuation identical.
"""

import numpy as np
import pytest

from gpu_util import jittered_lattice, say

pytestmark = pytest.mark.gpu


def _dpd(snap, seed=9, timestep=None, from_gsd=None):
    import azplugins_amd.harness as hoomd

    sim = hoomd.Simulation(seed=seed)
    if from_gsd is not None:
        sim.create_state_from_gsd(from_gsd)
    else:
        if timestep is not None:
            sim.timestep = timestep
        sim.create_state_from_snapshot(snap)
    integ = hoomd.md.Integrator(dt=0.01)
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    p = hoomd.azplugins.pair.DPDGeneralWeight(nlist=nl, kT=1.0, default_r_cut=1.0)
    p.params[("A", "A")] = dict(A=25.0, gamma=4.5, s=0.5)
    integ.forces = [p]
    integ.methods = [hoomd.md.methods.ConstantVolume(hoomd.filter.All())]
    sim.operations.integrator = integ
    return sim


def _state(sim):
    s = sim.state.get_snapshot().particles
    return s.position.copy(), s.velocity.copy(), s.image.copy()


def test_dpd_restart_from_gsd_is_exact(tmp_path):
    import azplugins_amd.harness as hoomd

    L = 8.0
    pos4, rng = jittered_lattice(12, L, 0.05, 1, seed=41)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = pos4.shape[0]
    snap.particles.position[:] = pos4[:, :3]
    snap.particles.velocity[:] = rng.normal(size=(pos4.shape[0], 3))
    fn = str(tmp_path / "traj.gsd")
    say(f"[gsd] DPD N={pos4.shape[0]} restart at step 40")
    sim = _dpd(snap)
    writer = hoomd.write.GSD(hoomd.trigger.Periodic(20), fn, mode="wb", dynamic=["property", "momentum"])
    sim.operations.writers.append(writer)
    sim.run(40)
    writer.close()
    ref_pos, ref_vel, ref_img = _state(sim)
    with hoomd.gsd.open(fn, "rb") as t:
        assert len(t) == 2
        assert [int(f.configuration.step) for f in t] == [20, 40]
        last = t[-1]
    # the frame is the state at step 40 in float32, in tag order
    assert np.array_equal(last.particles.position, ref_pos.astype(np.float32))
    assert np.array_equal(last.particles.velocity, ref_vel.astype(np.float32))
    assert np.array_equal(last.particles.image, ref_img)

    a = _dpd(None, from_gsd=fn)
    assert a.timestep == 40
    a.run(30)
    b = _dpd(hoomd.Snapshot.from_gsd_frame(last), timestep=40)
    b.run(30)
    c = _dpd(hoomd.Snapshot.from_gsd_frame(last), timestep=0)
    c.run(30)
    sa, sb, sc = _state(a), _state(b), _state(c)
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    # the step matters: the same state at step 0 draws other random forces
    assert not np.array_equal(sa[1], sc[1])
    # and the rounded restart stays close to the unrounded run
    sim.run(30)
    full = _state(sim)
    d = full[0] - sa[0] + L * (full[2] - sa[2])
    say(f"[gsd] restart vs uninterrupted after 30 steps: max |dx| {np.abs(d).max():.3e}")
    assert np.abs(d).max() < 1e-3


def test_write_static_and_create_state(tmp_path):
    import azplugins_amd.harness as hoomd

    L = 9.0
    pos4, rng = jittered_lattice(9, L, 0.1, 2, seed=5)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = pos4.shape[0]
    snap.particles.types = ["A", "B"]
    snap.particles.position[:] = pos4[:, :3]
    snap.particles.typeid[:] = pos4[:, 3].astype(np.uint32)
    snap.particles.mass[:] = rng.uniform(1, 2, pos4.shape[0])
    snap.particles.diameter[:] = rng.uniform(0.5, 1.5, pos4.shape[0])
    sim = hoomd.Simulation(seed=1)
    sim.timestep = 1234
    sim.create_state_from_snapshot(snap)
    fn = str(tmp_path / "one.gsd")
    hoomd.write.GSD.write(state=sim.state, filename=fn)
    sim2 = hoomd.Simulation(seed=1)
    sim2.create_state_from_gsd(fn)
    assert sim2.timestep == 1234
    s2 = sim2.state.get_snapshot().particles
    assert s2.types == ["A", "B"]
    assert np.array_equal(s2.typeid, snap.particles.typeid)
    for name in ("position", "mass", "diameter"):
        assert np.array_equal(getattr(s2, name), getattr(snap.particles, name).astype(np.float32)), name


def test_writer_dynamic_categories_and_truncate(tmp_path):
    """Later frames hold the step and the dynamic categories only; other
    chunks read back from frame 0.  truncate=True keeps one frame."""
    import azplugins_amd.harness as hoomd

    L = 8.0
    pos4, rng = jittered_lattice(8, L, 0.05, 1, seed=7)
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = pos4.shape[0]
    snap.particles.position[:] = pos4[:, :3]
    snap.particles.velocity[:] = rng.normal(size=(pos4.shape[0], 3))
    snap.particles.mass[:] = 2.0
    sim = _dpd(snap)
    fn = str(tmp_path / "dyn.gsd")
    fn_t = str(tmp_path / "last.gsd")
    w = hoomd.write.GSD(hoomd.trigger.Periodic(5), fn, mode="wb")  # dynamic: property only
    wt = hoomd.write.GSD(hoomd.trigger.Periodic(5), fn_t, mode="wb", truncate=True,
                         dynamic=["property", "momentum"])
    sim.operations.writers += [w, wt]
    say(f"[gsd] writer categories N={pos4.shape[0]}")
    sim.run(15)
    w.close()
    wt.close()
    last = sim.state.get_snapshot().particles
    with hoomd.gsd.open(fn, "rb") as t:
        assert len(t) == 3
        f = t.file
        assert f.chunk_exists(0, "particles/velocity") and f.chunk_exists(0, "particles/mass")
        assert not f.chunk_exists(2, "particles/velocity") and not f.chunk_exists(2, "particles/mass")
        assert f.chunk_exists(2, "particles/position")
        fr0, fr2 = t[0], t[2]
    assert fr2.configuration.step == 15
    assert np.array_equal(fr2.particles.position, last.position.astype(np.float32))
    # velocity is not a "property": frame 2 reads frame 0's
    assert np.array_equal(fr2.particles.velocity, fr0.particles.velocity)
    assert np.all(fr2.particles.mass == np.float32(2.0))
    with hoomd.gsd.open(fn_t, "rb") as t:
        assert len(t) == 1
        only = t[0]
    assert only.configuration.step == 15
    assert np.array_equal(only.particles.velocity, last.velocity.astype(np.float32))
