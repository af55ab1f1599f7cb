"""Drop-in harness behaviour around the pair forces.

Force computes fill the per-particle virial only when something reads it (a
ThermodynamicQuantities compute, or a force's ``virials``), as HOOMD computes
it only under its pressure-tensor flag; reading ``virials`` of a force that was
computed without it computes it for the current state.
"""

import numpy as np
import pytest

from gpu_util import GpuSystem, gpu_pair, jittered_lattice, say

pytestmark = pytest.mark.gpu


def _sim(pos4, L, precision):
    import azplugins_amd.harness as hoomd

    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = pos4.shape[0]
    snap.particles.position[:] = pos4[:, :3]
    sim = hoomd.Simulation(seed=3)
    sim.create_state_from_snapshot(snap)
    integ = hoomd.md.Integrator(dt=0.005)
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    plj = hoomd.azplugins.pair.PerturbedLennardJones(nlist=nl, default_r_cut=3.0)
    plj.params[("A", "A")] = dict(epsilon=1.0, sigma=1.0, attraction_scale_factor=0.5)
    plj.precision = precision
    integ.forces = [plj]
    integ.methods = [hoomd.md.methods.ConstantVolume(hoomd.filter.All())]
    sim.operations.integrator = integ
    return sim, plj


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_virials_on_demand(precision):
    L = 12.0
    pos4, _ = jittered_lattice(12, L, 0.1, 1, seed=17)
    sim, plj = _sim(pos4, L, precision)
    sim.run(0)
    assert not sim._needs_virial()
    got = plj.virials
    assert sim._needs_virial()
    ref = gpu_pair("PotentialPairPerturbedLennardJones", GpuSystem(pos4, np.array([L, L, L])), 1,
                   {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5}}, 3.0,
                   precision=precision, tiled=(precision == "mixed"))
    say(f"[harness] {precision}: on-demand virial max |v| {np.abs(got).max():.3e}")
    assert np.array_equal(got, ref.virial.T)
    assert np.array_equal(plj.forces, ref.force[:, :3])


def test_pressure_tensor():
    """P_ab = (sum of m v_a v_b + sum of W_ab) / V, and P = trace / 3."""
    import azplugins_amd.harness as hoomd

    L = 10.0
    pos4, rng = jittered_lattice(10, L, 0.1, 1, seed=23)
    sim, plj = _sim(pos4, L, "fp64")
    vel = rng.normal(size=(pos4.shape[0], 3))
    snap = sim.state.get_snapshot()
    snap.particles.velocity[:] = vel
    state_vel = np.zeros((pos4.shape[0], 4))
    state_vel[:, :3] = vel
    state_vel[:, 3] = 1.0
    sim.state.vel.copy_(hoomd.md.torch.tensor(state_vel, device="cuda"))
    thermo = hoomd.md.compute.ThermodynamicQuantities(hoomd.filter.All())
    sim.operations.computes.append(thermo)
    sim.run(0)
    got = np.array(thermo.pressure_tensor)
    K = np.array([(vel[:, a] * vel[:, b]).sum() for a, b in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))])
    W = plj.virials.sum(axis=0)
    ref = (K + W) / L**3
    say(f"[harness] pressure tensor {got}")
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    assert thermo.pressure == pytest.approx((got[0] + got[3] + got[5]) / 3.0, rel=1e-10)


@pytest.mark.parametrize("method", ["nve", "langevin"])
def test_fused_half_steps_are_bit_identical(method):
    """Steps whose second half runs into the next step's first half use one
    fused kernel; the trajectory equals the unfused one bit for bit, also
    across the steps where the sorter acts in between."""
    import azplugins_amd.harness as hoomd

    L = 12.0
    pos4, rng = jittered_lattice(12, L, 0.1, 1, seed=31)
    pos4 = pos4[rng.permutation(pos4.shape[0])]
    runs = []
    for fuse in (True, False):
        sim, plj = _sim(pos4, L, "mixed")
        integ = sim.operations.integrator
        if method == "langevin":
            integ.methods = [hoomd.md.methods.Langevin(hoomd.filter.All(), kT=1.0)]
        integ.fuse_half_steps = fuse
        sim.operations.tuners[0].trigger = hoomd.trigger.Periodic(7)
        sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.0)
        sim.run(20)
        sim.run(5)
        snap = sim.state.get_snapshot()
        runs.append((snap.particles.position.copy(), snap.particles.velocity.copy(), snap.particles.image.copy()))
    say(f"[harness] fused half steps ({method}): max |dx| {np.abs(runs[0][0] - runs[1][0]).max():.3e}")
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision,method", [("mixed", "langevin"), ("mixed", "nve"), ("fp64", "langevin"),
                                              ("fp64", "nve")])
def test_step_with_force_is_bit_identical(precision, method):
    """With one isotropic pair force and one translational method, the fused
    half steps go to the force's library call, whose step kernel also makes
    the next step's rebuild check.  Trajectories and the steps that rebuild
    the list equal those of the separate calls bit for bit, through rebuilds
    and sorter reorders."""
    import azplugins_amd.harness as hoomd

    L = 16.0
    pos4, rng = jittered_lattice(16, L, 0.1, 1, seed=37)
    pos4 = pos4[rng.permutation(pos4.shape[0])]
    runs = []
    for joined in (True, False):
        sim, plj = _sim(pos4, L, precision)
        integ = sim.operations.integrator
        if method == "langevin":
            integ.methods = [hoomd.md.methods.Langevin(hoomd.filter.All(), kT=1.2)]
        integ.step_with_force = joined
        sim.operations.tuners[0].trigger = hoomd.trigger.Periodic(23)
        sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.2)
        sim.run(0)
        cpp, rebuilt = plj._cpp, []
        calls = {"step": 0}

        def counted(fn, key=None):
            def wrapped(*a, **k):
                r = fn(*a, **k)
                rebuilt.append(bool(r))
                if key:
                    calls[key] += 1
                return r
            return wrapped

        cpp.update_compute = counted(cpp.update_compute)
        cpp.update_compute_step = counted(cpp.update_compute_step, "step")
        sim.run(60)
        sim.run(5)
        assert (calls["step"] > 50) == joined
        snap = sim.state.get_snapshot()
        runs.append((snap.particles.position.copy(), snap.particles.velocity.copy(), snap.particles.image.copy(),
                     np.array(rebuilt)))
    say(f"[harness] step with the force ({precision}, {method}): {int(runs[0][3].sum())} rebuilds in "
        f"{runs[0][3].size} updates, max |dx| {np.abs(runs[0][0] - runs[1][0]).max():.3e}")
    assert runs[0][3].sum() >= 3
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def test_buffer_tuner_in_md():
    """The buffer tuner picks one of its grid values during a run; the forces
    with the tuned buffer equal those of a fresh list."""
    import azplugins_amd.harness as hoomd

    L = 16.0
    pos4, rng = jittered_lattice(16, L, 0.1, 1, seed=41)
    sim, plj = _sim(pos4, L, "fp64")
    integ = sim.operations.integrator
    integ.methods = [hoomd.md.methods.Langevin(hoomd.filter.All(), kT=1.0)]
    nl = plj.nlist
    tuner = hoomd.md.tune.NeighborListBuffer.with_grid(trigger=30, nlist=nl, maximum_buffer=0.6,
                                                       minimum_buffer=0.2, n_bins=3, n_rounds=0)
    sim.operations.tuners.append(tuner)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.0)
    sim.run(150)
    say(f"[harness] buffer tuner: {tuner.results} -> {tuner.best_buffer}")
    assert tuner.tuned and tuner.best_buffer in tuner.results
    assert nl.buffer == tuner.best_buffer
    f_run = plj.forces
    snap = sim.state.get_snapshot()
    p4 = np.zeros((snap.particles.N, 4))
    p4[:, :3] = snap.particles.position
    ref = gpu_pair("PotentialPairPerturbedLennardJones", GpuSystem(p4, np.array([L, L, L])), 1,
                   {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5}}, 3.0,
                   r_buff=tuner.best_buffer)
    np.testing.assert_allclose(f_run, ref.force[:, :3], rtol=1e-10, atol=1e-10 * np.abs(ref.force).max())
