"""Wall potentials on the GPU against the oracle, and through the drop-in API.

The kernel is FP64 with the evaluators' operations in the reference's order;
it differs from the oracle only where the device sqrt/log round differently
from the host's (relative 1e-12 below)."""

import numpy as np
import pytest

import oracle
from gpu_util import GpuSystem, say

pytestmark = pytest.mark.gpu


def _geometry():
    spheres = [[0.3, -0.2, 0.1, 5.5, 1.0, 1.0], [2.0, 2.0, 2.0, 1.0, 0.0, 0.0]]
    cylinders = [[0.0, 0.5, 0.0, 0.2, 0.1, 1.0, 4.8, 1.0, 1.0]]
    planes = [[0.0, 0.0, -4.0, 0.0, 0.0, 1.0, 1.0], [0.0, 0.0, 4.5, 0.1, 0.0, -1.0, 0.0]]
    return spheres, cylinders, planes


@pytest.mark.parametrize("kind,name,values", [
    (oracle.WALL_LJ93, "WallsPotentialLJ93GPU", [[1.3, 0.9, 2.5, 0.0], [0.7, 1.1, 3.0, 0.4]]),
    (oracle.WALL_COLLOID, "WallsPotentialColloidGPU", [[2.0, 0.3, 0.8, 2.2, 0.5], [1.0, 0.25, 1.0, 2.5, 0.6]]),
])
def test_walls_match_oracle(kind, name, values):
    from azplugins_amd import _azp, _azplugins

    torch = _azp.torch
    rng = np.random.default_rng(kind + 3)
    N, L = 6000, 12.0
    pos4 = np.zeros((N, 4))
    pos4[:, :3] = rng.uniform(-L / 2, L / 2, size=(N, 3))
    pos4[:, 3] = rng.integers(0, 2, size=N)
    geo = _geometry()
    ref_f, ref_v = oracle.wall_forces(kind, pos4, values, *geo)
    w = getattr(_azplugins, name)(2)
    for t in range(2):
        w.setParams(t, values[t])
    w.setGeometry(*geo)
    sysm = GpuSystem(pos4, np.array([L, L, L]))
    pitch = (N + 63) // 64 * 64
    force = torch.zeros((N, 4), dtype=torch.float64, device="cuda")
    virial = torch.zeros((6, pitch), dtype=torch.float64, device="cuda")
    out = _azp.Outputs()
    out.d_force, out.d_virial, out.virial_pitch, out.d_torque = force.data_ptr(), virial.data_ptr(), pitch, None
    say(f"[walls] {name} N={N}")
    w.compute(sysm.particles(), out, compute_virial=True)
    got_f = force.cpu().numpy()
    got_v = virial[:, :N].cpu().numpy()
    assert np.count_nonzero(ref_f[:, 3]) > N // 10
    np.testing.assert_allclose(got_f, ref_f, rtol=1e-12, atol=1e-12 * np.nanmax(np.abs(ref_f)), equal_nan=True)
    np.testing.assert_allclose(got_v, ref_v, rtol=1e-12, atol=1e-12 * np.nanmax(np.abs(ref_v)), equal_nan=True)


GOLDEN = [
    ("Colloid", {"A": 1.0, "a": 1.0, "sigma": 1.0, "r_cut": 3.0}, (0, 0, 1.5), -0.0292, 0.8940),
    ("Colloid", {"A": 1.0, "a": 1.0, "sigma": 1.0, "r_cut": 3.0}, (0, 0, 3.5), 0, 0),
    ("LJ93", {"sigma": 1.0, "A": 1.0, "r_cut": 3.0}, (0, 0, 1.5), -0.2558, -0.5718),
    ("LJ93", {"sigma": 1.0, "A": 1.0, "r_cut": 3.0}, (0, 0, 3.5), 0, 0),
]


@pytest.mark.parametrize("case", GOLDEN, ids=lambda c: c[0])
def test_energy_and_force(case):
    """The reference's test_wall.py:59-101 through the drop-in API."""
    import azplugins_amd.harness as hoomd

    cls, params, position, energy, force = case
    snap = hoomd.Snapshot()
    snap.configuration.box = [20, 20, 20, 0, 0, 0]
    snap.particles.N = 1
    snap.particles.position[:] = [position]
    sim = hoomd.Simulation(seed=1)
    sim.create_state_from_snapshot(snap)
    integrator = hoomd.md.Integrator(dt=0.001)
    integrator.methods = [hoomd.md.methods.ConstantVolume(hoomd.filter.All())]
    wall = hoomd.wall.Plane(origin=(0, 0, 0), normal=(0, 0, 1))
    potential = getattr(hoomd.azplugins.wall, cls)(walls=[wall])
    potential.params["A"] = params
    integrator.forces = [potential]
    sim.operations.integrator = integrator
    sim.run(0)
    for attr in params:
        assert potential.params["A"][attr] == params[attr]
    np.testing.assert_array_almost_equal(potential.energies, energy, decimal=4)
    np.testing.assert_array_almost_equal(potential.forces, [[0, 0, force]], decimal=4)


def test_walls_confine_md():
    """Particles between two LJ 9-3 walls with a pair force: NVE keeps the
    total energy, and the walls' forces are what the oracle gives."""
    import azplugins_amd.harness as hoomd

    rng = np.random.default_rng(8)
    n, L = 8, 12.0
    g = (np.arange(n) + 0.5) * (L / n) - L / 2
    x, y, z = np.meshgrid(g, g, g * 0.75, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) + rng.uniform(-0.05, 0.05, size=(n**3, 3))
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = n**3
    snap.particles.position[:] = pos
    snap.particles.velocity[:] = rng.normal(scale=0.5, size=(n**3, 3))
    sim = hoomd.Simulation(seed=2)
    sim.create_state_from_snapshot(snap)
    integ = hoomd.md.Integrator(dt=0.001)
    integ.methods = [hoomd.md.methods.ConstantVolume(hoomd.filter.All())]
    walls = [hoomd.wall.Plane(origin=(0, 0, -5.0), normal=(0, 0, 1)),
             hoomd.wall.Plane(origin=(0, 0, 5.0), normal=(0, 0, -1))]
    lj93 = hoomd.azplugins.wall.LJ93(walls=walls)
    lj93.params["A"] = dict(A=2.0, sigma=1.0, r_cut=2.5)
    # shifted, so that pairs crossing r_cut do not change the energy
    plj = hoomd.azplugins.pair.PerturbedLennardJones(nlist=hoomd.md.nlist.Cell(buffer=0.4), default_r_cut=2.5,
                                                     mode="shift")
    plj.params[("A", "A")] = dict(epsilon=1.0, sigma=1.0, attraction_scale_factor=0.5)
    integ.forces = [lj93, plj]
    sim.operations.integrator = integ
    thermo = hoomd.md.compute.ThermodynamicQuantities(hoomd.filter.All())
    sim.operations.computes.append(thermo)
    sim.run(0)
    e0 = thermo.kinetic_energy + thermo.potential_energy
    snap0 = sim.state.get_snapshot()
    pos4 = np.zeros((n**3, 4))
    pos4[:, :3] = snap0.particles.position
    ref, _ = oracle.wall_forces(oracle.WALL_LJ93, pos4, [[2.0, 1.0, 2.5, 0.0]],
                                planes=[[0, 0, -5.0, 0, 0, 1, 1], [0, 0, 5.0, 0, 0, -1, 1]])
    np.testing.assert_allclose(lj93.forces, ref[:, :3], rtol=1e-12, atol=1e-12)
    sim.run(300)
    e1 = thermo.kinetic_energy + thermo.potential_energy
    say(f"[walls] NVE energy {e0:.6f} -> {e1:.6f}")
    assert abs(e1 - e0) < 2e-3 * abs(e0) + 1e-3 * n**3
    assert np.abs(sim.state.get_snapshot().particles.position[:, 2]).max() < 5.0


def test_unset_type_parameters_fail():
    import azplugins_amd.harness as hoomd

    snap = hoomd.Snapshot()
    snap.configuration.box = [10, 10, 10, 0, 0, 0]
    snap.particles.N = 2
    snap.particles.types = ["A", "B"]
    snap.particles.typeid[:] = [0, 1]
    sim = hoomd.Simulation(seed=1)
    sim.create_state_from_snapshot(snap)
    integ = hoomd.md.Integrator(dt=0.001)
    w = hoomd.azplugins.wall.LJ93(walls=[hoomd.wall.Plane(origin=(0, 0, 0), normal=(0, 0, 1))])
    w.params["A"] = dict(A=1.0, sigma=1.0, r_cut=3.0)
    integ.forces = [w]
    sim.operations.integrator = integ
    with pytest.raises(RuntimeError):
        sim.run(0)


def test_bad_type_is_reported_by_a_later_call():
    """The wall compute does not wait for the stream; a type index outside the
    table is reported once the kernel's flag reaches the host."""
    from azplugins_amd import _azp, _azplugins

    torch = _azp.torch
    pos4 = np.array([[0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 1.5, 3.0]])
    w = _azplugins.WallsPotentialLJ93GPU(1)
    w.setParams(0, [1.0, 1.0, 3.0, 0.0])
    w.setGeometry([], [], [[0, 0, 0, 0, 0, 1, 1]])
    sysm = GpuSystem(pos4, np.array([10.0, 10.0, 10.0]))
    force = torch.zeros((2, 4), dtype=torch.float64, device="cuda")
    out = _azp.Outputs()
    out.d_force, out.d_virial, out.virial_pitch, out.d_torque = force.data_ptr(), None, 64, None
    w.compute(sysm.particles(), out, compute_virial=False)
    torch.cuda.synchronize()
    with pytest.raises(_azp.AzpError):
        w.compute(sysm.particles(), out, compute_virial=False)
    assert force[1].abs().sum().item() == 0.0 and force[0, 2].item() != 0.0
    # the report clears the flag: with the type fixed the calls run normally
    sysm.pos[1, 3] = 0.0
    w.compute(sysm.particles(), out, compute_virial=False)
    w.compute(sysm.particles(), out, compute_virial=False)
    w.check()
    assert force[1, 2].item() != 0.0


def test_bad_type_on_the_last_call_is_reported_by_check():
    """Calls that run while an earlier flag copy is in flight still raise the
    sticky flag; the end-of-run check waits for the stream and reports it."""
    from azplugins_amd import _azp, _azplugins

    torch = _azp.torch
    pos4 = np.array([[0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 1.5, 0.0]])
    w = _azplugins.WallsPotentialLJ93GPU(1)
    w.setParams(0, [1.0, 1.0, 3.0, 0.0])
    w.setGeometry([], [], [[0, 0, 0, 0, 0, 1, 1]])
    sysm = GpuSystem(pos4, np.array([10.0, 10.0, 10.0]))
    force = torch.zeros((2, 4), dtype=torch.float64, device="cuda")
    out = _azp.Outputs()
    out.d_force, out.d_virial, out.virial_pitch, out.d_torque = force.data_ptr(), None, 64, None
    w.compute(sysm.particles(), out, compute_virial=False)
    w.check()
    sysm.pos[1, 3] = 3.0
    w.compute(sysm.particles(), out, compute_virial=False)
    with pytest.raises(_azp.AzpError, match="type index"):
        w.check()
    w.check()  # reported once, then cleared
