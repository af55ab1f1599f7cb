"""The integrators around the pair forces against the oracle and closed forms.

* The CPU reference workload (Hertz, N = 4096, simple cubic at density 0.8
  with +-0.1 jitter, r_cut = 1): FP64 force, energy and virial of the
  snapshot through the drop-in API against the oracle (1e-12), then 1000
  ConstantVolume (velocity Verlet) steps at dt = 0.005 from Maxwell-Boltzmann
  velocities at kT = 1 against ``oracle.run_nve`` (the same two half steps,
  a half list with Newton's third law).  The GPU sums each particle's pair
  terms in another order, so the trajectories separate from rounding: the
  oracle run against itself with positions perturbed by 1e-15 (relative)
  separates by 3e-14 after 100 steps and 1.4e-11 after 1000 (measured), and
  the bounds below are 30-100 times those figures.
* Langevin at kT = 0 (drag only, no pair forces) against the closed form of
  the discrete update: the half-step velocity decays by (1 - dt gamma / m)
  per step, the positions advance by the geometric sum of it; the angular
  momentum decays by (1 - dt gamma_r / I) per step and the orientation turns
  about the fixed axis by dt |L_k| / I per step.
"""

import numpy as np
import pytest

import oracle
from azplugins_amd import configs
from gpu_util import say

pytestmark = pytest.mark.gpu


def _hertz_sim(pos4, L, vel, dt):
    import azplugins_amd.harness as hoomd

    snap = hoomd.Snapshot()
    snap.configuration.box = [L[0], L[1], L[2], 0, 0, 0]
    snap.particles.N = pos4.shape[0]
    snap.particles.position[:] = pos4[:, :3]
    snap.particles.velocity[:] = vel
    sim = hoomd.Simulation(seed=1)
    sim.create_state_from_snapshot(snap)
    integ = hoomd.md.Integrator(dt=dt)
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    hz = hoomd.azplugins.pair.Hertz(nlist=nl, default_r_cut=1.0)
    hz.params[("A", "A")] = dict(epsilon=1.0)
    integ.forces = [hz]
    integ.methods = [hoomd.md.methods.ConstantVolume(hoomd.filter.All())]
    sim.operations.integrator = integ
    return sim, hz


def _min_image(d, L):
    return d - L * np.rint(d / L)


@pytest.mark.timeout(300)
def test_hertz_4096_snapshot_and_nve_vs_oracle():
    c = configs.hertz_4096()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    rng = np.random.default_rng(1)
    v = rng.normal(size=(N, 3)) * np.sqrt(c["kT"])
    v -= v.mean(axis=0)
    vel4 = np.zeros((N, 4))
    vel4[:, :3] = v
    vel4[:, 3] = 1.0
    packed = oracle.pack_matrix(oracle.HERTZ, c["params"], 1)
    # snapshot
    pairs = oracle.nlist(pos4, L, 1, oracle.r_listsq_matrix([1.0], 0.4), "cell")
    fo, vo = oracle.pair_forces(oracle.HERTZ, pos4, L, pairs, 1, packed, np.array([1.0]))
    sim, hz = _hertz_sim(pos4, L, v, c["dt"])
    sim.run(0)
    F = hz.forces
    E = hz.energies
    W = hz.virials
    den = np.abs(fo[:, :3]).max()
    ef = np.abs(F - fo[:, :3]).max() / den
    ee = np.abs(E - fo[:, 3]).max() / np.abs(fo[:, 3]).max()
    ew = np.abs(W - vo.T).max() / np.abs(vo).max()
    say(f"[hertz 4096] snapshot rel err force {ef:.2e} energy {ee:.2e} virial {ew:.2e}")
    assert ef < 1e-12 and ee < 1e-12 and ew < 1e-12
    E0 = 0.5 * (v * v).sum() + fo[:, 3].sum()
    # NVE: 100 steps, then 900 more
    done = 0
    for steps, bound in ((100, 1e-12), (900, 1e-9)):
        sim.run(steps)
        done += steps
        po, vo_, fo_, _, nb = oracle.run_nve(oracle.HERTZ, pos4, vel4, L, 1, packed, np.array([1.0]), 0.4,
                                             "none", c["dt"], done)
        snap = sim.state.get_snapshot()
        dpos = np.abs(_min_image(snap.particles.position - po[:, :3], L)).max()
        dvel = np.abs(snap.particles.velocity - vo_[:, :3]).max()
        ke_g = 0.5 * (snap.particles.velocity ** 2).sum()
        e_g = ke_g + hz.energy
        e_o = 0.5 * (vo_[:, :3] ** 2).sum() + fo_[:, 3].sum()
        say(f"  after {done} steps ({nb} oracle builds): max |dx| {dpos:.2e} max |dv| {dvel:.2e} "
            f"E gpu {e_g:.12f} oracle {e_o:.12f} drift {e_o - E0:.3e}")
        assert dpos < bound and dvel < bound
        assert abs(e_g - e_o) < bound * abs(E0)
    # the energy is conserved to the integrator's accuracy
    assert abs(e_o - E0) < 1e-5 * abs(E0)


@pytest.mark.timeout(300)
def test_langevin_zero_temperature_closed_form():
    import azplugins_amd.harness as hoomd

    n, a = 4, 6.0  # particles 6 apart, r_cut 1: no pair forces
    L = n * a
    g = (np.arange(n) + 0.5) * a - L / 2
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    N = pos.shape[0]
    rng = np.random.default_rng(9)
    v0 = rng.normal(size=(N, 3))
    L0 = rng.normal(size=(N, 3))
    mass = rng.uniform(0.5, 2.0, size=N)
    inertia = rng.uniform(0.5, 2.0, size=N)
    q0 = rng.normal(size=(N, 4))
    q0 /= np.linalg.norm(q0, axis=1)[:, None]
    snap = hoomd.Snapshot()
    snap.configuration.box = [L, L, L, 0, 0, 0]
    snap.particles.N = N
    snap.particles.position[:] = pos
    snap.particles.velocity[:] = v0
    snap.particles.mass[:] = mass
    snap.particles.moment_inertia[:] = inertia[:, None]
    snap.particles.orientation[:] = q0
    ang = np.zeros((N, 4))
    ang[:, 1:] = L0
    snap.particles.angmom[:] = ang
    sim = hoomd.Simulation(seed=5)
    sim.create_state_from_snapshot(snap)
    dt, gamma, gamma_r = 0.005, 2.0, 1.5
    integ = hoomd.md.Integrator(dt=dt, integrate_rotational_dof=True)
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    hz = hoomd.azplugins.pair.Hertz(nlist=nl, default_r_cut=1.0)
    hz.params[("A", "A")] = dict(epsilon=1.0)
    integ.forces = [hz]
    integ.methods = [hoomd.md.methods.Langevin(hoomd.filter.All(), kT=0.0, default_gamma=gamma,
                                               default_gamma_r=(gamma_r, gamma_r, gamma_r))]
    sim.operations.integrator = integ
    steps = 300
    sim.run(steps)
    s = sim.state.get_snapshot()
    # translation
    r = 1.0 - dt * gamma / mass
    x_cf = pos + (v0 * (mass / gamma)[:, None]) * (1.0 - r ** steps)[:, None]
    v_cf = v0 * (r ** (steps - 1) * (1.0 - 0.5 * dt * gamma / mass))[:, None]
    dx = np.abs(_min_image(s.particles.position - x_cf, L)).max()
    dv = np.abs(s.particles.velocity - v_cf).max()
    # rotation
    sr = 1.0 - dt * gamma_r / inertia
    L_cf = L0 * (sr ** steps)[:, None]
    Lnorm = np.linalg.norm(L0, axis=1)
    theta = dt * Lnorm / inertia * (1.0 - sr ** steps) / (1.0 - sr)
    axis = L0 / Lnorm[:, None]
    rq = np.concatenate([np.cos(0.5 * theta)[:, None], np.sin(0.5 * theta)[:, None] * axis], axis=1)
    w0, w1, w2, w3 = rq.T
    q = q0.T
    q_cf = np.stack([w0 * q[0] - w1 * q[1] - w2 * q[2] - w3 * q[3],
                     w0 * q[1] + w1 * q[0] + w2 * q[3] - w3 * q[2],
                     w0 * q[2] - w1 * q[3] + w2 * q[0] + w3 * q[1],
                     w0 * q[3] + w1 * q[2] - w2 * q[1] + w3 * q[0]], axis=1)
    dL = np.abs(s.particles.angmom[:, 1:] - L_cf).max()
    # q and -q are the same orientation
    dq = np.minimum(np.abs(s.particles.orientation - q_cf).max(axis=1),
                    np.abs(s.particles.orientation + q_cf).max(axis=1)).max()
    say(f"[langevin kT=0] {steps} steps: |dx| {dx:.2e} |dv| {dv:.2e} |dL| {dL:.2e} |dq| {dq:.2e}")
    assert np.all(hz.forces == 0.0)
    assert dx < 1e-12 and dv < 1e-13 and dL < 1e-13 and dq < 1e-11
