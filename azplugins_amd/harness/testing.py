"""Snapshot and simulation factories (HOOMD's pytest fixtures, restated).

``two_particle_snapshot_factory`` places particles at (-d/2, 0, .1) and
(d/2, 0, .1) in a cube of side L, as the fixture the reference's
test_pair.py:320-322 uses; ``lattice_snapshot_factory`` builds a simple-cubic
lattice of n^3 particles with spacing a (test_pair_dpd.py:184).
"""

import numpy as np

from . import Simulation, Snapshot


def two_particle_snapshot_factory(particle_types=("A",), dimensions=3, d=1.0, L=20.0):
    s = Snapshot()
    s.configuration.box = [L, L, L, 0, 0, 0]
    s.particles.N = 2
    s.particles.types = list(particle_types)
    s.particles.position[:] = [[-d / 2, 0, 0.1], [d / 2, 0, 0.1]]
    return s


def lattice_snapshot_factory(particle_types=("A",), dimensions=3, a=1.0, n=7, r=0.0, seed=0):
    s = Snapshot()
    n = int(n)
    L = n * a
    s.configuration.box = [L, L, L, 0, 0, 0]
    s.particles.N = n**3
    s.particles.types = list(particle_types)
    g = (np.arange(n) + 0.5) * a - L / 2
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    if r:
        pos += np.random.default_rng(seed).uniform(-r, r, size=pos.shape)
    s.particles.position[:] = pos
    return s


def simulation_factory(snapshot=None, seed=1):
    sim = Simulation(seed=seed)
    if snapshot is not None:
        sim.create_state_from_snapshot(snapshot)
    return sim
