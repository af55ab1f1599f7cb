"""Minimal HOOMD-compatible simulation harness on the gfx950 library.

HOOMD-blue (which the reference plugs into) is not required: this harness
provides the subset the pair-force path and its tests use — ``Snapshot``,
``Simulation`` with ``create_state_from_snapshot``/``run``/``timestep``/
``seed``, ``md.nlist.Cell``, ``md.Integrator``, ``md.methods.ConstantVolume``
and ``md.methods.Langevin``, ``md.compute.ThermodynamicQuantities``,
``tune.ParticleSorter``, ``trigger.Periodic``, ``wall`` (Sphere,
Cylinder, Plane), ``md.external.wall.WallPotential``, ``write.GSD`` and
``create_state_from_gsd`` (GSD files, module ``gsd``), ``filter.All`` and
``variant.Constant``.  Particle data lives in HBM as torch
tensors (device memory only; all arithmetic runs in libazp.so kernels).

Conventions that differ from HOOMD-blue are documented where they occur:
boxes are orthorhombic and periodic; angular momenta are space-frame vectors
of spherical tops (isotropic moment of inertia).
"""

import types as _types
import warnings

import numpy as np

from . import filter, gsd, trigger, tune, variant, wall, write  # noqa: F401
from .state import Box, Snapshot, State  # noqa: F401


class _Communicator:
    rank = 0
    num_ranks = 1


class Device:
    """Single-GPU device descriptor (one process per GPU)."""

    def __init__(self):
        self.communicator = _Communicator()


class GPU(Device):
    pass


class Operations:
    """The simulation's operations.  A new Operations holds one
    ``tune.ParticleSorter`` with its default trigger, so simulations keep
    their particles sorted unless it is removed from ``tuners``."""

    def __init__(self, sim):
        self._sim = sim
        self._integrator = None
        self.computes = []
        self.tuners = [tune.ParticleSorter()]
        self.writers = []

    @property
    def integrator(self):
        return self._integrator

    @integrator.setter
    def integrator(self, value):
        self._integrator = value
        self._sim._attached = False


class Simulation:
    """Owns the state, the operations and the timestep counter."""

    def __init__(self, device=None, seed=None):
        self.device = device if device is not None else GPU()
        self._seed = None if seed is None else int(seed) & 0xFFFF
        self.timestep = 0
        self.state = None
        self.operations = Operations(self)
        self._attached = False

    @property
    def seed(self):
        return 0 if self._seed is None else self._seed

    @seed.setter
    def seed(self, v):
        self._seed = int(v) & 0xFFFF

    def _warn_if_seed_unset(self):
        if self._seed is None:
            warnings.warn("Simulation.seed is not set, using default seed=0", stacklevel=3)

    def create_state_from_snapshot(self, snapshot):
        if self.state is not None:
            raise RuntimeError("state already created")
        self.state = State(self, snapshot)
        return self.state

    def create_state_from_gsd(self, filename, frame=-1):
        """The state from one frame of a GSD file; the timestep becomes the
        frame's step, so counter-based random forces continue."""
        if self.state is not None:
            raise RuntimeError("state already created")
        with gsd.open(filename, "r") as traj:
            fr = traj[frame]
        self.timestep = int(fr.configuration.step)
        return self.create_state_from_snapshot(Snapshot.from_gsd_frame(fr))

    def _needs_virial(self):
        """Whether force computes fill the virial: a thermodynamic compute is
        attached or a force's virials were read (HOOMD requests the pressure
        tensor the same way)."""
        from .md.compute import ThermodynamicQuantities

        return getattr(self, "_virial_requested", False) or any(
            isinstance(c, ThermodynamicQuantities) for c in self.operations.computes)

    def _attach(self):
        if self.state is None:
            raise RuntimeError("create the state before running")
        integ = self.operations.integrator
        if integ is not None:
            integ._attach(self)
        for c in self.operations.computes:
            c._attach(self)
        self._attached = True

    def _reorder(self, order):
        """A tuner moved particle order[k] to index k: every per-particle
        array follows, force outputs included, and neighbor lists rebuild."""
        scratch = self.state._reorder(order)
        integ = self.operations.integrator
        if integ is not None and integ._sim is self:
            integ._reorder(order, scratch)

    def run(self, steps):
        """Advance `steps` timesteps; forces are (re)computed at the start.

        Nothing but the steps runs inside run(), so parameters, the box and
        the particle arrays are fixed for its duration: the ctypes views of
        the state and the neighbor-list settings are built once per call."""
        if not self._attached:
            self._attach()
        integ = self.operations.integrator
        self.state._freeze_view()
        try:
            if integ is not None:
                integ._prepare_run(self.timestep)
            tuners = list(self.operations.tuners)
            writers = list(self.operations.writers)
            for t in tuners + writers:
                if t._sim is not self:
                    t._attach(self)
            n = int(steps)
            for k in range(n):
                for t in tuners:
                    t._act(self.timestep)
                nxt = self.timestep + 1
                if integ is not None:
                    # the next step's first half step may join this step's
                    # second half when no tuner or writer acts in between
                    fuse = k + 1 < n and not any(t.trigger(nxt) for t in tuners + writers)
                    integ._update(self.timestep, fuse_next=fuse)
                self.timestep = nxt
                for w in writers:
                    if w.trigger(nxt):
                        w._act(nxt)
            for w in writers:
                w.flush()
            if integ is not None:
                integ._finish_run()
        finally:
            self.state._thaw_view()


from . import md  # noqa: E402
from .. import pair as _pair  # noqa: E402
from .. import wall as _wall  # noqa: E402

# ``hoomd.azplugins.pair``, ``hoomd.azplugins.wall`` and ``hoomd.device.GPU`` spellings
azplugins = _types.SimpleNamespace(pair=_pair, wall=_wall)
device = _types.SimpleNamespace(GPU=GPU, Device=Device)
_ = np
