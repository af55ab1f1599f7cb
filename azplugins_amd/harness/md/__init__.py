"""Molecular dynamics: integrator, methods, neighbor lists, pair forces."""

from ... import _azp
from . import compute, external, methods, nlist, pair, tune  # noqa: F401

torch = _azp.torch


class Integrator:
    """Two-step integrator (HOOMD's ``md.Integrator``).

    update(t): every method's first half step, the net force at t + 1, every
    method's second half step — the order of upstream IntegratorTwoStep.
    """

    def __init__(self, dt, integrate_rotational_dof=False, forces=None, methods=None):
        self.dt = float(dt)
        self.integrate_rotational_dof = bool(integrate_rotational_dof)
        self.forces = list(forces) if forces is not None else []
        self.methods = list(methods) if methods is not None else []
        self._sim = None
        self._one_done = False

    def _attach(self, sim):
        self._sim = sim
        N = sim.state.N
        self._pitch = max(64, (N + 63) // 64 * 64)
        dev = torch.device("cuda")
        self._sum_force = torch.zeros((max(N, 1), 4), dtype=torch.float64, device=dev)
        self._sum_virial = torch.zeros((6, self._pitch), dtype=torch.float64, device=dev)
        self._sum_torque = torch.zeros((max(N, 1), 4), dtype=torch.float64, device=dev)
        for f in self.forces:
            f._attach(sim)
        for m in self.methods:
            m._attach(sim)
        self._net_force = self._sum_force
        self._net_virial = self._sum_virial
        self._net_torque = self._sum_torque

    def _compute_net_force(self, timestep):
        for f in self.forces:
            f._compute(timestep)
        if len(self.forces) == 1:
            f = self.forces[0]
            self._net_force, self._net_virial, self._net_torque = f._force, f._virial, f._torque
            return
        self._net_force, self._net_virial, self._net_torque = (
            self._sum_force, self._sum_virial, self._sum_torque)
        self._sum_force.zero_()
        self._sum_virial.zero_()
        self._sum_torque.zero_()
        for f in self.forces:
            self._sum_force.add_(f._force)
            self._sum_virial.add_(f._virial)
            self._sum_torque.add_(f._torque)

    def _reorder(self, order, scratch):
        from ..state import permute_rows

        N = self._sim.state.N
        for f in self.forces:
            f._reorder(order, scratch)
        if len(self.forces) != 1:
            permute_rows(self._sum_force, order, N, scratch)
            permute_rows(self._sum_torque, order, N, scratch)
            for r in range(6):
                permute_rows(self._sum_virial[r], order, N, scratch)

    def _prepare_run(self, timestep):
        self._compute_net_force(timestep)
        # accelerations of the current forces (second half step with dt = 0)
        for m in self.methods:
            m._step_two(self, timestep, dt=0.0)
        self._one_done = False

    def _finish_run(self):
        """End of run(): forces that report errors late (walls) check now."""
        for f in self.forces:
            f._finish_run()

    # Steps whose second half runs straight into the next step's first half
    # (nothing acts between them) fuse the two in one kernel per particle.
    fuse_half_steps = True
    # With one isotropic pair force and one translational method, those fused
    # half steps go to the force's library call, whose step kernel also makes
    # the next step's rebuild check (no displacement kernel of its own).
    step_with_force = True

    def _steps_with_force(self):
        return (self.step_with_force and self.fuse_half_steps and len(self.forces) == 1
                and len(self.methods) == 1 and not self.integrate_rotational_dof
                and hasattr(self.forces[0], "_takes_step") and self.forces[0]._takes_step())

    def _update(self, timestep, fuse_next=False):
        if not self._one_done:
            for m in self.methods:
                m._step_one(self, timestep)
        if fuse_next and self._steps_with_force():
            f = self.forces[0]
            f._compute_step(timestep + 1, self.methods[0], self)
            self._net_force, self._net_virial, self._net_torque = f._force, f._virial, f._torque
            self._one_done = True
            return
        self._compute_net_force(timestep + 1)
        if fuse_next and self.fuse_half_steps and len(self.methods) == 1:
            self.methods[0]._step_two_one(self, timestep + 1)
            self._one_done = True
        else:
            for m in self.methods:
                m._step_two(self, timestep + 1)
            self._one_done = False
