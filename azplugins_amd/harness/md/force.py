"""Force base class with HOOMD's per-particle output properties."""

import numpy as np

from ... import _azp

torch = _azp.torch


class Force:
    def __init__(self):
        self._attached = False
        self._sim = None
        self._force = None
        self._virial = None
        self._torque = None

    def _allocate(self, N):
        dev = torch.device("cuda")
        self._pitch = max(64, (N + 63) // 64 * 64)
        self._force = torch.zeros((max(N, 1), 4), dtype=torch.float64, device=dev)
        self._virial = torch.zeros((6, self._pitch), dtype=torch.float64, device=dev)
        self._torque = torch.zeros((max(N, 1), 4), dtype=torch.float64, device=dev)

    def _outputs(self):
        """azp_outputs view of the force arrays, rebuilt when they are reallocated."""
        o = getattr(self, "_out_c", None)
        if o is not None and o.d_force == self._force.data_ptr():
            return o
        o = _azp.Outputs()
        o.d_force = self._force.data_ptr()
        o.d_virial = self._virial.data_ptr()
        o.virial_pitch = self._pitch
        o.d_torque = self._torque.data_ptr()
        self._out_c = o
        return o

    def _reorder(self, order, scratch):
        """Particles were reordered: move the outputs with them and make the
        neighbor list rebuild."""
        from ..state import permute_rows

        if self._force is None:
            return
        N = self._sim.state.N
        permute_rows(self._force, order, N, scratch)
        permute_rows(self._torque, order, N, scratch)
        for r in range(6):
            permute_rows(self._virial[r], order, N, scratch)
        nl = getattr(self, "nlist", None)
        if nl is not None:
            nl._invalidate()

    def _finish_run(self):
        """Called at the end of every run(); errors a compute flagged on the
        device without waiting are reported here at the latest."""

    def _require(self):
        if not self._attached:
            raise RuntimeError("the force is not attached to a running simulation")

    @property
    def forces(self):
        self._require()
        st = self._sim.state
        return st._by_tag(self._force[:st.N, :3].cpu().numpy())

    @property
    def energies(self):
        self._require()
        st = self._sim.state
        return st._by_tag(self._force[:st.N, 3].cpu().numpy())

    @property
    def energy(self):
        return float(np.sum(self.energies))

    @property
    def virials(self):
        self._require()
        if not getattr(self, "_virial_current", True):
            # computed without the virial (nothing asked for it): from now on
            # every step computes it, and this step's is computed here
            self._sim._virial_requested = True
            self._compute(self._sim.timestep, virial=True)
        st = self._sim.state
        return st._by_tag(self._virial[:, :st.N].T.cpu().numpy())

    @property
    def torques(self):
        self._require()
        st = self._sim.state
        return st._by_tag(self._torque[:st.N, :3].cpu().numpy())
