"""ThermodynamicQuantities: FP64 reductions on the GPU (deterministic order)."""

from ... import _azp

torch = _azp.torch


class ThermodynamicQuantities:
    def __init__(self, filter):
        self.filter = filter
        self._sim = None
        self._work = None
        self._out = None

    def _attach(self, sim):
        self._sim = sim
        n = _azp.lib().azp_thermo_workspace_size()
        self._work = torch.zeros(n, dtype=torch.float64, device="cuda")
        self._out = torch.zeros(10, dtype=torch.float64, device="cuda")

    def _reduce(self):
        st = self._sim.state
        integ = self._sim.operations.integrator
        f = integ._net_force if integ is not None else None
        v = integ._net_virial if integ is not None else None
        pitch = integ._pitch if integ is not None else 0
        rot = integ is not None and integ.integrate_rotational_dof
        _azp.call("azp_gpu_thermo_reduce", _azp.dptr(st.vel), _azp.dptr(f), _azp.dptr(v), pitch,
                  _azp.dptr(st.angmom) if rot else None, _azp.dptr(st.inertia) if rot else None,
                  st.N, _azp.dptr(self._work), _azp.dptr(self._out), _azp.stream_ptr())
        return self._out.cpu().numpy()

    @property
    def translational_degrees_of_freedom(self):
        N = self._sim.state.N
        return max(3 * N - 3, 1)

    @property
    def rotational_degrees_of_freedom(self):
        integ = self._sim.operations.integrator
        if integ is None or not integ.integrate_rotational_dof:
            return 0
        return 3 * int((self._sim.state.inertia > 0).sum().item())

    @property
    def degrees_of_freedom(self):
        return self.translational_degrees_of_freedom + self.rotational_degrees_of_freedom

    @property
    def translational_kinetic_energy(self):
        return float(self._reduce()[0])

    @property
    def rotational_kinetic_energy(self):
        return float(self._reduce()[1])

    @property
    def kinetic_energy(self):
        r = self._reduce()
        return float(r[0] + r[1])

    @property
    def kinetic_temperature(self):
        r = self._reduce()
        return float(2.0 * (r[0] + r[1]) / self.degrees_of_freedom)

    @property
    def potential_energy(self):
        return float(self._reduce()[2])

    @property
    def pressure_tensor(self):
        """(sum of m v_a v_b + W_ab) / V in the order xx, xy, xz, yy, yz, zz."""
        r = self._reduce()
        st = self._sim.state
        _azp.call("azp_gpu_kinetic_tensor", _azp.dptr(st.vel), st.N, _azp.dptr(self._work),
                  _azp.dptr(self._out), _azp.stream_ptr())
        K = self._out[:6].cpu().numpy()
        V = st.box.volume
        W = r[3:9]
        return [float((K[c] + W[c]) / V) for c in range(6)]

    @property
    def pressure(self):
        r = self._reduce()
        V = self._sim.state.box.volume
        return float((2.0 * r[0] / 3.0 + (r[3] + r[6] + r[8]) / 3.0) / V)
