"""Integration methods: ConstantVolume (NVE) and Langevin, both on the GPU."""

from ... import _azp
from .. import variant


class Method:
    def __init__(self, filter):
        self.filter = filter
        self._sim = None

    def _attach(self, sim):
        self._sim = sim

    def _step_one(self, integ, timestep):
        st = self._sim.state
        box = st.box.to_c()
        _azp.call("azp_gpu_integrate_vv_one", _azp.dptr(st.pos), _azp.dptr(st.vel), _azp.dptr(st.accel),
                  _azp.dptr(st.image), st.N, _azp.ctypes.byref(box), integ.dt, _azp.stream_ptr())
        if integ.integrate_rotational_dof:
            _azp.call("azp_gpu_integrate_rot_one", _azp.dptr(st.orientation), _azp.dptr(st.angmom),
                      _azp.dptr(integ._net_torque), _azp.dptr(st.inertia), st.N, integ.dt, _azp.stream_ptr())

    def _langevin(self, timestep):
        return 0, 0.0, 0.0, 0.0

    def _step_two(self, integ, timestep, dt=None):
        st = self._sim.state
        dt = integ.dt if dt is None else dt
        lang, gamma, gamma_r, kT = self._langevin(timestep)
        if dt == 0.0:
            # accelerations of the current forces only (no thermostat terms)
            lang = 0
        seed = self._sim.seed
        _azp.call("azp_gpu_integrate_vv_two", _azp.dptr(st.vel), _azp.dptr(st.accel),
                  _azp.dptr(integ._net_force), _azp.dptr(st.tag), st.N, dt, lang, gamma, kT, seed,
                  timestep, _azp.stream_ptr())
        if integ.integrate_rotational_dof:
            _azp.call("azp_gpu_integrate_rot_two", _azp.dptr(st.angmom), _azp.dptr(integ._net_torque),
                      _azp.dptr(st.inertia), _azp.dptr(st.tag), st.N, dt, lang, gamma_r, kT, seed,
                      timestep, _azp.stream_ptr())


    def _step_two_one(self, integ, timestep):
        """_step_two(timestep) then _step_one(timestep), the translational
        part in one kernel (bit-identical to the two)."""
        st = self._sim.state
        lang, gamma, gamma_r, kT = self._langevin(timestep)
        seed = self._sim.seed
        box = st.box.to_c()
        _azp.call("azp_gpu_integrate_vv_two_one", _azp.dptr(st.pos), _azp.dptr(st.vel), _azp.dptr(st.image),
                  _azp.dptr(integ._net_force), _azp.dptr(st.tag), st.N, _azp.ctypes.byref(box), integ.dt, lang,
                  gamma, kT, seed, timestep, _azp.stream_ptr())
        if integ.integrate_rotational_dof:
            _azp.call("azp_gpu_integrate_rot_two", _azp.dptr(st.angmom), _azp.dptr(integ._net_torque),
                      _azp.dptr(st.inertia), _azp.dptr(st.tag), st.N, integ.dt, lang, gamma_r, kT, seed,
                      timestep, _azp.stream_ptr())
            _azp.call("azp_gpu_integrate_rot_one", _azp.dptr(st.orientation), _azp.dptr(st.angmom),
                      _azp.dptr(integ._net_torque), _azp.dptr(st.inertia), st.N, integ.dt, _azp.stream_ptr())


class ConstantVolume(Method):
    """Velocity Verlet at constant volume (NVE when thermostat is None)."""

    def __init__(self, filter, thermostat=None):
        super().__init__(filter)
        if thermostat is not None:
            raise NotImplementedError("thermostats other than None are not provided")


class Langevin(Method):
    """Langevin dynamics: drag gamma and uniform noise of variance 2 gamma kT/dt.

    Uses one drag coefficient for all types (``default_gamma``) and one
    rotational drag (``default_gamma_r``, isotropic).
    """

    def __init__(self, filter, kT, tally_reservoir_energy=False, default_gamma=1.0,
                 default_gamma_r=(1.0, 1.0, 1.0)):
        super().__init__(filter)
        self.kT = variant.as_variant(kT)
        self.default_gamma = float(default_gamma)
        gr = default_gamma_r
        self.default_gamma_r = float(gr[0] if isinstance(gr, (tuple, list)) else gr)

    def _langevin(self, timestep):
        return 1, self.default_gamma, self.default_gamma_r, self.kT(timestep)
