"""Pair force base classes (HOOMD's ``md.pair.Pair``).

Subclasses set ``_ext_module`` / ``_cpp_class_name`` / ``_accepted_modes`` and
add a ``params`` TypeParameter exactly as the reference's ``pair.py`` does; on
attach the base class instantiates ``_ext_module.<_cpp_class_name>GPU`` with
the number of particle types, pushes params, r_cut, r_on and the mode, and
registers the r_cut matrix with the neighbor list.
"""

from ...data import RCutDict, TypeParameter, TypeParameterDict  # noqa: F401
from ..force import Force

_PRECISIONS = ("fp64", "mixed", "mixed_acc64")


class Pair(Force):
    _accepted_modes = ("none", "shift", "xplor")
    _ext_module = None
    _cpp_class_name = None

    def __init__(self, nlist, default_r_cut=None, default_r_on=0.0, mode="none"):
        super().__init__()
        self._cpp = None
        self.nlist = nlist
        self.r_cut = RCutDict(default_r_cut)
        self.r_cut._owner = self
        self.r_on = RCutDict(default_r_on)
        self.r_on._owner = self
        self._mode = None
        self.mode = mode
        self._precision = "fp64"
        self._typeparams = {}
        self._param_dict = {}

    # ---- HOOMD-style attributes ----
    @property
    def mode(self):
        return self._mode

    @mode.setter
    def mode(self, value):
        if value not in self._accepted_modes:
            raise ValueError(f"mode {value!r} not in accepted modes {self._accepted_modes}")
        self._mode = value
        if self._cpp is not None:
            self._cpp.setShiftMode(value)

    @property
    def precision(self):
        """'fp64' (plain form), 'mixed' (FP32 pair math, FP32 sums of 32 entries
        added to FP64 accumulators) or 'mixed_acc64' (FP32 pair math, every
        entry added to FP64 accumulators; isotropic potentials)."""
        return self._precision

    @precision.setter
    def precision(self, value):
        if value not in _PRECISIONS:
            raise ValueError(f"precision must be one of {_PRECISIONS}")
        self._precision = value
        if self._cpp is not None:
            self._cpp.setPrecision(value)

    def _add_typeparam(self, tp):
        tp._owner = self
        self._typeparams[tp.name] = tp
        setattr(self, tp.name, tp)

    # ---- attach / push parameters ----
    def _types(self):
        return self._sim.state.types

    def _schema(self):
        return self._typeparams["params"].param_dict._schema

    def _push_pair(self, a, b):
        types = self._types()
        key = (types[a], types[b])
        tp = self._typeparams["params"]
        if key not in tp.param_dict:
            raise RuntimeError(f"params for type pair {key} are not set")
        vals = tp.param_dict[key]
        self._cpp.setParams(a, b, [float(vals[k]) for k in self._schema()])
        self._cpp.setRCut(a, b, self.r_cut[key])
        self._cpp.setROn(a, b, self.r_on.get(key, 0.0))

    def _attach(self, sim):
        self._sim = sim
        nt = len(sim.state.types)
        cls = getattr(self._ext_module, self._cpp_class_name + "GPU")
        self._cpp = cls(nt)
        for a in range(nt):
            for b in range(a, nt):
                self._push_pair(a, b)
        self._cpp.setShiftMode(self._mode)
        self._cpp.setPrecision(self._precision)
        self.nlist._attach(sim)
        self.nlist._register(self)
        self._allocate(sim.state.N)
        self._attached = True

    def _typeparam_changed(self, name, key):
        if self._cpp is not None:
            types = self._types()
            from ...data import _expand

            for ka, kb in _expand(key):
                if ka in types and kb in types:
                    self._push_pair(types.index(ka), types.index(kb))

    def _rcut_changed(self):
        if self._cpp is not None:
            nt = len(self._types())
            for a in range(nt):
                for b in range(a, nt):
                    self._push_pair(a, b)

    def _typeparam_readback(self, name, key):
        types = self._types()
        a, b = types.index(key[0]), types.index(key[1])
        vals = self._cpp.getParams(a, b)
        out = {}
        for (k, typ), v in zip(self._schema().items(), vals):
            out[k] = bool(v) if typ is bool else float(v)
        return out

    def _r_cut_matrix(self):
        return self._cpp.getRCutMatrix()

    def _thermostat(self, timestep):
        """Hook for pair forces with a thermostat (DPD)."""

    def _compute(self, timestep, virial=None):
        self._thermostat(timestep)
        # the virial only when something reads it (HOOMD's pressure_tensor flag)
        virial = self._sim._needs_virial() if virial is None else virial
        nl = self.nlist
        if nl._needs_update(timestep):
            # the rebuild check and the force in one library call: the force
            # launches as soon as the check's flag arrives
            self._cpp.update_compute(nl.handle, self._sim.state.particles(), self._outputs(), timestep,
                                     compute_virial=virial, force_rebuild=nl._take_force_flag())
            nl._last = timestep
        else:
            self._cpp.compute(nl.handle, self._sim.state.particles(), self._outputs(), timestep,
                              compute_virial=virial)
        self._virial_current = bool(virial)

    def _takes_step(self):
        return self._cpp is not None and self._cpp.takes_step

    def _compute_step(self, timestep, method, integ):
        """_compute(timestep) followed by method._step_two_one(integ,
        timestep) in one library call, whose step kernel also makes the next
        update's rebuild check (bit-identical to the separate calls)."""
        from .... import _azp

        self._thermostat(timestep)
        virial = self._sim._needs_virial()
        nl = self.nlist
        nl._needs_update(timestep)  # pushes the list settings when due
        st = self._sim.state
        lang, gamma, _gamma_r, kT = method._langevin(timestep)
        step = _azp.StepArgs(d_pos=st.pos.data_ptr(), d_vel=st.vel.data_ptr(), d_image=st.image.data_ptr(),
                             d_tag=st.tag.data_ptr(), dt=integ.dt, langevin=lang, gamma=gamma, kT=kT,
                             seed=int(self._sim.seed) & 0xFFFF, timestep=int(timestep))
        self._cpp.update_compute_step(nl.handle, st.particles(), self._outputs(), timestep, step,
                                      compute_virial=virial, force_rebuild=nl._take_force_flag())
        nl._last = timestep
        self._virial_current = bool(virial)


from . import aniso  # noqa: E402,F401
