"""External potentials (HOOMD's ``md.external``): the wall potential base."""

import types as _types

from ..wall import Cylinder, Plane, Sphere
from .force import Force


class WallPotential(Force):
    """Potential between every particle and a list of walls (HOOMD's
    ``md.external.wall.WallPotential``).

    Subclasses set ``_ext_module`` / ``_cpp_class_name`` and add a per-type
    ``params`` TypeParameter; on attach the base class instantiates
    ``_ext_module.<_cpp_class_name>GPU`` with the number of particle types
    and pushes the walls and the parameters.
    """

    _ext_module = None
    _cpp_class_name = None

    def __init__(self, walls):
        super().__init__()
        self.walls = list(walls)
        self._cpp = None
        self._typeparams = {}

    def _add_typeparam(self, tp):
        tp._owner = self
        self._typeparams[tp.name] = tp
        setattr(self, tp.name, tp)

    def _schema(self):
        return self._typeparams["params"].param_dict._schema

    def _push_type(self, t):
        name = self._sim.state.types[t]
        tp = self._typeparams["params"]
        if name not in tp.param_dict:
            raise RuntimeError(f"wall parameters for type {name!r} are not set")
        vals = tp.param_dict[name]
        self._cpp.setParams(t, [float(vals[k]) for k in self._schema()])

    def _push_walls(self):
        rows = {Sphere: [], Cylinder: [], Plane: []}
        for w in self.walls:
            for cls in rows:
                if isinstance(w, cls):
                    rows[cls].append(w._row())
                    break
            else:
                raise TypeError(f"unsupported wall geometry {type(w).__name__}")
        self._cpp.setGeometry(rows[Sphere], rows[Cylinder], rows[Plane])
        self._pushed_walls = [tuple(w._row()) for w in self.walls]

    def _attach(self, sim):
        self._sim = sim
        nt = len(sim.state.types)
        self._cpp = getattr(self._ext_module, self._cpp_class_name + "GPU")(nt)
        for t in range(nt):
            self._push_type(t)
        self._push_walls()
        self._allocate(sim.state.N)
        self._attached = True

    def _typeparam_changed(self, name, key):
        if self._cpp is not None:
            types = self._sim.state.types
            for k in ([key] if isinstance(key, str) else list(key)):
                if k in types:
                    self._push_type(types.index(k))

    def _typeparam_readback(self, name, key):
        vals = self._cpp.getParams(self._sim.state.types.index(key))
        return {k: float(v) for k, v in zip(self._schema(), vals)}

    def _compute(self, timestep, virial=None):
        if [tuple(w._row()) for w in self.walls] != self._pushed_walls:
            self._push_walls()
        virial = self._sim._needs_virial() if virial is None else virial
        self._cpp.compute(self._sim.state.particles(), self._outputs(), compute_virial=virial)
        self._virial_current = bool(virial)

    def _finish_run(self):
        if self._cpp is not None:
            self._cpp.check()


wall = _types.SimpleNamespace(WallPotential=WallPotential)
