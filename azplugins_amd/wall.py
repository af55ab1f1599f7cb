"""Wall potentials — drop-in for ``hoomd.azplugins.wall`` (reference src/wall.py).

Same class names, constructor signature, parameter keys (per particle type,
``r_extrap`` defaulting to 0) and ``_cpp_class_name`` strings as the
reference; the C++ objects they select are the gfx950 implementations in
``_azplugins`` (``azp_wall_*``).  Energies are shifted to zero at ``r_cut``
and go wholly to the particle.
"""

from . import _azplugins
from .harness.data import TypeParameter, TypeParameterDict
from .harness.md.external import WallPotential

__all__ = ["Colloid", "LJ93"]


class Colloid(WallPotential):
    r"""Colloid wall potential (reference wall.py:13-86, WallEvaluatorColloid.h).

    Lennard-Jones integrated between a sphere of radius a and a half space:
    U(r) = A sigma^6 / 7560 [(7a - r)/(r - a)^7 + (7a + r)/(r + a)^7]
           - A/6 [2 a r/(r^2 - a^2) + ln((r - a)/(r + a))].

    params keys (per particle type): ``A``, ``a``, ``sigma``, ``r_cut``,
    ``r_extrap`` (default 0).
    """

    _ext_module = _azplugins
    _cpp_class_name = "WallsPotentialColloid"

    def __init__(self, walls):
        super().__init__(walls)
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(A=float, a=float, sigma=float, r_cut=float, r_extrap=0.0, len_keys=1),
        )
        self._add_typeparam(params)


class LJ93(WallPotential):
    r"""Lennard-Jones 9-3 wall potential (reference wall.py:89-146, WallEvaluatorLJ93.h).

    U(r) = A [2/15 (sigma/r)^9 - (sigma/r)^3].

    params keys (per particle type): ``A``, ``sigma``, ``r_cut``, ``r_extrap``
    (default 0).
    """

    _ext_module = _azplugins
    _cpp_class_name = "WallsPotentialLJ93"

    def __init__(self, walls):
        super().__init__(walls)
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(A=float, sigma=float, r_cut=float, r_extrap=0.0, len_keys=1),
        )
        self._add_typeparam(params)
