"""Pair potentials — drop-in for ``hoomd.azplugins.pair`` (reference src/pair.py).

Same class names, constructor signatures, parameter keys, accepted modes and
``_cpp_class_name`` strings as the reference; the C++ objects they select are
the gfx950 implementations in ``_azplugins``.  One addition: every class has a
``precision`` attribute: ``"fp64"`` (default, the plain form), ``"mixed"``
(FP32 pair arithmetic; the terms of 32 list entries are summed in FP32 and
each such sum is added to FP64 accumulators) or, for the isotropic
potentials, ``"mixed_acc64"`` (FP32 pair arithmetic, every entry's terms added
to FP64 accumulators).
"""

from . import _azplugins
from .harness import variant
from .harness.data import TypeParameter, TypeParameterDict
from .harness.md import pair

__all__ = [
    "Colloid",
    "DPDGeneralWeight",
    "ExpandedYukawa",
    "Hertz",
    "PerturbedLennardJones",
    "TwoPatchMorse",
]


class Colloid(pair.Pair):
    r"""Colloid pair potential (reference pair.py:14-118).

    Lennard-Jones integrated over zero, one or two spheres: solvent-solvent
    LJ with A = 144 epsilon, colloid-solvent, and the Everaers-Ejtehadi
    colloid-colloid form, selected by whether ``a_1`` / ``a_2`` are zero.

    params keys: ``A``, ``a_1``, ``a_2``, ``sigma``.  Modes: none, shift, xplor.
    Runs in the plain FP64 form only.
    """

    _ext_module = _azplugins
    _cpp_class_name = "PotentialPairColloid"
    _accepted_modes = ("none", "shift", "xplor")

    def __init__(self, nlist, default_r_cut=None, default_r_on=0, mode="none"):
        super().__init__(nlist, default_r_cut, default_r_on, mode)
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(A=float, a_1=float, a_2=float, sigma=float, len_keys=2),
        )
        self._add_typeparam(params)


class DPDGeneralWeight(pair.Pair):
    r"""DPD with the generalized weight w(r) = (1 - r/r_cut)^s (reference pair.py:121-239).

    Conservative A (1 - r/r_cut), dissipative -gamma w (v_ij . r_hat) and
    random sqrt(gamma w) theta forces; the random variate is Philox keyed by
    the simulation seed, the timestep and the (ordered) particle tags.

    params keys: ``A``, ``gamma``, ``s``.  ``kT``: variant or float.  Mode: none.
    """

    _ext_module = _azplugins
    _cpp_class_name = "PotentialPairDPDThermoGeneralWeight"
    _accepted_modes = ("none",)

    def __init__(self, nlist, kT, default_r_cut=None):
        super().__init__(nlist=nlist, default_r_cut=default_r_cut, default_r_on=0, mode="none")
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(A=float, gamma=float, s=float, len_keys=2),
        )
        self._add_typeparam(params)
        self._param_dict["kT"] = variant.as_variant(kT)

    @property
    def kT(self):
        return self._param_dict["kT"]

    @kT.setter
    def kT(self, value):
        self._param_dict["kT"] = variant.as_variant(value)

    def _attach(self, sim):
        """DPD uses RNGs. Warn the user if they did not set the seed."""
        sim._warn_if_seed_unset()
        super()._attach(sim)

    def _thermostat(self, timestep):
        integ = self._sim.operations.integrator
        dt = integ.dt if integ is not None else 0.0
        self._cpp.setThermostat(self._sim.seed, dt, self.kT(timestep))


class ExpandedYukawa(pair.Pair):
    r"""Expanded Yukawa U = epsilon exp(-kappa (r - Delta)) / (r - Delta) (reference pair.py:242-297).

    params keys: ``epsilon``, ``kappa``, ``delta``.  Modes: none, shift, xplor.
    """

    _ext_module = _azplugins
    _cpp_class_name = "PotentialPairExpandedYukawa"

    def __init__(self, nlist, default_r_cut=None, default_r_on=0, mode="none"):
        super().__init__(nlist, default_r_cut, default_r_on, mode)
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(epsilon=float, kappa=float, delta=float, len_keys=2),
        )
        self._add_typeparam(params)


class Hertz(pair.Pair):
    r"""Hertz U = epsilon (1 - r/r_cut)^(5/2) (reference pair.py:300-351).

    params key: ``epsilon``.  Modes: none, shift, xplor.
    """

    _ext_module = _azplugins
    _cpp_class_name = "PotentialPairHertz"
    _accepted_modes = ("none", "shift", "xplor")

    def __init__(self, nlist, default_r_cut=None, default_r_on=0, mode="none"):
        super().__init__(nlist, default_r_cut, default_r_on, mode)
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(epsilon=float, len_keys=2),
        )
        self._add_typeparam(params)


class PerturbedLennardJones(pair.Pair):
    r"""Perturbed LJ: WCA core plus lambda-scaled attraction (reference pair.py:354-426).

    U = U_LJ + (1 - lambda) epsilon for r <= 2^(1/6) sigma, lambda U_LJ beyond.

    params keys: ``epsilon``, ``sigma``, ``attraction_scale_factor``.
    Modes: none, shift, xplor.
    """

    _ext_module = _azplugins
    _cpp_class_name = "PotentialPairPerturbedLennardJones"
    _accepted_modes = ("none", "shift", "xplor")

    def __init__(self, nlist, default_r_cut=None, default_r_on=0, mode="none"):
        super().__init__(nlist, default_r_cut, default_r_on, mode)
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(epsilon=float, sigma=float, attraction_scale_factor=float, len_keys=2),
        )
        self._add_typeparam(params)


class TwoPatchMorse(pair.aniso.AnisotropicPair):
    r"""Two-patch Morse: Morse well modulated by the patch orientations (reference pair.py:429-525).

    U = U_M(r) Omega(r_hat . n_i) Omega(r_hat . n_j), Omega(g) = 1/(1 + exp(-omega (g^2 - alpha))),
    n = the body x axis rotated by the particle's orientation.

    params keys: ``M_d``, ``M_r``, ``r_eq``, ``omega``, ``alpha``, ``repulsion`` (bool).
    Modes: none, shift.  Outputs include ``torques``.
    """

    _ext_module = _azplugins
    _cpp_class_name = "AnisoPotentialPairTwoPatchMorse"

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut, mode)
        params = TypeParameter(
            "params",
            "particle_types",
            TypeParameterDict(
                M_d=float,
                M_r=float,
                r_eq=float,
                omega=float,
                alpha=float,
                repulsion=bool,
                len_keys=2,
            ),
        )
        self._add_typeparam(params)
