"""Anisotropic pair force base class (HOOMD's ``md.pair.aniso.AnisotropicPair``)."""

from . import Pair


class AnisotropicPair(Pair):
    _accepted_modes = ("none", "shift")

    def __init__(self, nlist, default_r_cut=None, mode="none"):
        super().__init__(nlist, default_r_cut=default_r_cut, default_r_on=0.0, mode=mode)
