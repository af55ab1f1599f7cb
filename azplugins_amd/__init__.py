"""azplugins_amd — MI355X (gfx950) short-range pair forces of azplugins.

``azplugins_amd.pair`` and ``azplugins_amd.wall`` keep the class names,
constructor signatures, parameter keys and outputs of ``hoomd.azplugins.pair``
and ``hoomd.azplugins.wall``; the forces are computed by hand-written HIP
kernels in ``lib/libazp.so`` through the C ABI of ``include/azp.h``.
``azplugins_amd.harness`` is a minimal HOOMD-compatible simulation harness
(Simulation, State, nlist.Cell, Integrator, ConstantVolume, Langevin,
ThermodynamicQuantities, wall geometries, the particle sorter and the
neighbor-list buffer tuner) for machines without HOOMD-blue;
``use_as_hoomd()`` registers it under the name ``hoomd`` when no real HOOMD is
importable.
"""

__version__ = "0.1.0"


def use_as_hoomd():
    """Expose the harness as ``hoomd`` and this package as ``hoomd.azplugins``.

    Does nothing (and returns the real module) when HOOMD-blue is importable.
    """
    import importlib
    import sys

    try:
        return importlib.import_module("hoomd")
    except ImportError:
        pass
    from . import harness, pair, wall

    sys.modules["hoomd"] = harness
    sys.modules["hoomd.azplugins"] = sys.modules[__name__]
    sys.modules["hoomd.azplugins.pair"] = pair
    sys.modules["hoomd.azplugins.wall"] = wall
    for name in ("md", "md.external", "md.nlist", "md.methods", "md.compute", "md.tune", "wall", "tune",
                 "trigger", "filter", "variant", "write"):
        sys.modules["hoomd." + name] = importlib.import_module(harness.__name__ + "." + name)
    return harness
