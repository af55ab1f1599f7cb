"""azplugins_amd — MI355X (gfx950) short-range pair forces of azplugins.

``azplugins_amd.pair`` keeps the class names, constructor signatures,
parameter keys and outputs of ``hoomd.azplugins.pair``; the forces are
computed by hand-written HIP kernels in ``lib/libazp.so`` through the C ABI of
``include/azp.h``.  ``azplugins_amd.harness`` is a minimal HOOMD-compatible
simulation harness (Simulation, State, nlist.Cell, Integrator, ConstantVolume,
Langevin, ThermodynamicQuantities) for machines without HOOMD-blue;
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
    from . import harness, pair

    sys.modules["hoomd"] = harness
    sys.modules["hoomd.azplugins"] = sys.modules[__name__]
    sys.modules["hoomd.azplugins.pair"] = pair
    return harness
