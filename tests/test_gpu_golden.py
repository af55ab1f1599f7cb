"""The reference's own pair tests, run through the drop-in API on the GPU.

Mirrors src/pytest/test_pair.py:309-363 and src/pytest/test_pair_aniso.py:113-168
line for line (snapshot factory, nlist.Cell(buffer=0.4), dummy NVE integrator,
``sim.run(0)``, params read back after attach, energies/forces/torques to 4
decimals), using the harness in place of HOOMD-blue.  Both precision forms run
where the potential supports them; Colloid and DPD keep the plain FP64 form
only, and asking them for the mixed form must fail at attach with that
message.
"""

import json
import os

import numpy
import pytest

import azplugins_amd.harness as hoomd
from azplugins_amd.harness import testing
from gpu_util import say

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _cases(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)["cases"]


PAIR_CASES = _cases("pair_vectors.json")
ANISO_CASES = _cases("aniso_vectors.json")


def _precisions(name):
    return ["fp64"] if name in ("Colloid", "DPDGeneralWeight") else ["fp64", "mixed"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("potential_test", PAIR_CASES, ids=lambda c: f"{c['potential']}-L{c['source_line']}")
def test_energy_and_force(potential_test, precision):
    refused = precision not in _precisions(potential_test["potential"])
    say(f"[golden] {potential_test['potential']} line {potential_test['source_line']} {precision}")
    r_cut = potential_test["r_cut"]
    r_buff = 0.4
    L_domain_min = 2 * (r_cut + r_buff)
    sim = testing.simulation_factory(
        testing.two_particle_snapshot_factory(d=potential_test["distance"], L=2.1 * L_domain_min)
    )
    integrator = hoomd.md.Integrator(dt=0.001)
    nve = hoomd.md.methods.ConstantVolume(hoomd.filter.All())
    integrator.methods = [nve]
    cls = getattr(hoomd.azplugins.pair, potential_test["potential"])
    extra_args = {}
    if cls is hoomd.azplugins.pair.DPDGeneralWeight:
        extra_args["kT"] = 0.0
    else:
        extra_args["mode"] = "shift" if potential_test["shift"] else "none"
    potential = cls(nlist=hoomd.md.nlist.Cell(buffer=r_buff), default_r_cut=r_cut, **extra_args)
    potential.params[("A", "A")] = potential_test["params"]
    potential.precision = precision
    integrator.forces = [potential]
    sim.operations.integrator = integrator
    if refused:
        with pytest.raises(RuntimeError, match="runs in the plain FP64 form only"):
            sim.run(0)
        return
    sim.run(0)
    assert potential.params[("A", "A")] == potential_test["params"]
    e = potential_test["energy"]
    f = potential_test["force"]
    _check(potential.energies, [0.5 * e, 0.5 * e], precision)
    _check(potential.forces, [[-f, 0, 0], [f, 0, 0]], precision)


def _check(actual, desired, precision):
    """The reference's decimal=4 for the FP64 form; for the mixed form the same
    absolute bar or 2e-6 relative, whichever is looser (decimal=4 on |F| ~ 1200
    asks for 1e-7 relative, below FP32 resolution)."""
    if precision == "fp64":
        numpy.testing.assert_array_almost_equal(actual, desired, decimal=4)
    else:
        numpy.testing.assert_allclose(actual, desired, rtol=2e-6, atol=1.5e-4)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("potential_test", ANISO_CASES, ids=lambda c: f"TPM-L{c['source_line']}")
def test_energy_force_and_torque(potential_test, precision):
    say(f"[golden aniso] line {potential_test['source_line']} {precision}")
    snap = testing.two_particle_snapshot_factory()
    snap.particles.position[:] = [[-0.5, -0.10, -0.15], [0.5, 0.10, 0.15]]
    snap.particles.orientation[:] = [[1, 0, 0, 0], [1, 0, 0, 0]]
    snap.particles.moment_inertia[:] = [0.1, 0.1, 0.1]
    sim = testing.simulation_factory(snap)
    integrator = hoomd.md.Integrator(dt=0.001)
    nve = hoomd.md.methods.ConstantVolume(hoomd.filter.All())
    integrator.methods = [nve]
    potential = hoomd.azplugins.pair.TwoPatchMorse(
        nlist=hoomd.md.nlist.Cell(buffer=0.4),
        default_r_cut=potential_test["r_cut"],
        mode="shift" if potential_test["shift"] else "none",
    )
    potential.params[("A", "A")] = potential_test["params"]
    potential.precision = precision
    integrator.forces = [potential]
    sim.operations.integrator = integrator
    sim.run(0)
    ref_values = (list(potential_test["params"].values()),)
    test_values = [potential.params[("A", "A")][k] for k in potential_test["params"]]
    assert numpy.allclose(test_values, ref_values)
    e = potential_test["energy"]
    numpy.testing.assert_array_almost_equal(potential.energies, [0.5 * e, 0.5 * e], decimal=4)
    f = potential_test["force"]
    if f is not None:
        f = numpy.array(f)
        numpy.testing.assert_array_almost_equal(potential.forces, [-f, f], decimal=4)
    T = potential_test["torque"]
    if T is not None:
        numpy.testing.assert_array_almost_equal(potential.torques, [T, T], decimal=4)


def test_modes_and_errors():
    """Accepted modes per class (pair.py:108,215,342,415 + aniso defaults) and clean errors."""
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    with pytest.raises(ValueError):
        hoomd.azplugins.pair.DPDGeneralWeight(nlist=nl, kT=1.0, default_r_cut=1.0).mode = "shift"
    with pytest.raises(ValueError):
        hoomd.azplugins.pair.TwoPatchMorse(nlist=nl, default_r_cut=2.0, mode="xplor")
    for cls in (hoomd.azplugins.pair.Hertz, hoomd.azplugins.pair.ExpandedYukawa,
                hoomd.azplugins.pair.PerturbedLennardJones, hoomd.azplugins.pair.Colloid):
        for m in ("none", "shift", "xplor"):
            cls(nlist=nl, default_r_cut=2.0, mode=m)
    # missing parameters for a type pair are reported at attach, before any launch
    sim = testing.simulation_factory(testing.two_particle_snapshot_factory(particle_types=["A", "B"], d=1.0))
    integ = hoomd.md.Integrator(dt=0.001)
    p = hoomd.azplugins.pair.Hertz(nlist=hoomd.md.nlist.Cell(buffer=0.4), default_r_cut=1.5)
    p.params[("A", "A")] = dict(epsilon=1.0)
    integ.forces = [p]
    sim.operations.integrator = integ
    with pytest.raises(RuntimeError, match="not set"):
        sim.run(0)
    # a box too small for r_cut + r_buff is an error, not a silent self-interaction
    sim = testing.simulation_factory(testing.two_particle_snapshot_factory(d=1.0, L=5.0))
    integ = hoomd.md.Integrator(dt=0.001)
    p = hoomd.azplugins.pair.Hertz(nlist=hoomd.md.nlist.Cell(buffer=0.4), default_r_cut=2.5)
    p.params[("A", "A")] = dict(epsilon=1.0)
    integ.forces = [p]
    sim.operations.integrator = integ
    with pytest.raises(RuntimeError, match="AZP_ERR_BOX"):
        sim.run(0)
