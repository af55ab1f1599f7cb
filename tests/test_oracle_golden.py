"""Pin the oracle against the reference's own golden vectors and published KATs.

* 26 two-particle cases of ``src/pytest/test_pair.py:22-306`` driven exactly as
  ``test_energy_and_force`` does (``:312-363``): particles at (-d/2, 0, .1) and
  (d/2, 0, .1), box 2.1 * 2 (r_cut + 0.4), energies [e/2, e/2] and forces
  [[-f,0,0],[f,0,0]] to 4 decimals, params read back ``==`` after packing.
* 5 TwoPatchMorse cases of ``src/pytest/test_pair_aniso.py:21-168``.
* Philox4x32-10 known-answer vectors (Random123).
"""

import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

NAME_TO_CODE = {
    "Colloid": 3,
    "DPDGeneralWeight": 4,
    "ExpandedYukawa": 1,
    "Hertz": 0,
    "PerturbedLennardJones": 2,
    "TwoPatchMorse": 5,
}


def _load(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)["cases"]


PAIR_CASES = _load("pair_vectors.json")
ANISO_CASES = _load("aniso_vectors.json")


def two_particle(d, L):
    pos = np.array([[-d / 2, 0.0, 0.1, 0.0], [d / 2, 0.0, 0.1, 0.0]])
    return pos, np.array([L, L, L])


@pytest.mark.parametrize("case", PAIR_CASES, ids=lambda c: f"{c['potential']}-L{c['source_line']}")
def test_pair_golden(orc, case):
    ev = NAME_TO_CODE[case["potential"]]
    r_cut = case["r_cut"]
    r_buff = 0.4
    pos, L = two_particle(case["distance"], 2.1 * 2 * (r_cut + r_buff))
    packed = orc.pack_matrix(ev, {(0, 0): case["params"]}, 1)
    # params read back equal after attach (test_pair.py:349)
    assert orc.unpack(ev, packed[0]) == {k: float(v) for k, v in case["params"].items()}
    rl = orc.r_listsq_matrix([r_cut], r_buff)
    pairs = orc.nlist(pos, L, 1, rl, method="brute")
    rcutsq = np.array([r_cut * r_cut])
    if ev == orc.DPD_GENERAL_WEIGHT:
        vel = np.zeros((2, 4))
        vel[:, 3] = 1.0
        f, _ = orc.dpd_forces(pos, vel, [0, 1], L, pairs, 1, packed, rcutsq, seed=0,
                              timestep=0, dt=0.001, kT=0.0)
    else:
        f, _ = orc.pair_forces(ev, pos, L, pairs, 1, packed, rcutsq,
                               mode="shift" if case["shift"] else "none")
    e = case["energy"]
    fr = case["force"]
    np.testing.assert_array_almost_equal(f[:, 3], [0.5 * e, 0.5 * e], decimal=4)
    np.testing.assert_array_almost_equal(f[:, :3], [[-fr, 0, 0], [fr, 0, 0]], decimal=4)


def test_yukawa_full_precision(orc):
    """ExpandedYukawa literals in test_pair.py:119-160 are full FP64 values."""
    for case in PAIR_CASES:
        if case["potential"] != "ExpandedYukawa" or case["distance"] > case["r_cut"]:
            continue
        p = orc.pack(orc.EXPANDED_YUKAWA, case["params"])
        ok, fdivr, e = orc.eval_iso(orc.EXPANDED_YUKAWA, case["distance"] ** 2,
                                    case["r_cut"] ** 2, p, case["shift"])
        assert ok
        assert e == pytest.approx(case["energy"], rel=1e-14, abs=0)
        assert fdivr * case["distance"] == pytest.approx(case["force"], rel=1e-9)


@pytest.mark.parametrize("case", ANISO_CASES, ids=lambda c: f"TPM-L{c['source_line']}")
def test_aniso_golden(orc, case):
    ev = orc.TWO_PATCH_MORSE
    pos = np.array([[-0.5, -0.10, -0.15, 0.0], [0.5, 0.10, 0.15, 0.0]])
    quat = np.array([[1.0, 0, 0, 0], [1.0, 0, 0, 0]])
    r_cut = case["r_cut"]
    L = np.array([20.0, 20.0, 20.0])  # HOOMD two_particle_snapshot_factory default
    packed = orc.pack_matrix(ev, {(0, 0): case["params"]}, 1)
    back = orc.unpack(ev, packed[0])
    assert np.allclose([back[k] for k in case["params"]], list(case["params"].values()))
    pairs = orc.nlist(pos, L, 1, orc.r_listsq_matrix([r_cut], 0.4), method="brute")
    f, t, _ = orc.aniso_forces(pos, quat, L, pairs, 1, packed, np.array([r_cut**2]),
                               mode="shift" if case["shift"] else "none")
    e = case["energy"]
    np.testing.assert_array_almost_equal(f[:, 3], [0.5 * e, 0.5 * e], decimal=4)
    if case["force"] is not None:
        fr = np.array(case["force"])
        np.testing.assert_array_almost_equal(f[:, :3], [-fr, fr], decimal=4)
    if case["torque"] is not None:
        T = case["torque"]
        np.testing.assert_array_almost_equal(t[:, :3], [T, T], decimal=4)


def test_philox_kat(orc):
    for case in _load("philox_kat.json"):
        ctr = [int(x, 16) for x in case["ctr"]]
        key = [int(x, 16) for x in case["key"]]
        out = orc.philox(ctr, key)
        assert [f"{int(x):08x}" for x in out] == case["out"]
