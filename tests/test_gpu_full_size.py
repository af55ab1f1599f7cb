"""Parity at full workload sizes.

PerturbedLJ, N = 1,048,576 (FCC 64^3, density 0.85, r_cut 3, r_buff 0.4):
* the plain list and the tiled list (expanded through its neighbor sets)
  equal the oracle's list as sets of unordered tag pairs (73M pairs),
  symmetric and free of duplicates and self pairs,
* plain FP64 form within 1e-12 of the oracle (normwise), mixed form over the
  plain and over the tiled list within 1e-5,
* Newton's third law on the GPU's full-list sums.

DPD general weight, N = 2,000,000 (density 3, A/B 50/50 at random,
A = 25 / 40 / 25, gamma 4.5, s 0.5, r_cut 1, kT 1, dt 0.02):
* neighbor sets equal the oracle's,
* forces (conservative + dissipative + random), energies and virials within
  1e-12 of the oracle (the pair terms are bit-identical; only the order of
  each particle's sum differs),
* the total force vanishes (every random and dissipative pair force is
  antisymmetric),
* the kernel over the tiled list (neighbors staged in LDS) gives the same
  bits as over the plain list.
"""

import numpy as np
import pytest

import oracle
from azplugins_amd import configs
from gpu_util import GpuSystem, gpu_pair, rel_err, say

pytestmark = pytest.mark.gpu


def list_keys(nn, lst, N):
    """Sorted keys min(i, j) * N + max(i, j) of a full GPU list, after checking
    that it holds no self pairs, no duplicates and every pair both ways."""
    mx = lst.shape[0]
    mask = np.arange(mx, dtype=np.uint32)[:, None] < nn[None, :N]
    jj = lst[:, :N][mask].astype(np.uint64)
    ii = np.broadcast_to(np.arange(N, dtype=np.uint64), (mx, N))[mask]
    del mask
    assert not np.any(ii == jj), "self pair in the list"
    fwd = ii < jj
    k_f = ii[fwd] * np.uint64(N) + jj[fwd]
    k_b = jj[~fwd] * np.uint64(N) + ii[~fwd]
    del ii, jj, fwd
    k_f.sort()
    k_b.sort()
    assert np.all(k_f[1:] != k_f[:-1]), "duplicate entry in the list"
    assert np.array_equal(k_f, k_b), "the full list is not symmetric"
    return k_f


def oracle_keys(pairs, N):
    pi = pairs[0].astype(np.uint64)
    pj = pairs[1].astype(np.uint64)
    k = np.minimum(pi, pj) * np.uint64(N) + np.maximum(pi, pj)
    k.sort()
    return k


@pytest.mark.timeout(900)
def test_plj_1m_sets_fp64_and_mixed():
    c = configs.plj_1m()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    say(f"[full] plj_1m N={N} oracle nlist ...")
    pairs = oracle.nlist(pos4, L, 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
    ko = oracle_keys(pairs, N)
    packed = oracle.pack_matrix(oracle.PERTURBED_LJ, c["params"], 1)
    say("[full] oracle forces ...")
    fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, L, pairs, 1, packed, np.array([9.0]))
    del pairs
    sysm = GpuSystem(pos4, L)
    say("[full] gpu fp64 (plain list) ...")
    r = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, c["params"], 3.0)
    nn, lst = r.nl.download(N)
    kg = list_keys(nn, lst, N)
    del lst
    say(f"[full] plain list: {kg.shape[0]} pairs, oracle {ko.shape[0]}")
    assert np.array_equal(kg, ko)
    del kg
    e64 = (rel_err(r.force[:, :3], fo[:, :3]), rel_err(r.force[:, 3], fo[:, 3]), rel_err(r.virial, vo))
    say("[full] fp64 rel err", e64)
    assert max(e64) < 1e-12
    # third law: the force sum is rounding noise of the largest forces
    assert np.abs(r.force[:, :3].sum(axis=0)).max() < 1e-14 * np.abs(r.force[:, :3]).max() * np.sqrt(N)
    for tiled in (False, True):
        say(f"[full] gpu mixed tiled={tiled} ...")
        rm = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, c["params"], 3.0, precision="mixed",
                      tiled=tiled)
        if tiled:
            assert rm.nl.get_tiles()[3] in (256, 512), "the tiled list was not built"
            nn, lst = rm.nl.download(N)
            kt = list_keys(nn, lst, N)
            del lst
            say(f"[full] tiled list: {kt.shape[0]} pairs")
            assert np.array_equal(kt, ko)
            del kt
        em = (rel_err(rm.force[:, :3], fo[:, :3]), rel_err(rm.force[:, 3], fo[:, 3]), rel_err(rm.virial, vo))
        l2 = float(np.linalg.norm(rm.force[:, :3] - fo[:, :3]) / np.linalg.norm(fo[:, :3]))
        say(f"[full] mixed tiled={tiled} rel err", em, "l2", l2)
        assert max(em) < 1e-5


@pytest.mark.timeout(900)
def test_dpd_2m_full_size():
    c = configs.dpd_2m()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    rng = np.random.default_rng(c["seed"])
    vel4 = np.zeros((N, 4))
    vel4[:, :3] = rng.normal(size=(N, 3)) * np.sqrt(c["kT"])
    vel4[:, 3] = 1.0
    tags = np.arange(N)
    seed, t = 3, 123457
    say(f"[full] dpd_2m N={N} oracle nlist ...")
    rl = oracle.r_listsq_matrix([c["r_cut"]] * 4, c["r_buff"])
    pairs = oracle.nlist(pos4, L, 2, rl, "cell")
    ko = oracle_keys(pairs, N)
    say("[full] oracle DPD forces ...")
    fo, vo = oracle.dpd_forces(pos4, vel4, tags, L, pairs, 2, oracle.pack_matrix(4, c["params"], 2),
                               np.full(4, c["r_cut"] ** 2), seed, t, c["dt"], c["kT"])
    del pairs
    say("[full] gpu DPD ...")
    r = gpu_pair("PotentialPairDPDThermoGeneralWeight", GpuSystem(pos4, L, vel4, tag=tags), 2, c["params"],
                 c["r_cut"], thermostat=(seed, c["dt"], c["kT"]), timestep=t)
    nn, lst = r.nl.download(N)
    kg = list_keys(nn, lst, N)
    del lst
    say(f"[full] dpd list: {kg.shape[0]} pairs, oracle {ko.shape[0]}")
    assert np.array_equal(kg, ko)
    del kg
    err = (rel_err(r.force[:, :3], fo[:, :3]), rel_err(r.force[:, 3], fo[:, 3]), rel_err(r.virial, vo))
    say("[full] dpd rel err force/energy/virial", err)
    assert max(err) < 1e-12
    assert np.abs(r.force[:, :3].sum(axis=0)).max() < 1e-14 * np.abs(r.force[:, :3]).max() * np.sqrt(N)
    # over the tiled list (neighbors staged in LDS): the same entries in the
    # same order, so the same bits
    say("[full] gpu DPD over the tiled list ...")
    rt = gpu_pair("PotentialPairDPDThermoGeneralWeight", GpuSystem(pos4, L, vel4, tag=tags), 2, c["params"],
                  c["r_cut"], thermostat=(seed, c["dt"], c["kT"]), timestep=t, tiled=True)
    assert rt.nl.get_tiles()[3] in (256, 512), "the tiled list was not built"
    assert np.array_equal(rt.force, r.force) and np.array_equal(rt.virial, r.virial)
