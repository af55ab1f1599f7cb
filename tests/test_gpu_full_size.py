"""Parity at the headline size: PerturbedLJ, N = 1,048,576 (FCC 64^3, density 0.85, r_cut 3).

The oracle finishes this size in about half a minute (cell list + one force
evaluation), so the full workload is compared directly:
* neighbor counts equal to the oracle's (every particle), total pair count equal,
* plain FP64 form within 1e-12 of the oracle (normwise), mixed form within 1e-5,
* Newton's third law on the GPU's full-list sums (size-independent).
"""

import numpy as np
import pytest

import oracle
from azplugins_amd import configs
from gpu_util import GpuSystem, gpu_pair, rel_err, say

pytestmark = pytest.mark.gpu


def test_plj_1m_fp64_and_mixed():
    c = configs.plj_1m()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    say(f"[full] plj_1m N={N} oracle nlist ...")
    pairs = oracle.nlist(pos4, L, 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
    packed = oracle.pack_matrix(oracle.PERTURBED_LJ, c["params"], 1)
    say("[full] oracle forces ...")
    fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, L, pairs, 1, packed, np.array([9.0]))
    sysm = GpuSystem(pos4, L)
    say("[full] gpu fp64 ...")
    r = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, c["params"], 3.0)
    nn, _ = r.nl.download(N)
    counts = np.bincount(pairs[0], minlength=N) + np.bincount(pairs[1], minlength=N)
    assert np.array_equal(nn.astype(np.int64), counts)
    assert int(nn.sum()) == 2 * pairs[0].shape[0]
    e64 = (rel_err(r.force[:, :3], fo[:, :3]), rel_err(r.force[:, 3], fo[:, 3]), rel_err(r.virial, vo))
    say("[full] fp64 rel err", e64)
    assert max(e64) < 1e-12
    assert np.abs(r.force[:, :3].sum(axis=0)).max() < 1e-8
    say("[full] gpu mixed ...")
    rm = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, c["params"], 3.0, precision="mixed")
    em = (rel_err(rm.force[:, :3], fo[:, :3]), rel_err(rm.force[:, 3], fo[:, 3]), rel_err(rm.virial, vo))
    l2 = float(np.linalg.norm(rm.force[:, :3] - fo[:, :3]) / np.linalg.norm(fo[:, :3]))
    say("[full] mixed rel err", em, "l2", l2)
    assert max(em) < 1e-5
