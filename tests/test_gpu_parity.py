"""GPU parity of the pair-force path against the FP64 oracle (through the C ABI).

Bars:
* neighbor sets identical to the oracle's as unordered pair sets,
* plain FP64 form: forces, energies, virials (and torques) within 1e-12 of the
  oracle, normwise (max |delta| / max |ref|),
* mixed form (FP32 pair arithmetic, FP64 accumulation): within 1e-5 normwise,
* DPD: per-pair forces (conservative + dissipative + random) bit-exact on
  isolated pairs, random variates bit-exact, many-body sums within 1e-12.
"""

import numpy as np
import pytest

import oracle
from gpu_util import (CLASS, CONSERVATIVE_DPD, GpuSystem, gpu_pair, jittered_lattice, nlist_pairs,
                      rel_err, say)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5

ISO_PARAMS = {
    0: {(0, 0): {"epsilon": 1.0}, (0, 1): {"epsilon": 2.0}, (1, 1): {"epsilon": 1.5}},
    1: {(0, 0): {"epsilon": 1.0, "kappa": 2.0, "delta": 0.2},
        (0, 1): {"epsilon": 0.6, "kappa": 2.0, "delta": 0.35},
        (1, 1): {"epsilon": 1.4, "kappa": 2.0, "delta": 0.05}},
    2: {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5},
        (0, 1): {"epsilon": 0.8, "sigma": 0.9, "attraction_scale_factor": 0.3},
        (1, 1): {"epsilon": 1.2, "sigma": 0.95, "attraction_scale_factor": 0.7}},
    3: {(0, 0): {"A": 144.0, "a_1": 0.0, "a_2": 0.0, "sigma": 1.0},
        (0, 1): {"A": 100.0, "a_1": 0.0, "a_2": 1.2, "sigma": 1.0},
        (1, 1): {"A": 40.0, "a_1": 1.2, "a_2": 1.2, "sigma": 1.0}},
    4: {(0, 0): {"A": 25.0, "gamma": 4.5, "s": 0.5},
        (0, 1): {"A": 40.0, "gamma": 4.5, "s": 0.5},
        (1, 1): {"A": 25.0, "gamma": 4.5, "s": 0.5}},
}
ISO_RCUT = {0: {(0, 0): 1.0, (0, 1): 1.3, (1, 1): 1.1},
            1: 3.0,
            2: {(0, 0): 2.5, (0, 1): 2.2, (1, 1): 3.0},
            3: {(0, 0): 3.0, (0, 1): 4.5, (1, 1): 5.0},
            4: 1.0}
ISO_LATTICE = {0: (18, 15.0, 0.25), 1: (18, 21.0, 0.15), 2: (18, 19.0, 0.2), 3: (12, 26.0, 0.1),
               4: (18, 11.0, 0.3)}


def _rc_matrix(r_cut, nt):
    m = np.zeros(nt * nt)
    for a in range(nt):
        for b in range(nt):
            m[a * nt + b] = r_cut[(min(a, b), max(a, b))] if isinstance(r_cut, dict) else r_cut
    return m


def _oracle_iso(ev, pos4, L, nt, params, r_cut, mode, r_on=None, r_buff=0.4):
    rc = _rc_matrix(r_cut, nt)
    ron = _rc_matrix(r_on, nt) if r_on is not None else np.zeros(nt * nt)
    pairs = oracle.nlist(pos4, L, nt, oracle.r_listsq_matrix(rc, r_buff), "cell")
    packed = oracle.pack_matrix(ev, params, nt)
    f, v = oracle.pair_forces(ev, pos4, L, pairs, nt, packed, rc * rc, ron * ron, mode)
    return f, v, pairs


def _colloid_positions(n, L, seed):
    """Solvent (type 0) and colloids (type 1) kept apart by > 2 a + 0.4."""
    pos4, rng = jittered_lattice(n, L, 0.1, 1, seed)
    # make every 9th site a colloid, but only where no other colloid is near
    colloid = []
    for i in range(0, pos4.shape[0], 9):
        ok = True
        for j in colloid:
            d = pos4[i, :3] - pos4[j, :3]
            d -= L * np.rint(d / L)
            if np.dot(d, d) < 3.0**2:
                ok = False
                break
        if ok:
            colloid.append(i)
    pos4[colloid, 3] = 1
    # remove solvent overlapping a colloid core
    keep = np.ones(pos4.shape[0], dtype=bool)
    for j in colloid:
        d = pos4[:, :3] - pos4[j, :3]
        d -= L * np.rint(d / L)
        r = np.sqrt((d * d).sum(axis=1))
        keep &= ~((pos4[:, 3] == 0) & (r < 1.2 + 0.75))
    return pos4[keep].copy()


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
@pytest.mark.parametrize("ev", [0, 1, 2, 3, 4])
def test_iso_fp64_and_mixed(ev, mode):
    n, L, jit = ISO_LATTICE[ev]
    if ev == 3:
        pos4 = _colloid_positions(n, L, seed=10 + ev)
    else:
        pos4, _ = jittered_lattice(n, L, jit, 2, seed=10 + ev)
    Lv = np.array([L, L * 1.03, L * 0.98])
    pos4[:, :3] -= Lv * np.floor(pos4[:, :3] / Lv + 0.5)
    nt = 2
    r_on = {k: 0.7 * v for k, v in ISO_RCUT[ev].items()} if isinstance(ISO_RCUT[ev], dict) else 0.7 * ISO_RCUT[ev]
    cls = CONSERVATIVE_DPD if ev == 4 else CLASS[ev]
    say(f"[iso] evaluator={ev} mode={mode} N={pos4.shape[0]}")
    fo, vo, pairs = _oracle_iso(ev, pos4, Lv, nt, ISO_PARAMS[ev], ISO_RCUT[ev], mode, r_on)
    sysm = GpuSystem(pos4, Lv)
    r = gpu_pair(cls, sysm, nt, ISO_PARAMS[ev], ISO_RCUT[ev], mode=mode, r_on=r_on)
    # neighbor sets identical to the oracle's
    sg = nlist_pairs(r.nl, sysm.N)
    so = set(zip(pairs[0].tolist(), pairs[1].tolist()))
    assert sg == so
    ef = rel_err(r.force[:, :3], fo[:, :3])
    ee = rel_err(r.force[:, 3], fo[:, 3])
    ev_ = rel_err(r.virial, vo)
    say(f"  fp64 rel err force={ef:.2e} energy={ee:.2e} virial={ev_:.2e}")
    assert ef < FP64_TOL and ee < FP64_TOL and ev_ < FP64_TOL
    if ev == 3:
        return  # Colloid: plain FP64 form only
    rm = gpu_pair(cls, sysm, nt, ISO_PARAMS[ev], ISO_RCUT[ev], mode=mode, r_on=r_on, precision="mixed")
    mf = rel_err(rm.force[:, :3], fo[:, :3])
    me = rel_err(rm.force[:, 3], fo[:, 3])
    mv = rel_err(rm.virial, vo)
    say(f"  mixed rel err force={mf:.2e} energy={me:.2e} virial={mv:.2e}")
    assert mf < MIXED_TOL and me < MIXED_TOL and mv < MIXED_TOL


@pytest.mark.parametrize("L,rc,jit", [(7.0, 2.0, 0.3), (9.5, 1.5, 0.45), (5.0, 2.0, 0.35)])
def test_nlist_small_boxes_and_types(L, rc, jit):
    """Grids narrower than 3 cells (deduplicated stencil) and per-type cutoffs."""
    pos4, _ = jittered_lattice(9, L, jit, 3, seed=int(L * 7))
    Lv = np.array([L, L, L])
    nt = 3
    r_cut = {(a, b): rc * (1.0 - 0.15 * (a + b)) for a in range(nt) for b in range(a, nt)}
    r_cut[(0, 2)] = 0.0  # excluded type pair
    params = {k: {"epsilon": 1.0} for k in r_cut}
    say(f"[nlist] L={L} rc={rc}")
    r = gpu_pair("PotentialPairHertz", GpuSystem(pos4, Lv), nt, params, r_cut, r_buff=0.3)
    rcm = _rc_matrix(r_cut, nt)
    pi, pj = oracle.nlist(pos4, Lv, nt, oracle.r_listsq_matrix(rcm, 0.3), "brute")
    assert nlist_pairs(r.nl, pos4.shape[0]) == set(zip(pi.tolist(), pj.tolist()))


def test_nlist_overflow_grows_and_matches():
    """Dense clump: the first capacity estimate overflows; the list grows and retries."""
    rng = np.random.default_rng(3)
    L = 12.0
    pos = rng.uniform(-L / 2, L / 2, size=(3000, 4))
    pos[:1500, :3] = rng.uniform(-1.0, 1.0, size=(1500, 3))  # clump
    pos[:, 3] = 0
    Lv = np.array([L, L, L])
    params = {(0, 0): {"epsilon": 1.0}}
    say("[nlist overflow]")
    r = gpu_pair("PotentialPairHertz", GpuSystem(pos, Lv), 1, params, 1.5, r_buff=0.4)
    pi, pj = oracle.nlist(pos, Lv, 1, oracle.r_listsq_matrix([1.5], 0.4), "cell")
    assert nlist_pairs(r.nl, pos.shape[0]) == set(zip(pi.tolist(), pj.tolist()))
    fo, _ = oracle.pair_forces(0, pos, Lv, (pi, pj), 1, oracle.pack_matrix(0, params, 1), np.array([2.25]))
    assert rel_err(r.force[:, :3], fo[:, :3]) < FP64_TOL


def test_empty_and_isolated():
    """N with no neighbors at all: zero forces, empty list."""
    pos4 = np.array([[0.0, 0.0, 0.0, 0.0], [5.0, 5.0, 5.0, 0.0], [-5.0, 2.0, 1.0, 0.0]])
    Lv = np.array([20.0, 20.0, 20.0])
    r = gpu_pair("PotentialPairPerturbedLennardJones", GpuSystem(pos4, Lv), 1,
                 {(0, 0): ISO_PARAMS[2][(0, 0)]}, 3.0)
    assert np.all(r.force == 0.0)
    assert nlist_pairs(r.nl, 3) == set()


@pytest.mark.parametrize("tiled", [False, True])
def test_coincident_particles_match_the_oracle(tiled):
    """Two particles at one point: r^2 = 0 and the FP64 reciprocal gives
    1/0 = +inf as IEEE division does, so the pair's energy is +inf and its
    force and virial are NaN, as in the oracle.  Every other value is finite
    and equals the oracle's to 1e-12."""
    pos4, _ = jittered_lattice(8, 9.0, 0.1, 1, seed=51)
    pos4 = np.concatenate([pos4, pos4[:1]])
    N = pos4.shape[0]
    Lv = np.array([9.0, 9.0, 9.0])
    params = {(0, 0): ISO_PARAMS[2][(0, 0)]}
    rc = 2.5
    r = gpu_pair("PotentialPairPerturbedLennardJones", GpuSystem(pos4, Lv), 1, params, rc, tiled=tiled)
    pairs = oracle.nlist(pos4, Lv, 1, oracle.r_listsq_matrix([rc], 0.4), "cell")
    fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, Lv, pairs, 1,
                                oracle.pack_matrix(oracle.PERTURBED_LJ, params, 1), np.array([rc * rc]))
    say(f"[coincident] tiled={tiled}: energy {r.force[0, 3]}, force {r.force[0, :3]}")
    assert np.isposinf(r.force[0, 3]) and np.isposinf(r.force[N - 1, 3])
    assert np.all(np.isnan(r.force[[0, N - 1], :3]))
    for got, ref in ((r.force, fo), (r.virial, vo)):
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(np.isposinf(got), np.isposinf(ref))
        assert np.array_equal(np.isneginf(got), np.isneginf(ref))
    fin = np.isfinite(fo).all(axis=1)
    assert fin.sum() == N - 2
    assert rel_err(r.force[fin], fo[fin]) < 1e-12
    vfin = np.isfinite(vo).all(axis=0)
    assert rel_err(r.virial[:, vfin], vo[:, vfin]) < 1e-12


def test_invalid_inputs_are_errors():
    """Non-finite positions and type indices without parameters are reported
    as errors by the neighbor-list build and the rebuild check (no silent
    self-interaction, no out-of-range table reads)."""
    from azplugins_amd import _azp

    L = 10.0
    pos4, _ = jittered_lattice(6, L, 0.1, 1, seed=1)
    Lv = np.array([L, L, L])
    params = {(0, 0): {"epsilon": 1.0}}
    bad = pos4.copy()
    bad[5, 1] = np.nan
    with pytest.raises(_azp.AzpError, match="not finite"):
        gpu_pair("PotentialPairHertz", GpuSystem(bad, Lv), 1, params, 1.5)
    bad = pos4.copy()
    bad[7, 3] = 3.0
    with pytest.raises(_azp.AzpError, match="type index"):
        gpu_pair("PotentialPairHertz", GpuSystem(bad, Lv), 1, params, 1.5)
    # NaN appearing after the build is caught by the rebuild check
    sysm = GpuSystem(pos4, Lv)
    r = gpu_pair("PotentialPairHertz", sysm, 1, params, 1.5)
    sysm.pos[3, 0] = float("nan")
    with pytest.raises(_azp.AzpError, match="not finite"):
        r.nl.update(sysm.particles())


@pytest.mark.parametrize("tiled", [False, True])
def test_failed_build_forces_a_rebuild(tiled):
    """A build that fails (a type index without parameters) leaves no list.
    The particles move by more than r_buff / 2 after a good build, the rebuild
    fails on a bad type; once the type is valid again, the next update must
    rebuild although the positions did not move since the failed attempt, and
    the forces equal the oracle's."""
    from azplugins_amd import _azp, _azplugins

    L = 12.0
    pos4, rng = jittered_lattice(12, L, 0.1, 1, seed=5)
    Lv = np.array([L, L, L])
    params = {(0, 0): ISO_PARAMS[2][(0, 0)]}
    sysm = GpuSystem(pos4, Lv)
    nl = _azplugins.NeighborListGPU(0.4)
    pp = _azplugins.PotentialPairPerturbedLennardJonesGPU(1)
    pp.setParams(0, 0, [1.0, 1.0, 0.5])
    pp.setRCut(0, 0, 3.0)
    nl.set_r_cut(1, pp.getRCutMatrix())
    if tiled:
        nl.set_tiled(True)
    nl.update(sysm.particles())
    assert nl.num_builds() == 1
    moved = pos4.copy()
    moved[:, :3] += rng.uniform(-0.35, 0.35, size=(pos4.shape[0], 3))
    moved[:, :3] -= Lv * np.floor(moved[:, :3] / Lv + 0.5)
    bad = moved.copy()
    bad[7, 3] = 3.0
    sysm.pos.copy_(_azp.torch.tensor(bad, device="cuda"))
    with pytest.raises(_azp.AzpError, match="type index"):
        nl.update(sysm.particles())
    sysm.pos[7, 3] = 0.0
    nl.update(sysm.particles())
    assert nl.num_builds() == 2
    N = pos4.shape[0]
    pitch = (N + 63) // 64 * 64
    force = _azp.torch.zeros((N, 4), dtype=_azp.torch.float64, device="cuda")
    vir = _azp.torch.zeros((6, pitch), dtype=_azp.torch.float64, device="cuda")
    pp.compute(nl.handle, sysm.particles(), _azp.Outputs(force.data_ptr(), vir.data_ptr(), pitch, None), 0)
    _azp.torch.cuda.synchronize()
    fo, _, _ = _oracle_iso(2, moved, Lv, 1, params, 3.0, "none")
    err = rel_err(force.cpu().numpy()[:, :3], fo[:, :3])
    say(f"[rebuild after error] tiled={tiled} fp64 rel err {err:.2e}")
    assert err < FP64_TOL
