"""Pin the oracle where the reference's golden vectors leave it open.

Each check is chosen so that a plausible mistake (dropped term, wrong sign,
transposed index) fails it:

* F = -dU/dr by central finite differences for every evaluator and every mode
  (none / shift / xplor; xplor has no reference test),
* Newton's third law and the per-particle virial against the affine-strain
  derivative of the total energy (no reference test covers virials),
* TwoPatchMorse torques against the energy change under infinitesimal rotation,
* neighbor lists: cell-binned == brute-force pair sets, including small boxes,
* DPD: dissipative and random terms against their closed forms, the random
  variate's symmetry and moments, the shared portable pow against libm.
"""

import numpy as np
import pytest

RNG = np.random.default_rng(12345)


def min_image(d, L):
    return d - L * np.rint(d / L)


def random_fluid(N, L, ntypes=1, min_sep=0.0, seed=0):
    rng = np.random.default_rng(seed)
    pos = np.zeros((N, 4))
    k = 0
    while k < N:
        x = rng.uniform(-L / 2, L / 2, size=3)
        if min_sep > 0 and k > 0:
            d = min_image(pos[:k, :3] - x, L)
            if np.min(np.sum(d * d, axis=1)) < min_sep**2:
                continue
        pos[k, :3] = x
        pos[k, 3] = rng.integers(ntypes)
        k += 1
    return pos


ISO = {
    0: ({"epsilon": 1.7}, 1.5),
    1: ({"epsilon": 1.3, "kappa": 1.5, "delta": 0.4}, 3.0),
    2: ({"epsilon": 1.2, "sigma": 0.95, "attraction_scale_factor": 0.4}, 2.5),
}
COLLOID_CASES = [
    ({"A": 100.0, "a_1": 0.0, "a_2": 0.0, "sigma": 1.0}, 3.0, 0.95, 2.9),
    ({"A": 100.0, "a_1": 1.5, "a_2": 0.0, "sigma": 1.05}, 6.0, 2.3, 5.9),
    ({"A": 100.0, "a_1": 1.5, "a_2": 0.75, "sigma": 1.05}, 6.0, 2.7, 5.9),
]


def pair_energy(orc, ev, packed, rsq, rcutsq, ronsq, mode):
    """Two particles through the full driver (shift/xplor logic included)."""
    d = np.sqrt(rsq)
    pos = np.array([[0.0, 0.0, 0.0, 0.0], [d, 0.0, 0.0, 0.0]])
    L = np.array([50.0, 50.0, 50.0])
    pairs = (np.array([0], dtype=np.uint32), np.array([1], dtype=np.uint32))
    f, _ = orc.pair_forces(ev, pos, L, pairs, 1, packed[None, :], np.array([rcutsq]),
                           np.array([ronsq]), mode)
    return f[0, 3] + f[1, 3], -f[0, 0]  # total energy, force on 1 along +x


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
@pytest.mark.parametrize("ev", [0, 1, 2, 3])
def test_force_is_minus_energy_gradient(orc, ev, mode):
    cases = []
    if ev == 3:
        for params, rc, r0, r1 in COLLOID_CASES:
            cases.append((params, rc, np.linspace(r0, r1, 7)))
    else:
        params, rc = ISO[ev]
        lo = 0.5 if ev != 1 else 0.9
        cases.append((params, rc, np.linspace(lo, rc * 0.98, 9)))
    for params, rc, rs in cases:
        p = orc.pack(ev, params)
        ron = 0.6 * rc
        for r in rs:
            h = 1e-6 * r
            ep, _ = pair_energy(orc, ev, p, (r + h) ** 2, rc * rc, ron * ron, mode)
            em, _ = pair_energy(orc, ev, p, (r - h) ** 2, rc * rc, ron * ron, mode)
            _, f = pair_energy(orc, ev, p, r * r, rc * rc, ron * ron, mode)
            # the particle at +r feels F = -dU/dr along +x
            fd = -(ep - em) / (2 * h)
            assert f == pytest.approx(fd, rel=2e-6, abs=1e-7 * max(1.0, abs(f)))


def test_shift_zeroes_energy_at_cutoff(orc):
    for ev, (params, rc) in ISO.items():
        if ev == 0:
            continue  # Hertz has no shift branch; U(rc) = 0 already
        p = orc.pack(ev, params)
        r = rc * (1 - 1e-9)
        e, _ = pair_energy(orc, ev, p, r * r, rc * rc, 0.0, "shift")
        assert abs(e) < 1e-7


def test_xplor_switches_smoothly_and_falls_back_to_shift(orc):
    params, rc = ISO[2]
    p = orc.pack(2, params)
    ron = 0.7 * rc
    # S(r_on) = 1: unchanged inside r_on
    e0, f0 = pair_energy(orc, 2, p, (ron * 0.999) ** 2, rc * rc, ron * ron, "none")
    e1, f1 = pair_energy(orc, 2, p, (ron * 0.999) ** 2, rc * rc, ron * ron, "xplor")
    assert (e0, f0) == (e1, f1)
    # energy and force go to 0 at r_cut
    e2, f2 = pair_energy(orc, 2, p, (rc * (1 - 1e-7)) ** 2, rc * rc, ron * ron, "xplor")
    assert abs(e2) < 1e-9 and abs(f2) < 1e-6
    # r_on > r_cut: xplor degenerates to shift
    e3, f3 = pair_energy(orc, 2, p, 1.3**2, rc * rc, (1.1 * rc) ** 2, "xplor")
    e4, f4 = pair_energy(orc, 2, p, 1.3**2, rc * rc, 0.0, "shift")
    assert (e3, f3) == (e4, f4)


def _xplor_cases():
    cases = [(ev, params, rc, 0.6 * rc) for ev, (params, rc) in ISO.items()]
    cases += [(3, params, rc, r0 + 0.45 * (r1 - r0)) for params, rc, r0, r1 in COLLOID_CASES]
    return cases


@pytest.mark.parametrize("case", range(6))
def test_xplor_is_the_unique_cubic_switch(orc, case):
    """xplor multiplies E by a switch S(r) on [r_on, r_cut) and sets
    F = S f - (dS/dr) E.  The reference has no test of it.  Four conditions
    fix a polynomial of degree 3 in r^2 uniquely: S(r_on) = 1 and
    S'(r_on) = 0 (E and F continuous across r_on), S(r_cut) = 0 and
    S'(r_cut) = 0 (E and F vanish at r_cut).  The ratio E_xplor / E_none
    sampled inside the switching range must be such a cubic (its fourth
    divided difference in r^2 vanishes); with FD consistency of F (tested
    above) that pins the restated switch for every evaluator."""
    ev, params, rc, ron = _xplor_cases()[case]
    p = orc.pack(ev, params)
    rcsq, ronsq = rc * rc, ron * ron
    eps = 1e-9
    # continuity across r_on
    e_in, f_in = pair_energy(orc, ev, p, (ron * (1 - eps)) ** 2, rcsq, ronsq, "xplor")
    e_out, f_out = pair_energy(orc, ev, p, (ron * (1 + eps)) ** 2, rcsq, ronsq, "xplor")
    scale_e = max(abs(e_in), 1e-12)
    scale_f = max(abs(f_in), 1e-12)
    assert abs(e_out - e_in) < 1e-6 * scale_e + 1e-12
    assert abs(f_out - f_in) < 1e-5 * scale_f + 1e-10
    # below r_on the switch is off
    e_none, f_none = pair_energy(orc, ev, p, (ron * (1 - eps)) ** 2, rcsq, ronsq, "none")
    assert (e_in, f_in) == (e_none, f_none)
    # E and F vanish at r_cut
    e_c, f_c = pair_energy(orc, ev, p, (rc * (1 - 1e-8)) ** 2, rcsq, ronsq, "xplor")
    e_mid, f_mid = pair_energy(orc, ev, p, (0.5 * (ron + rc)) ** 2, rcsq, ronsq, "xplor")
    assert abs(e_c) < 1e-6 * abs(e_mid) + 1e-12
    assert abs(f_c) < 1e-4 * abs(f_mid) + 1e-9
    # the switch is a cubic in r^2: fourth divided difference of S(r^2) is 0
    xs = np.linspace(ronsq, rcsq, 7)[1:-1]
    S = []
    for x in xs:
        ex, _ = pair_energy(orc, ev, p, x, rcsq, ronsq, "xplor")
        en, _ = pair_energy(orc, ev, p, x, rcsq, ronsq, "none")
        assert en != 0.0
        S.append(ex / en)
    S = np.array(S)
    assert np.all((S > 0.0) & (S < 1.0))
    d = S.copy()
    for order in range(1, 5):
        d = (d[1:] - d[:-1]) / (xs[order:] - xs[:-order])
    coef3 = ((S[1:] - S[:-1]) / (xs[1:] - xs[:-1]))  # first divided differences, for scale
    assert abs(d[0]) < 1e-9 * np.abs(coef3).max() / (rcsq - ronsq) ** 3


def _many_body(orc, ev, params_by_pair, rc, N=160, L=7.0, ntypes=2, mode="none", seed=1):
    pos = random_fluid(N, L, ntypes, min_sep=0.85 if ev != 1 else 0.95, seed=seed)
    Lv = np.array([L, L, L * 1.1])
    packed = orc.pack_matrix(ev, params_by_pair, ntypes)
    rcut = np.full(ntypes * ntypes, rc)
    pairs = orc.nlist(pos, Lv, ntypes, orc.r_listsq_matrix(rcut, 0.4), method="brute")
    rcutsq = rcut**2
    ronsq = (0.6 * rcut) ** 2
    f, v = orc.pair_forces(ev, pos, Lv, pairs, ntypes, packed, rcutsq, ronsq, mode)
    return pos, Lv, packed, pairs, rcutsq, ronsq, f, v


def _total_energy_strained(orc, ev, pos, Lv, packed, pairs, ntypes, rcutsq, ronsq, mode, eps):
    """U with every pair separation mapped dx -> (I + eps) dx."""
    pi, pj = pairs
    d = min_image(pos[pi, :3] - pos[pj, :3], Lv) @ (np.eye(3) + eps).T
    # place each pair as an isolated two-particle system through the driver
    U = 0.0
    for k in range(d.shape[0]):
        rsq = float(d[k] @ d[k])
        tp = int(pos[pi[k], 3]) * ntypes + int(pos[pj[k], 3])
        e, _ = pair_energy(orc, ev, packed[tp], rsq, rcutsq[tp], ronsq[tp], mode)
        U += e
    return U


@pytest.mark.parametrize("mode", ["none", "xplor"])
@pytest.mark.parametrize("ev", [1, 2])
def test_many_body_third_law_and_virial(orc, ev, mode):
    if ev == 1:
        pp = {(0, 0): {"epsilon": 1.0, "kappa": 2.0, "delta": 0.2},
              (0, 1): {"epsilon": 0.7, "kappa": 2.0, "delta": 0.35},
              (1, 1): {"epsilon": 1.4, "kappa": 2.0, "delta": 0.1}}
        rc = 2.5
    else:
        pp = {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5},
              (0, 1): {"epsilon": 0.8, "sigma": 0.9, "attraction_scale_factor": 0.3},
              (1, 1): {"epsilon": 1.2, "sigma": 0.85, "attraction_scale_factor": 0.7}}
        rc = 2.5
    pos, Lv, packed, pairs, rcutsq, ronsq, f, v = _many_body(orc, ev, pp, rc, mode=mode)
    # Newton's third law: net force vanishes
    assert np.allclose(f[:, :3].sum(axis=0), 0.0, atol=1e-9 * np.abs(f[:, :3]).max())
    # W_ab (total) = -dU/d eps_ab under affine strain of the pair separations
    W = v.sum(axis=1)
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    h = 1e-6
    for n, (a, b) in enumerate(comps):
        eps = np.zeros((3, 3))
        eps[a, b] = h
        up = _total_energy_strained(orc, ev, pos, Lv, packed, pairs, 2, rcutsq, ronsq, mode, eps)
        um = _total_energy_strained(orc, ev, pos, Lv, packed, pairs, 2, rcutsq, ronsq, mode, -eps)
        dU = (up - um) / (2 * h)
        assert W[n] == pytest.approx(-dU, rel=1e-5, abs=1e-6 * np.abs(W).max())
    # per-particle force = -dU/dx_i by finite differences for a few particles
    U0 = lambda P: orc.pair_forces(ev, P, Lv, pairs, 2, packed, rcutsq, ronsq, mode)[0][:, 3].sum()
    for i in (0, 7, 42):
        for a in range(3):
            P = pos.copy()
            hh = 1e-6
            P[i, a] += hh
            up = U0(P)
            P[i, a] -= 2 * hh
            um = U0(P)
            assert f[i, a] == pytest.approx(-(up - um) / (2 * hh), rel=1e-5, abs=1e-6)


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
@pytest.mark.parametrize("ev", [0, 1, 2])
def test_threaded_full_list_equals_half_list(orc, ev, mode):
    """The CPU baseline's threaded full-list loop (orc_pair_forces_threaded over
    orc_full_list) gives the half-list third-law forces, energies and virials
    up to the order of the FP64 sums."""
    pp = {0: {(0, 0): {"epsilon": 1.0}, (0, 1): {"epsilon": 2.0}, (1, 1): {"epsilon": 0.5}},
          1: {(0, 0): {"epsilon": 1.0, "kappa": 2.0, "delta": 0.2},
              (0, 1): {"epsilon": 0.7, "kappa": 2.0, "delta": 0.35},
              (1, 1): {"epsilon": 1.4, "kappa": 2.0, "delta": 0.1}},
          2: {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5},
              (0, 1): {"epsilon": 0.8, "sigma": 0.9, "attraction_scale_factor": 0.3},
              (1, 1): {"epsilon": 1.2, "sigma": 0.85, "attraction_scale_factor": 0.7}}}[ev]
    pos, Lv, packed, pairs, rcutsq, ronsq, f, v = _many_body(orc, ev, pp, 2.5, mode=mode)
    csr = orc.full_list(pos.shape[0], pairs)
    assert csr[0][-1] == 2 * pairs[0].shape[0]
    ft, vt = orc.pair_forces_threaded(ev, pos, Lv, csr, 2, packed, rcutsq, ronsq, mode, nthreads=4)
    assert np.abs(ft - f).max() <= 1e-12 * np.abs(f).max()
    assert np.abs(vt - v).max() <= 1e-12 * np.abs(v).max()


@pytest.mark.parametrize("L,rc,ntypes", [(9.0, 1.2, 2), (4.1, 1.5, 1), (6.0, 2.5, 2)])
def test_nlist_cell_equals_brute(orc, L, rc, ntypes):
    pos = random_fluid(600, L, ntypes, seed=int(L * 10))
    Lv = np.array([L, L * 1.05, L * 0.97])
    rcut = np.array([rc, 0.8 * rc, 0.8 * rc, 0.6 * rc][: ntypes * ntypes])
    rl = orc.r_listsq_matrix(rcut, 0.3)
    a = orc.nlist(pos, Lv, ntypes, rl, "brute")
    b = orc.nlist(pos, Lv, ntypes, rl, "cell")
    sa = set(zip(a[0].tolist(), a[1].tolist()))
    sb = set(zip(b[0].tolist(), b[1].tolist()))
    assert sa == sb and len(sa) == a[0].shape[0] == b[0].shape[0]
    # independent numpy predicate
    d = min_image(pos[:, None, :3] - pos[None, :, :3], Lv)
    rsq = np.einsum("ijk,ijk->ij", d, d)
    t = pos[:, 3].astype(int)
    thr = rl.reshape(ntypes, ntypes)[t[:, None], t[None, :]]
    ii, jj = np.nonzero(np.triu(rsq <= thr, 1))
    assert set(zip(ii.tolist(), jj.tolist())) == sa


def test_dpd_terms_closed_form(orc):
    A, gamma, s = 25.0, 4.5, 0.5
    p = orc.pack(4, {"A": A, "gamma": gamma, "s": s})
    rc = 1.0
    dt, kT = 0.02, 1.3
    for trial in range(200):
        r = RNG.uniform(0.05, 0.999)
        rdotv = RNG.normal()
        ti, tj = (int(x) for x in RNG.integers(0, 2**32 - 1, size=2))
        ok, f, fc, e = orc.eval_dpd(r * r, rc * rc, p, 1234, ti, tj, 77, dt, kT, rdotv,
                                    portable_pow=False)
        assert ok
        w = (1 - r / rc) ** (0.5 * s) / r
        alpha = orc.dpd_alpha(1234, 77, ti, tj)[0]
        expect_c = A * (1 / r - 1 / rc)
        expect = expect_c - gamma * w * w * rdotv + np.sqrt(6 * gamma * kT / dt) * w * alpha
        assert fc == pytest.approx(expect_c, rel=1e-13)
        assert f == pytest.approx(expect, rel=1e-11, abs=1e-11)
        assert e == pytest.approx(A * (rc - r) - 0.5 * A / rc * (rc * rc - r * r), rel=1e-12, abs=1e-13)
    # kT = 0 switches off the random term exactly (test_pair.py:77)
    ok, f, fc, e = orc.eval_dpd(0.25, 1.0, p, 1, 3, 4, 0, dt, 0.0, 0.0)
    assert f == fc


def test_dpd_alpha_symmetry_and_moments(orc):
    n = 400000
    ti = RNG.integers(0, 2**32 - 1, size=n, dtype=np.uint64).astype(np.uint32)
    tj = RNG.integers(0, 2**32 - 1, size=n, dtype=np.uint64).astype(np.uint32)
    a = orc.dpd_alpha(42, 1000, ti, tj)
    b = orc.dpd_alpha(42, 1000, tj, ti)
    assert np.array_equal(a, b)
    assert np.all(np.abs(a) < 1.0)
    assert abs(a.mean()) < 5 * np.sqrt(1 / 3 / n)
    assert a.var() == pytest.approx(1 / 3, rel=0.01)
    # different timestep / seed give a different stream
    assert not np.array_equal(a[:100], orc.dpd_alpha(42, 1001, ti[:100], tj[:100]))
    assert not np.array_equal(a[:100], orc.dpd_alpha(43, 1000, ti[:100], tj[:100]))


def test_portable_pow_matches_libm(orc):
    xs = np.concatenate([np.linspace(1e-6, 1.0, 3001), RNG.uniform(0, 1, 3000), [1e-300, 0.5, 1.0]])
    ys = [0.0, 0.25, 0.5, 1.0, 1.5, 2.0, 0.1234, 3.7]
    worst = 0.0
    for y in ys:
        for x in xs:
            a = orc.ppow(float(x), y)
            b = float(np.power(x, y))
            if b == 0.0:
                assert a == 0.0 or abs(a) < 1e-300
                continue
            # exp(y log x) carries the conditioning |y ln x| of the product
            bound = (abs(y * np.log(x)) + 4.0) * 2.2205e-16
            worst = max(worst, abs(a - b) / abs(b) / bound)
    assert worst < 1.0
    for x in np.linspace(-700, 700, 2001):
        assert orc.pexp(float(x)) == pytest.approx(np.exp(x), rel=5e-16)
    for x in np.concatenate([np.logspace(-300, 300, 1001), RNG.uniform(0.5, 2, 1000)]):
        v = np.log(x)
        assert orc.plog(float(x)) == pytest.approx(v, rel=5e-16, abs=5e-16)


def _quat_mul(a, b):
    s1, v1 = a[0], a[1:]
    s2, v2 = b[0], b[1:]
    return np.concatenate([[s1 * s2 - v1 @ v2], s1 * v2 + s2 * v1 + np.cross(v1, v2)])


def test_two_patch_morse_force_and_torque_fd(orc):
    params = {"M_d": 1.3, "M_r": 0.3, "r_eq": 1.1, "omega": 6.0, "alpha": 0.3, "repulsion": True}
    p = orc.pack(5, params)
    rc = 2.5
    for trial in range(25):
        dr = RNG.normal(size=3)
        dr *= RNG.uniform(0.8, 2.3) / np.linalg.norm(dr)
        qi = RNG.normal(size=4)
        qi /= np.linalg.norm(qi)
        qj = RNG.normal(size=4)
        qj /= np.linalg.norm(qj)
        ok, F, e, ti, tj = orc.eval_two_patch_morse(dr, qi, qj, rc * rc, p)
        assert ok
        h = 1e-6
        for a in range(3):
            d1 = dr.copy()
            d1[a] += h
            d2 = dr.copy()
            d2[a] -= h
            ep = orc.eval_two_patch_morse(d1, qi, qj, rc * rc, p)[2]
            em = orc.eval_two_patch_morse(d2, qi, qj, rc * rc, p)[2]
            # dr = r_i - r_j, F acts on i: F = -dU/d(dr)
            assert F[a] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-7)
        for a in range(3):
            axis = np.zeros(3)
            axis[a] = 1.0
            rot = lambda th: np.concatenate([[np.cos(th / 2)], np.sin(th / 2) * axis])
            ep = orc.eval_two_patch_morse(dr, _quat_mul(rot(h), qi), qj, rc * rc, p)[2]
            em = orc.eval_two_patch_morse(dr, _quat_mul(rot(-h), qi), qj, rc * rc, p)[2]
            assert ti[a] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-7)
            ep = orc.eval_two_patch_morse(dr, qi, _quat_mul(rot(h), qj), rc * rc, p)[2]
            em = orc.eval_two_patch_morse(dr, qi, _quat_mul(rot(-h), qj), rc * rc, p)[2]
            assert tj[a] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-7)


def test_aniso_driver_third_law_and_virial_trace(orc):
    params = {"M_d": 1.0, "M_r": 0.2, "r_eq": 1.0, "omega": 20.0, "alpha": 0.4, "repulsion": True}
    L = 8.0
    pos = random_fluid(250, L, 1, min_sep=0.8, seed=5)
    Lv = np.array([L, L, L])
    q = RNG.normal(size=(250, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    packed = orc.pack_matrix(5, {(0, 0): params}, 1)
    pairs = orc.nlist(pos, Lv, 1, orc.r_listsq_matrix([2.5], 0.4), "brute")
    f, t, v = orc.aniso_forces(pos, q, Lv, pairs, 1, packed, np.array([6.25]), "shift")
    assert np.allclose(f[:, :3].sum(axis=0), 0, atol=1e-9 * np.abs(f).max())
    # trace of the virial = sum over pairs of dx . F (independent of index order)
    pi, pj = pairs
    d = min_image(pos[pi, :3] - pos[pj, :3], Lv)
    tr = 0.0
    for k in range(pi.shape[0]):
        ok, F, e, _, _ = orc.eval_two_patch_morse(d[k], q[pi[k]], q[pj[k]], 6.25, packed[0], True)
        if ok:
            tr += d[k] @ F
    assert v[0].sum() + v[3].sum() + v[5].sum() == pytest.approx(tr, rel=1e-12)


def test_aniso_virial_is_the_strain_derivative(orc):
    """Every stored component of the anisotropic virial, off-diagonal ones
    included, against the affine strain derivative of the total energy:
    W_ab = -dU/d eps_ab where eps_ab moves coordinate a of each separation by
    eps_ab times coordinate b (orientations held fixed).  With non-central
    forces W is not symmetric, and this fixes which of sum F_a dx_b and
    sum F_b dx_a the stored xy, xz and yz hold.  Mode none, where the force is
    the energy's gradient (the shift term of mode shift depends on the angles
    but is not differentiated, as in the reference's evaluator); the pairs
    inside r_cut are evaluated without the cut for the differences, so no
    pair crosses it."""
    params = {"M_d": 1.0, "M_r": 0.2, "r_eq": 1.0, "omega": 20.0, "alpha": 0.4, "repulsion": True}
    L = 7.0
    pos = random_fluid(150, L, 1, min_sep=0.85, seed=9)
    Lv = np.array([L, L, L])
    rng = np.random.default_rng(31)
    q = rng.normal(size=(150, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    packed = orc.pack_matrix(5, {(0, 0): params}, 1)
    pairs = orc.nlist(pos, Lv, 1, orc.r_listsq_matrix([2.5], 0.4), "brute")
    f, t, v = orc.aniso_forces(pos, q, Lv, pairs, 1, packed, np.array([6.25]), "none")
    W = v.sum(axis=1)
    pi, pj = pairs
    d0 = min_image(pos[pi, :3] - pos[pj, :3], Lv)
    inside = (d0 * d0).sum(axis=1) <= 6.25

    def energy(eps):
        d = d0[inside] @ (np.eye(3) + eps).T
        a, b = pi[inside], pj[inside]
        U = 0.0
        for k in range(d.shape[0]):
            ok, F, e, _, _ = orc.eval_two_patch_morse(d[k], q[a[k]], q[b[k]], 100.0, packed[0], False)
            assert ok
            U += e
        return U

    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    h = 1e-6
    asym = 0.0
    for n, (a, b) in enumerate(comps):
        eps = np.zeros((3, 3))
        eps[a, b] = h
        dU = (energy(eps) - energy(-eps)) / (2 * h)
        assert W[n] == pytest.approx(-dU, rel=1e-6, abs=1e-7 * np.abs(W).max())
        if a != b:
            # the transposed strain gives another number: the test tells the orders apart
            epsT = np.zeros((3, 3))
            epsT[b, a] = h
            dUT = (energy(epsT) - energy(-epsT)) / (2 * h)
            asym = max(asym, abs(dU - dUT) / np.abs(W).max())
    assert asym > 1e-2
