"""Oracle of the wall potentials, pinned without a GPU.

* the reference's golden values (test_wall.py:22-56): one particle above a
  plane at z = 0, energy and force along z to 4 decimals;
* F = -dE/dx by central differences for spheres, cylinders and planes, inside
  and outside, with and without extrapolation;
* E(r_cut) = 0 (the shift), continuity of E and F at r_extrap, and the open /
  closed surface rule;
* parameter read-back through the library's host code (the reference's asDict
  formulas: cbrt(sigma^3), 6 c_2, (7560 c_1 / A)^(1/6)).
"""

import ctypes

import numpy as np
import pytest

import oracle

PLANE_Z = [[0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0]]

# test_wall.py:22-56: (kind, values, z, energy, force_z)
GOLDEN = [
    (oracle.WALL_COLLOID, [1.0, 1.0, 1.0, 3.0, 0.0], 1.5, -0.0292, 0.8940),
    (oracle.WALL_COLLOID, [1.0, 1.0, 1.0, 3.0, 0.0], 3.5, 0.0, 0.0),
    (oracle.WALL_LJ93, [1.0, 1.0, 3.0, 0.0], 1.5, -0.2558, -0.5718),
    (oracle.WALL_LJ93, [1.0, 1.0, 3.0, 0.0], 3.5, 0.0, 0.0),
]


@pytest.mark.parametrize("case", GOLDEN)
def test_golden_plane(case):
    kind, vals, z, e, fz = case
    f, _ = oracle.wall_forces(kind, np.array([[0.0, 0.0, z, 0.0]]), [vals], planes=PLANE_Z)
    np.testing.assert_array_almost_equal(f[0, 3], e, decimal=4)
    np.testing.assert_array_almost_equal(f[0, :3], [0.0, 0.0, fz], decimal=4)


def _geometry():
    spheres = [[0.3, -0.2, 0.1, 5.0, 1.0, 1.0],   # inside a sphere
               [4.0, 4.0, 4.0, 1.2, 0.0, 1.0]]    # outside a small sphere
    cylinders = [[0.0, 0.0, 0.0, 0.2, 0.1, 1.0, 4.5, 1.0, 1.0]]
    planes = [[0.0, 0.0, -3.0, 0.1, -0.2, 1.0, 1.0]]
    return spheres, cylinders, planes


def _energy(kind, vals, x, geo):
    f, _ = oracle.wall_forces(kind, np.array([[*x, 0.0]]), [vals], *geo, with_virial=False)
    return f[0, 3], f[0, :3]


@pytest.mark.parametrize("kind,vals", [
    (oracle.WALL_LJ93, [1.3, 0.9, 2.5, 0.0]),
    (oracle.WALL_LJ93, [1.3, 0.9, 2.5, 0.6]),
    (oracle.WALL_COLLOID, [2.0, 0.5, 0.8, 2.2, 0.0]),
    (oracle.WALL_COLLOID, [2.0, 0.5, 0.8, 2.2, 0.8]),
])
def test_force_is_minus_gradient(kind, vals):
    geo = _geometry()
    rng = np.random.default_rng(4)
    h = 1e-6
    checked = 0
    for _ in range(200):
        x = rng.uniform(-4.0, 4.0, size=3)
        e0, f0 = _energy(kind, vals, x, geo)
        if e0 == 0.0 and not f0.any():
            continue
        g = np.zeros(3)
        for a in range(3):
            dx = np.zeros(3)
            dx[a] = h
            g[a] = (_energy(kind, vals, x + dx, geo)[0] - _energy(kind, vals, x - dx, geo)[0]) / (2 * h)
        # the colloid form is undefined closer than a to a wall (as in the
        # reference); skip such points, and points within h of a kink (a
        # cutoff, r_extrap or a surface)
        if not (np.isfinite(g).all() and np.isfinite(f0).all()):
            continue
        if np.abs(f0 + g).max() > 1e-4 * (1.0 + np.abs(f0).max()):
            e_near = [_energy(kind, vals, x + d, geo)[0] for d in np.eye(3) * 10 * h]
            if max(abs(en - e0) for en in e_near) > 1e3 * h * (1 + np.abs(f0).max()):
                continue
        np.testing.assert_allclose(f0, -g, rtol=1e-5, atol=1e-6 * (1.0 + np.abs(f0).max()))
        checked += 1
    assert checked > 60


@pytest.mark.parametrize("kind,vals", [
    (oracle.WALL_LJ93, [1.3, 0.9, 2.5, 0.6]),
    (oracle.WALL_COLLOID, [2.0, 0.5, 0.8, 2.2, 0.8]),
])
def test_shift_extrapolation_and_surface(kind, vals):
    r_cut, r_e = vals[-2], vals[-1]
    plane = [[0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0]]
    E = lambda z, pl=plane: oracle.wall_forces(kind, np.array([[0.0, 0.0, z, 0.0]]), [vals], planes=pl)[0][0]
    # zero energy at the cutoff, nothing beyond it
    assert abs(E(r_cut * (1 - 1e-9))[3]) < 1e-6
    assert not E(r_cut * (1 + 1e-9)).any()
    # continuity at r_extrap, and the linear form below it (also outside: z < 0)
    z0 = r_e * (1 - 1e-10)
    lo, hi = E(z0), E(r_e * (1 + 1e-10))
    np.testing.assert_allclose(lo, hi, rtol=1e-6, atol=1e-9)
    f_e = lo[2]
    for z in (0.5 * r_e, 0.0, -0.7):
        got = E(z)
        np.testing.assert_allclose(got[2], f_e, rtol=1e-14)
        np.testing.assert_allclose(got[3], lo[3] + f_e * (z0 - z), rtol=1e-12, atol=1e-12)
    # without extrapolation: outside and on an open surface nothing acts,
    # on a closed surface the (singular) evaluation at r = 0 is skipped by the cutoff test only
    v0 = list(vals[:-1]) + [0.0]
    E0 = lambda z, openflag: oracle.wall_forces(kind, np.array([[0.0, 0.0, z, 0.0]]), [v0],
                                                planes=[[0, 0, 0, 0, 0, 1, openflag]])[0][0]
    assert not E0(-0.3, 1.0).any() and not E0(0.0, 1.0).any()


def test_sphere_inside_and_outside_are_complementary():
    vals = [1.0, 1.0, 3.0, 0.0]
    x = np.array([[0.0, 0.0, 4.0, 0.0], [0.0, 0.0, 6.0, 0.0]])
    fin, _ = oracle.wall_forces(oracle.WALL_LJ93, x, [vals], spheres=[[0, 0, 0, 5.0, 1.0, 1.0]])
    fout, _ = oracle.wall_forces(oracle.WALL_LJ93, x, [vals], spheres=[[0, 0, 0, 5.0, 0.0, 1.0]])
    fpl, _ = oracle.wall_forces(oracle.WALL_LJ93, np.array([[0, 0, 1.0, 0]]), [vals], planes=PLANE_Z)
    # distance 1 from the surface on either side: the plane's value, pointing away from the surface
    np.testing.assert_allclose(fin[0], [0, 0, -fpl[0, 2], fpl[0, 3]], rtol=1e-14)
    np.testing.assert_allclose(fout[1], fpl[0], rtol=1e-14)
    assert not fin[1].any() and not fout[0].any()


@pytest.mark.parametrize("name,vals", [
    ("WallsPotentialLJ93", [1.0, 1.0, 3.0, 0.0]),
    ("WallsPotentialLJ93GPU", [2.0 * np.pi / 3.0, 1.05, 2.5, 0.2]),
    ("WallsPotentialColloid", [1.0, 1.0, 1.0, 3.0, 0.0]),
    ("WallsPotentialColloidGPU", [4.0 * np.pi**2, 1.0, 1.0, 3.0, 0.0]),
])
def test_param_readback(name, vals):
    from azplugins_amd import _azplugins

    w = getattr(_azplugins, name)(2)
    w.setParams(1, vals)
    got = w.getParams(1)
    # the asDict formulas return the inputs exactly for these values (as the
    # reference's test asserts after attach, test_wall.py:84-86)
    assert got == pytest.approx(vals, rel=1e-15, abs=0.0)
    if name.startswith("WallsPotentialLJ93"):
        assert got == vals


def test_wall_argument_checks():
    from azplugins_amd import _azp, _azplugins

    L = _azp.lib()
    h = ctypes.c_void_p()
    assert L.azp_wall_create(ctypes.byref(h), b"WallsPotentialNope", 1) == 1
    w = _azplugins.WallsPotentialLJ93GPU(1)
    with pytest.raises(_azp.AzpError):
        w.setParams(1, [1.0, 1.0, 3.0, 0.0])   # type out of range
    with pytest.raises(_azp.AzpError):
        w.setGeometry([], [], [[0, 0, 0, 0, 0, 0, 1]])  # zero normal
    with pytest.raises(_azp.AzpError):
        w.setGeometry([[0, 0, 0, -1.0, 1, 1]], [], [])  # negative radius
