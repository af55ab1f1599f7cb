"""The kernel-ABI launchers as a HOOMD caller binds them.

* The mixed form over the plain list with d_pos_fixed = NULL (the binding
  shown for HOOMD's PotentialPairGPU): the launcher packs the fixed-point
  positions into its own scratch; the result equals, bit for bit, the call
  that passes positions packed by azp_gpu_pack_fixed, also after the scratch
  has to grow for a larger system.
* HOOMD's head-list layout (d_head_list, unpadded rows separated by gaps)
  against the padded column layout: the same entries in the same order, so
  forces, energies, virials and torques are bit-identical, for the isotropic
  FP64 and mixed forms, DPD with its thermostat, and TwoPatchMorse.
* The DPD launcher over the tiled list (with and without d_nlist) against
  the column layout of the same build: bit-identical.
"""

import ctypes

import numpy as np
import pytest

from azplugins_amd import _azp
from gpu_util import GpuSystem, gpu_pair, jittered_lattice, raw_args, say

torch = _azp.torch
pytestmark = pytest.mark.gpu

PLJ = {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5}}


def _pod(evaluator, values):
    size = _azp.lib().azp_param_size(evaluator)
    pod = (ctypes.c_char * size)()
    _azp.call("azp_param_pack", evaluator, _azp.dbl_array(values), pod)
    return np.frombuffer(bytes(pod), dtype=np.uint8).copy()


def _params_dev(evaluator, rows):
    return torch.tensor(np.concatenate([_pod(evaluator, r) for r in rows]), device="cuda")


def _outputs(N):
    pitch = (N + 63) // 64 * 64
    f = torch.full((N, 4), np.nan, dtype=torch.float64, device="cuda")
    v = torch.full((6, pitch), np.nan, dtype=torch.float64, device="cuda")
    t = torch.full((N, 4), np.nan, dtype=torch.float64, device="cuda")
    return f, v, t, pitch


def _host(*ts):
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in ts]


@pytest.mark.parametrize("n", [10, 14])
def test_mixed_plain_list_without_pos_fixed(n):
    L = 1.07 * n
    pos4, _ = jittered_lattice(n, L, 0.15, 1, seed=40 + n)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    ref = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0, precision="mixed")
    d_params = _params_dev(_azp.EVAL_PERTURBED_LJ, [[1.0, 1.0, 0.5]])
    rcsq = torch.tensor([9.0], dtype=torch.float64, device="cuda")
    N = sysm.N
    f1, v1, _, pitch = _outputs(N)
    a = raw_args(sysm, ref.nl, rcsq, None, 1, f1, v1, pitch, precision=_azp.PRECISIONS["mixed"])
    assert not a.particles.d_pos_fixed
    say(f"[launcher] mixed, plain list, d_pos_fixed NULL, N={N}")
    _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(a), _azp.EVAL_PERTURBED_LJ, _azp.dptr(d_params))
    f2, v2, _, _ = _outputs(N)
    pf = torch.empty((N, 4), dtype=torch.int32, device="cuda")
    box = a.particles.box
    _azp.call("azp_gpu_pack_fixed", sysm.pos.data_ptr(), N, ctypes.byref(box), pf.data_ptr(), _azp.stream_ptr())
    b = raw_args(sysm, ref.nl, rcsq, None, 1, f2, v2, pitch, precision=_azp.PRECISIONS["mixed"])
    b.particles.d_pos_fixed = pf.data_ptr()
    _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(b), _azp.EVAL_PERTURBED_LJ, _azp.dptr(d_params))
    h1, w1, h2, w2 = _host(f1, v1, f2, v2)
    assert np.array_equal(h1, h2) and np.array_equal(w1[:, :N], w2[:, :N])
    assert np.array_equal(h1, ref.force) and np.array_equal(w1[:, :N], ref.virial)


def _head_layout(nl, N, seed):
    """HOOMD-style flat list + head list from the column list: each row holds
    n_neigh[i] entries followed by a gap of 0-5 slots filled with index 0 (a
    valid particle: an entry read past n_neigh[i] changes the result)."""
    nn, lst = nl.download(N)
    rng = np.random.default_rng(seed)
    gaps = rng.integers(0, 6, size=N)
    cap = nn.astype(np.int64) + gaps
    head = np.zeros(N, dtype=np.uint64)
    head[1:] = np.cumsum(cap)[:-1]
    flat = np.zeros(int(cap.sum()) + 1, dtype=np.uint32)
    for i in range(N):
        flat[int(head[i]):int(head[i]) + nn[i]] = lst[: nn[i], i]
    d_flat = torch.tensor(flat.astype(np.int64).astype(np.int32), device="cuda")
    d_head = torch.tensor(head.astype(np.int64), device="cuda")
    return d_flat, d_head


def _to_head(a, d_flat, d_head):
    a.d_nlist = d_flat.data_ptr()
    a.d_head_list = d_head.data_ptr()
    a.nlist_pitch = 0
    return a


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("ev", [_azp.EVAL_PERTURBED_LJ, _azp.EVAL_EXPANDED_YUKAWA])
def test_head_list_iso_bit_identical(ev, precision):
    L = 13.0
    pos4, _ = jittered_lattice(12, L, 0.2, 2, seed=7)
    Lv = np.array([L, L * 1.02, L])
    sysm = GpuSystem(pos4, Lv)
    if ev == _azp.EVAL_PERTURBED_LJ:
        cls, rows = "PotentialPairPerturbedLennardJones", [[1.0, 1.0, 0.5], [0.8, 0.9, 0.3], [0.8, 0.9, 0.3],
                                                           [1.2, 0.95, 0.7]]
        params = {(0, 0): dict(zip(("epsilon", "sigma", "attraction_scale_factor"), rows[0])),
                  (0, 1): dict(zip(("epsilon", "sigma", "attraction_scale_factor"), rows[1])),
                  (1, 1): dict(zip(("epsilon", "sigma", "attraction_scale_factor"), rows[3]))}
    else:
        cls, rows = "PotentialPairExpandedYukawa", [[1.0, 2.0, 0.2], [0.6, 2.0, 0.35], [0.6, 2.0, 0.35],
                                                    [1.4, 2.0, 0.05]]
        params = {(0, 0): dict(zip(("epsilon", "kappa", "delta"), rows[0])),
                  (0, 1): dict(zip(("epsilon", "kappa", "delta"), rows[1])),
                  (1, 1): dict(zip(("epsilon", "kappa", "delta"), rows[3]))}
    ref = gpu_pair(cls, sysm, 2, params, 2.5, precision=precision)
    d_params = _params_dev(ev, rows)
    rcsq = torch.full((4,), 6.25, dtype=torch.float64, device="cuda")
    N = sysm.N
    d_flat, d_head = _head_layout(ref.nl, N, seed=ev)
    results = []
    for head in (False, True):
        f, v, _, pitch = _outputs(N)
        a = raw_args(sysm, ref.nl, rcsq, None, 2, f, v, pitch, precision=_azp.PRECISIONS[precision])
        if head:
            _to_head(a, d_flat, d_head)
        say(f"[head list] evaluator={ev} {precision} head={head}")
        _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(a), ev, _azp.dptr(d_params))
        results.append(_host(f, v))
    assert np.array_equal(results[0][0], results[1][0])
    assert np.array_equal(results[0][1][:, :N], results[1][1][:, :N])
    assert np.array_equal(results[0][0], ref.force)


def test_head_list_dpd_and_aniso_bit_identical():
    L = 10.0
    rng = np.random.default_rng(3)
    N = int(3.0 * L**3)
    pos4 = np.zeros((N, 4))
    pos4[:, :3] = rng.uniform(-L / 2, L / 2, size=(N, 3))
    pos4[:, 3] = rng.integers(0, 2, size=N)
    vel4 = np.zeros((N, 4))
    vel4[:, :3] = rng.normal(size=(N, 3))
    vel4[:, 3] = 1.0
    Lv = np.array([L, L, L])
    dpd = {(0, 0): {"A": 25.0, "gamma": 4.5, "s": 0.5}, (0, 1): {"A": 40.0, "gamma": 4.5, "s": 0.5},
           (1, 1): {"A": 25.0, "gamma": 4.5, "s": 0.5}}
    sysm = GpuSystem(pos4, Lv, vel4)
    seed, t, dt, kT = 11, 4242, 0.02, 1.0
    ref = gpu_pair("PotentialPairDPDThermoGeneralWeight", sysm, 2, dpd, 1.0, thermostat=(seed, dt, kT), timestep=t)
    rows = [[25.0, 4.5, 0.5], [40.0, 4.5, 0.5], [40.0, 4.5, 0.5], [25.0, 4.5, 0.5]]
    d_params = _params_dev(_azp.EVAL_DPD_GENERAL_WEIGHT, rows)
    rcsq = torch.ones(4, dtype=torch.float64, device="cuda")
    d_flat, d_head = _head_layout(ref.nl, N, seed=1)
    results = []
    for head in (False, True):
        f, v, _, pitch = _outputs(N)
        a = raw_args(sysm, ref.nl, rcsq, None, 2, f, v, pitch)
        a.seed, a.timestep, a.dt, a.kT = seed, t, dt, kT
        if head:
            _to_head(a, d_flat, d_head)
        say(f"[head list] dpd head={head}")
        _azp.call("azp_gpu_compute_dpd_forces", ctypes.byref(a), _azp.dptr(d_params))
        results.append(_host(f, v))
    assert np.array_equal(results[0][0], results[1][0])
    assert np.array_equal(results[0][1][:, :N], results[1][1][:, :N])
    assert np.array_equal(results[0][0], ref.force)

    # TwoPatchMorse on a jittered lattice with random orientations
    pos4, rng = jittered_lattice(11, 12.0, 0.2, 1, seed=8)
    q = rng.normal(size=(pos4.shape[0], 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    Lv = np.array([12.0, 12.0, 12.0])
    sysm = GpuSystem(pos4, Lv, quat=q)
    tpm = {(0, 0): {"M_d": 1.0, "M_r": 0.2, "r_eq": 1.0, "omega": 20.0, "alpha": 0.4, "repulsion": True}}
    d_params = _params_dev(_azp.EVAL_TWO_PATCH_MORSE, [[1.0, 0.2, 1.0, 20.0, 0.4, 1.0]])
    rcsq = torch.tensor([6.25], dtype=torch.float64, device="cuda")
    N = sysm.N
    for precision in ("fp64", "mixed"):
        ref = gpu_pair("AnisoPotentialPairTwoPatchMorse", sysm, 1, tpm, 2.5, precision=precision)
        d_flat, d_head = _head_layout(ref.nl, N, seed=2)
        results = []
        for head in (False, True):
            f, v, tq, pitch = _outputs(N)
            a = raw_args(sysm, ref.nl, rcsq, None, 1, f, v, pitch, precision=_azp.PRECISIONS[precision])
            a.out.d_torque = tq.data_ptr()
            if head:
                _to_head(a, d_flat, d_head)
            say(f"[head list] two-patch Morse {precision} head={head}")
            _azp.call("azp_gpu_compute_pair_aniso_forces", ctypes.byref(a), _azp.EVAL_TWO_PATCH_MORSE,
                      _azp.dptr(d_params))
            results.append(_host(f, v, tq))
        for x, y in zip(results[0], results[1]):
            assert np.array_equal(x[:, :N] if x.shape[0] == 6 else x, y[:, :N] if y.shape[0] == 6 else y)
        assert np.array_equal(results[0][0], ref.force)
        assert np.array_equal(results[0][2][:, :3], ref.torque[:, :3])


def test_dpd_launcher_tiled_list_bit_identical():
    """azp_gpu_compute_dpd_forces given the tiled list (azp_nlist_get_tiles,
    azp_nlist_get_tile_order; the neighbor positions staged in LDS) gives the
    same bits as the column layout of the same build, and as the launcher
    with d_nlist NULL."""
    from azplugins_amd import _azplugins

    L = 12.0
    rng = np.random.default_rng(21)
    N = int(3.0 * L**3)
    pos4 = np.zeros((N, 4))
    pos4[:, :3] = rng.uniform(-L / 2, L / 2, size=(N, 3))
    pos4[:, 3] = rng.integers(0, 2, size=N)
    vel4 = np.zeros((N, 4))
    vel4[:, :3] = rng.normal(size=(N, 3))
    vel4[:, 3] = 1.0
    sysm = GpuSystem(pos4, np.array([L, L, L]), vel4)
    rows = [[25.0, 4.5, 0.5], [40.0, 4.5, 0.5], [40.0, 4.5, 0.5], [25.0, 4.5, 0.5]]
    d_params = _params_dev(_azp.EVAL_DPD_GENERAL_WEIGHT, rows)
    rcsq = torch.ones(4, dtype=torch.float64, device="cuda")
    nl = _azplugins.NeighborListGPU(0.4)
    nl.set_r_cut(2, [1.0] * 4)
    nl.set_tiled(True)
    nl.update(sysm.particles())
    tiles = nl.get_tiles()
    assert tiles[3] in (256, 512)
    order = nl.get_tile_order()
    results = []
    for form in ("column", "tiled", "tiled without d_nlist"):
        f, v, _, pitch = _outputs(N)
        a = raw_args(sysm, nl, rcsq, None, 2, f, v, pitch)
        a.seed, a.timestep, a.dt, a.kT = 7, 321, 0.02, 1.0
        if form != "column":
            a.d_tile_count, a.d_tile_jset, a.d_tile_list, a.tile_size, a.tile_slots = tiles
            a.d_tile_order, a.d_tile_nn = order
        if form == "tiled without d_nlist":
            a.d_nlist = None
        say(f"[launcher] dpd {form}")
        _azp.call("azp_gpu_compute_dpd_forces", ctypes.byref(a), _azp.dptr(d_params))
        results.append(_host(f, v))
    for r in results[1:]:
        assert np.array_equal(results[0][0], r[0])
        assert np.array_equal(results[0][1][:, :N], r[1][:, :N])
