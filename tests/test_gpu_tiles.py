"""Tiled neighbor list (azp_nlist_set_tiled) for the mixed-precision pair kernel.

The tiled list stores each block's entries as 16-bit ranks into the block's
sorted set of distinct neighbors, and the kernel reads neighbor positions from
LDS.  The per-pair arithmetic and the order of the terms are those of the
plain list, so forces, energies and virials must be bit-identical to the
mixed form over the plain list (and within 1e-5 of the FP64 oracle); blocks
with too many distinct neighbors fall back to the plain list.
"""

import numpy as np
import pytest

import oracle
from azplugins_amd import _azp
from gpu_util import GpuSystem, gpu_pair, jittered_lattice, rel_err, say

torch = _azp.torch
pytestmark = pytest.mark.gpu

PLJ = {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5}}


def _both(cls, sysm, nt, params, rc, tiled=True, **kw):
    plain = gpu_pair(cls, sysm, nt, params, rc, precision="mixed", **kw)
    tiled = gpu_pair(cls, sysm, nt, params, rc, precision="mixed", tiled=tiled, **kw)
    return plain, tiled


@pytest.mark.parametrize("mode", [256, 512])
def test_tiled_sizes_equal_plain(mode):
    L = 20.0
    pos4, _ = jittered_lattice(20, L, 0.1, 1, seed=3)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0, tiled=mode)
    size, slots = tiled.nl.get_tiles()[3:]
    say(f"[tiles] forced tile size {mode}: got {size}, LDS slots {slots}")
    assert size == mode and 0 < slots <= 3968
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)


@pytest.mark.parametrize("n,L", [(7, 7.3), (10, 10.4), (18, 18.9), (24, 25.3)])
def test_tiled_equals_plain_plj(n, L):
    """N = 343 .. 13824: one partial tile up to many tiles with a partial last one."""
    pos4, _ = jittered_lattice(n, L, 0.1, 1, seed=n)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
    _, _, _, size, slots = tiled.nl.get_tiles()
    say(f"[tiles] PLJ N={sysm.N}: tile size {size}, LDS slots {slots}")
    assert size in (256, 512)
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)
    pairs = oracle.nlist(pos4, Lv, 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
    fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, Lv, pairs, 1,
                                oracle.pack_matrix(oracle.PERTURBED_LJ, PLJ, 1), np.array([9.0]))
    assert rel_err(tiled.force[:, :3], fo[:, :3]) < 1e-5
    assert rel_err(tiled.virial, vo) < 1e-5


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
def test_tiled_equals_plain_two_types(mode):
    L = 16.0
    pos4, _ = jittered_lattice(16, L, 0.15, 2, seed=41)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    yk = {(0, 0): {"epsilon": 1.0, "kappa": 2.0, "delta": 0.2}, (0, 1): {"epsilon": 0.7, "kappa": 1.5, "delta": 0.3},
          (1, 1): {"epsilon": 1.3, "kappa": 2.5, "delta": 0.1}}
    hz = {(0, 0): {"epsilon": 1.0}, (0, 1): {"epsilon": 2.0}, (1, 1): {"epsilon": 0.5}}
    for cls, params, rc in (("PotentialPairExpandedYukawa", yk, 3.0), ("PotentialPairHertz", hz, 1.3)):
        if cls == "PotentialPairHertz" and mode != "none":
            continue
        r_on = 2.0 if mode == "xplor" else None
        plain, tiled = _both(cls, sysm, 2, params, rc, mode=mode, r_on=r_on)
        size = tiled.nl.get_tiles()[3]
        say(f"[tiles] {cls} {mode}: tile size {size}")
        assert size in (256, 512)
        assert np.array_equal(plain.force, tiled.force), cls
        assert np.array_equal(plain.virial, tiled.virial), cls


def test_tiles_fall_back_when_a_block_has_too_many_neighbors():
    """r_cut 7 at density 1.37: ~2300 neighbors per particle; the distinct neighbors of
    256 consecutive particles (~7000) exceed what a tile holds."""
    L = 18.0
    pos4, _ = jittered_lattice(20, L, 0.1, 1, seed=5)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 7.0)
    assert tiled.nl.get_tiles()[3] == 0
    assert np.array_equal(plain.force, tiled.force)


def test_tiled_launcher_entry_point():
    """azp_gpu_compute_pair_forces with the tile pointers in azp_pair_args."""
    from gpu_util import raw_args
    import ctypes

    L = 20.0
    pos4, _ = jittered_lattice(19, L, 0.1, 1, seed=9)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    ref = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0, precision="mixed", tiled=True)
    pod = (ctypes.c_char * 32)()
    _azp.call("azp_param_pack", 2, _azp.dbl_array([1.0, 1.0, 0.5]), pod)
    d_params = torch.tensor(np.frombuffer(bytes(pod), dtype=np.uint8).copy(), device="cuda")
    rcsq = torch.tensor([9.0], dtype=torch.float64, device="cuda")
    N = sysm.N
    pitch = (N + 63) // 64 * 64
    force = torch.zeros((N, 4), dtype=torch.float64, device="cuda")
    vir = torch.zeros((6, pitch), dtype=torch.float64, device="cuda")
    pf = torch.zeros((N, 4), dtype=torch.int32, device="cuda")
    box = _azp.Box(*Lv)
    _azp.call("azp_gpu_pack_fixed", _azp.dptr(sysm.pos), N, ctypes.byref(box), _azp.dptr(pf), _azp.stream_ptr())
    a = raw_args(sysm, ref.nl, rcsq, None, 1, force, vir, pitch, precision=1)
    a.particles.d_pos_fixed = pf.data_ptr()
    a.d_tile_count, a.d_tile_jset, a.d_tile_list, a.tile_size, a.tile_slots = ref.nl.get_tiles()
    a.particles.d_pos_fixed = None  # the tiled kernel converts positions itself
    _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(a), 2, _azp.dptr(d_params))
    torch.cuda.synchronize()
    assert np.array_equal(force.cpu().numpy(), ref.force)
    a.tile_size = 128
    with pytest.raises(_azp.AzpError):
        _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(a), 2, _azp.dptr(d_params))
    a.tile_size = 512
    a.tile_slots = 5000
    with pytest.raises(_azp.AzpError):
        _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(a), 2, _azp.dptr(d_params))
