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
from gpu_util import GpuSystem, gpu_pair, jittered_lattice, nlist_pairs, rel_err, say

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
    assert tiled.nl.get_tiles()[3] == 0 and tiled.nl.walk_kind() == 0
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
    a.d_tile_order, a.d_tile_nn = ref.nl.get_tile_order()
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


# ---------------------------------------------------------------------------
# Fused per-tile walk (k_tile_walk.hip): with tiles enabled the build stages
# each tile's surrounding cells in LDS and writes the tiled list directly.
# The plain list expanded from it must equal the plain walk's list entry for
# entry (same order), and the pair set must be the oracle's.
# ---------------------------------------------------------------------------

def _morton(pos4, L, cell=3.4):
    from azplugins_amd import configs

    return pos4[configs.morton_order(pos4, np.array([L, L, L]), cell)].copy()


def _lists(sysm, nt, rc_matrix, tiled, r_buff=0.4):
    from azplugins_amd import _azplugins

    nl = _azplugins.NeighborListGPU(r_buff)
    nl.set_r_cut(nt, rc_matrix)
    if tiled:
        nl.set_tiled(tiled)
    nl.update(sysm.particles())
    kind = nl.walk_kind()
    nn, lst = nl.download(sysm.N)
    return nl, kind, nn, lst


def _same_lists(nn_a, lst_a, nn_b, lst_b):
    assert np.array_equal(nn_a, nn_b)
    for i in range(nn_a.shape[0]):
        k = int(nn_a[i])
        if not np.array_equal(lst_a[:k, i], lst_b[:k, i]):
            return False
    return True


@pytest.mark.parametrize("n,L,mode", [(24, 25.3, 1), (32, 33.7, 1), (24, 25.3, 256), (18, 18.9, 512),
                                      (7, 7.3, 1)])
def test_fused_walk_equals_plain_walk(n, L, mode):
    """Morton-ordered lattices (the order a space-filling-curve sort leaves):
    N = 343 (two cells per axis: repeated stencil cells) .. 32768."""
    pos4, _ = jittered_lattice(n, L, 0.1, 1, seed=n + 5)
    pos4 = _morton(pos4, L)
    sysm = GpuSystem(pos4, np.array([L, L, L]))
    rc = [3.0]
    nl_t, kind, nn_t, lst_t = _lists(sysm, 1, rc, mode)
    _, kind_p, nn_p, lst_p = _lists(sysm, 1, rc, 0)
    size, slots = nl_t.get_tiles()[3:]
    say(f"[fused] N={sysm.N} L={L} mode {mode}: walk kind {kind}, tile size {size}, slots {slots}")
    assert kind == 1 and kind_p == 0
    assert size == (512 if mode in (1, 512) else 256)
    assert _same_lists(nn_t, lst_t, nn_p, lst_p)
    pairs = oracle.nlist(pos4, np.array([L, L, L]), 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
    got = nlist_pairs(nl_t, sysm.N)
    assert got == set(zip(pairs[0].tolist(), pairs[1].tolist()))


def test_fused_walk_two_types_excluded_pair():
    """Two types with the (1, 1) pair excluded (r_cut 0) and unequal cutoffs."""
    L = 22.0
    pos4, _ = jittered_lattice(22, L, 0.15, 2, seed=77)
    pos4 = _morton(pos4, L)
    sysm = GpuSystem(pos4, np.array([L, L, L]))
    rc = [2.5, 3.0, 3.0, 0.0]
    _, kind, nn_t, lst_t = _lists(sysm, 2, rc, 1)
    _, _, nn_p, lst_p = _lists(sysm, 2, rc, 0)
    say(f"[fused] two types, excluded (1,1): walk kind {kind}")
    assert kind == 1
    assert _same_lists(nn_t, lst_t, nn_p, lst_p)
    types = pos4[:, 3].astype(int)
    for i in range(0, sysm.N, 97):
        js = lst_t[: nn_t[i], i]
        assert not (types[i] == 1 and np.any(types[js] == 1))


def test_fused_walk_any_particle_order():
    """Tiles are runs of cell-list slots (cells in Morton order), so a random
    particle order still gives compact tiles: the per-tile walk runs and the
    lists and forces equal the plain ones."""
    L = 25.3
    pos4, _ = jittered_lattice(24, L, 0.1, 1, seed=9)
    pos4 = pos4[np.random.default_rng(1).permutation(pos4.shape[0])].copy()
    sysm = GpuSystem(pos4, np.array([L, L, L]))
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
    kind = tiled.nl.walk_kind()
    size = tiled.nl.get_tiles()[3]
    say(f"[fused] random order: walk kind {kind}, tile size {size}")
    assert kind == 1 and size == 512
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)
    nn_t, lst_t = tiled.nl.download(sysm.N)
    nn_p, lst_p = plain.nl.download(sysm.N)
    assert _same_lists(nn_t, lst_t, nn_p, lst_p)


def test_fused_walk_forces_and_rebuild():
    """Mixed forces over the fused tiled list equal those over the plain list;
    a second build after moving particles is fused again and still equal."""
    L = 25.3
    pos4, _ = jittered_lattice(24, L, 0.1, 1, seed=21)
    pos4 = _morton(pos4, L)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
    assert tiled.nl.walk_kind() == 1
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)
    # move every particle by up to 0.3 (more than r_buff / 2): the list rebuilds
    rng = np.random.default_rng(5)
    moved = pos4.copy()
    moved[:, :3] += rng.uniform(-0.3, 0.3, (pos4.shape[0], 3))
    moved[:, :3] -= L * np.round(moved[:, :3] / L)
    sysm.pos.copy_(torch.tensor(moved, device=sysm.pos.device))
    parts = sysm.particles()
    tiled.nl.update(parts)
    assert tiled.nl.walk_kind() == 1 and tiled.nl.num_builds() == 2
    pairs = oracle.nlist(moved, Lv, 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
    assert nlist_pairs(tiled.nl, sysm.N) == set(zip(pairs[0].tolist(), pairs[1].tolist()))


def test_fused_walk_relative_offsets_large_box():
    """N = 262,144 at density 0.85 (19 cells per axis): most tiles lie within
    L/4 and 8 r_list of their anchor and are walked in FP32 offsets.  Their
    lists must still equal the plain walk's entry for entry, and the mixed
    forces over them the plain-list ones."""
    n = 64
    L = (n**3 / 0.85) ** (1.0 / 3.0)
    pos4, _ = jittered_lattice(n, L, 0.1, 1, seed=64)
    sysm = GpuSystem(pos4, np.array([L, L, L]))
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
    ntiles, rel = tiled.nl.tile_stats()
    say(f"[fused] N={sysm.N} L={L:.2f}: {rel} of {ntiles} tiles in FP32 offsets")
    assert tiled.nl.walk_kind() == 1 and rel > ntiles // 4
    nn_t, lst_t = tiled.nl.download(sysm.N)
    nn_p, lst_p = plain.nl.download(sysm.N)
    assert np.array_equal(nn_t, nn_p)
    mask = np.arange(lst_t.shape[0])[:, None] < nn_t[None, :]
    assert np.array_equal(lst_t[:, : sysm.N][mask], lst_p[:, : sysm.N][mask])
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)


@pytest.mark.parametrize("Lv", [(26.0, 22.0, 19.0), (70.0, 64.0, 58.0), (26.0, 22.0, 6.8)])
def test_fused_walk_non_cubic_box(Lv):
    """Orthorhombic boxes: the per-tile walk in fixed point with per-axis
    scales (small box) and in FP32 offsets (large box); lists equal the plain
    walk's, the pair set equals the oracle's (small box), and the mixed forces
    over the tiled list equal the plain-list ones and the FP64 oracle to 1e-5.
    The third box is twice r_list in z, the smallest box allowed: one cell
    thick, where a particle's own cell is the only stencil entry on that
    axis."""
    Lv = np.array(Lv)
    rng = np.random.default_rng(int(Lv[0]))
    n = [int(round(l / 1.056)) for l in Lv]
    g = [(np.arange(k) + 0.5) * (l / k) - l / 2 for k, l in zip(n, Lv)]
    x, y, z = np.meshgrid(*g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) + rng.uniform(-0.1, 0.1, (x.size, 3))
    p -= Lv * np.floor(p / Lv + 0.5)
    pos4 = np.zeros((p.shape[0], 4))
    pos4[:, :3] = p
    sysm = GpuSystem(pos4, Lv)
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
    ntiles, rel = tiled.nl.tile_stats()
    say(f"[fused] box {Lv.tolist()} N={sysm.N}: walk kind {tiled.nl.walk_kind()}, {rel} of {ntiles} tiles in FP32 offsets")
    assert tiled.nl.walk_kind() == 1
    nn_t, lst_t = tiled.nl.download(sysm.N)
    nn_p, lst_p = plain.nl.download(sysm.N)
    assert np.array_equal(nn_t, nn_p)
    mask = np.arange(lst_t.shape[0])[:, None] < nn_t[None, :]
    assert np.array_equal(lst_t[:, : sysm.N][mask], lst_p[:, : sysm.N][mask])
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)
    if sysm.N < 20000:
        pairs = oracle.nlist(pos4, Lv, 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
        assert nlist_pairs(tiled.nl, sysm.N) == set(zip(pairs[0].tolist(), pairs[1].tolist()))
        fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, Lv, pairs, 1,
                                    oracle.pack_matrix(oracle.PERTURBED_LJ, PLJ, 1), np.array([9.0]))
        assert rel_err(tiled.force[:, :3], fo[:, :3]) < 1e-5
        assert rel_err(tiled.virial, vo) < 1e-5
    else:
        assert rel > 0


def test_fused_walk_fallbacks():
    """A dilute system (density 0.02: a tile spans more than 128 cells) and a
    nine-type system (above the walk's type table) take the plain walk; the
    tiled list, when built, and the forces still equal the plain ones."""
    rng = np.random.default_rng(11)
    L = 60.0
    pos4 = np.zeros((4320, 4))
    pos4[:, :3] = rng.uniform(-L / 2, L / 2, (pos4.shape[0], 3))
    sysm = GpuSystem(pos4, np.array([L, L, L]))
    plain, tiled = _both("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
    say(f"[fused] dilute: walk kind {tiled.nl.walk_kind()}, tile size {tiled.nl.get_tiles()[3]}")
    assert tiled.nl.walk_kind() == 0
    assert np.array_equal(plain.force, tiled.force)

    L = 16.0
    pos4, _ = jittered_lattice(16, L, 0.1, 9, seed=19)
    sysm = GpuSystem(pos4, np.array([L, L, L]))
    hz = {(a, b): {"epsilon": 1.0 + 0.1 * (a + b)} for a in range(9) for b in range(a, 9)}
    plain, tiled = _both("PotentialPairHertz", sysm, 9, hz, 1.3)
    say(f"[fused] nine types: walk kind {tiled.nl.walk_kind()}, tile size {tiled.nl.get_tiles()[3]}")
    assert tiled.nl.walk_kind() == 0
    assert np.array_equal(plain.force, tiled.force)


def _fp64_pair(cls, sysm, nt, params, rc, **kw):
    plain = gpu_pair(cls, sysm, nt, params, rc, **kw)
    tiled = gpu_pair(cls, sysm, nt, params, rc, tiled=True, **kw)
    return plain, tiled


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
def test_fp64_tiled_equals_plain_and_oracle(mode):
    """The FP64 form over the tiled list reads FP64 positions from LDS (32-byte
    slots with two types): the same terms in the same order as over the plain
    list, so the results are bit-identical, and within 1e-12 of the oracle."""
    L = 16.0
    pos4, _ = jittered_lattice(16, L, 0.15, 2, seed=43)
    Lv = np.array([L, L * 1.02, L])
    pos4[:, :3] -= Lv * np.floor(pos4[:, :3] / Lv + 0.5)
    sysm = GpuSystem(pos4, Lv)
    yk = {(0, 0): {"epsilon": 1.0, "kappa": 2.0, "delta": 0.2}, (0, 1): {"epsilon": 0.7, "kappa": 1.5, "delta": 0.3},
          (1, 1): {"epsilon": 1.3, "kappa": 2.5, "delta": 0.1}}
    plj = {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5},
           (0, 1): {"epsilon": 0.8, "sigma": 0.9, "attraction_scale_factor": 0.3},
           (1, 1): {"epsilon": 1.2, "sigma": 0.95, "attraction_scale_factor": 0.7}}
    hz = {(0, 0): {"epsilon": 1.0}, (0, 1): {"epsilon": 2.0}, (1, 1): {"epsilon": 0.5}}
    cases = (("PotentialPairExpandedYukawa", oracle.EXPANDED_YUKAWA, yk, 3.0),
             ("PotentialPairPerturbedLennardJones", oracle.PERTURBED_LJ, plj, 2.5),
             ("PotentialPairHertz", oracle.HERTZ, hz, 1.3))
    for cls, ev, params, rc in cases:
        if cls == "PotentialPairHertz" and mode != "none":
            continue
        r_on = 0.7 * rc if mode == "xplor" else None
        plain, tiled = _fp64_pair(cls, sysm, 2, params, rc, mode=mode, r_on=r_on)
        size = tiled.nl.get_tiles()[3]
        say(f"[tiles fp64] {cls} {mode}: tile size {size}")
        assert size in (256, 512)
        assert np.array_equal(plain.force, tiled.force), cls
        assert np.array_equal(plain.virial, tiled.virial), cls
        rcm = np.full(4, rc)
        pairs = oracle.nlist(pos4, Lv, 2, oracle.r_listsq_matrix(rcm, 0.4), "cell")
        fo, vo = oracle.pair_forces(ev, pos4, Lv, pairs, 2, oracle.pack_matrix(ev, params, 2), rcm * rcm,
                                    (0.7 * rcm) ** 2, mode)
        err = max(rel_err(tiled.force[:, :3], fo[:, :3]), rel_err(tiled.force[:, 3], fo[:, 3]),
                  rel_err(tiled.virial, vo))
        say(f"  fp64 tiled vs oracle {err:.2e}")
        assert err < 1e-12


@pytest.mark.parametrize("mode", ["none", "shift", "xplor"])
def test_fp64_tiled_one_type(mode):
    """With one particle type the FP64 form stages only x, y, z (24-byte slots,
    the entry's byte offset times 3/2): bit-identical to the plain list and
    within 1e-12 of the oracle, for tiles of both sizes."""
    L = 18.0
    pos4, _ = jittered_lattice(18, L, 0.15, 1, seed=44)
    Lv = np.array([L, L * 1.01, L * 0.99])
    pos4[:, :3] -= Lv * np.floor(pos4[:, :3] / Lv + 0.5)
    sysm = GpuSystem(pos4, Lv)
    params = {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5}}
    rc = 2.5
    r_on = 0.7 * rc if mode == "xplor" else None
    pairs = oracle.nlist(pos4, Lv, 1, oracle.r_listsq_matrix([rc], 0.4), "cell")
    fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, Lv, pairs, 1,
                                oracle.pack_matrix(oracle.PERTURBED_LJ, params, 1), np.array([rc * rc]),
                                np.array([(0.7 * rc) ** 2]), mode)
    plain = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, params, rc, mode=mode, r_on=r_on)
    for size in (256, 512):
        tiled = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, params, rc, mode=mode, r_on=r_on,
                         tiled=size)
        say(f"[tiles fp64 one type] {mode}: tile size {tiled.nl.get_tiles()[3]}")
        assert tiled.nl.get_tiles()[3] == size
        assert np.array_equal(plain.force, tiled.force)
        assert np.array_equal(plain.virial, tiled.virial)
        err = max(rel_err(tiled.force[:, :3], fo[:, :3]), rel_err(tiled.force[:, 3], fo[:, 3]),
                  rel_err(tiled.virial, vo))
        assert err < 1e-12


def test_fp64_tiled_colloid():
    """Colloid (FP64 only) over the tiled list equals the plain list."""
    from test_gpu_parity import ISO_PARAMS, ISO_RCUT, _colloid_positions

    L = 26.0
    pos4 = _colloid_positions(12, L, seed=13)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    plain, tiled = _fp64_pair("PotentialPairColloid", sysm, 2, ISO_PARAMS[3], ISO_RCUT[3])
    say(f"[tiles fp64] colloid tile size {tiled.nl.get_tiles()[3]}")
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)


@pytest.mark.parametrize("ntypes", [1, 2])
@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, 0.5)])
def test_fp64_tiled_images_per_axis(shift, ntypes):
    """The FP64 form over the tiled list takes the periodic image only on the
    axes where the tile's positions span 0.45 L or more.  A 12^3 lattice blob
    in a box of 30 spans 12.3 < 13.5 on every axis at the centre (no image
    anywhere), and straddles the x face (shift L/2 on x) or every face (shift
    L/2 on all axes): in each case the forces and virials equal the plain
    list's bit for bit, and the oracle's to 1e-12."""
    L = 30.0
    Lv = np.array([L, L, L])
    pos4, rng = jittered_lattice(12, 12.0, 0.15, ntypes, seed=45)
    pos4[:, :3] += Lv * np.array(shift)
    pos4[:, :3] -= Lv * np.floor(pos4[:, :3] / Lv + 0.5)
    sysm = GpuSystem(pos4, Lv)
    plj = {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5},
           (0, 1): {"epsilon": 0.8, "sigma": 0.9, "attraction_scale_factor": 0.3},
           (1, 1): {"epsilon": 1.2, "sigma": 0.95, "attraction_scale_factor": 0.7}}
    params = {k: v for k, v in plj.items() if max(k) < ntypes}
    rc = 2.5
    plain, tiled = _fp64_pair("PotentialPairPerturbedLennardJones", sysm, ntypes, params, rc, mode="shift")
    say(f"[tiles fp64 images] shift {shift} ntypes {ntypes}: tile size {tiled.nl.get_tiles()[3]}")
    assert tiled.nl.get_tiles()[3] in (256, 512)
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)
    rcm = np.full(ntypes * ntypes, rc)
    pairs = oracle.nlist(pos4, Lv, ntypes, oracle.r_listsq_matrix(rcm, 0.4), "cell")
    fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, Lv, pairs, ntypes,
                                oracle.pack_matrix(oracle.PERTURBED_LJ, params, ntypes), rcm * rcm,
                                (0.7 * rcm) ** 2, "shift")
    err = max(rel_err(tiled.force[:, :3], fo[:, :3]), rel_err(tiled.force[:, 3], fo[:, 3]),
              rel_err(tiled.virial, vo))
    say(f"  fp64 tiled vs oracle {err:.2e}")
    assert err < 1e-12


def test_mixed_acc64_tiled_equals_plain():
    """FP32 pair arithmetic with every entry's terms added to the FP64
    accumulators: tiled and plain lists give the same bits; the force error
    against the oracle is within the mixed bound and not larger than the
    FP32-partial-sum form's."""
    L = 20.0
    pos4, _ = jittered_lattice(20, L, 0.1, 1, seed=21)
    Lv = np.array([L, L, L])
    sysm = GpuSystem(pos4, Lv)
    plain = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0, precision="mixed_acc64")
    tiled = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0, precision="mixed_acc64", tiled=True)
    part = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0, precision="mixed", tiled=True)
    assert tiled.nl.get_tiles()[3] in (256, 512)
    assert np.array_equal(plain.force, tiled.force)
    assert np.array_equal(plain.virial, tiled.virial)
    pairs = oracle.nlist(pos4, Lv, 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
    fo, vo = oracle.pair_forces(oracle.PERTURBED_LJ, pos4, Lv, pairs, 1,
                                oracle.pack_matrix(oracle.PERTURBED_LJ, PLJ, 1), np.array([9.0]))
    e64 = rel_err(tiled.force[:, :3], fo[:, :3])
    e32 = rel_err(part.force[:, :3], fo[:, :3])
    say(f"[tiles] mixed_acc64 force rel err {e64:.2e}, mixed (32-entry FP32 sums) {e32:.2e}")
    assert e64 < 1e-5 and rel_err(tiled.virial, vo) < 1e-5
    assert e64 <= 1.5 * e32
