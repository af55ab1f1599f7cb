"""Seeded synthetic workloads of the pair-force path.

Each generator returns host arrays ``pos4`` (N, 4: x, y, z, type) with
positions in [-L/2, L/2), the box lengths ``L`` (3,) and whatever else the
workload needs.  Particles are ordered along a Morton (Z-order) curve of cells
of width ~r_list, the order a space-filling-curve sort (HOOMD's SFCPackTuner)
leaves them in, so that neighbors are close in memory.

Workloads (names used by bench.py and the tests):
  hertz_4096         Hertz, N=4096, 16^3 simple cubic in L=17.235, jitter +-0.1
  plj_1m             PerturbedLJ, N=1,048,576 FCC 64^3, density 0.85, jitter +-0.1
  dpd_2m             DPD general weight, N=2,000,000 uniform, density 3, A/B 50/50
  yukawa_262k        ExpandedYukawa, N=262,144 (64^3 SC), density 0.5, A/B 70/30
  tpm_batch_512      TwoPatchMorse, 512 replicas, N_r log-uniform in [500, 20000],
                     density 0.6, random unit quaternions
Smaller instances of the same recipes are produced with the ``n``/``N``
arguments.
"""

import numpy as np


def _wrap(x, L):
    return x - L * np.floor(x / L + 0.5)


def _part1by2(v):
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_order(pos, L, cell):
    """Permutation sorting positions by the Morton code of their cell."""
    n = np.maximum(1, np.floor(L / cell)).astype(np.int64)
    c = np.floor((pos[:, :3] + 0.5 * L) / L * n).astype(np.int64)
    c = np.clip(c, 0, n - 1)
    key = _part1by2(c[:, 0]) | (_part1by2(c[:, 1]) << np.uint64(1)) | (_part1by2(c[:, 2]) << np.uint64(2))
    return np.argsort(key, kind="stable")


def sc_lattice(n, L, jitter, rng):
    a = L / n
    g = (np.arange(n) + 0.5) * a - 0.5 * L
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    p += rng.uniform(-jitter, jitter, size=p.shape)
    return _wrap(p, L)


def fcc_lattice(n, L, jitter, rng):
    a = L / n
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    g = np.arange(n) * a
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    p = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) - 0.5 * L + 0.25 * a
    p += rng.uniform(-jitter, jitter, size=p.shape)
    return _wrap(p, L)


def _finish(p, types, L, r_sort):
    pos4 = np.zeros((p.shape[0], 4))
    pos4[:, :3] = p
    pos4[:, 3] = types
    order = morton_order(pos4, L, r_sort)
    return pos4[order].copy()


def hertz_4096(n=16, seed=1):
    rng = np.random.default_rng(seed)
    a = 17.235 / 16
    L = n * a
    p = sc_lattice(n, L, 0.1, rng)
    Lv = np.array([L, L, L])
    pos4 = _finish(p, np.zeros(p.shape[0]), Lv, 1.4)
    return dict(name="hertz_4096", pos4=pos4, L=Lv, ntypes=1, r_cut=1.0, r_buff=0.4,
                params={(0, 0): {"epsilon": 1.0}}, dt=0.005, kT=1.0)


def plj_1m(n=64, density=0.85, seed=2):
    rng = np.random.default_rng(seed)
    N = 4 * n**3
    L = (N / density) ** (1.0 / 3.0)
    p = fcc_lattice(n, L, 0.1, rng)
    Lv = np.array([L, L, L])
    pos4 = _finish(p, np.zeros(N), Lv, 3.4)
    return dict(name="plj_1m", pos4=pos4, L=Lv, ntypes=1, r_cut=3.0, r_buff=0.4,
                params={(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5}},
                dt=0.005, kT=1.2, gamma=1.0)


def dpd_2m(N=2_000_000, density=3.0, seed=3):
    rng = np.random.default_rng(seed)
    L = (N / density) ** (1.0 / 3.0)
    p = rng.uniform(-0.5 * L, 0.5 * L, size=(N, 3))
    types = rng.integers(0, 2, size=N)
    Lv = np.array([L, L, L])
    pos4 = _finish(p, types, Lv, 1.4)
    params = {
        (0, 0): {"A": 25.0, "gamma": 4.5, "s": 0.5},
        (1, 1): {"A": 25.0, "gamma": 4.5, "s": 0.5},
        (0, 1): {"A": 40.0, "gamma": 4.5, "s": 0.5},
    }
    return dict(name="dpd_2m", pos4=pos4, L=Lv, ntypes=2, r_cut=1.0, r_buff=0.4, params=params,
                dt=0.02, kT=1.0, seed=3)


def yukawa_262k(n=64, density=0.5, seed=4):
    rng = np.random.default_rng(seed)
    N = n**3
    L = (N / density) ** (1.0 / 3.0)
    p = sc_lattice(n, L, 0.1, rng)
    types = (rng.uniform(size=N) >= 0.7).astype(np.float64)
    Lv = np.array([L, L, L])
    pos4 = _finish(p, types, Lv, 4.4)
    params = {}
    for key in ((0, 0), (0, 1), (1, 1)):
        params[key] = {"epsilon": float(rng.uniform(0.5, 1.5)), "kappa": 2.0,
                       "delta": float(rng.uniform(0.0, 0.5))}
    return dict(name="yukawa_262k", pos4=pos4, L=Lv, ntypes=2, r_cut=4.0, r_buff=0.4, params=params)


def random_unit_quaternions(n, rng):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1)[:, None]


def tpm_replica(Nr, density, rng):
    """One TwoPatchMorse replica: jittered simple-cubic fill of a cube."""
    L = (Nr / density) ** (1.0 / 3.0)
    n = int(np.ceil(Nr ** (1.0 / 3.0)))
    p = sc_lattice(n, L, 0.05 * L / n, rng)[:Nr]
    Lv = np.array([L, L, L])
    pos4 = _finish(p, np.zeros(Nr), Lv, 2.9)
    return pos4, Lv


def tpm_batch(n_replicas=512, n_min=500, n_max=20000, density=0.6, seed=5):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=n_replicas)).astype(np.int64)
    reps = []
    for Nr in sizes:
        pos4, Lv = tpm_replica(int(Nr), density, rng)
        reps.append(dict(pos4=pos4, L=Lv, quat=random_unit_quaternions(int(Nr), rng),
                         kT=float(rng.uniform(0.2, 1.0))))
    params = {(0, 0): {"M_d": 1.0, "M_r": 0.2, "r_eq": 1.0, "omega": 20.0, "alpha": 0.4,
                       "repulsion": True}}
    return dict(name="tpm_batch_512", replicas=reps, ntypes=1, r_cut=2.5, r_buff=0.4,
                params=params, dt=0.005)
