"""Helpers for the GPU parity tests: run one force evaluation through the C ABI.

Every call goes through libazp.so (``azp_nlist_*``, ``azp_pair_*`` or the
``azp_gpu_compute_*`` launchers); the oracle is only used by the tests to
check the results.
"""

import ctypes
import sys

import numpy as np

from azplugins_amd import _azp, _azplugins

torch = _azp.torch

CLASS = {
    0: "PotentialPairHertz",
    1: "PotentialPairExpandedYukawa",
    2: "PotentialPairPerturbedLennardJones",
    3: "PotentialPairColloid",
    4: "PotentialPairDPDThermoGeneralWeight",
    5: "AnisoPotentialPairTwoPatchMorse",
}
CONSERVATIVE_DPD = "PotentialPairConservativeGeneralWeight"

KEYS = {
    0: ("epsilon",),
    1: ("epsilon", "kappa", "delta"),
    2: ("epsilon", "sigma", "attraction_scale_factor"),
    3: ("A", "a_1", "a_2", "sigma"),
    4: ("A", "gamma", "s"),
    5: ("M_d", "M_r", "r_eq", "omega", "alpha", "repulsion"),
}


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


class GpuSystem:
    def __init__(self, pos4, L, vel4=None, quat=None, tag=None):
        dev = torch.device("cuda")
        self.N = pos4.shape[0]
        self.L = np.asarray(L, dtype=np.float64)
        self.pos = torch.tensor(np.ascontiguousarray(pos4, dtype=np.float64), device=dev)
        v = vel4 if vel4 is not None else np.tile([0.0, 0.0, 0.0, 1.0], (self.N, 1))
        self.vel = torch.tensor(np.ascontiguousarray(v, dtype=np.float64), device=dev)
        q = quat if quat is not None else np.tile([1.0, 0.0, 0.0, 0.0], (self.N, 1))
        self.quat = torch.tensor(np.ascontiguousarray(q, dtype=np.float64), device=dev)
        t = tag if tag is not None else np.arange(self.N)
        self.tag = torch.tensor(np.asarray(t, dtype=np.int64).astype(np.int32), device=dev)

    def particles(self):
        p = _azp.Particles()
        p.N = self.N
        p.box = _azp.Box(*[float(x) for x in self.L])
        p.d_pos = self.pos.data_ptr()
        p.d_pos_fixed = None
        p.d_vel = self.vel.data_ptr()
        p.d_tag = self.tag.data_ptr()
        p.d_orientation = self.quat.data_ptr()
        return p


class Result:
    pass


def gpu_pair(cls_name, system, ntypes, params, r_cut, r_buff=0.4, mode="none", precision="fp64",
             r_on=None, thermostat=None, timestep=0, virial=True, evaluator=None, tiled=False):
    """One force evaluation through azp_nlist_update + azp_pair_compute."""
    pp = getattr(_azplugins, cls_name + "GPU")(ntypes)
    ev = pp.evaluator
    for (a, b), v in params.items():
        pp.setParams(a, b, [float(v[k]) for k in KEYS[ev]])
    for a in range(ntypes):
        for b in range(a, ntypes):
            rc = r_cut[(a, b)] if isinstance(r_cut, dict) else r_cut
            pp.setRCut(a, b, rc)
            if r_on is not None:
                ro = r_on[(a, b)] if isinstance(r_on, dict) else r_on
                pp.setROn(a, b, ro)
    pp.setShiftMode(mode)
    pp.setPrecision(precision)
    if thermostat is not None:
        pp.setThermostat(*thermostat)
    nl = _azplugins.NeighborListGPU(r_buff)
    nl.set_r_cut(ntypes, pp.getRCutMatrix())
    if tiled:
        nl.set_tiled(tiled)
    parts = system.particles()
    nl.update(parts)
    N = system.N
    pitch = max(64, (N + 63) // 64 * 64)
    dev = torch.device("cuda")
    force = torch.full((N, 4), np.nan, dtype=torch.float64, device=dev)
    vir = torch.full((6, pitch), np.nan, dtype=torch.float64, device=dev)
    torque = torch.full((N, 4), np.nan, dtype=torch.float64, device=dev)
    out = _azp.Outputs()
    out.d_force = force.data_ptr()
    out.d_virial = vir.data_ptr()
    out.virial_pitch = pitch
    out.d_torque = torque.data_ptr()
    pp.compute(nl.handle, parts, out, timestep, compute_virial=virial)
    torch.cuda.synchronize()
    r = Result()
    r.force = force.cpu().numpy()
    r.virial = vir[:, :N].cpu().numpy()
    r.torque = torque.cpu().numpy()
    r.nl = nl
    r.pp = pp
    return r


def nlist_pairs(nl, N):
    """Full GPU list -> set of unordered (i, j) pairs, plus consistency checks."""
    nn, lst = nl.download(N)
    assert lst.shape[0] >= (int(nn.max()) if N else 0)
    ii = np.repeat(np.arange(N), nn)
    jj = np.concatenate([lst[: nn[i], i] for i in range(N)]) if N else np.zeros(0, dtype=np.uint32)
    directed = set(zip(ii.tolist(), jj.tolist()))
    assert len(directed) == ii.shape[0], "duplicate entries in the GPU list"
    for i, j in directed:
        assert (j, i) in directed, "GPU full list is not symmetric"
    return {(min(i, j), max(i, j)) for i, j in directed}


def rel_err(a, b):
    """max |a - b| / max |b| (normwise, robust to cancellation in sums)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def jittered_lattice(n, L, jitter, ntypes=1, seed=0, frac=None):
    rng = np.random.default_rng(seed)
    a = L / n
    g = (np.arange(n) + 0.5) * a - L / 2
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    p += rng.uniform(-jitter, jitter, size=p.shape)
    p -= L * np.floor(p / L + 0.5)
    pos4 = np.zeros((p.shape[0], 4))
    pos4[:, :3] = p
    if ntypes > 1:
        pos4[:, 3] = rng.integers(0, ntypes, size=p.shape[0])
    return pos4, rng


def raw_args(system, nl, rcutsq_dev, ronsq_dev, ntypes, out_force, out_virial, pitch, mode=0,
             precision=0, virial=1, block=0):
    """azp_pair_args for the launcher entry points."""
    a = _azp.PairArgs()
    a.particles = system.particles()
    nn, lst, npitch, mx = nl.get()
    a.out.d_force = out_force.data_ptr()
    a.out.d_virial = out_virial.data_ptr()
    a.out.virial_pitch = pitch
    a.out.d_torque = None
    a.d_n_neigh = nn
    a.d_nlist = lst
    a.nlist_pitch = npitch
    a.max_neighbors = mx
    a.d_rcutsq = rcutsq_dev.data_ptr()
    a.d_ronsq = ronsq_dev.data_ptr() if ronsq_dev is not None else None
    a.ntypes = ntypes
    a.shift_mode = mode
    a.compute_virial = virial
    a.precision = precision
    a.block_size = block
    a.stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return a
