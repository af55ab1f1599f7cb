"""Batched TwoPatchMorse replicas on one GPU (``azp_batch_*`` in the C ABI).

Many independent periodic systems of uneven size, each with its own box and
Langevin temperature, share one set of device arrays.  Forces and torques of
all replicas come from one persistent kernel launch whose launch shape is
chosen by :meth:`ReplicaBatch.autotune`; time steps are replayed from captured
HIP graphs (:meth:`ReplicaBatch.run`).  The pair term is the one of
``AnisoPotentialPair<AnisoPairEvaluatorTwoPatchMorse>``
(AnisoPairEvaluatorTwoPatchMorse.h:127-216); parameters use the keys of
``hoomd.azplugins.pair.TwoPatchMorse`` (pair.py:430-510).
"""

import ctypes

import numpy as np

from . import _azp

torch = _azp.torch

TPM_KEYS = ("M_d", "M_r", "r_eq", "omega", "alpha", "repulsion")


class ReplicaBatch:
    """R replicas of sizes ``sizes`` in boxes ``boxes`` ((R, 3) lengths).

    Particles of replica r occupy rows ``offsets[r]:offsets[r + 1]`` of every
    array; positions are in the replica's own box frame ([-L/2, L/2)).
    """

    def __init__(self, sizes, boxes, ntypes=1, r_cut=2.5, r_buff=0.4, precision="mixed", mode="none",
                 compute_virial=True, device="cuda"):
        sizes = np.asarray(sizes, dtype=np.uint32)
        boxes = np.asarray(boxes, dtype=np.float64).reshape(len(sizes), 3)
        self.R = int(len(sizes))
        self.sizes = sizes
        self.boxes = boxes
        self.offsets = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))])
        self.N = int(self.offsets[-1])
        self.ntypes = int(ntypes)
        self._h = ctypes.c_void_p()
        c_sizes = (ctypes.c_uint32 * self.R)(*[int(v) for v in sizes])
        c_boxes = (_azp.Box * self.R)(*[_azp.Box(*map(float, b)) for b in boxes])
        _azp.call("azp_batch_create", ctypes.byref(self._h), self.R, c_sizes, c_boxes, self.ntypes,
                  _azp.EVAL_TWO_PATCH_MORSE)
        for a in range(self.ntypes):
            for b in range(a, self.ntypes):
                _azp.call("azp_batch_set_r_cut", self._h, a, b, float(r_cut))
        _azp.call("azp_batch_set_r_buff", self._h, float(r_buff))
        self.precision = precision
        self.mode = mode
        self.compute_virial = bool(compute_virial)
        self._set_mode()
        self.device = torch.device(device)
        self._alloc()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _azp.lib().azp_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---- configuration ----
    def _set_mode(self):
        _azp.call("azp_batch_set_mode", self._h, _azp.SHIFT_MODES[self.mode], _azp.PRECISIONS[self.precision],
                  1 if self.compute_virial else 0)

    def set_precision(self, precision):
        self.precision = precision
        self._set_mode()

    def set_params(self, ti, tj, **p):
        vals = [float(p[k]) for k in TPM_KEYS]
        _azp.call("azp_batch_set_params", self._h, int(ti), int(tj), _azp.dbl_array(vals))

    def set_langevin(self, kT, dt=0.005, gamma=1.0, gamma_r=1.0, inertia=1.0, seed=0):
        kT = np.asarray(kT, dtype=np.float64).reshape(self.R)
        _azp.call("azp_batch_set_langevin", self._h, _azp.dbl_array(kT), float(dt), float(gamma), float(gamma_r),
                  float(inertia), int(seed) & 0xFFFF)
        self.inertia = float(inertia)

    def set_timestep(self, t):
        _azp.call("azp_batch_set_timestep", self._h, int(t), _azp.stream_ptr())

    # ---- device arrays ----
    def _alloc(self):
        N, dev = self.N, self.device
        z4 = lambda dt=torch.float64: torch.zeros((max(N, 1), 4), dtype=dt, device=dev)  # noqa: E731
        self.pos = z4()
        self.pos_fixed = z4(torch.int32)
        self.orientation = z4()
        self.orientation[:, 0] = 1.0
        self.vel = z4()
        self.vel[:, 3] = 1.0  # mass
        self.angmom = z4()
        self.accel = z4()
        self.force = z4()
        self.torque = z4()
        self.vpitch = (N + 63) // 64 * 64
        self.virial = torch.zeros((6, self.vpitch), dtype=torch.float64, device=dev)
        self.bind()

    def bind(self):
        a = _azp.BatchArrays()
        a.d_pos = self.pos.data_ptr()
        a.d_pos_fixed = self.pos_fixed.data_ptr()
        a.d_orientation = self.orientation.data_ptr()
        a.d_vel = self.vel.data_ptr()
        a.d_angmom = self.angmom.data_ptr()
        a.d_accel = self.accel.data_ptr()
        a.d_force = self.force.data_ptr()
        a.d_torque = self.torque.data_ptr()
        a.d_virial = self.virial.data_ptr()
        a.virial_pitch = self.vpitch
        _azp.call("azp_batch_bind", self._h, ctypes.byref(a))

    def load(self, pos4, quat=None, vel=None, angmom=None):
        """Copy concatenated host arrays ((N, 4) positions+type, (N, 4) quaternions, ...)."""
        self.pos.copy_(torch.as_tensor(np.asarray(pos4, dtype=np.float64)))
        if quat is not None:
            self.orientation.copy_(torch.as_tensor(np.asarray(quat, dtype=np.float64)))
        if vel is not None:
            self.vel[:, :3].copy_(torch.as_tensor(np.asarray(vel, dtype=np.float64)[:, :3]))
        if angmom is not None:
            self.angmom[:, :3].copy_(torch.as_tensor(np.asarray(angmom, dtype=np.float64)[:, :3]))
        self.bind()

    def replica(self, r):
        return slice(int(self.offsets[r]), int(self.offsets[r + 1]))

    # ---- work ----
    def compute(self, force_rebuild=False):
        _azp.call("azp_batch_compute", self._h, 1 if force_rebuild else 0, _azp.stream_ptr())

    def run(self, nsteps, use_graph=True):
        _azp.call("azp_batch_run", self._h, int(nsteps), 1 if use_graph else 0, _azp.stream_ptr())

    def thermo(self):
        """(R, 4): translational KE, rotational KE, potential energy, virial trace."""
        out = torch.zeros((self.R, 4), dtype=torch.float64, device=self.device)
        _azp.call("azp_batch_thermo", self._h, ctypes.c_void_p(out.data_ptr()), _azp.stream_ptr())
        return out.cpu().numpy()

    def autotune(self):
        t = _azp.BatchTuning()
        _azp.call("azp_batch_autotune", self._h, ctypes.byref(t), _azp.stream_ptr())
        return t.as_dict()

    @property
    def tuning(self):
        t = _azp.BatchTuning()
        _azp.call("azp_batch_get_tuning", self._h, ctypes.byref(t))
        return t.as_dict()

    def set_tuning(self, block_size, grab, workgroups_per_cu, walk_block=256):
        t = _azp.BatchTuning(block_size, grab, workgroups_per_cu, walk_block, 0.0)
        _azp.call("azp_batch_set_tuning", self._h, ctypes.byref(t))

    def stats(self):
        s = _azp.BatchStats()
        _azp.call("azp_batch_stats", self._h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def nlist(self):
        """(n_neigh[N], list[max_neighbors, pitch]) on the host."""
        pitch = ctypes.c_uint32()
        mx = ctypes.c_uint32()
        nn = np.zeros(self.N, dtype=np.uint32)
        one = np.zeros(1, dtype=np.uint32)
        try:
            _azp.call("azp_batch_nlist_download", self._h, nn.ctypes.data_as(ctypes.c_void_p),
                      one.ctypes.data_as(ctypes.c_void_p), 0, ctypes.byref(pitch), ctypes.byref(mx))
        except _azp.AzpError:
            pass  # capacity query
        # pinned host memory: the list arrives in one DMA copy
        lst_t = torch.empty(max(int(mx.value) * int(pitch.value), 1), dtype=torch.int32, pin_memory=True)
        _azp.call("azp_batch_nlist_download", self._h, nn.ctypes.data_as(ctypes.c_void_p),
                  lst_t.data_ptr(), int(mx.value) * int(pitch.value), ctypes.byref(pitch), ctypes.byref(mx))
        lst = lst_t.numpy().view(np.uint32)[: int(mx.value) * int(pitch.value)]
        return nn, lst.reshape(int(mx.value), int(pitch.value))


def from_config(cfg, precision="mixed", device="cuda"):
    """ReplicaBatch loaded with a configs.tpm_batch()-style dict."""
    reps = cfg["replicas"]
    sizes = [r["pos4"].shape[0] for r in reps]
    boxes = [r["L"] for r in reps]
    b = ReplicaBatch(sizes, boxes, ntypes=cfg["ntypes"], r_cut=cfg["r_cut"], r_buff=cfg["r_buff"],
                     precision=precision, device=device)
    for (ti, tj), p in cfg["params"].items():
        b.set_params(ti, tj, **p)
    b.load(np.concatenate([r["pos4"] for r in reps]), np.concatenate([r["quat"] for r in reps]))
    b.set_langevin([r["kT"] for r in reps], dt=cfg.get("dt", 0.005))
    return b
