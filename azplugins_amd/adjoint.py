"""Differentiable pair forces for force-matching fits (torch.autograd).

``PairForceField`` evaluates per-particle forces F (N, 3) and energies E (N,)
of ExpandedYukawa or PerturbedLennardJones as a differentiable function of the
positions and the per-type-pair parameters theta, on the gfx950 kernels:

* forward:  ``azp_gpu_compute_pair_forces`` (FP64 or mixed form),
* backward: ``azp_gpu_pair_adjoint``, a hand-written adjoint kernel giving
  dL/dr and dL/dtheta from dL/dF and dL/dE in one pass over the neighbor list.

theta is a tensor of shape (n_pairs, 3) with one row per unordered type pair
(a, b), a <= b, in the order ``pair_keys``; columns follow the reference's
parameter keys: (epsilon, kappa, delta) for ExpandedYukawa and (epsilon, sigma,
attraction_scale_factor) for PerturbedLennardJones.

The forward runs in ``precision`` ("fp64" or "mixed"), the backward pass (the
adjoint kernel) in ``adjoint_precision``, FP64 by default.

Example::

    ff = PairForceField("ExpandedYukawa", ntypes=2, box=L, r_cut=4.0)
    F, E = ff(pos, typeid, theta)        # pos (N,3) float64 cuda, requires_grad
    loss = ((F - F_target) ** 2).mean()
    loss.backward()                      # pos.grad, theta.grad
"""

import ctypes

import numpy as np

from . import _azp, _azplugins

torch = _azp.torch

_EVALUATORS = {
    "ExpandedYukawa": (_azp.EVAL_EXPANDED_YUKAWA, "PotentialPairExpandedYukawa"),
    "PerturbedLennardJones": (_azp.EVAL_PERTURBED_LJ, "PotentialPairPerturbedLennardJones"),
}
NPAR = 3


class PairForceField:
    """Differentiable forces/energies of one isotropic pair potential."""

    def __init__(self, potential, ntypes, box, r_cut, r_buff=0.4, mode="none", precision="fp64",
                 adjoint_precision="fp64"):
        if potential not in _EVALUATORS:
            raise ValueError(f"potential must be one of {sorted(_EVALUATORS)}")
        if mode not in ("none", "shift"):
            raise ValueError("mode must be 'none' or 'shift'")
        self.potential = potential
        self.evaluator, cls = _EVALUATORS[potential]
        self.ntypes = int(ntypes)
        self.box = _azp.Box(*[float(x) for x in np.broadcast_to(np.asarray(box, dtype=float), (3,))])
        if precision not in ("fp64", "mixed", "mixed_acc64") or adjoint_precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64', 'mixed' or 'mixed_acc64', adjoint_precision 'fp64' "
                             "or 'mixed'")
        self.mode = mode
        self.precision = precision
        # The backward pass evaluates the per-pair partials in FP64 by default.
        # The forward's own FP32 error still reaches the gradients of a force-
        # matching loss, scaled by the inverse relative force residual: at the
        # 262,144-particle workload with targets from 5 % parameter changes the
        # mixed forward leaves the parameter gradients ~1.7e-6 from FP64 finite
        # differences, the FP64 forward ~2e-10.
        self.adjoint_precision = adjoint_precision
        self.pair_keys = [(a, b) for a in range(self.ntypes) for b in range(a, self.ntypes)]
        rc = np.zeros(self.ntypes * self.ntypes)
        for a in range(self.ntypes):
            for b in range(self.ntypes):
                key = (min(a, b), max(a, b))
                rc[a * self.ntypes + b] = r_cut[key] if isinstance(r_cut, dict) else float(r_cut)
        self._rcut = rc
        dev = torch.device("cuda")
        self._rcutsq = torch.tensor(rc * rc, dtype=torch.float64, device=dev)
        self._nl = _azplugins.NeighborListGPU(r_buff)
        self._nl.set_r_cut(self.ntypes, rc.tolist())
        self._psize = _azp.lib().azp_param_size(self.evaluator)
        self._pos4 = None

    # ---- parameter packing (the reference's param_type constructor) ----
    def _pack(self, theta):
        vals = theta.detach().to("cpu", torch.float64).numpy()
        if vals.shape != (len(self.pair_keys), NPAR):
            raise ValueError(f"theta must have shape ({len(self.pair_keys)}, {NPAR})")
        nt = self.ntypes
        buf = np.zeros(nt * nt * self._psize, dtype=np.uint8)
        pod = (ctypes.c_char * self._psize)()
        for k, (a, b) in enumerate(self.pair_keys):
            _azp.call("azp_param_pack", self.evaluator, _azp.dbl_array(vals[k].tolist()), pod)
            raw = np.frombuffer(bytes(pod), dtype=np.uint8)
            for t in {a * nt + b, b * nt + a}:
                buf[t * self._psize:(t + 1) * self._psize] = raw
        return torch.tensor(buf, device="cuda")

    def _args(self, pos4, force, precision):
        N = pos4.shape[0]
        a = _azp.PairArgs()
        a.particles = _azp.Particles(N, self.box, pos4.data_ptr(), None, None, None, None)
        nn, lst, pitch, mx = self._nl.get()
        a.out.d_force = force.data_ptr() if force is not None else None
        a.out.d_virial = None
        a.out.virial_pitch = 0
        a.out.d_torque = None
        a.d_n_neigh = nn
        a.d_nlist = lst
        a.nlist_pitch = pitch
        a.max_neighbors = mx
        a.d_rcutsq = self._rcutsq.data_ptr()
        a.d_ronsq = None
        a.ntypes = self.ntypes
        a.shift_mode = _azp.SHIFT_MODES[self.mode]
        a.compute_virial = 0
        a.precision = _azp.PRECISIONS[precision]
        a.block_size = 0
        a.stream = _azp.stream_ptr()
        return a

    def __call__(self, pos, typeid, theta):
        return _PairForceFunction.apply(pos, theta, typeid, self)


class _PairForceFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, theta, typeid, ff):
        pos = pos.detach()
        N = pos.shape[0]
        pos4 = torch.empty((N, 4), dtype=torch.float64, device=pos.device)
        pos4[:, :3] = pos
        pos4[:, 3] = typeid.to(torch.float64)
        pos4 = pos4.contiguous()
        parts = _azp.Particles(N, ff.box, pos4.data_ptr(), None, None, None, None)
        ff._nl.update(parts)
        d_params = ff._pack(theta)
        force = torch.zeros((N, 4), dtype=torch.float64, device=pos.device)
        args = ff._args(pos4, force, ff.precision)
        if ff.precision != "fp64":
            pos_fixed = torch.empty((N, 4), dtype=torch.int32, device=pos.device)
            _azp.call("azp_gpu_pack_fixed", pos4.data_ptr(), N, ctypes.byref(ff.box), pos_fixed.data_ptr(),
                      _azp.stream_ptr())
            args.particles.d_pos_fixed = pos_fixed.data_ptr()
            ctx._keep = pos_fixed
        _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(args), ff.evaluator, _azp.dptr(d_params))
        ctx.save_for_backward(pos4, d_params)
        ctx.ff = ff
        ctx.theta_shape = theta.shape
        ctx.theta_dtype = theta.dtype
        ctx.theta_device = theta.device
        return force[:, :3].clone(), force[:, 3].clone()

    @staticmethod
    def backward(ctx, gF, gE):
        pos4, d_params = ctx.saved_tensors
        ff = ctx.ff
        N = pos4.shape[0]
        dev = pos4.device
        grad = torch.zeros((N, 4), dtype=torch.float64, device=dev)
        if gF is not None:
            grad[:, :3] = gF
        if gE is not None:
            grad[:, 3] = gE
        grad = grad.contiguous()
        gpos = torch.zeros((N, 4), dtype=torch.float64, device=dev)
        nt = ff.ntypes
        gpar = torch.zeros(nt * nt * NPAR, dtype=torch.float64, device=dev)
        wlen = _azp.lib().azp_adjoint_workspace_size(N, nt)
        work = torch.zeros(wlen, dtype=torch.float64, device=dev)
        args = ff._args(pos4, None, ff.adjoint_precision)
        _azp.call("azp_gpu_pair_adjoint", ctypes.byref(args), ff.evaluator, _azp.dptr(d_params),
                  _azp.dptr(grad), _azp.dptr(gpos), _azp.dptr(gpar), _azp.dptr(work), wlen)
        g = gpar.view(nt, nt, NPAR)
        rows = []
        for a, b in ff.pair_keys:
            rows.append(g[a, b] + g[b, a] if a != b else g[a, a])
        gtheta = torch.stack(rows).to(device=ctx.theta_device, dtype=ctx.theta_dtype)
        return gpos[:, :3].clone(), gtheta, None, None
