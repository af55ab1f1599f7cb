"""azp_gpu_compute_pair_forces_noalloc against azp_gpu_compute_pair_forces.

The allocation-free launcher packs the mixed form's fixed-point positions
into a file-scope device array instead of library-allocated memory.  Two
calls in a row (the second on a smaller system, so the array keeps the first
call's longer contents beyond N) must equal the allocating launcher bit for
bit, as must the FP64 form, which uses no scratch.  A system beyond the
array's 2^20 particles is rejected on the host before anything is launched.
"""

import ctypes

import numpy as np
import pytest

from azplugins_amd import _azp
from gpu_util import GpuSystem, gpu_pair, jittered_lattice, raw_args, say

torch = _azp.torch
pytestmark = pytest.mark.gpu

PLJ = {(0, 0): {"epsilon": 1.0, "sigma": 1.0, "attraction_scale_factor": 0.5}}


def _d_params():
    size = _azp.lib().azp_param_size(_azp.EVAL_PERTURBED_LJ)
    pod = (ctypes.c_char * size)()
    _azp.call("azp_param_pack", _azp.EVAL_PERTURBED_LJ, _azp.dbl_array([1.0, 1.0, 0.5]), pod)
    return torch.tensor(np.frombuffer(bytes(pod), dtype=np.uint8).copy(), device="cuda")


def _run(entry, sysm, nl, precision, d_params, rcsq):
    N = sysm.N
    pitch = (N + 63) // 64 * 64
    f = torch.full((N, 4), np.nan, dtype=torch.float64, device="cuda")
    v = torch.full((6, pitch), np.nan, dtype=torch.float64, device="cuda")
    a = raw_args(sysm, nl, rcsq, None, 1, f, v, pitch, precision=_azp.PRECISIONS[precision])
    assert not a.particles.d_pos_fixed
    _azp.call(entry, ctypes.byref(a), _azp.EVAL_PERTURBED_LJ, _azp.dptr(d_params))
    torch.cuda.synchronize()
    return f.cpu().numpy(), v[:, :N].cpu().numpy()


def test_noalloc_matches_allocating_launcher():
    d_params = _d_params()
    rcsq = torch.tensor([9.0], dtype=torch.float64, device="cuda")
    for n in (16, 11):  # the second call is smaller
        L = 1.07 * n
        pos4, _ = jittered_lattice(n, L, 0.15, 1, seed=60 + n)
        sysm = GpuSystem(pos4, np.array([L, L, L]))
        ref = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
        for precision in ("mixed", "fp64"):
            say(f"[noalloc] N={sysm.N} {precision}")
            fa, va = _run("azp_gpu_compute_pair_forces", sysm, ref.nl, precision, d_params, rcsq)
            fn, vn = _run("azp_gpu_compute_pair_forces_noalloc", sysm, ref.nl, precision, d_params, rcsq)
            assert np.array_equal(fa, fn) and np.array_equal(va, vn)
            assert np.all(np.isfinite(fn))


def test_noalloc_rejects_more_particles_than_its_scratch():
    """N = 2^20 + 1 in the form that needs the scratch: a host-side error.
    The arguments point at small arrays, which nothing reads because the size
    check comes before any launch."""
    d_params = _d_params()
    rcsq = torch.tensor([9.0], dtype=torch.float64, device="cuda")
    pos4, _ = jittered_lattice(8, 8.6, 0.1, 1, seed=3)
    sysm = GpuSystem(pos4, np.array([8.6, 8.6, 8.6]))
    ref = gpu_pair("PotentialPairPerturbedLennardJones", sysm, 1, PLJ, 3.0)
    N = sysm.N
    pitch = (N + 63) // 64 * 64
    f = torch.zeros((N, 4), dtype=torch.float64, device="cuda")
    v = torch.zeros((6, pitch), dtype=torch.float64, device="cuda")
    a = raw_args(sysm, ref.nl, rcsq, None, 1, f, v, pitch, precision=_azp.PRECISIONS["mixed"])
    big = (1 << 20) + 1
    a.particles.N = big
    a.nlist_pitch = big
    a.compute_virial = 0
    say("[noalloc] N = 2^20 + 1 is rejected")
    with pytest.raises(_azp.AzpError, match="at most 1048576 particles"):
        _azp.call("azp_gpu_compute_pair_forces_noalloc", ctypes.byref(a), _azp.EVAL_PERTURBED_LJ,
                  _azp.dptr(d_params))
    torch.cuda.synchronize()
