"""ORACLE — test infrastructure only.

ctypes bindings for ``oracle/lib/liboracle.so``, the FP64 CPU restatement of
azplugins' short-range pair-force path (see ``oracle/oracle.h``).  Only
``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, always as the checker or the timed CPU
baseline; the ``azplugins_amd`` product never imports it.

Evaluator codes and the packed-parameter layout (8 doubles per type pair, the
reference's stored ``param_type`` fields in declaration order) match
``oracle.h``.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

HERTZ = 0
EXPANDED_YUKAWA = 1
PERTURBED_LJ = 2
COLLOID = 3
DPD_GENERAL_WEIGHT = 4
TWO_PATCH_MORSE = 5

SHIFT_NONE = 0
SHIFT_SHIFT = 1
SHIFT_XPLOR = 2

MODE_CODES = {"none": SHIFT_NONE, "shift": SHIFT_SHIFT, "xplor": SHIFT_XPLOR}

# dict key order of the reference's Python TypeParameterDicts (pair.py)
PARAM_KEYS = {
    HERTZ: ("epsilon",),
    EXPANDED_YUKAWA: ("epsilon", "kappa", "delta"),
    PERTURBED_LJ: ("epsilon", "sigma", "attraction_scale_factor"),
    COLLOID: ("A", "a_1", "a_2", "sigma"),
    DPD_GENERAL_WEIGHT: ("A", "gamma", "s"),
    TWO_PATCH_MORSE: ("M_d", "M_r", "r_eq", "omega", "alpha", "repulsion"),
}

PSTRIDE = 8

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"oracle library missing at {_LIB_PATH}; run `make -C oracle`"
            )
        L = ctypes.CDLL(_LIB_PATH)
        d = ctypes.c_double
        pd = ctypes.POINTER(ctypes.c_double)
        pu = ctypes.POINTER(ctypes.c_uint32)
        i64 = ctypes.c_int64
        u32 = ctypes.c_uint32
        ci = ctypes.c_int
        L.orc_param_pack.argtypes = [ci, pd, pd]
        L.orc_param_unpack.argtypes = [ci, pd, pd]
        L.orc_eval_iso.argtypes = [ci, d, d, pd, ci, pd, pd]
        L.orc_eval_dpd.argtypes = [
            d, d, pd, ctypes.c_uint16, u32, u32, u32, d, d, d, ci, pd, pd, pd,
        ]
        L.orc_eval_two_patch_morse.argtypes = [pd, pd, pd, d, pd, ci, pd, pd, pd, pd]
        L.orc_philox4x32_10.argtypes = [pu, pu, pu]
        L.orc_dpd_alpha.argtypes = [ctypes.c_uint16, u32, u32, u32]
        L.orc_dpd_alpha.restype = d
        L.orc_dpd_alpha_many.argtypes = [ctypes.c_uint16, u32, ctypes.c_uint64, pu, pu, pd]
        for fn in ("orc_ppow",):
            getattr(L, fn).argtypes = [d, d]
            getattr(L, fn).restype = d
        for fn in ("orc_pexp", "orc_plog"):
            getattr(L, fn).argtypes = [d]
            getattr(L, fn).restype = d
        L.orc_min_image.argtypes = [pd, pd]
        for fn in ("orc_nlist_brute", "orc_nlist_cell"):
            getattr(L, fn).argtypes = [u32, pd, pd, u32, pd, i64, pu, pu]
            getattr(L, fn).restype = i64
        L.orc_pair_forces.argtypes = [ci, u32, pd, pd, i64, pu, pu, u32, pd, pd, pd, ci, pd, pd]
        L.orc_full_list.argtypes = [u32, i64, pu, pu, ctypes.POINTER(ctypes.c_int64), pu]
        L.orc_pair_forces_threaded.argtypes = [ci, u32, pd, pd, ctypes.POINTER(ctypes.c_int64), pu, u32, pd, pd, pd,
                                               ci, pd, pd, ci]
        L.orc_dpd_forces.argtypes = [
            u32, pd, pd, pu, pd, i64, pu, pu, u32, pd, pd,
            ctypes.c_uint16, u32, d, d, ci, pd, pd,
        ]
        L.orc_aniso_forces.argtypes = [u32, pd, pd, pd, i64, pu, pu, u32, pd, pd, ci, pd, pd, pd]
        L.orc_count_pairs_in_cut.argtypes = [u32, pd, pd, i64, pu, pu, u32, pd]
        L.orc_count_pairs_in_cut.restype = i64
        L.orc_run_nve.argtypes = [ci, u32, pd, pd, pd, u32, pd, pd, d, ci, d, ctypes.c_uint64, pd, pd]
        L.orc_run_nve.restype = i64
        L.orc_wall_forces.argtypes = [ci, u32, pd, u32, pd, u32, pd, u32, pd, u32, pd, pd, pd]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _pu(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def pack(evaluator, values):
    """dict (or value sequence in key order) -> stored param fields (8 doubles)."""
    if isinstance(values, dict):
        values = [float(values[k]) for k in PARAM_KEYS[evaluator]]
    v = np.zeros(PSTRIDE)
    v[: len(values)] = values
    p = np.zeros(PSTRIDE)
    lib().orc_param_pack(evaluator, _pd(v), _pd(p))
    return p


def unpack(evaluator, packed):
    p = _f64(packed)
    v = np.zeros(PSTRIDE)
    lib().orc_param_unpack(evaluator, _pd(p), _pd(v))
    keys = PARAM_KEYS[evaluator]
    out = {k: float(v[n]) for n, k in enumerate(keys)}
    if evaluator == TWO_PATCH_MORSE:
        out["repulsion"] = bool(out["repulsion"])
    return out


def pack_matrix(evaluator, params, ntypes):
    """params: {(ti, tj): dict} with integer type indices (symmetric)."""
    m = np.zeros((ntypes * ntypes, PSTRIDE))
    for (a, b), v in params.items():
        p = pack(evaluator, v)
        m[a * ntypes + b] = p
        m[b * ntypes + a] = p
    return m


def eval_iso(evaluator, rsq, rcutsq, packed, shift=False):
    f = ctypes.c_double(0.0)
    e = ctypes.c_double(0.0)
    p = _f64(packed)
    ok = lib().orc_eval_iso(evaluator, rsq, rcutsq, _pd(p), int(shift), ctypes.byref(f), ctypes.byref(e))
    return bool(ok), f.value, e.value


def eval_dpd(rsq, rcutsq, packed, seed, tag_i, tag_j, timestep, dt, kT, rdotv, portable_pow=True):
    f = ctypes.c_double(0.0)
    fc = ctypes.c_double(0.0)
    e = ctypes.c_double(0.0)
    p = _f64(packed)
    ok = lib().orc_eval_dpd(
        rsq, rcutsq, _pd(p), seed, tag_i, tag_j, timestep, dt, kT, rdotv,
        int(portable_pow), ctypes.byref(f), ctypes.byref(fc), ctypes.byref(e),
    )
    return bool(ok), f.value, fc.value, e.value


def eval_two_patch_morse(dr, qi, qj, rcutsq, packed, shift=False):
    dr = _f64(dr)
    qi = _f64(qi)
    qj = _f64(qj)
    p = _f64(packed)
    F = np.zeros(3)
    ti = np.zeros(3)
    tj = np.zeros(3)
    e = ctypes.c_double(0.0)
    ok = lib().orc_eval_two_patch_morse(
        _pd(dr), _pd(qi), _pd(qj), rcutsq, _pd(p), int(shift), _pd(F), ctypes.byref(e), _pd(ti), _pd(tj)
    )
    return bool(ok), F, e.value, ti, tj


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_pu(c), _pu(k), _pu(out))
    return out


def dpd_alpha(seed, timestep, tag_i, tag_j):
    ti = np.ascontiguousarray(np.atleast_1d(tag_i), dtype=np.uint32)
    tj = np.ascontiguousarray(np.atleast_1d(tag_j), dtype=np.uint32)
    out = np.zeros(ti.shape[0])
    lib().orc_dpd_alpha_many(seed, timestep, ti.shape[0], _pu(ti), _pu(tj), _pd(out))
    return out


def ppow(x, y):
    return lib().orc_ppow(x, y)


def pexp(x):
    return lib().orc_pexp(x)


def plog(x):
    return lib().orc_plog(x)


def r_listsq_matrix(r_cut, r_buff):
    r_cut = np.asarray(r_cut, dtype=np.float64).reshape(-1)
    rl = r_cut + r_buff
    return np.where(r_cut > 0.0, rl * rl, 0.0)


def nlist(pos4, L, ntypes, r_listsq, method="cell"):
    """Half neighbor list as (pair_i, pair_j) arrays, each unordered pair once."""
    pos4 = _f64(pos4)
    L = _f64(L)
    rl = _f64(r_listsq)
    N = pos4.shape[0]
    fn = lib().orc_nlist_cell if method == "cell" else lib().orc_nlist_brute
    # half-list size of a uniform fluid at this density and the largest list
    # radius, with 30 % headroom, so that one walk usually suffices
    rmax = float(np.sqrt(np.max(rl))) if rl.size else 0.0
    est = 0.5 * N * (N / float(np.prod(L))) * 4.0 / 3.0 * np.pi * rmax**3
    cap = int(max(1024, min(1.3 * est + 16 * N, N * (N - 1) / 2 + 1)))
    while True:
        pi = np.zeros(cap, dtype=np.uint32)
        pj = np.zeros(cap, dtype=np.uint32)
        n = fn(N, _pd(pos4), _pd(L), ntypes, _pd(rl), cap, _pu(pi), _pu(pj))
        if n >= 0:
            return pi[:n].copy(), pj[:n].copy()
        cap = -n + 1024


def pair_forces(evaluator, pos4, L, pairs, ntypes, packed, rcutsq, ronsq=None, mode="none",
                with_virial=True):
    pos4 = _f64(pos4)
    L = _f64(L)
    pi, pj = pairs
    N = pos4.shape[0]
    f = np.zeros((N, 4))
    v = np.zeros((6, N)) if with_virial else None
    packed = _f64(packed)
    rc = _f64(rcutsq)
    ron = _f64(ronsq if ronsq is not None else np.zeros_like(rc))
    lib().orc_pair_forces(
        evaluator, N, _pd(pos4), _pd(L), pi.shape[0], _pu(pi), _pu(pj), ntypes,
        _pd(packed), _pd(rc), _pd(ron), MODE_CODES.get(mode, mode), _pd(f),
        _pd(v) if v is not None else None,
    )
    return f, v


def full_list(N, pairs):
    """CSR full (directed) list from a half list: (offsets[N + 1], neighbors)."""
    pi = np.ascontiguousarray(pairs[0], dtype=np.uint32)
    pj = np.ascontiguousarray(pairs[1], dtype=np.uint32)
    offsets = np.zeros(N + 1, dtype=np.int64)
    nbrs = np.zeros(2 * pi.shape[0], dtype=np.uint32)
    lib().orc_full_list(N, pi.shape[0], _pu(pi), _pu(pj), offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                        _pu(nbrs))
    return offsets, nbrs


def pair_forces_threaded(evaluator, pos4, L, csr, ntypes, packed, rcutsq, ronsq=None, mode="none",
                         nthreads=1, with_virial=True):
    """orc_pair_forces over a full CSR list, OpenMP-threaded over particles."""
    pos4 = _f64(pos4)
    L = _f64(L)
    offsets, nbrs = csr
    N = pos4.shape[0]
    f = np.zeros((N, 4))
    v = np.zeros((6, N)) if with_virial else None
    packed = _f64(packed)
    rc = _f64(rcutsq)
    ron = _f64(ronsq if ronsq is not None else np.zeros_like(rc))
    lib().orc_pair_forces_threaded(
        evaluator, N, _pd(pos4), _pd(L), offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _pu(nbrs), ntypes,
        _pd(packed), _pd(rc), _pd(ron), MODE_CODES.get(mode, mode), _pd(f),
        _pd(v) if v is not None else None, int(nthreads),
    )
    return f, v


def dpd_forces(pos4, vel4, tag, L, pairs, ntypes, packed, rcutsq, seed, timestep, dt, kT,
               portable_pow=True):
    pos4 = _f64(pos4)
    vel4 = _f64(vel4)
    tag = np.ascontiguousarray(tag, dtype=np.uint32)
    L = _f64(L)
    pi, pj = pairs
    N = pos4.shape[0]
    f = np.zeros((N, 4))
    v = np.zeros((6, N))
    packed = _f64(packed)
    rc = _f64(rcutsq)
    lib().orc_dpd_forces(
        N, _pd(pos4), _pd(vel4), _pu(tag), _pd(L), pi.shape[0], _pu(pi), _pu(pj), ntypes,
        _pd(packed), _pd(rc), seed, timestep, dt, kT, int(portable_pow), _pd(f), _pd(v),
    )
    return f, v


def aniso_forces(pos4, quat4, L, pairs, ntypes, packed, rcutsq, mode="none"):
    pos4 = _f64(pos4)
    quat4 = _f64(quat4)
    L = _f64(L)
    pi, pj = pairs
    N = pos4.shape[0]
    f = np.zeros((N, 4))
    t = np.zeros((N, 4))
    v = np.zeros((6, N))
    packed = _f64(packed)
    rc = _f64(rcutsq)
    lib().orc_aniso_forces(
        N, _pd(pos4), _pd(quat4), _pd(L), pi.shape[0], _pu(pi), _pu(pj), ntypes,
        _pd(packed), _pd(rc), MODE_CODES.get(mode, mode), _pd(f), _pd(t), _pd(v),
    )
    return f, t, v


def count_pairs_in_cut(pos4, L, pairs, ntypes, rcutsq):
    pos4 = _f64(pos4)
    L = _f64(L)
    pi, pj = pairs
    rc = _f64(rcutsq)
    return lib().orc_count_pairs_in_cut(
        pos4.shape[0], _pd(pos4), _pd(L), pi.shape[0], _pu(pi), _pu(pj), ntypes, _pd(rc)
    )


def run_nve(evaluator, pos4, vel4, L, ntypes, packed, r_cut, r_buff, mode, dt, steps):
    pos4 = np.array(pos4, dtype=np.float64, order="C")
    vel4 = np.array(vel4, dtype=np.float64, order="C")
    L = _f64(L)
    N = pos4.shape[0]
    f = np.zeros((N, 4))
    v = np.zeros((6, N))
    rc = _f64(r_cut)
    packed = _f64(packed)
    nb = lib().orc_run_nve(
        evaluator, N, _pd(pos4), _pd(vel4), _pd(L), ntypes, _pd(packed), _pd(rc), r_buff,
        MODE_CODES.get(mode, mode), dt, steps, _pd(f), _pd(v),
    )
    return pos4, vel4, f, v, nb


WALL_LJ93 = 0
WALL_COLLOID = 1
WALL_KEYS = {WALL_LJ93: ("A", "sigma", "r_cut", "r_extrap"),
             WALL_COLLOID: ("A", "a", "sigma", "r_cut", "r_extrap")}


def wall_forces(kind, pos4, values, spheres=(), cylinders=(), planes=(), with_virial=True):
    """orc_wall_forces: values[t] in wall.py schema order; geometry rows as in
    the C header.  Returns (force4 (N, 4), virial (6, N) or None)."""
    pos4 = _f64(pos4)
    N = pos4.shape[0]
    vals = _f64(np.asarray(values, dtype=np.float64).reshape(-1))
    ntypes = vals.shape[0] // len(WALL_KEYS[kind])
    sp = _f64(np.asarray(spheres, dtype=np.float64).reshape(-1, 6))
    cy = _f64(np.asarray(cylinders, dtype=np.float64).reshape(-1, 9))
    pl = _f64(np.asarray(planes, dtype=np.float64).reshape(-1, 7))
    f = np.zeros((N, 4))
    v = np.zeros((6, N)) if with_virial else None
    lib().orc_wall_forces(kind, N, _pd(pos4), ntypes, _pd(vals), sp.shape[0], _pd(sp), cy.shape[0], _pd(cy),
                          pl.shape[0], _pd(pl), _pd(f), _pd(v) if v is not None else None)
    return f, v
