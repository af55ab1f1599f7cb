// TwoPatchMorse pair term shared by the single-system and the batched
// anisotropic kernels.  The evaluator restates
// AnisoPairEvaluatorTwoPatchMorse.h:127-216; the per-entry driver restates the
// force / torque / energy / virial accumulation of AnisoPotentialPair (full
// list: particle i keeps F, tau_i and half the pair energy).
#pragma once

#include <type_traits>

#include "device_common.h"

namespace azp
{

template<typename R> struct TPMPrep
    {
    R M_d, M_rinv, r_eq, omega, alpha;
    R rcsq, rcsq_lo, rcsq_hi;
    R ushift; // U_Morse(rc), used when shifting
    double rcsq64;
    int repulsion;
    int shift;
    };

template<typename R>
__device__ TPMPrep<R> tpm_prepare(const azp_param_two_patch_morse& p, double rcsq, bool shift)
    {
    TPMPrep<R> q;
    q.M_d = (R)p.M_d;
    q.M_rinv = (R)p.M_rinv;
    q.r_eq = (R)p.r_eq;
    q.omega = (R)p.omega;
    q.alpha = (R)p.alpha;
    q.rcsq = (R)rcsq;
    q.rcsq_lo = (R)(rcsq * (1.0 - 1e-5));
    q.rcsq_hi = (R)(rcsq * (1.0 + 1e-5));
    q.rcsq64 = rcsq;
    q.repulsion = p.repulsion ? 1 : 0;
    q.shift = shift ? 1 : 0;
    const double rcut = sqrt(rcsq);
    const double me = exp(-(rcut - p.r_eq) * p.M_rinv);
    const double ome = 1.0 - me;
    q.ushift = (R)(p.M_d * (ome * ome - 1.0));
    return q;
    }

// n = rotate(q, (1,0,0)) = (s^2 - |v|^2) e_x + 2 s (v x e_x) + 2 (v . e_x) v
template<typename R> __device__ __forceinline__ void director(R s, R vx, R vy, R vz, R n[3])
    {
    const R a = s * s - (vx * vx + vy * vy + vz * vz);
    n[0] = a + R(2) * vx * vx;
    n[1] = R(2) * s * vz + R(2) * vx * vy;
    n[2] = -R(2) * s * vy + R(2) * vx * vz;
    }

template<typename R>
__device__ __forceinline__ void tpm_eval(const TPMPrep<R>& q,
                                         const R dr[3],
                                         R rsq,
                                         const R n_i[3],
                                         const R n_j[3],
                                         R f[3],
                                         R t_i[3],
                                         R& e)
    {
    // reciprocals through f_rcp: the hardware reciprocal in FP32 (about one
    // ulp), the correctly rounded 1/x in FP64
    const R rinv = f_rsqrt(rsq);
    const R r = f_rcp(rinv);
    const R u[3] = {dr[0] * rinv, dr[1] * rinv, dr[2] * rinv};
    R UMorse = R(-1) * q.M_d;
    R dUMorse_dr = R(0);
    if (r > q.r_eq || q.repulsion)
        {
        const R me = f_exp(-(r - q.r_eq) * q.M_rinv);
        const R ome = R(1) - me;
        UMorse = q.M_d * (ome * ome - R(1));
        dUMorse_dr = R(2) * q.M_d * q.M_rinv * me * ome;
        }
    const R gi = u[0] * n_i[0] + u[1] * n_i[1] + u[2] * n_i[2];
    const R gie = f_exp(-q.omega * (gi * gi - q.alpha));
    const R Oi = f_rcp(R(1) + gie);
    const R gj = u[0] * n_j[0] + u[1] * n_j[1] + u[2] * n_j[2];
    const R gje = f_exp(-q.omega * (gj * gj - q.alpha));
    const R Oj = f_rcp(R(1) + gje);
    e = UMorse * Oi * Oj;
    const R dU_dr = dUMorse_dr * Oi * Oj;
    const R dOi = R(2) * q.omega * gi * gie * Oi * Oi;
    const R dOj = R(2) * q.omega * gj * gje * Oj * Oj;
    const R dU_dgi = dOi * UMorse * Oj;
    const R dU_dgj = dOj * UMorse * Oi;
    // c = u x n ; n_perp = (-u) x c
    const R ci[3] = {u[1] * n_i[2] - u[2] * n_i[1], u[2] * n_i[0] - u[0] * n_i[2], u[0] * n_i[1] - u[1] * n_i[0]};
    const R cj[3] = {u[1] * n_j[2] - u[2] * n_j[1], u[2] * n_j[0] - u[0] * n_j[2], u[0] * n_j[1] - u[1] * n_j[0]};
    const R pi[3] = {-u[1] * ci[2] + u[2] * ci[1], -u[2] * ci[0] + u[0] * ci[2], -u[0] * ci[1] + u[1] * ci[0]};
    const R pj[3] = {-u[1] * cj[2] + u[2] * cj[1], -u[2] * cj[0] + u[0] * cj[2], -u[0] * cj[1] + u[1] * cj[0]};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        {
        f[k] = -dU_dr * u[k] - rinv * (dU_dgi * pi[k] + dU_dgj * pj[k]);
        t_i[k] = dU_dgi * ci[k];
        }
    if (q.shift)
        e -= q.ushift * Oi * Oj;
    }

struct AnisoAcc
    {
    double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0, pe = 0;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    };

// Geometry of the periodic box particle i lives in, as one lane sees it.
struct LaneBox
    {
    double L[3];
    double Linv[3];
    float fx_scale[3]; // L / 2^32 (mixed form)
    };

// Particle i's own data, loaded once per lane.
template<typename R> struct AnisoSelf
    {
    double4 pos;
    int4 pfix;
    R n[3];
    uint32_t type;
    };

template<bool MIXED>
__device__ __forceinline__ AnisoSelf<std::conditional_t<MIXED, float, double>>
aniso_self(const double4* pos, const int4* pos_fixed, const double4* orientation, uint32_t i)
    {
    using R = std::conditional_t<MIXED, float, double>;
    AnisoSelf<R> s;
    s.pos = pos[i];
    if constexpr (MIXED)
        s.pfix = pos_fixed[i];
    else
        s.pfix = make_int4(0, 0, 0, 0);
    const double4 o = orientation[i];
    director<R>((R)o.x, (R)o.y, (R)o.z, (R)o.w, s.n);
    s.type = (uint32_t)s.pos.w;
    return s;
    }

// Director of particle j from its quaternion (s, x, y, z), or a director
// computed beforehand (x, y, z of a float4).
template<typename R> __device__ __forceinline__ void director_of(const double4& o, R n[3])
    {
    director<R>((R)o.x, (R)o.y, (R)o.z, (R)o.w, n);
    }
template<typename R> __device__ __forceinline__ void director_of(const float4& d, R n[3])
    {
    n[0] = (R)d.x;
    n[1] = (R)d.y;
    n[2] = (R)d.z;
    }

// One list entry (i, j) from j's prefetched position (int4 fixed point in
// the mixed form, double4 in the plain form) and orientation: the cutoff test
// (rsq > rcsq is out; the mixed form re-tests pairs within 1e-5 of the cutoff
// in FP64), the evaluator and the accumulation of F, tau_i, E and the virial
// dx_b F_a.  Padding entries (valid false) contribute nothing.
template<bool MIXED, bool VIRIAL, bool SINGLE, typename PosT, typename OriT>
__device__ __forceinline__ void tpm_term(bool valid,
                                         uint32_t j,
                                         const PosT& pj,
                                         const OriT& oj,
                                         const double4* pos,
                                         const AnisoSelf<std::conditional_t<MIXED, float, double>>& me,
                                         const LaneBox& box,
                                         uint32_t nt,
                                         const TPMPrep<std::conditional_t<MIXED, float, double>>& q0,
                                         const TPMPrep<std::conditional_t<MIXED, float, double>>* s_q,
                                         AnisoAcc& acc)
    {
    using R = std::conditional_t<MIXED, float, double>;
    R dr[3];
    R rsq;
    const uint32_t tj = (uint32_t)pj.w;
    const TPMPrep<R>& q = SINGLE ? q0 : s_q[me.type * nt + (valid ? tj : me.type)];
    bool in;
    if constexpr (!MIXED)
        {
        dr[0] = min_image_1d(me.pos.x - pj.x, box.L[0], box.Linv[0]);
        dr[1] = min_image_1d(me.pos.y - pj.y, box.L[1], box.Linv[1]);
        dr[2] = min_image_1d(me.pos.z - pj.z, box.L[2], box.Linv[2]);
        rsq = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        in = valid && !(rsq > q.rcsq); // AnisoPairEvaluatorTwoPatchMorse.h:135
        }
    else
        {
        dr[0] = (R)fixed_delta(me.pfix.x, pj.x) * box.fx_scale[0];
        dr[1] = (R)fixed_delta(me.pfix.y, pj.y) * box.fx_scale[1];
        dr[2] = (R)fixed_delta(me.pfix.z, pj.z) * box.fx_scale[2];
        rsq = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        in = valid && !(rsq > q.rcsq_hi);
        if (in && rsq > q.rcsq_lo)
            {
            const double4 pj64 = pos[j];
            const double ex = min_image_1d(me.pos.x - pj64.x, box.L[0], box.Linv[0]);
            const double ey = min_image_1d(me.pos.y - pj64.y, box.L[1], box.Linv[1]);
            const double ez = min_image_1d(me.pos.z - pj64.z, box.L[2], box.Linv[2]);
            const double r64 = rsq3_nofma(ex, ey, ez);
            in = !(r64 > q.rcsq64);
            }
        }
    if (!in)
        return;
    R n_j[3];
    director_of<R>(oj, n_j);
    R f[3], t[3], e;
    tpm_eval<R>(q, dr, rsq, me.n, n_j, f, t, e);
    const double Fx = f[0], Fy = f[1], Fz = f[2];
    acc.fx += Fx;
    acc.fy += Fy;
    acc.fz += Fz;
    acc.tx += (double)t[0];
    acc.ty += (double)t[1];
    acc.tz += (double)t[2];
    acc.pe += (double)e;
    if (VIRIAL)
        {
        const double dx = dr[0], dy = dr[1], dz = dr[2];
        acc.v0 = fma(dx, Fx, acc.v0);
        acc.v1 = fma(dy, Fx, acc.v1);
        acc.v2 = fma(dz, Fx, acc.v2);
        acc.v3 = fma(dy, Fy, acc.v3);
        acc.v4 = fma(dz, Fy, acc.v4);
        acc.v5 = fma(dz, Fz, acc.v5);
        }
    }

__device__ __forceinline__ uint32_t tpm_wave_max(uint32_t v)
    {
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
    }

// Particle ie's list walked in a wave-uniform loop of `trips` entries (the
// wavefront's longest list, rounded up to 4), four entries at a time: their
// list words, positions and orientations (or precomputed directors, `dirs`)
// are loaded together before the terms are evaluated.  Entries past the
// lane's own count are padding.
template<bool MIXED, bool VIRIAL, bool SINGLE, bool DIRS>
__device__ __forceinline__ void tpm_list(uint32_t ie,
                                         uint32_t nn,
                                         uint32_t trips,
                                         const uint32_t* row,
                                         size_t stride,
                                         const double4* pos,
                                         const int4* pos_fixed,
                                         const double4* orientation,
                                         const float4* dirs,
                                         const AnisoSelf<std::conditional_t<MIXED, float, double>>& me,
                                         const LaneBox& box,
                                         uint32_t nt,
                                         const TPMPrep<std::conditional_t<MIXED, float, double>>& q0,
                                         const TPMPrep<std::conditional_t<MIXED, float, double>>* s_q,
                                         AnisoAcc& acc)
    {
    constexpr uint32_t U = 4;
    using PosT = std::conditional_t<MIXED, int4, double4>;
    using OriT = std::conditional_t<DIRS, float4, double4>;
    // consecutive entries (HOOMD's head-list rows): below F every lane has
    // U entries left, read as one 16-byte load
    uint32_t F = 0;
    if (stride == 1)
        {
        uint32_t m = nn;
        for (int off = 32; off > 0; off >>= 1)
            m = min(m, (uint32_t)__shfl_xor((int)m, off));
        F = __builtin_amdgcn_readfirstlane(m & ~(U - 1));
        }
    for (uint32_t k = 0; k < trips; k += U)
        {
        uint32_t jj[U];
        if (k < F)
            {
            uint4 w;
            __builtin_memcpy(&w, row + k, sizeof(w));
            jj[0] = w.x;
            jj[1] = w.y;
            jj[2] = w.z;
            jj[3] = w.w;
            }
        else
            {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                jj[u] = k + u < nn ? row[(size_t)(k + u) * stride] : ie;
            }
        PosT pj[U];
        OriT oj[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            {
            if constexpr (MIXED)
                pj[u] = pos_fixed[jj[u]];
            else
                pj[u] = pos[jj[u]];
            if constexpr (DIRS)
                oj[u] = dirs[jj[u]];
            else
                oj[u] = orientation[jj[u]];
            }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            tpm_term<MIXED, VIRIAL, SINGLE>(k + u < nn, jj[u], pj[u], oj[u], pos, me, box, nt, q0, s_q, acc);
        }
    }

template<bool VIRIAL>
__device__ __forceinline__ void aniso_store(double4* force,
                                            double4* torque,
                                            double* virial,
                                            uint32_t vpitch,
                                            uint32_t i,
                                            const AnisoAcc& acc)
    {
    force[i] = make_double4(acc.fx, acc.fy, acc.fz, 0.5 * acc.pe);
    torque[i] = make_double4(acc.tx, acc.ty, acc.tz, 0.0);
    if (VIRIAL)
        {
        const size_t p = vpitch;
        virial[0 * p + i] = 0.5 * acc.v0;
        virial[1 * p + i] = 0.5 * acc.v1;
        virial[2 * p + i] = 0.5 * acc.v2;
        virial[3 * p + i] = 0.5 * acc.v3;
        virial[4 * p + i] = 0.5 * acc.v4;
        virial[5 * p + i] = 0.5 * acc.v5;
        }
    }

} // namespace azp
