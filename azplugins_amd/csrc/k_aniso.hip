// Anisotropic pair forces and torques (TwoPatchMorse) over a full neighbor list.
//
// Replaces gpu_compute_pair_aniso_forces<AnisoPairEvaluatorTwoPatchMorse>
// (AnisoPotentialPairGPUKernel.cu.inc:21-25).  The evaluator restates
// AnisoPairEvaluatorTwoPatchMorse.h:127-216.  With a full list each lane
// evaluates (i, j) with dr = r_i - r_j and keeps force and torque_i; particle
// j's torque comes from its own evaluation of (j, i), which equals torque_j of
// (i, j) exactly in exact arithmetic (both are dU/dgamma_j (r_hat x n_j)).
#include <type_traits>

#include "device_common.h"

namespace azp
{

struct AnisoKernelArgs
    {
    const double4* pos;
    const int4* pos_fixed;
    const double4* orientation;
    uint32_t N;
    DevBox box;
    float fx_scale[3];
    const uint32_t* n_neigh;
    const uint32_t* nlist;
    uint32_t nlist_pitch;
    const azp_param_two_patch_morse* params;
    const double* rcutsq;
    uint32_t ntypes;
    int shift_mode;
    double4* force;
    double4* torque;
    double* virial;
    uint32_t vpitch;
    };

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
    const R rinv = f_rsqrt(rsq);
    const R r = R(1) / rinv;
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
    const R Oi = R(1) / (R(1) + gie);
    const R gj = u[0] * n_j[0] + u[1] * n_j[1] + u[2] * n_j[2];
    const R gje = f_exp(-q.omega * (gj * gj - q.alpha));
    const R Oj = R(1) / (R(1) + gje);
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

template<bool MIXED, bool VIRIAL, bool SINGLE>
__global__ void __launch_bounds__(512) k_aniso(const AnisoKernelArgs a)
    {
    using R = std::conditional_t<MIXED, float, double>;
    extern __shared__ double smem[];
    TPMPrep<R>* s_q = reinterpret_cast<TPMPrep<R>*>(smem);
    const uint32_t nt = a.ntypes;
    TPMPrep<R> q0;
    if (SINGLE)
        {
        q0 = tpm_prepare<R>(a.params[0], a.rcutsq[0], a.shift_mode == AZP_SHIFT_SHIFT);
        }
    else
        {
        for (uint32_t t = threadIdx.x; t < nt * nt; t += blockDim.x)
            s_q[t] = tpm_prepare<R>(a.params[t], a.rcutsq[t], a.shift_mode == AZP_SHIFT_SHIFT);
        __syncthreads();
        }
    const uint32_t i = xcd_logical_block() * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;

    const double4 oi = a.orientation[i];
    R n_i[3];
    director<R>((R)oi.x, (R)oi.y, (R)oi.z, (R)oi.w, n_i);
    double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0, pe = 0;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    const uint32_t nn = a.n_neigh[i];
    const double4 pi64 = a.pos[i];
    int4 pfi;
    if constexpr (MIXED)
        pfi = a.pos_fixed[i];
    const uint32_t ti = (uint32_t)pi64.w;
    for (uint32_t k = 0; k < nn; ++k)
        {
        const uint32_t j = a.nlist[(size_t)k * a.nlist_pitch + i];
        R dr[3];
        R rsq;
        uint32_t tj;
        bool in;
        if constexpr (!MIXED)
            {
            const double4 pj = a.pos[j];
            dr[0] = min_image_1d(pi64.x - pj.x, a.box.L[0], a.box.Linv[0]);
            dr[1] = min_image_1d(pi64.y - pj.y, a.box.L[1], a.box.Linv[1]);
            dr[2] = min_image_1d(pi64.z - pj.z, a.box.L[2], a.box.Linv[2]);
            rsq = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            tj = (uint32_t)pj.w;
            const TPMPrep<R>& q = SINGLE ? q0 : s_q[ti * nt + tj];
            in = !(rsq > q.rcsq); // AnisoPairEvaluatorTwoPatchMorse.h:135
            }
        else
            {
            const int4 pj = a.pos_fixed[j];
            dr[0] = (R)fixed_delta(pfi.x, pj.x) * a.fx_scale[0];
            dr[1] = (R)fixed_delta(pfi.y, pj.y) * a.fx_scale[1];
            dr[2] = (R)fixed_delta(pfi.z, pj.z) * a.fx_scale[2];
            rsq = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            tj = (uint32_t)pj.w;
            const TPMPrep<R>& q = SINGLE ? q0 : s_q[ti * nt + tj];
            in = !(rsq > q.rcsq_hi);
            if (in && rsq > q.rcsq_lo)
                {
                const double4 pj64 = a.pos[j];
                const double ex = min_image_1d(pi64.x - pj64.x, a.box.L[0], a.box.Linv[0]);
                const double ey = min_image_1d(pi64.y - pj64.y, a.box.L[1], a.box.Linv[1]);
                const double ez = min_image_1d(pi64.z - pj64.z, a.box.L[2], a.box.Linv[2]);
                const double r64 = rsq3_nofma(ex, ey, ez);
                in = !(r64 > q.rcsq64);
                }
            }
        if (!in)
            continue;
        const TPMPrep<R>& q = SINGLE ? q0 : s_q[ti * nt + tj];
        const double4 oj = a.orientation[j];
        R n_j[3];
        director<R>((R)oj.x, (R)oj.y, (R)oj.z, (R)oj.w, n_j);
        R f[3], t[3], e;
        tpm_eval<R>(q, dr, rsq, n_i, n_j, f, t, e);
        const double Fx = f[0], Fy = f[1], Fz = f[2];
        fx += Fx;
        fy += Fy;
        fz += Fz;
        tx += (double)t[0];
        ty += (double)t[1];
        tz += (double)t[2];
        pe += (double)e;
        if (VIRIAL)
            {
            const double dx = dr[0], dy = dr[1], dz = dr[2];
            v0 = fma(dx, Fx, v0);
            v1 = fma(dy, Fx, v1);
            v2 = fma(dz, Fx, v2);
            v3 = fma(dy, Fy, v3);
            v4 = fma(dz, Fy, v4);
            v5 = fma(dz, Fz, v5);
            }
        }
    a.force[i] = make_double4(fx, fy, fz, 0.5 * pe);
    a.torque[i] = make_double4(tx, ty, tz, 0.0);
    if (VIRIAL)
        {
        const size_t p = a.vpitch;
        a.virial[0 * p + i] = 0.5 * v0;
        a.virial[1 * p + i] = 0.5 * v1;
        a.virial[2 * p + i] = 0.5 * v2;
        a.virial[3 * p + i] = 0.5 * v3;
        a.virial[4 * p + i] = 0.5 * v4;
        a.virial[5 * p + i] = 0.5 * v5;
        }
    }

template<bool MIXED, bool VIRIAL, bool SINGLE>
static int launch_aniso_t(const AnisoKernelArgs& ka, unsigned block, hipStream_t s)
    {
    using R = std::conditional_t<MIXED, float, double>;
    const size_t shmem = SINGLE ? 0 : (size_t)ka.ntypes * ka.ntypes * sizeof(TPMPrep<R>);
    if (shmem > 64 * 1024)
        return set_error(AZP_ERR_UNSUPPORTED, "too many particle types for LDS staging");
    hipLaunchKernelGGL((k_aniso<MIXED, VIRIAL, SINGLE>), dim3(grid_for(ka.N, block)), dim3(block), shmem, s, ka);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

template<bool MIXED>
static int launch_aniso(const AnisoKernelArgs& ka, bool vir, unsigned block, hipStream_t s)
    {
    const bool single = ka.ntypes == 1;
    if (vir)
        return single ? launch_aniso_t<MIXED, true, true>(ka, block, s)
                      : launch_aniso_t<MIXED, true, false>(ka, block, s);
    return single ? launch_aniso_t<MIXED, false, true>(ka, block, s)
                  : launch_aniso_t<MIXED, false, false>(ka, block, s);
    }

int validate_common(const azp_pair_args* args);

} // namespace azp

using namespace azp;

extern "C" int azp_gpu_compute_pair_aniso_forces(const azp_pair_args* args, int evaluator, const void* d_params)
    {
    AZP_TRY(validate_common(args));
    const azp_particles& p = args->particles;
    if (p.N == 0)
        return AZP_OK;
    if (evaluator != AZP_EVAL_TWO_PATCH_MORSE)
        return set_error(AZP_ERR_INVALID, "evaluator %d is not anisotropic", evaluator);
    if (!d_params || !p.d_orientation || !args->out.d_torque)
        return set_error(AZP_ERR_INVALID, "aniso needs params, orientations and a torque array");
    const bool mixed = args->precision == AZP_PREC_MIXED;
    if (mixed && !p.d_pos_fixed)
        return set_error(AZP_ERR_INVALID, "mixed precision needs d_pos_fixed");
    if (args->shift_mode == AZP_SHIFT_XPLOR)
        return set_error(AZP_ERR_UNSUPPORTED, "anisotropic pairs accept modes none and shift");
    AnisoKernelArgs ka;
    ka.pos = reinterpret_cast<const double4*>(p.d_pos);
    ka.pos_fixed = reinterpret_cast<const int4*>(p.d_pos_fixed);
    ka.orientation = reinterpret_cast<const double4*>(p.d_orientation);
    ka.N = p.N;
    ka.box = make_box(p.box);
    for (int k = 0; k < 3; ++k)
        ka.fx_scale[k] = (float)(ka.box.L[k] / 4294967296.0);
    ka.n_neigh = args->d_n_neigh;
    ka.nlist = args->d_nlist;
    ka.nlist_pitch = args->nlist_pitch;
    ka.params = static_cast<const azp_param_two_patch_morse*>(d_params);
    ka.rcutsq = args->d_rcutsq;
    ka.ntypes = args->ntypes;
    ka.shift_mode = args->shift_mode;
    ka.force = reinterpret_cast<double4*>(args->out.d_force);
    ka.torque = reinterpret_cast<double4*>(args->out.d_torque);
    ka.virial = args->out.d_virial;
    ka.vpitch = args->out.virial_pitch;
    const unsigned block = args->block_size ? args->block_size : 256;
    hipStream_t s = (hipStream_t)args->stream;
    return mixed ? launch_aniso<true>(ka, args->compute_virial != 0, block, s)
                 : launch_aniso<false>(ka, args->compute_virial != 0, block, s);
    }
