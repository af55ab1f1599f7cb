// Adjoint (vector-Jacobian product) of the isotropic pair forces and energies.
//
// Forward:  F_i = sum_j f(r_ij) dx_ij,  E_i = 1/2 sum_j u(r_ij),  dx_ij = r_i - r_j,
//           f = -u'(r)/r (the evaluators' force_divr), over the full list.
// Given upstream gradients G_i = dL/dF_i and w_i = dL/dE_i this computes
//   dL/dr_k = sum_j [ f dG + (f'/r)(dx . dG) dx - 1/2 (w_k + w_j) f dx ],
//             dG = G_k - G_j, dx = dx_kj  (H_kj = f I + (f'/r) dx dx^T),
//   dL/dtheta = sum over directed pairs (k, j) of (G_k . dx) df/dtheta
//               + 1/2 w_k du/dtheta, binned by the type pair (t_k, t_j),
// for theta = (epsilon, kappa, delta) of ExpandedYukawa
// (PairEvaluatorExpandedYukawa.h:92-115) and (epsilon, sigma,
// attraction_scale_factor) of PerturbedLennardJones
// (PairEvaluatorPerturbedLennardJones.h:117-155), in the energy-shifted and
// unshifted modes.  Pair arithmetic in R (FP64 or FP32), accumulation in FP64.
// Parameter gradients are reduced per workgroup in a fixed order, then over
// workgroups by a second kernel (deterministic, no atomics).
#include <type_traits>

#include "device_common.h"

namespace azp
{

constexpr int kAdjNpar = 3;
constexpr int kAdjMaxTypes = 4;

template<typename R> struct AdjOut
    {
    R f, fp, u;
    R du[kAdjNpar];
    R df[kAdjNpar];
    };

// ---------------------------------------------------------------- Yukawa
template<typename R> struct AdjYukawaPrep
    {
    R eps, kappa, delta;
    R u_c, du_c[kAdjNpar]; // energy shift and its derivatives (zero if unshifted)
    };

template<typename R>
__device__ AdjYukawaPrep<R> adj_yukawa_prepare(const azp_param_expanded_yukawa& p, double rcsq, bool shift)
    {
    AdjYukawaPrep<R> q;
    q.eps = (R)p.epsilon;
    q.kappa = (R)p.kappa;
    q.delta = (R)p.delta;
    double uc = 0, d0 = 0, d1 = 0, d2 = 0;
    if (shift)
        {
        const double sc = sqrt(rcsq) - p.delta;
        const double Ec = exp(-p.kappa * sc);
        uc = p.epsilon * Ec / sc;
        d0 = Ec / sc;
        d1 = -sc * uc;
        d2 = uc * (p.kappa + 1.0 / sc);
        }
    q.u_c = (R)uc;
    q.du_c[0] = (R)d0;
    q.du_c[1] = (R)d1;
    q.du_c[2] = (R)d2;
    return q;
    }

template<typename R> __device__ __forceinline__ void adj_yukawa_eval(const AdjYukawaPrep<R>& q, R rsq, AdjOut<R>& o)
    {
    const R r = f_sqrt(rsq);
    const R rinv = f_rcp(r);
    const R s = r - q.delta;
    const R sinv = f_rcp(s);
    const R E = f_exp(-q.kappa * s);
    const R Es = E * sinv;    // du/deps
    const R u = q.eps * Es;
    const R A = q.kappa + sinv;
    const R g = u * A;        // -du/dr
    o.f = g * rinv;
    const R gprime = -u * (A * A + sinv * sinv); // dg/dr
    o.fp = (gprime - o.f) * rinv * rinv;
    o.u = u - q.u_c;
    o.du[0] = Es - q.du_c[0];
    o.du[1] = -s * u - q.du_c[1];
    o.du[2] = g - q.du_c[2];
    o.df[0] = A * Es * rinv;
    o.df[1] = -u * s * q.kappa * rinv;
    o.df[2] = -gprime * rinv;
    }

// ---------------------------------------------------------------- PerturbedLJ
template<typename R> struct AdjPLJPrep
    {
    R eps, sigma, lam;
    R lj1e, lj2e; // 4 sigma^12, 4 sigma^6 (per unit epsilon)
    R rwcasq;
    R u_c, du_c[kAdjNpar];
    };

// (u, du/deps, du/dsigma, du/dlambda) of the PLJ pair energy at rsq
__device__ __forceinline__ void plj_energy_grads(double eps, double sigma, double lam, double rwcasq, double rsq,
                                                 double& u, double d[3])
    {
    const double s2 = sigma * sigma, s6 = s2 * s2 * s2;
    const double lj1e = 4.0 * s6 * s6, lj2e = 4.0 * s6;
    const double r2inv = 1.0 / rsq, r6 = r2inv * r2inv * r2inv;
    const double uLJe = r6 * (lj1e * r6 - lj2e);                   // u_LJ / eps
    const double duLJ_ds = eps * r6 * (12.0 * lj1e * r6 - 6.0 * lj2e) / sigma;
    if (rsq < rwcasq)
        {
        u = eps * uLJe + (1.0 - lam) * eps;
        d[0] = uLJe + (1.0 - lam);
        d[1] = duLJ_ds;
        d[2] = -eps;
        }
    else
        {
        u = lam * eps * uLJe;
        d[0] = lam * uLJe;
        d[1] = lam * duLJ_ds;
        d[2] = eps * uLJe;
        }
    }

template<typename R>
__device__ AdjPLJPrep<R> adj_plj_prepare(const azp_param_perturbed_lj& p, double rcsq, bool shift)
    {
    // user parameters from the stored fields (asDict, PairEvaluatorPerturbedLennardJones.h:47-54)
    const double eps = p.epsilon_x_4 / 4.0;
    const double sigma = pow(p.sigma_6, 1.0 / 6.0);
    AdjPLJPrep<R> q;
    q.eps = (R)eps;
    q.sigma = (R)sigma;
    q.lam = (R)p.attraction_scale_factor;
    const double s2 = sigma * sigma, s6 = s2 * s2 * s2;
    q.lj1e = (R)(4.0 * s6 * s6);
    q.lj2e = (R)(4.0 * s6);
    q.rwcasq = (R)p.rwcasq;
    double uc = 0, d[3] = {0, 0, 0};
    if (shift)
        plj_energy_grads(eps, sigma, p.attraction_scale_factor, p.rwcasq, rcsq, uc, d);
    q.u_c = (R)uc;
    for (int k = 0; k < 3; ++k)
        q.du_c[k] = (R)d[k];
    return q;
    }

template<typename R> __device__ __forceinline__ void adj_plj_eval(const AdjPLJPrep<R>& q, R rsq, AdjOut<R>& o)
    {
    const R r2inv = f_rcp(rsq);
    const R r6 = r2inv * r2inv * r2inv;
    const R sinv = f_rcp(q.sigma);
    const R uLJe = r6 * (q.lj1e * r6 - q.lj2e);
    const R fLJe = r2inv * r6 * (R(12) * q.lj1e * r6 - R(6) * q.lj2e);
    const R fpLJe = r2inv * r2inv * r6 * (R(-168) * q.lj1e * r6 + R(48) * q.lj2e);
    const R duLJ_ds = q.eps * r6 * (R(12) * q.lj1e * r6 - R(6) * q.lj2e) * sinv;
    const R dfLJ_ds = q.eps * r2inv * r6 * (R(144) * q.lj1e * r6 - R(36) * q.lj2e) * sinv;
    if (rsq < q.rwcasq)
        {
        o.f = q.eps * fLJe;
        o.fp = q.eps * fpLJe;
        o.u = q.eps * uLJe + (R(1) - q.lam) * q.eps;
        o.du[0] = uLJe + (R(1) - q.lam);
        o.du[1] = duLJ_ds;
        o.du[2] = -q.eps;
        o.df[0] = fLJe;
        o.df[1] = dfLJ_ds;
        o.df[2] = R(0);
        }
    else
        {
        o.f = q.lam * q.eps * fLJe;
        o.fp = q.lam * q.eps * fpLJe;
        o.u = q.lam * q.eps * uLJe;
        o.du[0] = q.lam * uLJe;
        o.du[1] = q.lam * duLJ_ds;
        o.du[2] = q.eps * uLJe;
        o.df[0] = q.lam * fLJe;
        o.df[1] = q.lam * dfLJ_ds;
        o.df[2] = q.eps * fLJe;
        }
    o.u -= q.u_c;
    for (int k = 0; k < kAdjNpar; ++k)
        o.du[k] -= q.du_c[k];
    }

struct YukawaAdj
    {
    using param_type = azp_param_expanded_yukawa;
    template<typename R> using Prep = AdjYukawaPrep<R>;
    template<typename R> __device__ static Prep<R> prepare(const param_type& p, double rcsq, bool shift)
        {
        return adj_yukawa_prepare<R>(p, rcsq, shift);
        }
    template<typename R> __device__ static bool active(const Prep<R>& q)
        {
        return q.eps != R(0);
        }
    template<typename R> __device__ static void eval(const Prep<R>& q, R rsq, AdjOut<R>& o)
        {
        adj_yukawa_eval<R>(q, rsq, o);
        }
    };

struct PLJAdj
    {
    using param_type = azp_param_perturbed_lj;
    template<typename R> using Prep = AdjPLJPrep<R>;
    template<typename R> __device__ static Prep<R> prepare(const param_type& p, double rcsq, bool shift)
        {
        return adj_plj_prepare<R>(p, rcsq, shift);
        }
    template<typename R> __device__ static bool active(const Prep<R>& q)
        {
        // the reference evaluator skips pairs whose lj1 = 4 eps sigma^12 is zero
        return q.eps * q.lj1e != R(0);
        }
    template<typename R> __device__ static void eval(const Prep<R>& q, R rsq, AdjOut<R>& o)
        {
        adj_plj_eval<R>(q, rsq, o);
        }
    };

struct AdjArgs
    {
    const double4* pos;
    const double4* grad; // (dL/dF_x, dL/dF_y, dL/dF_z, dL/dE) per particle
    uint32_t N;
    DevBox box;
    const uint32_t* n_neigh;
    const uint32_t* nlist;
    uint32_t nlist_pitch;
    const void* params;
    const double* rcsq;
    uint32_t ntypes;
    int shift;
    double4* grad_pos;   // dL/dr (x, y, z, 0)
    double* partial;     // [gridDim.x][ntypes^2 * kAdjNpar]
    };

template<class Adj, typename R, int NT>
__global__ void __launch_bounds__(256) k_adjoint(const AdjArgs a)
    {
    using Prep = typename Adj::template Prep<R>;
    __shared__ Prep s_q[NT * NT];
    __shared__ double s_rc[NT * NT];
    __shared__ double s_acc[256][NT * kAdjNpar];
    __shared__ uint32_t s_t[256];
    const auto* params = static_cast<const typename Adj::param_type*>(a.params);
    const uint32_t nt = a.ntypes;
    for (uint32_t t = threadIdx.x; t < nt * nt; t += blockDim.x)
        {
        s_q[t] = Adj::template prepare<R>(params[t], a.rcsq[t], a.shift != 0);
        s_rc[t] = a.rcsq[t];
        }
    __syncthreads();

    const uint32_t i = xcd_logical_block() * blockDim.x + threadIdx.x;
    double acc[NT * kAdjNpar];
#pragma unroll
    for (int k = 0; k < NT * kAdjNpar; ++k)
        acc[k] = 0.0;
    uint32_t ti = 0;
    if (i < a.N)
        {
        const double4 pi = a.pos[i];
        const double4 gi = a.grad[i];
        ti = (uint32_t)pi.w;
        double gx = 0, gy = 0, gz = 0;
        const uint32_t nn = a.n_neigh[i];
        for (uint32_t k = 0; k < nn; ++k)
            {
            const uint32_t j = a.nlist[(size_t)k * a.nlist_pitch + i];
            const double4 pj = a.pos[j];
            const double dx = min_image_1d(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
            const double dy = min_image_1d(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
            const double dz = min_image_1d(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
            const double rsq = dx * dx + dy * dy + dz * dz;
            const uint32_t tj = (uint32_t)pj.w;
            const uint32_t tp = ti * nt + tj;
            if (!(rsq < s_rc[tp]) || !Adj::active(s_q[tp]))
                continue;
            AdjOut<R> o;
            Adj::eval(s_q[tp], (R)rsq, o);
            const double4 gj = a.grad[j];
            const double f = o.f, fp = o.fp;
            const double dGx = gi.x - gj.x, dGy = gi.y - gj.y, dGz = gi.z - gj.z;
            const double dxdG = dx * dGx + dy * dGy + dz * dGz;
            const double hw = 0.5 * (gi.w + gj.w);
            gx += f * dGx + fp * dxdG * dx - hw * f * dx;
            gy += f * dGy + fp * dxdG * dy - hw * f * dy;
            gz += f * dGz + fp * dxdG * dz - hw * f * dz;
            const double gdx = gi.x * dx + gi.y * dy + gi.z * dz;
#pragma unroll
            for (int t = 0; t < NT; ++t)
                {
                const double sel = (tj == (uint32_t)t) ? 1.0 : 0.0;
#pragma unroll
                for (int p = 0; p < kAdjNpar; ++p)
                    acc[t * kAdjNpar + p] += sel * (gdx * (double)o.df[p] + 0.5 * gi.w * (double)o.du[p]);
                }
            }
        a.grad_pos[i] = make_double4(gx, gy, gz, 0.0);
        }
    // workgroup reduction of the parameter gradients, binned by (t_i, t_j)
#pragma unroll
    for (int k = 0; k < NT * kAdjNpar; ++k)
        s_acc[threadIdx.x][k] = acc[k];
    s_t[threadIdx.x] = (i < a.N) ? ti : 0xffffffffu;
    __syncthreads();
    const uint32_t nq = nt * nt * kAdjNpar;
    for (uint32_t q = threadIdx.x; q < nq; q += blockDim.x)
        {
        const uint32_t t = q / (nt * kAdjNpar);
        const uint32_t rem = q % (nt * kAdjNpar);
        double s = 0.0;
        for (uint32_t l = 0; l < blockDim.x; ++l)
            if (s_t[l] == t)
                s += s_acc[l][rem];
        a.partial[(size_t)blockIdx.x * nq + q] = s;
        }
    }

// Sum of the blocks' partial parameter gradients: one workgroup of 256
// threads per gradient component q.  Thread t adds partials t, t + 256, ...
// and the 256 sums are added in a fixed tree in LDS, so the result is the
// same on every run.
constexpr uint32_t kFinalThreads = 256;

__global__ void __launch_bounds__(kFinalThreads)
k_adjoint_final(const double* __restrict__ partial, uint32_t nblocks, uint32_t nq, double* __restrict__ out)
    {
    __shared__ double s_sum[kFinalThreads];
    const uint32_t q = blockIdx.x;
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += kFinalThreads)
        s += partial[(size_t)b * nq + q];
    s_sum[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = kFinalThreads / 2; w > 0; w >>= 1)
        {
        if (threadIdx.x < w)
            s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
        __syncthreads();
        }
    if (threadIdx.x == 0)
        out[q] = s_sum[0];
    }

template<class Adj, typename R>
static int launch_adj(const AdjArgs& ka, uint32_t nblocks, hipStream_t s)
    {
    if (ka.ntypes <= 1)
        hipLaunchKernelGGL((k_adjoint<Adj, R, 1>), dim3(nblocks), dim3(256), 0, s, ka);
    else if (ka.ntypes <= 2)
        hipLaunchKernelGGL((k_adjoint<Adj, R, 2>), dim3(nblocks), dim3(256), 0, s, ka);
    else
        hipLaunchKernelGGL((k_adjoint<Adj, R, 4>), dim3(nblocks), dim3(256), 0, s, ka);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

} // namespace azp

using namespace azp;

extern "C" size_t azp_adjoint_workspace_size(uint32_t N, uint32_t ntypes)
    {
    const size_t nblocks = (N + 255) / 256;
    return (nblocks > 0 ? nblocks : 1) * (size_t)ntypes * ntypes * kAdjNpar;
    }

extern "C" int azp_gpu_pair_adjoint(const azp_pair_args* args,
                                    int evaluator,
                                    const void* d_params,
                                    const double* d_grad,
                                    double* d_grad_pos,
                                    double* d_grad_params,
                                    double* d_work,
                                    size_t work_len)
    {
    if (!args)
        return set_error(AZP_ERR_INVALID, "null args");
    const azp_particles& p = args->particles;
    if (evaluator != AZP_EVAL_EXPANDED_YUKAWA && evaluator != AZP_EVAL_PERTURBED_LJ)
        return set_error(AZP_ERR_UNSUPPORTED, "the adjoint covers ExpandedYukawa and PerturbedLennardJones");
    if (args->ntypes == 0 || args->ntypes > (uint32_t)kAdjMaxTypes)
        return set_error(AZP_ERR_UNSUPPORTED, "the adjoint supports 1 to %d particle types", kAdjMaxTypes);
    if (args->shift_mode == AZP_SHIFT_XPLOR)
        return set_error(AZP_ERR_UNSUPPORTED, "the adjoint supports modes none and shift");
    if (!box_valid(p.box))
        return set_error(AZP_ERR_BOX, "invalid box");
    const uint32_t nq = args->ntypes * args->ntypes * kAdjNpar;
    if (p.N == 0)
        return AZP_OK;
    if (!p.d_pos || !d_params || !d_grad || !d_grad_pos || !d_grad_params || !d_work || !args->d_n_neigh
        || !args->d_nlist || !args->d_rcutsq)
        return set_error(AZP_ERR_INVALID, "null device pointer in adjoint arguments");
    const uint32_t nblocks = (p.N + 255) / 256;
    if (work_len < azp_adjoint_workspace_size(p.N, args->ntypes))
        return set_error(AZP_ERR_INVALID, "adjoint workspace needs %zu doubles", azp_adjoint_workspace_size(p.N, args->ntypes));
    AdjArgs ka;
    ka.pos = reinterpret_cast<const double4*>(p.d_pos);
    ka.grad = reinterpret_cast<const double4*>(d_grad);
    ka.N = p.N;
    ka.box = make_box(p.box);
    ka.n_neigh = args->d_n_neigh;
    ka.nlist = args->d_nlist;
    ka.nlist_pitch = args->nlist_pitch;
    ka.params = d_params;
    ka.rcsq = args->d_rcutsq;
    ka.ntypes = args->ntypes;
    ka.shift = args->shift_mode == AZP_SHIFT_SHIFT;
    ka.grad_pos = reinterpret_cast<double4*>(d_grad_pos);
    ka.partial = d_work;
    hipStream_t s = (hipStream_t)args->stream;
    if (args->precision == AZP_PREC_MIXED_ACC64)
        return set_error(AZP_ERR_UNSUPPORTED, "the adjoint has the FP64 and mixed forms");
    const bool mixed = args->precision == AZP_PREC_MIXED;
    int st;
    if (evaluator == AZP_EVAL_EXPANDED_YUKAWA)
        st = mixed ? launch_adj<YukawaAdj, float>(ka, nblocks, s) : launch_adj<YukawaAdj, double>(ka, nblocks, s);
    else
        st = mixed ? launch_adj<PLJAdj, float>(ka, nblocks, s) : launch_adj<PLJAdj, double>(ka, nblocks, s);
    if (st != AZP_OK)
        return st;
    hipLaunchKernelGGL(k_adjoint_final, dim3(nq), dim3(kFinalThreads), 0, s, d_work, nblocks, nq, d_grad_params);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }
