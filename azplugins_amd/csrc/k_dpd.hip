// DPD forces with the general weight function over a full neighbor list.
//
// Replaces gpu_compute_dpd_forces<DPDPairEvaluatorGeneralWeight>
// (PotentialPairDPDThermoGPUKernel.cu.inc:21-24).  The per-pair arithmetic
// restates DPDPairEvaluatorGeneralWeight.h:198-255 operation by operation.
// This translation unit is compiled with -ffp-contract=off and uses the
// portable pow (portable_math.h), so every pair's conservative, dissipative
// and random terms are bit-identical to the FP64 host oracle; only the order in
// which a particle's pair terms are summed differs (full list here, half list
// with Newton's third law on the host).
//
// The random variate is Philox4x32-10 keyed by (id 200, timestep, seed) with
// counter (min tag, max tag), so both directed evaluations of a pair draw the
// same alpha and the pair force stays antisymmetric.
#include "device_common.h"
#include "portable_math.h"

namespace azp
{

struct DPDKernelArgs
    {
    const double4* pos;
    const double4* vel;
    const uint32_t* tag;
    uint32_t N;
    DevBox box;
    const uint32_t* n_neigh;
    const uint32_t* nlist;
    uint32_t nlist_pitch;
    const azp_param_dpd_general_weight* params;
    const double* rcutsq;
    uint32_t ntypes;
    uint16_t seed;
    uint32_t timestep; // the evaluator stores the timestep as 32 bits
    double dt;
    double kT;
    double4* force;
    double* virial;
    uint32_t vpitch;
    };

struct DPDPair
    {
    double A, gamma, s, rcutsq;
    };

// DPDPairEvaluatorGeneralWeight::evalForceEnergyThermo (:198-255).
__device__ __forceinline__ bool dpd_eval(const DPDPair& q,
                                         double rsq,
                                         uint16_t seed,
                                         uint32_t ti,
                                         uint32_t tj,
                                         uint32_t timestep,
                                         double dt,
                                         double kT,
                                         double rdotv,
                                         double& fdivr,
                                         double& fdivr_cons,
                                         double& e)
    {
    if (!(rsq < q.rcutsq))
        return false;
    const double rinv = 1.0 / sqrt(rsq);
    const double r = 1.0 / rinv;
    const double rcutinv = 1.0 / sqrt(q.rcutsq);
    const double rcut = 1.0 / rcutinv;
    const double alpha = dpd_alpha(seed, timestep, ti, tj);
    fdivr = q.A * (rinv - rcutinv);
    fdivr_cons = fdivr;
    const double base = 1.0 - r * rcutinv;
    const double wR = azp_ppow(base, 0.5 * q.s) * rinv;
    fdivr -= q.gamma * wR * wR * rdotv;
    fdivr += (1.0 / sqrt(dt / (kT * q.gamma * 6.0))) * wR * alpha;
    e = q.A * (rcut - r) - 0.5 * q.A * rcutinv * (q.rcutsq - rsq);
    return true;
    }

template<bool VIRIAL>
__global__ void __launch_bounds__(512) k_dpd(const DPDKernelArgs a)
    {
    extern __shared__ double smem[];
    DPDPair* s_p = reinterpret_cast<DPDPair*>(smem);
    const uint32_t nt = a.ntypes;
    for (uint32_t t = threadIdx.x; t < nt * nt; t += blockDim.x)
        {
        s_p[t].A = a.params[t].A;
        s_p[t].gamma = a.params[t].gamma;
        s_p[t].s = a.params[t].s;
        s_p[t].rcutsq = a.rcutsq[t];
        }
    __syncthreads();
    const uint32_t i = xcd_logical_block() * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    const double4 pi = a.pos[i];
    const double4 vi = a.vel[i];
    const uint32_t tagi = a.tag[i];
    const uint32_t ti = (uint32_t)pi.w;
    double fx = 0, fy = 0, fz = 0, pe = 0;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    const uint32_t nn = a.n_neigh[i];
    for (uint32_t k = 0; k < nn; ++k)
        {
        const uint32_t j = a.nlist[(size_t)k * a.nlist_pitch + i];
        const double4 pj = a.pos[j];
        const double dx = min_image_1d(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
        const double dy = min_image_1d(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
        const double dz = min_image_1d(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
        const double rsq = dx * dx + dy * dy + dz * dz;
        const DPDPair q = s_p[ti * nt + (uint32_t)pj.w];
        if (!(rsq < q.rcutsq))
            continue;
        const double4 vj = a.vel[j];
        const double rdotv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
        double fdivr, fcons, e;
        dpd_eval(q, rsq, a.seed, tagi, a.tag[j], a.timestep, a.dt, a.kT, rdotv, fdivr, fcons, e);
        fx += dx * fdivr;
        fy += dy * fdivr;
        fz += dz * fdivr;
        pe += e * 0.5;
        if (VIRIAL)
            {
            const double fd2 = fcons * 0.5;
            v0 += fd2 * dx * dx;
            v1 += fd2 * dx * dy;
            v2 += fd2 * dx * dz;
            v3 += fd2 * dy * dy;
            v4 += fd2 * dy * dz;
            v5 += fd2 * dz * dz;
            }
        }
    a.force[i] = make_double4(fx, fy, fz, pe);
    if (VIRIAL)
        {
        const size_t p = a.vpitch;
        a.virial[0 * p + i] = v0;
        a.virial[1 * p + i] = v1;
        a.virial[2 * p + i] = v2;
        a.virial[3 * p + i] = v3;
        a.virial[4 * p + i] = v4;
        a.virial[5 * p + i] = v5;
        }
    }

__global__ void k_dpd_alpha(uint16_t seed,
                            uint32_t timestep,
                            uint64_t n,
                            const uint32_t* __restrict__ ti,
                            const uint32_t* __restrict__ tj,
                            double* __restrict__ out)
    {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    out[k] = dpd_alpha(seed, timestep, ti[k], tj[k]);
    }

int validate_common(const azp_pair_args* args);

} // namespace azp

using namespace azp;

extern "C" int azp_gpu_compute_dpd_forces(const azp_pair_args* args, const void* d_params)
    {
    AZP_TRY(validate_common(args));
    const azp_particles& p = args->particles;
    if (p.N == 0)
        return AZP_OK;
    if (!d_params || !p.d_vel || !p.d_tag)
        return set_error(AZP_ERR_INVALID, "DPD needs params, velocities and tags");
    if (args->precision != AZP_PREC_FP64)
        return set_error(AZP_ERR_UNSUPPORTED, "DPD runs in the plain FP64 form only");
    DPDKernelArgs ka;
    ka.pos = reinterpret_cast<const double4*>(p.d_pos);
    ka.vel = reinterpret_cast<const double4*>(p.d_vel);
    ka.tag = p.d_tag;
    ka.N = p.N;
    ka.box = make_box(p.box);
    ka.n_neigh = args->d_n_neigh;
    ka.nlist = args->d_nlist;
    ka.nlist_pitch = args->nlist_pitch;
    ka.params = static_cast<const azp_param_dpd_general_weight*>(d_params);
    ka.rcutsq = args->d_rcutsq;
    ka.ntypes = args->ntypes;
    ka.seed = args->seed;
    ka.timestep = (uint32_t)args->timestep;
    ka.dt = args->dt;
    ka.kT = args->kT;
    ka.force = reinterpret_cast<double4*>(args->out.d_force);
    ka.virial = args->out.d_virial;
    ka.vpitch = args->out.virial_pitch;
    const unsigned block = args->block_size ? args->block_size : 256;
    const size_t shmem = (size_t)args->ntypes * args->ntypes * sizeof(DPDPair);
    if (shmem > 64 * 1024)
        return set_error(AZP_ERR_UNSUPPORTED, "too many particle types for LDS staging");
    hipStream_t s = (hipStream_t)args->stream;
    if (args->compute_virial)
        hipLaunchKernelGGL(k_dpd<true>, dim3(grid_for(p.N, block)), dim3(block), shmem, s, ka);
    else
        hipLaunchKernelGGL(k_dpd<false>, dim3(grid_for(p.N, block)), dim3(block), shmem, s, ka);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

extern "C" int azp_gpu_dpd_alpha(uint16_t seed,
                                 uint64_t timestep,
                                 uint64_t n,
                                 const uint32_t* d_tag_i,
                                 const uint32_t* d_tag_j,
                                 double* d_alpha,
                                 void* stream)
    {
    if (n == 0)
        return AZP_OK;
    if (!d_tag_i || !d_tag_j || !d_alpha)
        return set_error(AZP_ERR_INVALID, "null pointer");
    const unsigned block = 256;
    const uint64_t grid = (n + block - 1) / block;
    if (grid > 0x7fffffffull)
        return set_error(AZP_ERR_INVALID, "too many variates in one call");
    hipLaunchKernelGGL(k_dpd_alpha,
                       dim3((unsigned)grid),
                       dim3(block),
                       0,
                       (hipStream_t)stream,
                       seed,
                       (uint32_t)timestep,
                       n,
                       d_tag_i,
                       d_tag_j,
                       d_alpha);
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }
