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
    const uint64_t* head_list; // HOOMD's layout when non-null
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
    uint32_t queue_len; // in-cut neighbor indices per lane held in LDS
    };

// One type pair's parameters with the quantities that depend on the pair
// type alone (1/r_cut, r_cut, s/2, the random prefactor) evaluated once per
// launch, by the same operations the evaluator applies per pair.
struct DPDPair
    {
    double A, gamma, half_s, rcutsq, rcutinv, rcut, rand_pre;
    };

// DPDPairEvaluatorGeneralWeight::evalForceEnergyThermo (:198-255) for a pair
// already known to be inside r_cut.
__device__ __forceinline__ void dpd_eval(const DPDPair& q,
                                         double rsq,
                                         uint16_t seed,
                                         uint32_t ti,
                                         uint32_t tj,
                                         uint32_t timestep,
                                         double rdotv,
                                         double& fdivr,
                                         double& fdivr_cons,
                                         double& e)
    {
    const double rinv = 1.0 / sqrt(rsq);
    const double r = 1.0 / rinv;
    const double alpha = dpd_alpha(seed, timestep, ti, tj);
    fdivr = q.A * (rinv - q.rcutinv);
    fdivr_cons = fdivr;
    const double base = 1.0 - r * q.rcutinv;
    const double wR = azp_ppow(base, q.half_s) * rinv;
    fdivr -= q.gamma * wR * wR * rdotv;
    fdivr += q.rand_pre * wR * alpha;
    e = q.A * (q.rcut - r) - 0.5 * q.A * q.rcutinv * (q.rcutsq - rsq);
    }

// One thread per particle over its full-list row.  Only about a third of
// the entries lie inside r_cut (the list radius is r_cut + r_buff), while a
// wave would pay the full evaluation on every entry that any of its lanes
// has in cut.  So the kernel alternates two phases.  A scan reads the row
// four entries at a time (four position gathers in flight per lane) and
// appends the indices of the entries inside r_cut to the lane's queue in LDS
// (column tid of a [queue_len][blockDim] array).  An evaluation pass then
// takes the queued neighbors in order; the wave evaluates as many pairs as
// the lane with the most in-cut entries has, not as many as its longest
// row.  Each particle still sums its pair terms in row order.
__device__ __forceinline__ DPDPair make_dpd_pair(const DPDKernelArgs& a, uint32_t t)
    {
    DPDPair q;
    q.A = a.params[t].A;
    q.gamma = a.params[t].gamma;
    q.half_s = 0.5 * a.params[t].s;
    q.rcutsq = a.rcutsq[t];
    q.rcutinv = 1.0 / sqrt(q.rcutsq);
    q.rcut = 1.0 / q.rcutinv;
    q.rand_pre = 1.0 / sqrt(a.dt / (a.kT * q.gamma * 6.0));
    return q;
    }

// LDS_TABLE: the type-pair table is staged in LDS (up to 16 KiB, 17 types);
// otherwise each in-cut pair derives its entry from the global parameter
// arrays by the same operations (bit-identical, slower).
template<bool VIRIAL, bool LDS_TABLE>
__global__ void __launch_bounds__(512) k_dpd(const DPDKernelArgs a)
    {
    extern __shared__ double smem[];
    DPDPair* s_p = reinterpret_cast<DPDPair*>(smem);
    const uint32_t nt = a.ntypes;
    uint32_t* s_q = reinterpret_cast<uint32_t*>(LDS_TABLE ? reinterpret_cast<char*>(s_p + nt * nt)
                                                          : reinterpret_cast<char*>(smem));
    if (LDS_TABLE)
        {
        for (uint32_t t = threadIdx.x; t < nt * nt; t += blockDim.x)
            s_p[t] = make_dpd_pair(a, t);
        }
    __syncthreads();
    const uint32_t i = xcd_logical_block() * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    const uint32_t bs = blockDim.x;
    uint32_t* q_col = s_q + threadIdx.x;
    const uint32_t qlen = a.queue_len;
    const double4 pi = a.pos[i];
    const double4 vi = a.vel[i];
    const uint32_t tagi = a.tag[i];
    const uint32_t ti = (uint32_t)pi.w;
    const double* rc_row = LDS_TABLE ? &s_p[ti * nt].rcutsq : a.rcutsq + ti * nt;
    constexpr uint32_t rc_stride = LDS_TABLE ? sizeof(DPDPair) / sizeof(double) : 1u;
    const uint32_t* row = a.head_list ? a.nlist + a.head_list[i] : a.nlist + i;
    const size_t pitch = a.head_list ? 1 : a.nlist_pitch;
    double fx = 0, fy = 0, fz = 0, pe = 0;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    const uint32_t nn = a.n_neigh[i];
    auto in_cut = [&](const double4& pj) {
        const double dx = min_image_1d(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
        const double dy = min_image_1d(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
        const double dz = min_image_1d(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
        const double rsq = dx * dx + dy * dy + dz * dz;
        return rsq < rc_row[(uint32_t)pj.w * rc_stride];
    };
    uint32_t k = 0;
    while (k < nn)
        {
        // scan: fill the queue with the next in-cut neighbors
        uint32_t cnt = 0;
        while (k + 4 <= nn && cnt + 4 <= qlen)
            {
            const uint32_t j0 = row[(size_t)k * pitch];
            const uint32_t j1 = row[(size_t)(k + 1) * pitch];
            const uint32_t j2 = row[(size_t)(k + 2) * pitch];
            const uint32_t j3 = row[(size_t)(k + 3) * pitch];
            const double4 p0 = a.pos[j0];
            const double4 p1 = a.pos[j1];
            const double4 p2 = a.pos[j2];
            const double4 p3 = a.pos[j3];
            const bool c0 = in_cut(p0), c1 = in_cut(p1), c2 = in_cut(p2), c3 = in_cut(p3);
            if (c0)
                q_col[(cnt++) * bs] = j0;
            if (c1)
                q_col[(cnt++) * bs] = j1;
            if (c2)
                q_col[(cnt++) * bs] = j2;
            if (c3)
                q_col[(cnt++) * bs] = j3;
            k += 4;
            }
        while (k < nn && cnt < qlen)
            {
            const uint32_t j = row[(size_t)k * pitch];
            ++k;
            if (in_cut(a.pos[j]))
                q_col[(cnt++) * bs] = j;
            }
        // evaluate the queued pairs in row order
        for (uint32_t c = 0; c < cnt; ++c)
            {
            const uint32_t j = q_col[c * bs];
            const double4 pj = a.pos[j];
            const double4 vj = a.vel[j];
            const uint32_t tagj = a.tag[j];
            const double dx = min_image_1d(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
            const double dy = min_image_1d(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
            const double dz = min_image_1d(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
            const double rsq = dx * dx + dy * dy + dz * dz;
            const DPDPair q = LDS_TABLE ? s_p[ti * nt + (uint32_t)pj.w] : make_dpd_pair(a, ti * nt + (uint32_t)pj.w);
            const double rdotv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
            double fdivr, fcons, e;
            dpd_eval(q, rsq, a.seed, tagi, tagj, a.timestep, rdotv, fdivr, fcons, e);
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
    ka.head_list = args->d_head_list;
    if (!args->d_nlist)
        return set_error(AZP_ERR_INVALID, "the DPD kernel reads the plain list (d_nlist)");
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
    // type-pair table in LDS up to 16 KiB (17 types); beyond that it is
    // derived per pair from the global arrays
    const size_t table_bytes = (size_t)args->ntypes * args->ntypes * sizeof(DPDPair);
    const bool lds_table = table_bytes <= 16 * 1024;
    const size_t table = lds_table ? table_bytes : 0;
    // queue of 32 indices per lane, fewer for blocks above 256 threads
    // (queue and table stay within 64 KiB of LDS)
    ka.queue_len = std::min<uint32_t>(32u, (uint32_t)((48 * 1024) / (block * sizeof(uint32_t))));
    if (ka.queue_len < 4)
        return set_error(AZP_ERR_INVALID, "DPD block size too large");
    const size_t shmem = table + (size_t)ka.queue_len * block * sizeof(uint32_t);
    hipStream_t s = (hipStream_t)args->stream;
    const dim3 grid(grid_for(p.N, block));
    if (args->compute_virial)
        {
        if (lds_table)
            hipLaunchKernelGGL((k_dpd<true, true>), grid, dim3(block), shmem, s, ka);
        else
            hipLaunchKernelGGL((k_dpd<true, false>), grid, dim3(block), shmem, s, ka);
        }
    else
        {
        if (lds_table)
            hipLaunchKernelGGL((k_dpd<false, true>), grid, dim3(block), shmem, s, ka);
        else
            hipLaunchKernelGGL((k_dpd<false, false>), grid, dim3(block), shmem, s, ka);
        }
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
