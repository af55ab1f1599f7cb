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
#include "tiles.h"

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
    // tiled list (k_dpd_tiled): per tile the neighbor set and its count,
    // 16-bit entries two per word; columns of cell-list slots when tile_order
    // is set
    const uint32_t* tile_count;
    const uint32_t* tile_jset;
    const uint32_t* tile_list;
    const uint32_t* tile_order;
    const uint32_t* tile_nn;
    uint32_t tile_slots;
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
    // f_rcp is the correctly rounded 1/x: the same bits as the divisions
    const double rinv = f_rcp(sqrt(rsq));
    const double r = f_rcp(rinv);
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
// At least four wavefronts per SIMD (at most 128 VGPRs, a few spilled): the
// kernel waits on its gathers more than it issues, and the fourth wavefront
// measured 4 % faster than three without the bound.
template<bool VIRIAL, bool LDS_TABLE>
__global__ void __launch_bounds__(512, 4) k_dpd(const DPDKernelArgs a)
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
        const double dx = min_image_fused(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
        const double dy = min_image_fused(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
        const double dz = min_image_fused(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
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
            uint32_t j0, j1, j2, j3;
            if (pitch == 1)
                {
                // consecutive entries (HOOMD's head-list rows): one 16-byte load
                uint4 w;
                __builtin_memcpy(&w, row + k, sizeof(w));
                j0 = w.x;
                j1 = w.y;
                j2 = w.z;
                j3 = w.w;
                }
            else
                {
                j0 = row[(size_t)k * pitch];
                j1 = row[(size_t)(k + 1) * pitch];
                j2 = row[(size_t)(k + 2) * pitch];
                j3 = row[(size_t)(k + 3) * pitch];
                }
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
            const double dx = min_image_fused(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
            const double dy = min_image_fused(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
            const double dz = min_image_fused(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
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

// DPD over the tiled list (tiles.h): one workgroup per tile stages the FP64
// positions of the tile's neighbor set in LDS (32 bytes a neighbor, the
// fourth word holding the type and the particle index), so the scan over
// every list entry reads its neighbors from LDS instead of gathering them
// from global memory; the velocities and tags of the in-cut neighbors (about
// a third of the entries) are still gathered.  Staging the
// velocities and tags too (64 bytes a neighbor) left room for one workgroup
// per CU and measured slower.  A list entry is 16 x the neighbor's rank in
// the set.  The scan, the in-cut queue (16-bit ranks) and the evaluation are
// those of k_dpd, entry for entry: the tiled columns hold the plain list's
// rows in the same order, so the results are bit-identical to k_dpd over the
// plain list.
// Four wavefronts per SIMD, i.e. two 512-lane workgroups per CU (their LDS
// fits twice), so one workgroup's staging overlaps the other's work; with
// three (the unbounded register count) it ran slower than the plain kernel.
template<bool VIRIAL>
__global__ void __launch_bounds__(512, 4) k_dpd_tiled(const DPDKernelArgs a)
    {
    extern __shared__ double smem[];
    const uint32_t nt = a.ntypes;
    const uint32_t slots = a.tile_slots;
    double4* s_pos = reinterpret_cast<double4*>(smem);
    DPDPair* s_p = reinterpret_cast<DPDPair*>(s_pos + slots);
    uint16_t* s_q = reinterpret_cast<uint16_t*>(s_p + nt * nt);
    const uint32_t lb = xcd_logical_block();
    const uint32_t bs = blockDim.x;
    const uint32_t* js = a.tile_jset + (size_t)lb * kTileCap;
    // slot m: the neighbor's x, y, z and, in the fourth word, its type (low
    // half) and particle index (high half), so the evaluation gathers its
    // velocity and tag without reading the neighbor set again
    stage_slots<8>(
        a.tile_count[lb], js,
        [&](uint32_t j) {
            double4 p = a.pos[j];
            p.w = __hiloint2double((int)j, (int)(uint32_t)p.w);
            return p;
        },
        [&](uint32_t m, const double4& p) { s_pos[m] = p; });
    for (uint32_t t = threadIdx.x; t < nt * nt; t += bs)
        s_p[t] = make_dpd_pair(a, t);
    __syncthreads();
    const uint32_t ie = lb * bs + threadIdx.x;
    if (ie >= a.N)
        return;
    const uint32_t pe = a.tile_order ? a.tile_order[ie] : ie;
    const uint32_t nn = a.tile_order ? a.tile_nn[ie] : a.n_neigh[ie];
    uint16_t* q_col = s_q + threadIdx.x;
    const uint32_t qlen = a.queue_len;
    const double4 pi = a.pos[pe];
    const double4 vi = a.vel[pe];
    const uint32_t tagi = a.tag[pe];
    const uint32_t ti = (uint32_t)pi.w;
    const double* rc_row = &s_p[ti * nt].rcutsq;
    constexpr uint32_t rc_stride = sizeof(DPDPair) / sizeof(double);
    const uint32_t* col = a.tile_list + ie;
    const size_t pitch = a.nlist_pitch;
    double fx = 0, fy = 0, fz = 0, pe_sum = 0;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    auto in_cut = [&](const double4& pj) {
        const double dx = min_image_fused(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
        const double dy = min_image_fused(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
        const double dz = min_image_fused(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
        const double rsq = dx * dx + dy * dy + dz * dz;
        return rsq < rc_row[(uint32_t)__double2loint(pj.w) * rc_stride];
    };
    auto rank_at = [&](uint32_t kk) {
        const uint32_t w = col[(size_t)(kk >> 1) * pitch];
        return ((kk & 1u) ? (w >> 16) : (w & 0xffffu)) >> 4;
    };
    uint32_t k = 0;
    while (k < nn)
        {
        // scan: fill the queue with the next in-cut neighbors.  A full queue
        // can stop the scan at an odd entry; one single step then realigns k
        // to a word, so the four-entry steps read entries k .. k + 3 as two
        // whole words.
        uint32_t cnt = 0;
        if (k & 1u)
            {
            const uint32_t r = rank_at(k);
            ++k;
            if (in_cut(s_pos[r]))
                q_col[(cnt++) * bs] = (uint16_t)r;
            }
        while (k + 4 <= nn && cnt + 4 <= qlen)
            {
            const uint32_t w0 = col[(size_t)(k >> 1) * pitch];
            const uint32_t w1 = col[(size_t)((k >> 1) + 1) * pitch];
            const uint32_t r0 = (w0 & 0xffffu) >> 4, r1 = w0 >> 20;
            const uint32_t r2 = (w1 & 0xffffu) >> 4, r3 = w1 >> 20;
            const bool c0 = in_cut(s_pos[r0]), c1 = in_cut(s_pos[r1]);
            const bool c2 = in_cut(s_pos[r2]), c3 = in_cut(s_pos[r3]);
            if (c0)
                q_col[(cnt++) * bs] = (uint16_t)r0;
            if (c1)
                q_col[(cnt++) * bs] = (uint16_t)r1;
            if (c2)
                q_col[(cnt++) * bs] = (uint16_t)r2;
            if (c3)
                q_col[(cnt++) * bs] = (uint16_t)r3;
            k += 4;
            }
        while (k < nn && cnt < qlen)
            {
            const uint32_t r = rank_at(k);
            ++k;
            if (in_cut(s_pos[r]))
                q_col[(cnt++) * bs] = (uint16_t)r;
            }
        // evaluate the queued pairs in row order
        for (uint32_t c = 0; c < cnt; ++c)
            {
            const uint32_t r = q_col[c * bs];
            const double4 pj = s_pos[r];
            const uint32_t j = (uint32_t)__double2hiint(pj.w);
            const double4 vj = a.vel[j];
            const uint32_t tagj = a.tag[j];
            const double dx = min_image_fused(pi.x - pj.x, a.box.L[0], a.box.Linv[0]);
            const double dy = min_image_fused(pi.y - pj.y, a.box.L[1], a.box.Linv[1]);
            const double dz = min_image_fused(pi.z - pj.z, a.box.L[2], a.box.Linv[2]);
            const double rsq = dx * dx + dy * dy + dz * dz;
            const DPDPair q = s_p[ti * nt + (uint32_t)__double2loint(pj.w)];
            const double rdotv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
            double fdivr, fcons, e;
            dpd_eval(q, rsq, a.seed, tagi, tagj, a.timestep, rdotv, fdivr, fcons, e);
            fx += dx * fdivr;
            fy += dy * fdivr;
            fz += dz * fdivr;
            pe_sum += e * 0.5;
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
    a.force[pe] = make_double4(fx, fy, fz, pe_sum);
    if (VIRIAL)
        {
        const size_t p = a.vpitch;
        a.virial[0 * p + pe] = v0;
        a.virial[1 * p + pe] = v1;
        a.virial[2 * p + pe] = v2;
        a.virial[3 * p + pe] = v3;
        a.virial[4 * p + pe] = v4;
        a.virial[5 * p + pe] = v5;
        }
    }

// LDS of the tiled kernel: the staged neighbors, the type-pair table and a
// queue of `qlen` 16-bit ranks per lane.
static size_t dpd_tile_lds(uint32_t ntypes, uint32_t slots, uint32_t tile_size, uint32_t qlen)
    {
    return (size_t)slots * sizeof(double4) + (size_t)ntypes * ntypes * sizeof(DPDPair)
           + (size_t)qlen * tile_size * sizeof(uint16_t);
    }

constexpr size_t kDpdTileLds = 160 * 1024; // a gfx950 workgroup's LDS
constexpr uint32_t kDpdTileQueue = 16;     // in-cut ranks per lane (12.6 on average at density 3)

bool dpd_tile_fits(uint32_t ntypes, uint32_t slots, uint32_t tile_size)
    {
    return tile_size != 0 && slots != 0 && slots <= kTileCap
           && dpd_tile_lds(ntypes, slots, tile_size, kDpdTileQueue) <= kDpdTileLds;
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
    DPDKernelArgs ka{};
    ka.pos = reinterpret_cast<const double4*>(p.d_pos);
    ka.vel = reinterpret_cast<const double4*>(p.d_vel);
    ka.tag = p.d_tag;
    ka.N = p.N;
    ka.box = make_box(p.box);
    ka.n_neigh = args->d_n_neigh;
    ka.nlist = args->d_nlist;
    ka.nlist_pitch = args->nlist_pitch;
    ka.head_list = args->d_head_list;
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
    hipStream_t s = (hipStream_t)args->stream;
    if (args->d_tile_list && !args->d_head_list && dpd_tile_fits(args->ntypes, args->tile_slots, args->tile_size))
        {
        if (!args->d_tile_count || !args->d_tile_jset || (!args->d_tile_order && !args->d_n_neigh)
            || (args->d_tile_order && !args->d_tile_nn) || (args->tile_size != 256 && args->tile_size != 512))
            return set_error(AZP_ERR_INVALID, "incomplete tiled list");
        ka.tile_count = args->d_tile_count;
        ka.tile_jset = args->d_tile_jset;
        ka.tile_list = args->d_tile_list;
        ka.tile_order = args->d_tile_order;
        ka.tile_nn = args->d_tile_nn;
        ka.tile_slots = args->tile_slots;
        ka.queue_len = kDpdTileQueue;
        const size_t shmem = dpd_tile_lds(args->ntypes, args->tile_slots, args->tile_size, kDpdTileQueue);
        if (shmem > 64 * 1024)
            {
            // more than 64 KiB of dynamic LDS (gfx950 gives a workgroup 160 KiB)
            static bool raised = false;
            if (!raised)
                {
                for (const void* f : {reinterpret_cast<const void*>(&k_dpd_tiled<true>),
                                      reinterpret_cast<const void*>(&k_dpd_tiled<false>)})
                    {
                    const hipError_t e
                        = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDpdTileLds);
                    if (e != hipSuccess)
                        return set_error(AZP_ERR_HIP, "hipFuncSetAttribute (dynamic LDS) failed: %s",
                                         hipGetErrorString(e));
                    }
                raised = true;
                }
            }
        const dim3 grid(grid_for(p.N, args->tile_size));
        if (args->compute_virial)
            hipLaunchKernelGGL((k_dpd_tiled<true>), grid, dim3(args->tile_size), shmem, s, ka);
        else
            hipLaunchKernelGGL((k_dpd_tiled<false>), grid, dim3(args->tile_size), shmem, s, ka);
        AZP_LAUNCH_CHECK();
        return AZP_OK;
        }
    if (!args->d_nlist)
        return set_error(AZP_ERR_INVALID, "the DPD kernel reads the plain list (d_nlist) when no tiled list fits");
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
