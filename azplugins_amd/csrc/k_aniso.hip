// Anisotropic pair forces and torques (TwoPatchMorse) over a full neighbor list.
//
// Replaces gpu_compute_pair_aniso_forces<AnisoPairEvaluatorTwoPatchMorse>
// (AnisoPotentialPairGPUKernel.cu.inc:21-25).  The evaluator restates
// AnisoPairEvaluatorTwoPatchMorse.h:127-216.  With a full list each lane
// evaluates (i, j) with dr = r_i - r_j and keeps force and torque_i; particle
// j's torque comes from its own evaluation of (j, i), which equals torque_j of
// (i, j) exactly in exact arithmetic (both are dU/dgamma_j (r_hat x n_j)).
#include <type_traits>

#include "tpm.h"

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
    const uint64_t* head_list; // HOOMD's layout when non-null
    const azp_param_two_patch_morse* params;
    const double* rcutsq;
    uint32_t ntypes;
    int shift_mode;
    double4* force;
    double4* torque;
    double* virial;
    uint32_t vpitch;
    };

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
    const uint32_t lb = xcd_logical_block();
    const uint32_t i = lb * blockDim.x + threadIdx.x;
    // wavefronts entirely past N leave together; lanes past N in the last
    // wavefront walk no entries and store nothing
    if (lb * blockDim.x + (threadIdx.x & ~63u) >= a.N)
        return;
    const uint32_t ie = min(i, a.N - 1);
    const AnisoSelf<R> me = aniso_self<MIXED>(a.pos, a.pos_fixed, a.orientation, ie);
    LaneBox box;
    for (int k = 0; k < 3; ++k)
        {
        box.L[k] = a.box.L[k];
        box.Linv[k] = a.box.Linv[k];
        box.fx_scale[k] = a.fx_scale[k];
        }
    AnisoAcc acc;
    const uint32_t nn = i < a.N ? a.n_neigh[i] : 0u;
    const uint32_t trips = __builtin_amdgcn_readfirstlane((tpm_wave_max(nn) + 3u) & ~3u);
    const uint32_t* row = a.head_list ? a.nlist + a.head_list[ie] : a.nlist + ie;
    const size_t stride = a.head_list ? 1 : a.nlist_pitch;
    tpm_list<MIXED, VIRIAL, SINGLE, false>(ie, nn, trips, row, stride, a.pos, a.pos_fixed, a.orientation,
                                           nullptr, me, box, nt, q0, s_q, acc);
    if (i < a.N)
        aniso_store<VIRIAL>(a.force, a.torque, a.virial, a.vpitch, i, acc);
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

__global__ void k_pack_fixed_aniso(const double4* __restrict__ pos, uint32_t N, DevBox box, int4* __restrict__ out)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const double4 p = pos[i];
    out[i] = make_int4(to_fixed(p.x, box.Linv[0]), to_fixed(p.y, box.Linv[1]), to_fixed(p.z, box.Linv[2]), (int)p.w);
    }

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
    if (args->precision == AZP_PREC_MIXED_ACC64)
        return set_error(AZP_ERR_UNSUPPORTED, "the anisotropic kernel has the FP64 and mixed forms");
    const bool mixed = args->precision == AZP_PREC_MIXED;
    if (!args->d_nlist)
        return set_error(AZP_ERR_INVALID, "the anisotropic kernel reads the plain list (d_nlist)");
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
    ka.head_list = args->d_head_list;
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
    if (mixed && !ka.pos_fixed)
        {
        // fixed-point positions packed into the launchers' scratch
        int4* pf = nullptr;
        AZP_TRY(launcher_heap_scratch(p.N, &pf));
        hipLaunchKernelGGL(k_pack_fixed_aniso, dim3(grid_for(p.N, 256)), dim3(256), 0, s, ka.pos, p.N, ka.box, pf);
        AZP_LAUNCH_CHECK();
        ka.pos_fixed = pf;
        }
    return mixed ? launch_aniso<true>(ka, args->compute_virial != 0, block, s)
                 : launch_aniso<false>(ka, args->compute_virial != 0, block, s);
    }
