// Isotropic pair forces over a full neighbor list (gfx950).
//
// Replaces the launcher instantiated per evaluator at
// PotentialPairGPUKernel.cu.inc:25-28 (gpu_compute_pair_forces<Eval>).
//
// Layout: one lane per particle i, walking its full neighbor list
// nlist[k * pitch + i] — consecutive lanes read consecutive words, so the
// list stream is fully coalesced.  Per-type-pair constants (cutoff, hoisted
// energy shift, parameter products) live in LDS; with one particle type they
// are held in registers.  Force, energy and the six virial components are
// accumulated in FP64 registers and written once (no atomics; every lane owns
// its particle).
//
// Two precision forms:
//   plain (FP64):  FP64 positions, min image, evaluator and accumulation.
//   mixed:         positions as wrapped 32-bit fixed point (16-byte gather),
//                  dx exact in integers then FP32, FP32 evaluator; every 32
//                  list entries are summed in FP32 and added to FP64
//                  accumulators.  Pairs whose FP32 rsq lies within a relative
//                  band of 1e-5 around rcutsq are re-tested on the FP64
//                  positions, so the in/out cutoff decision is the FP64 one.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "pair_kernel.h"

namespace azp
{

__global__ void k_pack_fixed(const double4* __restrict__ pos, uint32_t N, DevBox box, int4* __restrict__ out)
    {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    out[i] = fixed4(pos[i], box);
    }


int pack_fixed(const double* d_pos, uint32_t N, const azp_box& box, int32_t* d_out, hipStream_t stream)
    {
    if (N == 0)
        return AZP_OK;
    const unsigned block = 256;
    hipLaunchKernelGGL(k_pack_fixed,
                       dim3(grid_for(N, block)),
                       dim3(block),
                       0,
                       stream,
                       reinterpret_cast<const double4*>(d_pos),
                       N,
                       make_box(box),
                       reinterpret_cast<int4*>(d_out));
    AZP_LAUNCH_CHECK();
    return AZP_OK;
    }

// Whether the kernel of `evaluator` in the given form can run over a tiled
// list with `tile_slots` neighbor slots for `ntypes` types (its positions and
// type tables fit in LDS).
bool pair_tile_fits(int evaluator, uint32_t ntypes, uint32_t tile_slots, bool mixed)
    {
    switch (evaluator)
        {
    case AZP_EVAL_HERTZ:
        return mixed ? tile_fits<EvalHertz>(ntypes, tile_slots) : tile_fits64<EvalHertz>(ntypes, tile_slots);
    case AZP_EVAL_EXPANDED_YUKAWA:
        return mixed ? tile_fits<EvalExpandedYukawa>(ntypes, tile_slots)
                     : tile_fits64<EvalExpandedYukawa>(ntypes, tile_slots);
    case AZP_EVAL_PERTURBED_LJ:
        return mixed ? tile_fits<EvalPerturbedLJ>(ntypes, tile_slots)
                     : tile_fits64<EvalPerturbedLJ>(ntypes, tile_slots);
    case AZP_EVAL_COLLOID:
        return !mixed && tile_fits64<EvalColloid>(ntypes, tile_slots);
    case AZP_EVAL_DPD_GENERAL_WEIGHT:
        return mixed ? tile_fits<EvalDPDConservative>(ntypes, tile_slots)
                     : tile_fits64<EvalDPDConservative>(ntypes, tile_slots);
    default:
        return false;
        }
    }

int validate_common(const azp_pair_args* args)
    {
    if (!args)
        return set_error(AZP_ERR_INVALID, "null args");
    const azp_particles& p = args->particles;
    if (p.N == 0)
        return AZP_OK;
    if (!p.d_pos || !args->out.d_force || !args->d_n_neigh || !args->d_rcutsq)
        return set_error(AZP_ERR_INVALID, "missing device array in pair args");
    if (!args->d_nlist && !args->d_tile_list)
        return set_error(AZP_ERR_INVALID, "missing neighbor list in pair args");
    if (!box_valid(p.box))
        return set_error(AZP_ERR_BOX, "box lengths must be positive");
    if (args->ntypes == 0)
        return set_error(AZP_ERR_INVALID, "ntypes must be >= 1");
    if (!args->d_head_list && args->nlist_pitch < p.N)
        return set_error(AZP_ERR_INVALID, "nlist pitch %u < N %u", args->nlist_pitch, p.N);
    if (args->compute_virial && (!args->out.d_virial || args->out.virial_pitch < p.N))
        return set_error(AZP_ERR_INVALID, "virial requested without a virial array of pitch >= N");
    if (args->block_size != 0
        && (args->block_size % 64 != 0 || args->block_size > 512))
        return set_error(AZP_ERR_INVALID, "block size %u must be a multiple of 64, <= 512", args->block_size);
    return AZP_OK;
    }

} // namespace azp

using namespace azp;

extern "C" int azp_gpu_pack_fixed(const double* d_pos,
                                  uint32_t N,
                                  const azp_box* box,
                                  int32_t* d_pos_fixed,
                                  void* stream)
    {
    if (!box || !box_valid(*box))
        return set_error(AZP_ERR_BOX, "invalid box");
    if (N > 0 && (!d_pos || !d_pos_fixed))
        return set_error(AZP_ERR_INVALID, "null pointer");
    return pack_fixed(d_pos, N, *box, d_pos_fixed, (hipStream_t)stream);
    }

namespace azp
{

// Library-owned scratch of the launcher: the fixed-point positions of the
// mixed form over the plain list when the caller passes no d_pos_fixed.  One
// buffer per device, grown on demand (after a device synchronization) and
// kept until the process ends.  Launches that overlap on one device share it.
struct LauncherScratch
    {
    int4* ptr = nullptr;
    size_t n = 0;
    };
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static LauncherScratch g_scratch_heap[kScratchDevices];

int launcher_heap_scratch(uint32_t N, int4** out)
    {
    int dev = 0;
    AZP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return set_error(AZP_ERR_INVALID, "device ordinal %d is outside the launcher's table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    LauncherScratch& b = g_scratch_heap[dev];
    if (b.n < N)
        {
        if (b.ptr)
            {
            AZP_HIP(hipDeviceSynchronize());
            AZP_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.n = 0;
            }
        const size_t want = std::max<size_t>((size_t)N + N / 8, 4096);
        AZP_HIP(hipMalloc((void**)&b.ptr, want * sizeof(int4)));
        b.n = want;
        }
    *out = b.ptr;
    return AZP_OK;
    }

// Scratch of azp_gpu_compute_pair_forces_noalloc: a file-scope device array
// of kNoAllocMaxN int4 (16 MiB on every device the library runs on, held
// until the process ends).  In host code `scratch` names a host-side
// stand-in, not device memory: its device address comes from
// hipGetSymbolAddress, once per device with that device current, and only
// that address is passed to kernels.  Calls that overlap on one device share
// the array.  Nothing in it has to start at zero: the pack kernel writes the
// N entries that the pair kernel then reads.
constexpr uint32_t kNoAllocMaxN = 1u << 20;
static __device__ __align__(256) int4 scratch[kNoAllocMaxN];
static int4* g_scratch_static[kScratchDevices];

static int noalloc_scratch(uint32_t N, int4** out)
    {
    if (N > kNoAllocMaxN)
        return set_error(AZP_ERR_INVALID,
                         "the allocation-free launcher packs at most %u particles for the mixed form over the "
                         "plain list (N = %u): pass d_pos_fixed, a tiled list, or use azp_gpu_compute_pair_forces",
                         kNoAllocMaxN, N);
    int dev = 0;
    AZP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return set_error(AZP_ERR_INVALID, "device ordinal %d is outside the launcher's table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_static[dev])
        {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(AZP_ERR_HIP, "hipGetSymbolAddress of the launcher scratch failed: %s",
                             hipGetErrorString(e));
        g_scratch_static[dev] = static_cast<int4*>(ptr);
        }
    *out = g_scratch_static[dev];
    return AZP_OK;
    }

// Where the launcher puts fixed-point positions it has to pack itself.
enum class PackTarget
    {
    Heap,   // library-owned device buffer, grown with hipMalloc
    Static, // the file-scope array above (no allocation)
    };

static int compute_pair_forces(const azp_pair_args* args, int evaluator, const void* d_params, PackTarget target)
    {
    AZP_TRY(validate_common(args));
    const azp_particles& p = args->particles;
    if (p.N == 0)
        return AZP_OK;
    if (!d_params)
        return set_error(AZP_ERR_INVALID, "null params");
    if (args->precision != AZP_PREC_FP64 && args->precision != AZP_PREC_MIXED
        && args->precision != AZP_PREC_MIXED_ACC64)
        return set_error(AZP_ERR_INVALID, "unknown precision %d", args->precision);
    const bool mixed = args->precision != AZP_PREC_FP64;
    if (args->shift_mode == AZP_SHIFT_XPLOR && !args->d_ronsq)
        return set_error(AZP_ERR_INVALID, "xplor needs d_ronsq");

    PairKernelArgs ka;
    ka.pos = reinterpret_cast<const double4*>(p.d_pos);
    ka.pos_fixed = reinterpret_cast<const int4*>(p.d_pos_fixed);
    ka.N = p.N;
    ka.box = make_box(p.box);
    for (int k = 0; k < 3; ++k)
        {
        ka.fx_scale64[k] = ka.box.L[k] / 4294967296.0;
        ka.fx_scale[k] = (float)ka.fx_scale64[k];
        ka.fx_scale2[k] = (float)(ka.fx_scale64[k] * ka.fx_scale64[k]);
        }
    ka.n_neigh = args->d_n_neigh;
    ka.nlist = args->d_nlist;
    ka.nlist_pitch = args->nlist_pitch;
    ka.params = d_params;
    ka.rcutsq = args->d_rcutsq;
    ka.ronsq = args->d_ronsq;
    ka.ntypes = args->ntypes;
    ka.shift_mode = args->shift_mode;
    ka.force = reinterpret_cast<double4*>(args->out.d_force);
    ka.virial = args->out.d_virial;
    ka.vpitch = args->out.virial_pitch;
    ka.tile_count = args->d_tile_count;
    ka.tile_jset = args->d_tile_jset;
    ka.tile_list = args->d_tile_list;
    ka.tile_order = args->d_tile_list ? args->d_tile_order : nullptr;
    ka.tile_nn = args->d_tile_list ? args->d_tile_nn : nullptr;
    ka.head_list = args->d_head_list;
    ka.acc64 = args->precision == AZP_PREC_MIXED_ACC64;
    if (args->d_head_list && (args->d_tile_list || !args->d_nlist))
        return set_error(AZP_ERR_INVALID, "HOOMD's head list goes with d_nlist and without a tiled list");
    unsigned tile = 0, slots = 0;
    if (args->d_tile_list)
        {
        if (!args->d_tile_count || !args->d_tile_jset)
            return set_error(AZP_ERR_INVALID, "a tiled list needs its counts, neighbor sets and entries");
        if ((args->d_tile_order == nullptr) != (args->d_tile_nn == nullptr))
            return set_error(AZP_ERR_INVALID, "slot-ordered tiles need both d_tile_order and d_tile_nn");
        if (args->tile_size != 256 && args->tile_size != 512)
            return set_error(AZP_ERR_INVALID, "tile size %u must be 256 or 512", args->tile_size);
        if (args->tile_slots == 0 || args->tile_slots > kTileCap)
            return set_error(AZP_ERR_INVALID, "tile slots %u must be in [1, %u]", args->tile_slots, kTileCap);
        tile = args->tile_size;
        slots = args->tile_slots;
        }
    // 512-lane workgroups: measured fastest for both forms on the 1M-particle
    // PerturbedLJ workload (block sweep 64..512)
    const unsigned block = args->block_size ? args->block_size : 512;
    if (block < 64 || block > 512 || block % 64 != 0)
        return set_error(AZP_ERR_INVALID, "block size %u must be a multiple of 64 in [64, 512]", block);
    const bool vir = args->compute_virial != 0;
    hipStream_t s = (hipStream_t)args->stream;
    // the mixed form over the plain list reads fixed-point positions: packed
    // here into the launcher's scratch when the caller brings none
    const bool tiled_ok = tile != 0 && pair_tile_fits(evaluator, args->ntypes, slots, mixed);
    if (mixed && !p.d_pos_fixed && !tiled_ok && args->d_nlist && evaluator != AZP_EVAL_COLLOID)
        {
        int4* pf = nullptr;
        switch (target)
            {
        case PackTarget::Heap:
            AZP_TRY(launcher_heap_scratch(p.N, &pf));
            break;
        case PackTarget::Static:
            AZP_TRY(noalloc_scratch(p.N, &pf));
            break;
            }
        AZP_TRY(pack_fixed(p.d_pos, p.N, p.box, reinterpret_cast<int32_t*>(pf), s));
        ka.pos_fixed = pf;
        }
    if (args->d_head_list)
        return launch_pair_head(ka, evaluator, mixed, vir, block, s);
    if (tiled_ok)
        return launch_pair_tiled(ka, evaluator, mixed, vir, tile, slots, s);
    if (!args->d_nlist)
        return set_error(AZP_ERR_INVALID,
                         "the tiled list's %u neighbor slots and the type tables exceed the LDS of a workgroup; "
                         "pass the plain list",
                         slots);
    return launch_pair_column(ka, evaluator, mixed, vir, block, s);
    }

int launch_pair_column(const PairKernelArgs& ka, int evaluator, bool mixed, bool virial, unsigned block,
                       hipStream_t s)
    {
    return launch_pair_evaluator<kLayoutColumn>(ka, evaluator, mixed, virial, block, 0, s);
    }

} // namespace azp

extern "C" int azp_gpu_compute_pair_forces(const azp_pair_args* args, int evaluator, const void* d_params)
    {
    return compute_pair_forces(args, evaluator, d_params, PackTarget::Heap);
    }

extern "C" int azp_gpu_compute_pair_forces_noalloc(const azp_pair_args* args, int evaluator, const void* d_params)
    {
    return compute_pair_forces(args, evaluator, d_params, PackTarget::Static);
    }

