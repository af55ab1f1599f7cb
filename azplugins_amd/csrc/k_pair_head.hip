// Isotropic pair forces over HOOMD's own neighbor-list layout (gfx950).
//
// The reference's launcher receives HOOMD's list as pair_args_t
// (PotentialPairGPUKernel.cu.inc:25-28): particle i's neighbors at
// d_nlist[d_head_list[i] + k] for k < d_n_neigh[i], rows unpadded.  These
// instances of the pair kernel read that layout directly, so a HOOMD caller
// passes its arrays without transposing them.  Each lane reads its own row
// (consecutive entries of one lane are contiguous, the lanes of a wavefront
// are not), so the list stream is not coalesced across the wavefront as in
// the column layout; the pair terms and their order are the same, and the
// results equal the column layout's bit for bit.
#include "pair_kernel.h"

namespace azp
{

int launch_pair_head(const PairKernelArgs& ka, int evaluator, bool mixed, bool virial, unsigned block,
                     hipStream_t s)
    {
    switch (evaluator)
        {
    case AZP_EVAL_HERTZ:
        return mixed ? launch_eval_head<EvalHertz, true>(ka, virial, block, s)
                     : launch_eval_head<EvalHertz, false>(ka, virial, block, s);
    case AZP_EVAL_EXPANDED_YUKAWA:
        return mixed ? launch_eval_head<EvalExpandedYukawa, true>(ka, virial, block, s)
                     : launch_eval_head<EvalExpandedYukawa, false>(ka, virial, block, s);
    case AZP_EVAL_PERTURBED_LJ:
        return mixed ? launch_eval_head<EvalPerturbedLJ, true>(ka, virial, block, s)
                     : launch_eval_head<EvalPerturbedLJ, false>(ka, virial, block, s);
    case AZP_EVAL_COLLOID:
        if (mixed)
            return set_error(AZP_ERR_UNSUPPORTED,
                             "Colloid runs in the plain FP64 form only (its colloid-colloid "
                             "branch cancels catastrophically in FP32)");
        return launch_eval_head<EvalColloid, false>(ka, virial, block, s);
    case AZP_EVAL_DPD_GENERAL_WEIGHT:
        return mixed ? launch_eval_head<EvalDPDConservative, true>(ka, virial, block, s)
                     : launch_eval_head<EvalDPDConservative, false>(ka, virial, block, s);
    default:
        return set_error(AZP_ERR_INVALID, "evaluator %d is not an isotropic pair evaluator", evaluator);
        }
    }

} // namespace azp
