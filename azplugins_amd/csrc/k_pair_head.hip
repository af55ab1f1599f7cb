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
    return launch_pair_evaluator<kLayoutHead>(ka, evaluator, mixed, virial, block, 0, s);
    }

} // namespace azp
