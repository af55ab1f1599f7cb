// Isotropic pair forces over the tiled list (gfx950).
//
// One workgroup per tile of 512 (or 256) consecutive cell-list slots: the
// tile's distinct neighbors are staged in LDS (fixed-point int4 in the mixed
// form, the FP64 double4 positions in the plain form) and each lane walks its
// particle's 16-bit entries (16 x the neighbor's rank in the tile) reading
// positions from LDS.  The kernel template is in pair_kernel.h; this unit
// holds its tiled instances.
#include "pair_kernel.h"

namespace azp
{

int launch_pair_tiled(const PairKernelArgs& ka, int evaluator, bool mixed, bool virial, unsigned tile,
                      unsigned slots, hipStream_t s)
    {
    return launch_pair_evaluator<kLayoutTiled>(ka, evaluator, mixed, virial, tile, slots, s);
    }

} // namespace azp
