// Contracted fused Ac2 step (hsx_set_contract(1)): the tiled kernel with each
// weighted sum a fused multiply-add, in a translation unit of its own.
#include "hsx_tiled.h"

namespace hsx {

int tiled_ac2_step_fma(const hsx_grid *g, void *out, const void *prn, const void *prc, const void *coef,
                       hipStream_t s)
{
    return run_tiled<MODE_AC2_FMA>(g, out, prn, prc, coef, s);
}

} // namespace hsx
