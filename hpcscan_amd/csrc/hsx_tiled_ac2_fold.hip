// Fused Ac2 step with the free surface folded in (hsx_ac2_step_fs): the FOLD
// instances of the tiled kernel, in a translation unit of their own.
#include "hsx_tiled.h"

namespace hsx {

int tiled_ac2_step_fold(const hsx_grid *g, void *prn, const void *prc, const void *coef, int flags,
                        const int64_t fb[6], hipStream_t s)
{
    // the kernel keeps the byte distance from a row to its n2 mirror in 32 bits
    const int64_t es = g->storage == HSX_FP64 ? 8 : 4;
    if ((int64_t)g->fd_order * g->n1 * es >= ((int64_t)1 << 31))
        return fail(HSX_ERR_UNSUPPORTED, "folded free surface: rows too long");
    t_fold.flags = flags;
    for (int k = 0; k < 6; k++) t_fold.fb[k] = fb[k];
    const int rc = run_tiled<MODE_AC2, true>(g, prn, prn, prc, coef, s);
    t_fold.flags = 0;
    return rc;
}

} // namespace hsx
