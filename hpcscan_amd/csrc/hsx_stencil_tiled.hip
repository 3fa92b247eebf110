// LDS-tiled, n3-streaming stencil kernels for gfx950 (placeholder: the
// reference-order kernels serve every request until the tiled family lands).
#include "hsx_internal.h"

namespace hsx {

int tiled_ac2_step(const hsx_grid *, void *, const void *, const void *, hipStream_t)
{
    return fail(HSX_ERR_UNSUPPORTED, "tiled kernels not built");
}

int tiled_laplacian(const hsx_grid *, const void *, void *, hipStream_t)
{
    return fail(HSX_ERR_UNSUPPORTED, "tiled kernels not built");
}

int tiled_fd_d2(const hsx_grid *, int, const void *, void *, hipStream_t)
{
    return fail(HSX_ERR_UNSUPPORTED, "tiled kernels not built");
}

} // namespace hsx
