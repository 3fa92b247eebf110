// Contracted tiled 3D Laplacian (hsx_set_contract(1)), in a translation unit
// of its own.
#include "hsx_tiled.h"

namespace hsx {

int tiled_laplacian_fma(const hsx_grid *g, const void *u, void *w, hipStream_t s)
{
    return run_tiled<MODE_LAP_FMA>(g, w, nullptr, u, nullptr, s);
}

} // namespace hsx
