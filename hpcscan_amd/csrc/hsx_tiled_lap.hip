// Tiled 3D Laplacian (Grid::FD_LAPLACIAN).
#include "hsx_tiled.h"

namespace hsx {

int tiled_laplacian(const hsx_grid *g, const void *u, void *w, hipStream_t s)
{
    if (contract_mode()) return tiled_laplacian_fma(g, u, w, s);
    return run_tiled<MODE_LAP>(g, w, nullptr, u, nullptr, s);
}

} // namespace hsx
