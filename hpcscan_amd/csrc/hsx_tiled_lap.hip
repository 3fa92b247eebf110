// Tiled 3D Laplacian (Grid::FD_LAPLACIAN).
#include "hsx_tiled.h"

namespace hsx {

int tiled_laplacian(const hsx_grid *g, const void *u, void *w, hipStream_t s)
{
    return run_tiled<MODE_LAP>(g, w, nullptr, u, nullptr, s);
}

} // namespace hsx
