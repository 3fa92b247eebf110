// Tiled staggered first derivative along one axis (Grid::FD_D1_N*), MODE_D1X.
#include "hsx_tiled.h"

namespace hsx {

int tiled_fd_d1_x(const hsx_grid *g, const void *u, void *w, hipStream_t s)
{
    return run_tiled<MODE_D1X>(g, w, nullptr, u, nullptr, s);
}

} // namespace hsx
