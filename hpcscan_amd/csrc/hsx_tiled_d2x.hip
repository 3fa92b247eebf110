// Tiled D2 along one axis (Grid::FD_D2_N*), MODE_D2X.
#include "hsx_tiled.h"

namespace hsx {

int tiled_fd_d2_x(const hsx_grid *g, const void *u, void *w, hipStream_t s)
{
    return run_tiled<MODE_D2X>(g, w, nullptr, u, nullptr, s);
}

} // namespace hsx
