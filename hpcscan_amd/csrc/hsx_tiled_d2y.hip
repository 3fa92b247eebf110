// Tiled D2 along one axis (Grid::FD_D2_N*), MODE_D2Y.
#include "hsx_tiled.h"

namespace hsx {

int tiled_fd_d2_y(const hsx_grid *g, const void *u, void *w, hipStream_t s)
{
    return run_tiled<MODE_D2Y>(g, w, nullptr, u, nullptr, s);
}

} // namespace hsx
