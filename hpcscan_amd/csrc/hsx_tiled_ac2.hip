// Tiled fused Ac2 step (Grid::computePressureWithFD) and the tile-shape override.
#include "hsx_tiled.h"

namespace hsx {

int g_tile_lx = 0, g_tile_chunk = 0;

int set_tile(int lx, int chunk)
{
    if (lx != 0 && lx != 16 && lx != 32 && lx != 64)
        return fail(HSX_ERR_INVALID, "tile lanes %d not in {0,16,32,64}", lx);
    if (chunk < 0) return fail(HSX_ERR_INVALID, "chunk %d < 0", chunk);
    g_tile_lx = lx;
    g_tile_chunk = chunk;
    return HSX_OK;
}

int tiled_ac2_step(const hsx_grid *g, void *prn, const void *prc, const void *coef, hipStream_t s)
{
    return run_tiled<MODE_AC2>(g, prn, prc, coef, s);
}

} // namespace hsx
