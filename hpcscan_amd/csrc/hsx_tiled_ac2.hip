// Tiled fused Ac2 step (Grid::computePressureWithFD) and the tile-shape override.
#include "hsx_tiled.h"

namespace hsx {

TileCfg g_tile = {0, 0, 0, 0};

int set_tile(int lx, int by, int chunk, int flags)
{
    if (lx != 0 && lx != 16 && lx != 32 && lx != 64)
        return fail(HSX_ERR_INVALID, "tile lanes %d not in {0,16,32,64}", lx);
    if (by != 0 && by != 8 && by != 16) return fail(HSX_ERR_INVALID, "tile rows %d not in {0,8,16}", by);
    if (chunk < 0) return fail(HSX_ERR_INVALID, "chunk %d < 0", chunk);
    if (flags & ~(HSX_TILE_CACHED | HSX_TILE_NT | HSX_TILE_NO_RING))
        return fail(HSX_ERR_INVALID, "tile flags 0x%x", flags);
    if ((flags & HSX_TILE_CACHED) && (flags & HSX_TILE_NT))
        return fail(HSX_ERR_INVALID, "tile flags: cached and nontemporal streams together");
    g_tile = TileCfg{lx, by, chunk, flags};
    return HSX_OK;
}

int tiled_ac2_step(const hsx_grid *g, void *out, const void *prn, const void *prc, const void *coef, hipStream_t s)
{
    return run_tiled<MODE_AC2>(g, out, prn, prc, coef, s);
}

} // namespace hsx
