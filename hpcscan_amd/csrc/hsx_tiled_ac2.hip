// Tiled fused Ac2 step (Grid::computePressureWithFD), the tile-shape override
// and the tile autotuner.
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "hsx_tiled.h"

namespace hsx {

static std::mutex g_tile_mu;
static TileCfg g_tile = {0, 0, 0, 0};   // hsx_set_tile

TileCfg tile_forced(void)
{
    std::lock_guard<std::mutex> lk(g_tile_mu);
    return g_tile;
}

int set_tile(int lx, int by, int chunk, int flags)
{
    if (lx != 0 && lx != 16 && lx != 32 && lx != 64 && lx != 128)
        return fail(HSX_ERR_INVALID, "tile lanes %d not in {0,16,32,64,128}", lx);
    if (by != 0 && by != 8 && by != 16) return fail(HSX_ERR_INVALID, "tile rows %d not in {0,8,16}", by);
    if (lx == 128 && by == 16) return fail(HSX_ERR_INVALID, "128-lane tiles have 8 rows (1024 threads)");
    if (chunk < 0) return fail(HSX_ERR_INVALID, "chunk %d < 0", chunk);
    if (flags & ~(HSX_TILE_CACHED | HSX_TILE_NT | HSX_TILE_NO_RING))
        return fail(HSX_ERR_INVALID, "tile flags 0x%x", flags);
    if ((flags & HSX_TILE_CACHED) && (flags & HSX_TILE_NT))
        return fail(HSX_ERR_INVALID, "tile flags: cached and nontemporal streams together");
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile = TileCfg{lx, by, chunk, flags};
    return HSX_OK;
}

thread_local FoldReq t_fold = {0, {-1, -1, -1, -1, -1, -1}};

bool wide_tiles_fit(const hsx_grid *g)
{
    if (g->storage != HSX_FP32 || g->compute != HSX_FP32 || g->dim != 3) return false;
    return dispatch_R(g->fd_order / 2, [&](auto rc_) -> int {
               constexpr int RR = decltype(rc_)::value;
               using K = Tiled<float, float, RR, 4, 128, 8, MODE_AC2, true>;
               if constexpr (K::RING && K::LDS <= TILE_LDS_MAX) {
                   static const bool spills = tiled_spills(k_tiled<float, float, RR, 4, 128, 8, MODE_AC2, true>);
                   return spills ? 0 : 1;
               } else {
                   return 0;
               }
           }) == 1;
}


int tiled_ac2_step(const hsx_grid *g, void *out, const void *prn, const void *prc, const void *coef, hipStream_t s)
{
    if (contract_mode()) return tiled_ac2_step_fma(g, out, prn, prc, coef, s);
    return run_tiled<MODE_AC2>(g, out, prn, prc, coef, s);
}

// ---- autotuner ---------------------------------------------------------------------
// Shapes chosen by hsx_autotune, per operation, FD order, precisions and layout.
struct TuneKey {
    int mode, order, storage, compute;
    int64_t n[3], inner[6];
    bool operator<(const TuneKey &o) const
    {
        return std::memcmp(this, &o, sizeof(TuneKey)) < 0;
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TileCfg> g_tuned;

static TuneKey tune_key(const hsx_grid *g, int mode)
{
    TuneKey k;
    std::memset(&k, 0, sizeof(k));
    k.mode = mode;
    k.order = g->fd_order;
    k.storage = g->storage;
    k.compute = g->compute;
    k.n[0] = g->n1;
    k.n[1] = g->n2;
    k.n[2] = g->n3;
    for (int i = 0; i < 6; i++) k.inner[i] = g->inner[i];
    return k;
}

bool tuned_shape(const hsx_grid *g, int mode, TileCfg *out)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tuned.empty()) return false;
    auto it = g_tuned.find(tune_key(g, mode));
    if (it == g_tuned.end()) return false;
    *out = it->second;
    return true;
}

} // namespace hsx

using namespace hsx;

extern "C" int hsx_autotune(const hsx_grid *g, int reps, int shape[4], float *best_ms)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (reps < 1 || reps > 1000) return fail(HSX_ERR_INVALID, "reps %d not in 1..1000", reps);
    if (g->dim != 3) return fail(HSX_ERR_UNSUPPORTED, "the tiled kernels (and the tuner) need a 3D grid");
    if ((rc = ensure_init())) return rc;
    hipStream_t s = stream_of(nullptr);
    const size_t es = g->storage == HSX_FP64 ? 8 : 4;
    const size_t bytes = (size_t)(g->n1 * g->n2 * g->n3) * es;
    void *buf[3] = {nullptr, nullptr, nullptr};
    auto release = [&] {
        for (void *&p : buf)
            if (p) (void)hipFree(p);
    };
    for (void *&p : buf) {
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            release();
            return hip_fail(e, "hipMalloc (autotune)");
        }
        (void)hipMemsetAsync(p, 0, bytes, s);
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        release();
        return fail(HSX_ERR_HIP, "hipEventCreate (autotune)");
    }
    const int64_t N3 = g->inner[5] - g->inner[4] + 1;
    const int R = g->fd_order / 2;
    // candidates: lanes x rows, stream mode, halo ring, n3 chunk; the first is
    // the library's own choice, so the shape kept is never slower than it in
    // this sweep
    std::vector<TileCfg> cand = {
        {0, 0, 0, 0},
        {64, 16, 0, 0}, {64, 16, 0, HSX_TILE_NO_RING}, {64, 16, 0, HSX_TILE_CACHED},
        {32, 16, 0, 0}, {32, 16, 0, HSX_TILE_NO_RING},
        {64, 8, 0, 0}, {64, 8, 0, HSX_TILE_NO_RING}, {64, 8, 0, HSX_TILE_NO_RING | HSX_TILE_NT},
        {64, 8, (int)std::min<int64_t>(N3, 8 * R + 64), HSX_TILE_NO_RING},
        {32, 8, 0, 0}, {32, 8, 0, HSX_TILE_NO_RING}, {16, 8, 0, 0},
        {128, 8, 0, 0},
    };
    TileCfg best = {0, 0, 0, 0};
    float best_t = 1e30f;
    const bool wide = wide_tiles_fit(g);
    for (const TileCfg &c : cand) {
        if (c.lx == 128 && !wide) continue;      // would run another candidate's shape
        rc = run_tiled_shape<MODE_AC2>(g, buf[0], buf[0], buf[1], buf[2], s, c);   // warm-up (and fit check)
        if (rc == HSX_ERR_UNSUPPORTED) continue;
        if (rc) break;
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < reps && !rc; k++) rc = run_tiled_shape<MODE_AC2>(g, buf[0], buf[0], buf[1], buf[2], s, c);
        (void)hipEventRecord(e1, s);
        if (rc) break;
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= reps;
        if (ms < best_t) {
            best_t = ms;
            best = c;
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    release();
    if (rc) return rc;
    if (best_t >= 1e29f) return fail(HSX_ERR_UNSUPPORTED, "autotune: no tile shape fits this grid");
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        // the fused step only: the Laplacian keeps its own choice (at order 16,
        // 500^3, the fused step's tuned shape took the FD_D2 Laplacian from 355
        // to 306 GPoint/s; it is bound by VALU issue, the step by HBM)
        g_tuned[tune_key(g, MODE_AC2)] = best;
    }
    if (shape) {
        shape[0] = best.lx;
        shape[1] = best.by;
        shape[2] = best.chunk;
        shape[3] = best.flags;
    }
    if (best_ms) *best_ms = best_t;
    return HSX_OK;
}

extern "C" int hsx_autotune_clear(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
    return HSX_OK;
}
