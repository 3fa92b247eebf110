// Batched shots: many grids (one order and precision, any sizes) advanced
// together, one launch of each kernel per time step for all of them:
//   free surface of every shot (2 launches: n1 faces and mirrors, then the
//   boundary planes), the fused Ac2 step of every shot
//   (one k_tiled_batch launch over the tiles of all grids), then sources and
//   receivers of every shot (one launch, one workgroup per shot).
// The whole time loop can be captured once into a HIP graph and replayed.
// Each shot is the forward modeling of hsx_forward (same arithmetic, same
// results).  No counterpart in the reference, which runs one grid per process.
#include <functional>
#include <vector>

#include "hsx_tiled.h"

namespace hsx {

struct SrcRec {             // one shot, one step parity
    void *field;            // wavefield written by the step
    const int64_t *src;
    const void *wavelet;
    int64_t nsrc;
    const int64_t *rec;
    void *traces;
    int64_t nrec;
    int64_t npoint;         // points of the grid: indices outside [0, npoint) are skipped
};

// one workgroup per shot: add the source samples of step n, then record
template <typename TS>
__global__ void __launch_bounds__(256) k_srcrec_batch(const SrcRec *sr, int64_t n)
{
    const SrcRec &x = sr[blockIdx.x];
    TS *u = reinterpret_cast<TS *>(x.field);
    const TS *w = reinterpret_cast<const TS *>(x.wavelet);
    const uint64_t np = (uint64_t)x.npoint;
    for (int64_t i = threadIdx.x; i < x.nsrc; i += blockDim.x) {
        if ((uint64_t)x.src[i] >= np) continue;
        TS *p = u + x.src[i];
        *p = (TS)((double)*p + 1.0 * (double)w[n * x.nsrc + i]);
    }
    __syncthreads();
    TS *tr = reinterpret_cast<TS *>(x.traces);
    for (int64_t i = threadIdx.x; i < x.nrec; i += blockDim.x) {
        const int64_t j = x.rec[i];
        tr[n * x.nrec + i] = (uint64_t)j < np ? u[j] : (TS)0;
    }
}

template <typename TS>
__global__ void __launch_bounds__(256) k_zero_batch(TS *const *f, const int64_t *npoint)
{
    TS *u = f[blockIdx.y];
    const int64_t n = npoint[blockIdx.y];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        u[i] = (TS)0;
}

} // namespace hsx

using namespace hsx;

struct hsx_batch {
    int nshots = 0;
    int storage = HSX_FP32;
    int64_t nt = 0;
    std::vector<void *> owned;                 // device allocations to free
    void *fields = nullptr;                    // [2*nshots] A_s then B_s
    void *npoint = nullptr;                    // [2*nshots]
    int64_t most_npoint = 0;
    void *bc_fields[2] = {nullptr, nullptr};   // [nshots] the wavefield B applies to, per parity
    void *sets = nullptr;                      // FaceSet per shot (without the n1 faces k_bc_n1 takes)
    int64_t most = 0, most_plane = 0;
    void *n1 = nullptr;                        // N1Faces per shot (k_bc_n1), when n1_nh > 0
    int n1_nh = 0;                             // R / vector width
    int64_t n1_rows = 0;                       // largest number of rows of an n1 face
    void *sr[2] = {nullptr, nullptr};          // SrcRec per shot, per parity
    std::function<int(int, hipStream_t)> step; // batched fused step, per parity
    std::vector<void *> step_owned;            // device argument tables of step
    int shape[4] = {0, 0, 0, 0};               // lanes, rows, chunk (0: the rule), ring of step
    std::vector<hsx_grid> grids;               // the shots' descriptors and coef buffers,
    std::vector<const void *> coefs;           // to rebuild step (hsx_batch_autotune)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace hsx {

static int dalloc(hsx_batch *b, size_t bytes, void **p)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (batch)");
    b->owned.push_back(*p);
    return HSX_OK;
}

static int upload(hsx_batch *b, const void *host, size_t bytes, void **p)
{
    int rc = dalloc(b, bytes, p);
    if (rc) return rc;
    HSX_HIP(hipMemcpy(*p, host, bytes, hipMemcpyHostToDevice));
    return HSX_OK;
}

// one fused-step instance shared by every shot; force_chunk > 0: planes per
// workgroup, else whole columns when the tiles fill the device once and
// 32-plane chunks when they do not
template <typename TS, int R, int LX, int BY, bool ring>
static int batch_setup_shape(hsx_batch *b, const hsx_grid *grids, const void *const *coef, int64_t force_chunk = 0)
{
    constexpr int VW = 16 / sizeof(TS);
    using K = Tiled<TS, TS, R, VW, LX, BY, MODE_AC2, ring>;
    auto kern = k_tiled_batch<TS, TS, R, VW, LX, BY, MODE_AC2, ring>;
    static const int per_cu = tiled_per_cu(kern, LX * BY, K::LDS);
    const int ns = b->nshots;
    const int64_t resident = (int64_t)per_cu * device_cus();
    const int64_t slots = std::max<int64_t>(1, resident / ns);
    std::vector<TiledArgs<TS>> args[2];
    std::vector<int> off(ns + 1, 0);
    std::vector<void *> F(2 * ns);            // A_s then B_s
    HSX_HIP(hipMemcpy(F.data(), b->fields, 2 * ns * sizeof(void *), hipMemcpyDeviceToHost));
    for (int p = 0; p < 2; p++) args[p].resize(ns);
    // when the shots' tiles outnumber the resident workgroups the launch runs
    // in several rounds: short chunks then balance them (32 shots of
    // 192^3-320^3, order 8: 2.47 s for 1000 steps with 32-plane chunks, 2.75 s
    // with whole columns; 16-40 planes all within 1 %)
    int64_t tiles = 0, chunk = force_chunk;
    const TileCfg cfg = tile_forced();
    for (int pass = 0; pass < 2; pass++) {
        tiles = 0;
        for (int s = 0; s < ns; s++) {
            void *A = F[s], *B = F[ns + s];
            // parity 0: B is the current wavefield, A the previous one (and the output)
            int rc = fill_tiled_args<TS, TS, R, VW, LX, BY, MODE_AC2>(&grids[s], A, A, B, coef[s], slots, cfg,
                                                                     args[0][s], chunk);
            if (!rc)
                rc = fill_tiled_args<TS, TS, R, VW, LX, BY, MODE_AC2>(&grids[s], B, B, A, coef[s], slots, cfg,
                                                                     args[1][s], chunk);
            if (rc) return rc;
            off[s + 1] = off[s] + args[0][s].ntx * args[0][s].nty * args[0][s].nchunk;
            tiles += (int64_t)args[0][s].ntx * args[0][s].nty;
        }
        if (tiles <= resident || force_chunk > 0) break;
        chunk = 32;
    }
    // argument tables of this step; those of the step it replaces are freed
    std::vector<void *> mine;
    auto up = [&](const void *host, size_t bytes, void **p) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 1);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc (batch step)");
        mine.push_back(*p);
        HSX_HIP(hipMemcpy(*p, host, bytes, hipMemcpyHostToDevice));
        return HSX_OK;
    };
    void *dargs[2], *doff;
    int rc;
    if ((rc = up(args[0].data(), ns * sizeof(TiledArgs<TS>), &dargs[0])) ||
        (rc = up(args[1].data(), ns * sizeof(TiledArgs<TS>), &dargs[1])) ||
        (rc = up(off.data(), (ns + 1) * sizeof(int), &doff))) {
        for (void *q : mine) (void)hipFree(q);
        return rc;
    }
    if (!b->step_owned.empty()) {
        HSX_HIP(hipDeviceSynchronize());   // no launch of the old step may still read its tables
        for (void *q : b->step_owned) (void)hipFree(q);
    }
    b->step_owned = mine;
    b->shape[0] = LX;
    b->shape[1] = BY;
    b->shape[2] = (int)force_chunk;
    b->shape[3] = ring ? 1 : 0;
    const int total = off[ns];
    const unsigned grid = (unsigned)(((int64_t)total + 7) / 8 * 8);
    b->step = [=](int parity, hipStream_t s) -> int {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(LX, BY), K::LDS, s, (const TiledArgs<TS> *)dargs[parity],
                           (const int *)doff, ns);
        HSX_LAUNCHED("k_tiled_batch");
        return HSX_OK;
    };
    return HSX_OK;
}

// The batched step as one persistent launch (k_tiled_batch_seg): every shot's
// tiles as whole columns, the columns of all shots laid end to end and cut into
// one part of equal planes per resident workgroup (at least 4R planes a part).
template <typename TS, int R, int LX, int BY, bool ring>
static int batch_setup_persistent(hsx_batch *b, const hsx_grid *grids, const void *const *coef)
{
    constexpr int VW = 16 / sizeof(TS);
    using K = Tiled<TS, TS, R, VW, LX, BY, MODE_AC2, ring>;
    auto kern = k_tiled_batch_seg<TS, TS, R, VW, LX, BY, MODE_AC2, ring>;
    static const int per_cu = tiled_per_cu(kern, LX * BY, K::LDS);
    const int ns = b->nshots;
    const int64_t resident = (int64_t)per_cu * device_cus();
    std::vector<TiledArgs<TS>> args[2];
    std::vector<void *> F(2 * ns);
    HSX_HIP(hipMemcpy(F.data(), b->fields, 2 * ns * sizeof(void *), hipMemcpyDeviceToHost));
    for (int p = 0; p < 2; p++) args[p].resize(ns);
    const TileCfg cfg = tile_forced();
    int64_t planes = 0;
    for (int s = 0; s < ns; s++) {
        void *A = F[s], *B = F[ns + s];
        // whole columns: one chunk of all N3 planes per tile
        int rc = fill_tiled_args<TS, TS, R, VW, LX, BY, MODE_AC2>(&grids[s], A, A, B, coef[s], 1, cfg, args[0][s],
                                                                 (int64_t)1 << 30);
        if (!rc)
            rc = fill_tiled_args<TS, TS, R, VW, LX, BY, MODE_AC2>(&grids[s], B, B, A, coef[s], 1, cfg, args[1][s],
                                                                 (int64_t)1 << 30);
        if (rc) return rc;
        const TiledArgs<TS> &a = args[0][s];
        if (a.nchunk != 1) return fail(HSX_ERR_INVALID, "batch: shot %d not in whole columns", s);
        planes += (int64_t)a.ntx * a.nty * (a.e3 - a.s3 + 1);
    }
    const int64_t parts = std::max<int64_t>(1, std::min<int64_t>(resident, planes / (4 * R)));
    if (parts > ((int64_t)1 << 30) / 8) return fail(HSX_ERR_UNSUPPORTED, "batch: too many parts");
    // part g takes the planes [planes * g / parts, planes * (g + 1) / parts) of
    // the columns laid end to end (shot, then tile, in order)
    std::vector<BatchSeg> seg;
    std::vector<int> so(parts + 1, 0);
    int64_t pos = 0;      // first plane of the current column in the global order
    int64_t g = 0;
    for (int s = 0; s < ns && g < parts; s++) {
        const TiledArgs<TS> &a = args[0][s];
        const int64_t n3 = a.e3 - a.s3 + 1, tiles = (int64_t)a.ntx * a.nty;
        for (int64_t t = 0; t < tiles; t++, pos += n3) {
            // the parts that overlap [pos, pos + n3)
            while (g < parts) {
                const int64_t lo = planes * g / parts, hi = planes * (g + 1) / parts;
                const int64_t a0 = std::max(lo, pos), a1 = std::min(hi, pos + n3);
                if (a0 < a1)
                    seg.push_back({s, (int)t, (int)(a.s3 + (a0 - pos)), (int)(a.s3 + (a1 - pos))});
                if (hi <= pos + n3) {           // part g ends in this column: close it
                    so[g + 1] = (int)seg.size();
                    g++;
                } else {
                    break;                      // part g goes on in the next column
                }
            }
        }
    }
    if (g != parts || (int64_t)seg.size() >= ((int64_t)1 << 31))
        return fail(HSX_ERR_INVALID, "batch: %lld of %lld parts cut", (long long)g, (long long)parts);
    std::vector<void *> mine;
    auto up = [&](const void *host, size_t bytes, void **p) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 1);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc (batch step)");
        mine.push_back(*p);
        HSX_HIP(hipMemcpy(*p, host, bytes, hipMemcpyHostToDevice));
        return HSX_OK;
    };
    void *dargs[2], *dseg, *dso;
    int rc;
    if ((rc = up(args[0].data(), ns * sizeof(TiledArgs<TS>), &dargs[0])) ||
        (rc = up(args[1].data(), ns * sizeof(TiledArgs<TS>), &dargs[1])) ||
        (rc = up(seg.data(), seg.size() * sizeof(BatchSeg), &dseg)) ||
        (rc = up(so.data(), so.size() * sizeof(int), &dso))) {
        for (void *q : mine) (void)hipFree(q);
        return rc;
    }
    if (!b->step_owned.empty()) {
        HSX_HIP(hipDeviceSynchronize());   // no launch of the old step may still read its tables
        for (void *q : b->step_owned) (void)hipFree(q);
    }
    b->step_owned = mine;
    b->shape[0] = LX;
    b->shape[1] = BY;
    b->shape[2] = -1;
    b->shape[3] = ring ? 1 : 0;
    const int np = (int)parts;
    const unsigned grid = (unsigned)((parts + 7) / 8 * 8);
    b->step = [=](int parity, hipStream_t st) -> int {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(LX, BY), K::LDS, st, (const TiledArgs<TS> *)dargs[parity],
                           (const BatchSeg *)dseg, (const int *)dso, np);
        HSX_LAUNCHED("k_tiled_batch_seg");
        return HSX_OK;
    };
    return HSX_OK;
}

// the fused-step instance shared by every shot: 16 rows with the halo ring
// where its LDS fits and it keeps its values in registers, else 16 rows
// without the ring, else 8 rows (64 lanes; 32 for FP64)
template <typename TS, int R>
static int batch_setup_step(hsx_batch *b, const hsx_grid *grids, const void *const *coef)
{
    constexpr int VW = 16 / sizeof(TS);
    constexpr int LX = sizeof(TS) == 4 ? 64 : 32;
    constexpr bool ring = Tiled<TS, TS, R, VW, LX, 16, MODE_AC2, true>::LDS <= TILE_LDS_MAX;
    constexpr bool tall = Tiled<TS, TS, R, VW, LX, 16, MODE_AC2, false>::LDS <= TILE_LDS_MAX;
    if constexpr (ring) {
        if (!tiled_spills(k_tiled_batch<TS, TS, R, VW, LX, 16, MODE_AC2, true>))
            return batch_setup_shape<TS, R, LX, 16, true>(b, grids, coef);
    }
    if constexpr (tall) {
        if (!tiled_spills(k_tiled_batch<TS, TS, R, VW, LX, 16, MODE_AC2, false>))
            return batch_setup_shape<TS, R, LX, 16, false>(b, grids, coef);
    }
    // 64 lanes x 16 rows spill (FP32, order 16): 32 lanes x 16 rows with the ring
    if constexpr (LX == 64 && Tiled<TS, TS, R, VW, 32, 16, MODE_AC2, true>::LDS <= TILE_LDS_MAX) {
        if (!tiled_spills(k_tiled_batch<TS, TS, R, VW, 32, 16, MODE_AC2, true>))
            return batch_setup_shape<TS, R, 32, 16, true>(b, grids, coef);
    }
    return batch_setup_shape<TS, R, LX, 8, false>(b, grids, coef);
}

// a batch step of the given shape, if it fits this order and precision
// (LDS, registers without scratch); HSX_ERR_UNSUPPORTED otherwise
template <typename TS, int R, int LX, int BY, bool RING>
static int batch_try_shape(hsx_batch *b, int64_t chunk)
{
    constexpr int VW = 16 / sizeof(TS);
    if constexpr (Tiled<TS, TS, R, VW, LX, BY, MODE_AC2, RING>::LDS > TILE_LDS_MAX ||
                  (RING && !Tiled<TS, TS, R, VW, LX, BY, MODE_AC2, RING>::RING)) {
        return HSX_ERR_UNSUPPORTED;
    } else {
        if (chunk < 0) {
            if (tiled_spills(k_tiled_batch_seg<TS, TS, R, VW, LX, BY, MODE_AC2, RING>)) return HSX_ERR_UNSUPPORTED;
            return batch_setup_persistent<TS, R, LX, BY, RING>(b, b->grids.data(), b->coefs.data());
        }
        if (tiled_spills(k_tiled_batch<TS, TS, R, VW, LX, BY, MODE_AC2, RING>)) return HSX_ERR_UNSUPPORTED;
        return batch_setup_shape<TS, R, LX, BY, RING>(b, b->grids.data(), b->coefs.data(), chunk);
    }
}

template <typename TS, int R>
static int batch_shape_rt(hsx_batch *b, int lx, int by, bool ring, int64_t chunk)
{
    if (lx == 64 && by == 16) return ring ? batch_try_shape<TS, R, 64, 16, true>(b, chunk)
                                          : batch_try_shape<TS, R, 64, 16, false>(b, chunk);
    if (lx == 32 && by == 16) return ring ? batch_try_shape<TS, R, 32, 16, true>(b, chunk)
                                          : batch_try_shape<TS, R, 32, 16, false>(b, chunk);
    if (lx == 64 && by == 8) return batch_try_shape<TS, R, 64, 8, false>(b, chunk);
    if (lx == 32 && by == 8) return batch_try_shape<TS, R, 32, 8, false>(b, chunk);
    return HSX_ERR_UNSUPPORTED;
}

static int batch_record(hsx_batch *b, hipStream_t s)
{
    const int ns = b->nshots;
    const unsigned nbz = (unsigned)std::max<int64_t>(1, std::min<int64_t>((b->most_npoint + 255) / 256, 2048));
    if (b->storage == HSX_FP64)
        hipLaunchKernelGGL(k_zero_batch<double>, dim3(nbz, 2 * ns), dim3(256), 0, s, (double *const *)b->fields,
                           (const int64_t *)b->npoint);
    else
        hipLaunchKernelGGL(k_zero_batch<float>, dim3(nbz, 2 * ns), dim3(256), 0, s, (float *const *)b->fields,
                           (const int64_t *)b->npoint);
    HSX_LAUNCHED("k_zero_batch");
    for (int64_t n = 0; n < b->nt; n++) {
        const int p = (int)(n & 1);
        int rc = b->n1_nh ? bc_batch_fs_launch(b->storage, ns, (void *const *)b->bc_fields[p], (const N1Faces *)b->n1,
                                               b->n1_nh, b->n1_rows, (const FaceSet *)b->sets, b->most,
                                               b->most_plane, s)
                          : bc_batch_launch(b->storage, ns, (void *const *)b->bc_fields[p], (const FaceSet *)b->sets,
                                            b->most, b->most_plane, s);
        if (!rc) rc = b->step(p, s);
        if (rc) return rc;
        if (b->storage == HSX_FP64)
            hipLaunchKernelGGL(k_srcrec_batch<double>, dim3(ns), dim3(256), 0, s, (const SrcRec *)b->sr[p], n);
        else
            hipLaunchKernelGGL(k_srcrec_batch<float>, dim3(ns), dim3(256), 0, s, (const SrcRec *)b->sr[p], n);
        HSX_LAUNCHED("k_srcrec_batch");
    }
    return HSX_OK;
}

} // namespace hsx

extern "C" {

int hsx_batch_create(int nshots, const hsx_grid *grids, const void *const *coef, const hsx_shot *shots,
                     void *const *traces, hsx_batch **out)
{
    if (!out) return fail(HSX_ERR_INVALID, "null batch pointer");
    *out = nullptr;
    if (nshots < 1 || nshots > 4096) return fail(HSX_ERR_INVALID, "nshots %d not in 1..4096", nshots);
    int rc;
    if ((rc = check_ptr(grids, "grids")) || (rc = check_ptr(coef, "coef")) || (rc = check_ptr(shots, "shots")) ||
        (rc = check_ptr(traces, "traces")))
        return rc;
    if ((rc = ensure_init())) return rc;
    const hsx_grid &g0 = grids[0];
    for (int s = 0; s < nshots; s++) {
        const hsx_grid &g = grids[s];
        if ((rc = check_grid(&g))) return rc;
        if (g.dim != 3 || g.fd_order != g0.fd_order || g.storage != g0.storage || g.compute != g.storage)
            return fail(HSX_ERR_UNSUPPORTED, "batch: shot %d: every grid must be 3D with one order and precision", s);
        const int vw = g.storage == HSX_FP64 ? 2 : 4;
        if (g.n1 % vw) return fail(HSX_ERR_UNSUPPORTED, "batch: shot %d: n1 %% %d != 0", s, vw);
        if (shots[s].nt != shots[0].nt) return fail(HSX_ERR_UNSUPPORTED, "batch: every shot must have the same nt");
        if ((rc = check_ptr(coef[s], "coef[s]")) || (rc = check_ptr(traces[s], "traces[s]"))) return rc;
        if (((uintptr_t)coef[s]) % 16) return fail(HSX_ERR_UNSUPPORTED, "batch: coef of shot %d not 16-byte aligned", s);
    }
    hsx_batch *b = new hsx_batch();
    b->nshots = nshots;
    b->grids.assign(grids, grids + nshots);
    b->coefs.assign(coef, coef + nshots);
    b->storage = g0.storage;
    b->nt = shots[0].nt;
    const size_t es = g0.storage == HSX_FP64 ? 8 : 4;
    std::vector<void *> F(2 * nshots);
    std::vector<int64_t> np(2 * nshots);
    std::vector<FaceSet> sets(nshots);
    std::vector<N1Faces> n1(nshots);
    std::vector<SrcRec> sr[2];
    sr[0].resize(nshots);
    sr[1].resize(nshots);
    std::vector<void *> bcf[2];
    bcf[0].resize(nshots);
    bcf[1].resize(nshots);
    auto bail = [&](int code) {
        hsx_batch_destroy(b);
        return code;
    };
    for (int s = 0; s < nshots; s++) {
        const hsx_grid &g = grids[s];
        const int64_t n = g.n1 * g.n2 * g.n3;
        void *A, *B;
        if ((rc = dalloc(b, n * es, &A)) || (rc = dalloc(b, n * es, &B))) return bail(rc);
        F[s] = A;
        F[nshots + s] = B;
        np[s] = np[nshots + s] = n;
        b->most_npoint = std::max(b->most_npoint, n);
        bool closed;
        int64_t most, most_plane;
        if ((rc = make_faceset(&g, 0x3F, &sets[s], &closed, &most, &most_plane))) return bail(rc);
        if (!closed) return bail(fail(HSX_ERR_UNSUPPORTED, "batch: shot %d too small for the two-launch boundary", s));
        b->most = std::max(b->most, most);
        b->most_plane = std::max(b->most_plane, most_plane);
        // n1 faces in one 16-byte pass where they qualify; the rest stay with
        // the generic face kernels (the boundary-plane set zmask / zidx is kept)
        N1Faces &x = n1[s];
        sets[s].vec1 = n1_groups_ok(&g, A, sets[s]) && n1_groups_ok(&g, B, sets[s]);
        x.sides = n1_vector_sides(&g, A, sets[s], &x.lo, &x.hi);
        if (x.sides) {
            const Face &f = (x.sides & 1) ? x.lo : x.hi;
            b->n1_rows = std::max(b->n1_rows, (f.be - f.bs + 1) * (f.oe - f.os + 1));
            b->n1_nh = (int)((f.he - f.hs + 1) / (g.storage == HSX_FP64 ? 2 : 4));
            int k = 0;
            for (int j = 0; j < sets[s].n; j++)
                if (sets[s].f[j].axis != 0) sets[s].f[k++] = sets[s].f[j];
            sets[s].n = k;
        }
        // parity 0: free surface on B, step into A; parity 1: the other way round
        bcf[0][s] = B;
        bcf[1][s] = A;
        for (int p = 0; p < 2; p++) {
            SrcRec &x = sr[p][s];
            x.field = p == 0 ? A : B;
            x.src = shots[s].src_idx;
            x.wavelet = shots[s].wavelet;
            x.nsrc = shots[s].nsrc;
            x.rec = shots[s].rec_idx;
            x.traces = traces[s];
            x.nrec = shots[s].nrec;
            x.npoint = n;
        }
    }
    if ((rc = upload(b, F.data(), F.size() * sizeof(void *), &b->fields)) ||
        (rc = upload(b, np.data(), np.size() * sizeof(int64_t), &b->npoint)) ||
        (rc = upload(b, sets.data(), sets.size() * sizeof(FaceSet), &b->sets)))
        return bail(rc);
    if (b->n1_nh && (rc = upload(b, n1.data(), n1.size() * sizeof(N1Faces), &b->n1))) return bail(rc);
    for (int p = 0; p < 2; p++)
        if ((rc = upload(b, bcf[p].data(), nshots * sizeof(void *), &b->bc_fields[p])) ||
            (rc = upload(b, sr[p].data(), nshots * sizeof(SrcRec), &b->sr[p])))
            return bail(rc);
    rc = dispatch_R(g0.fd_order / 2, [&](auto rc_) -> int {
        constexpr int R = decltype(rc_)::value;
        if (g0.storage == HSX_FP64) return batch_setup_step<double, R>(b, grids, coef);
        return batch_setup_step<float, R>(b, grids, coef);
    });
    if (rc) return bail(rc);
    *out = b;
    return HSX_OK;
}

int hsx_batch_run(hsx_batch *b, int use_graph, void *stream)
{
    if (!b) return fail(HSX_ERR_INVALID, "null batch");
    int rc = ensure_init();
    if (rc) return rc;
    hipStream_t s = stream_of(stream);
    if (!use_graph) {
        if ((rc = batch_record(b, s))) return rc;
        HSX_HIP(hipStreamSynchronize(s));
        return HSX_OK;
    }
    if (!b->exec) {
        // capture the whole time loop once (kernels only: no memset nodes)
        HSX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = batch_record(b, s);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(s, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        b->graph = gr;
        HSX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HSX_HIP(hipGraphLaunch(b->exec, s));
    HSX_HIP(hipStreamSynchronize(s));
    return HSX_OK;
}

int hsx_batch_autotune(hsx_batch *b, int reps, int shape[4], float *best_ms)
{
    if (!b) return fail(HSX_ERR_INVALID, "null batch");
    if (reps < 1 || reps > 1000) return fail(HSX_ERR_INVALID, "reps %d not in 1..1000", reps);
    int rc = ensure_init();
    if (rc) return rc;
    hipStream_t s = stream_of(nullptr);
    const int R = b->grids[0].fd_order / 2;
    auto setup = [&](int lx, int by, bool ring, int64_t chunk) -> int {
        return dispatch_R(R, [&](auto rc_) -> int {
            constexpr int RR = decltype(rc_)::value;
            if (b->storage == HSX_FP64) return batch_shape_rt<double, RR>(b, lx, by, ring, chunk);
            return batch_shape_rt<float, RR>(b, lx, by, ring, chunk);
        });
    };
    struct Cand {
        int lx, by;
        bool ring;
        int chunk;
    };
    // the shape hsx_batch_create chose first, then the others
    std::vector<Cand> cand = {{b->shape[0], b->shape[1], b->shape[3] != 0, b->shape[2]}};
    for (int lx : {64, 32})
        for (int by : {16, 8})
            for (bool ring : {true, false})
                for (int chunk : {0, 16, 32, 64, -1}) {
                    if (by == 8 && ring) continue;
                    const Cand c = {lx, by, ring, chunk};
                    if (c.lx == cand[0].lx && c.by == cand[0].by && c.ring == cand[0].ring && c.chunk == cand[0].chunk)
                        continue;
                    cand.push_back(c);
                }
    hipEvent_t e0, e1;
    HSX_HIP(hipEventCreate(&e0));
    HSX_HIP(hipEventCreate(&e1));
    Cand best = cand[0];
    float best_t = 1e30f;
    for (const Cand &c : cand) {
        rc = setup(c.lx, c.by, c.ring, c.chunk);
        if (rc == HSX_ERR_UNSUPPORTED) continue;
        if (rc) break;
        if ((rc = b->step(0, s))) break;                         // warm-up
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < reps && !rc; k++) rc = b->step(k & 1, s);
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
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == HSX_ERR_UNSUPPORTED) rc = HSX_OK;
    if (!rc && best_t >= 1e29f) rc = fail(HSX_ERR_UNSUPPORTED, "batch autotune: no shape fits");
    if (!rc) rc = setup(best.lx, best.by, best.ring, best.chunk);
    if (rc) return rc;
    // the captured time loop (if any) launched the old step: capture again
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (shape)
        for (int k = 0; k < 4; k++) shape[k] = b->shape[k];
    if (best_ms) *best_ms = best_t;
    return HSX_OK;
}

int hsx_batch_set_shape(hsx_batch *b, int lanes, int rows, int chunk, int ring)
{
    if (!b) return fail(HSX_ERR_INVALID, "null batch");
    if (chunk < -1 || chunk > (1 << 20)) return fail(HSX_ERR_INVALID, "batch shape: chunk %d", chunk);
    int rc = ensure_init();
    if (rc) return rc;
    const int R = b->grids[0].fd_order / 2;
    rc = dispatch_R(R, [&](auto rc_) -> int {
        constexpr int RR = decltype(rc_)::value;
        if (b->storage == HSX_FP64) return batch_shape_rt<double, RR>(b, lanes, rows, ring != 0, chunk);
        return batch_shape_rt<float, RR>(b, lanes, rows, ring != 0, chunk);
    });
    if (rc == HSX_ERR_UNSUPPORTED)
        return fail(HSX_ERR_UNSUPPORTED, "batch shape %d x %d ring %d does not fit this order and precision", lanes,
                    rows, ring);
    if (rc) return rc;
    // a captured time loop launched the old step: capture again
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return HSX_OK;
}

int hsx_batch_destroy(hsx_batch *b)
{
    if (!b) return HSX_OK;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipDeviceSynchronize();
    for (void *p : b->owned) (void)hipFree(p);
    for (void *p : b->step_owned) (void)hipFree(p);
    delete b;
    return HSX_OK;
}

} // extern "C"
