// Stencil entry points of the C ABI and the reference-order kernels.
//
// Reference kernels: one thread per inner point, block 64 x 4 over (i1, i2),
// grid z over i3.  The per-point arithmetic is the reference macro order
// (src/fdm.h:281-545) with no contraction (built with -ffp-contract=off):
//   O2 : ((A0*u0 + u[+1]) + u[-1]) * inv2
//   O2R: ((A0*u0 + A1*(u[+1]+u[-1])) + ... + AR*(u[+R]+u[-R])) * inv2
// where each pair sum is done in the storage type (both operands are Myfloat).
// hsx_ac2_step / hsx_laplacian / hsx_fd_d2 use the LDS-tiled kernels of
// hsx_stencil_tiled.hip when the grid fits them, else these.
#include "hsx_internal.h"
#include "hsx_stencil_point.h"

namespace hsx {

constexpr int RX = 64, RY = 4;

template <typename TS, typename TC, int R>
__global__ void __launch_bounds__(RX *RY) k_fd_d2_ref(const TS *__restrict__ u, TS *__restrict__ w, Box b,
                                                      int64_t n1, int64_t n12, int64_t stride, Coefs<TC> c,
                                                      TC inv2)
{
    const int64_t i1 = b.s1 + (int64_t)blockIdx.x * RX + threadIdx.x;
    const int64_t i2 = b.s2 + (int64_t)blockIdx.y * RY + threadIdx.y;
    const int64_t i3 = b.s3 + blockIdx.z;
    if (i1 > b.e1 || i2 > b.e2) return;
    const int64_t ii = i1 + i2 * n1 + i3 * n12;
    w[ii] = (TS)d2_point<TS, TC, R>(u, ii, stride, c.A, inv2);
}

template <typename TS, typename TC, int R>
__global__ void __launch_bounds__(RX *RY) k_fd_d1_ref(const TS *__restrict__ u, TS *__restrict__ w, Box b,
                                                      int64_t n1, int64_t n12, int64_t stride, Coefs<TC> c,
                                                      TC inv)
{
    const int64_t i1 = b.s1 + (int64_t)blockIdx.x * RX + threadIdx.x;
    const int64_t i2 = b.s2 + (int64_t)blockIdx.y * RY + threadIdx.y;
    const int64_t i3 = b.s3 + blockIdx.z;
    if (i1 > b.e1 || i2 > b.e2) return;
    const int64_t ii = i1 + i2 * n1 + i3 * n12;
    w[ii] = (TS)d1_point<TS, TC, R>(u, ii, stride, c.D1, inv);
}

template <typename TS, typename TC, int R, int DIM>
__global__ void __launch_bounds__(RX *RY) k_laplacian_ref(const TS *__restrict__ u, TS *__restrict__ w, Box b,
                                                          int64_t n1, int64_t n12, Coefs<TC> c)
{
    const int64_t i1 = b.s1 + (int64_t)blockIdx.x * RX + threadIdx.x;
    const int64_t i2 = b.s2 + (int64_t)blockIdx.y * RY + threadIdx.y;
    const int64_t i3 = b.s3 + blockIdx.z;
    if (i1 > b.e1 || i2 > b.e2) return;
    const int64_t ii = i1 + i2 * n1 + i3 * n12;
    w[ii] = (TS)laplacian_point<TS, TC, R, DIM>(u, ii, n1, n12, c);
}

// out = 2 prc - prn + coef * Lap(prc); out may be prn (the in-place step)
template <typename TS, typename TC, int R, int DIM>
__global__ void __launch_bounds__(RX *RY) k_ac2_ref(TS *out, const TS *prn, const TS *__restrict__ prc,
                                                    const TS *__restrict__ coef, Box b, int64_t n1, int64_t n12,
                                                    Coefs<TC> c)
{
    const int64_t i1 = b.s1 + (int64_t)blockIdx.x * RX + threadIdx.x;
    const int64_t i2 = b.s2 + (int64_t)blockIdx.y * RY + threadIdx.y;
    const int64_t i3 = b.s3 + blockIdx.z;
    if (i1 > b.e1 || i2 > b.e2) return;
    const int64_t ii = i1 + i2 * n1 + i3 * n12;
    const TC lap = laplacian_point<TS, TC, R, DIM>(prc, ii, n1, n12, c);
    out[ii] = leapfrog<TS, TC>(prc[ii], prn[ii], coef[ii], lap);
}

static int ref_grid(const Box &b, dim3 &grid)
{
    if (b.n3() > 65535) return fail(HSX_ERR_UNSUPPORTED, "n3 inner %lld > 65535", (long long)b.n3());
    grid = dim3((unsigned)((b.n1() + RX - 1) / RX), (unsigned)((b.n2() + RY - 1) / RY), (unsigned)b.n3());
    return HSX_OK;
}

static int stencil_checks(const hsx_grid *g, std::initializer_list<std::pair<const void *, const char *>> ptrs)
{
    int rc = check_grid(g);
    for (auto &p : ptrs)
        if (!rc) rc = check_ptr(p.first, p.second);
    if (!rc) rc = ensure_init();
    return rc;
}

} // namespace hsx

using namespace hsx;

extern "C" {

int hsx_fd_d2(const hsx_grid *g, int axis, const void *u, void *w, void *stream)
{
    int rc = stencil_checks(g, {{u, "u"}, {w, "w"}});
    if (rc) return rc;
    if (axis < 0 || axis >= g->dim) return fail(HSX_ERR_INVALID, "axis %d with dim %d", axis, g->dim);
    hipStream_t s = stream_of(stream);
    if (kernel_family() != HSX_KERNEL_REF) {
        rc = tiled_fd_d2(g, axis, u, w, s);
        if (rc != HSX_ERR_UNSUPPORTED || kernel_family() == HSX_KERNEL_TILED) return rc;
    }
    const Box b = box_of(g->inner);
    dim3 grid;
    if ((rc = ref_grid(b, grid))) return rc;
    const int64_t stride = axis == 0 ? 1 : axis == 1 ? g->n1 : g->n1 * g->n2;
    return dispatch_types(g, [&](auto ts, auto tc) -> int {
        using TS = decltype(ts);
        using TC = decltype(tc);
        const Coefs<TC> c = make_coefs<TC>(g);
        return dispatch_R(g->fd_order / 2, [&](auto rc_) -> int {
            constexpr int R = decltype(rc_)::value;
            hipLaunchKernelGGL((k_fd_d2_ref<TS, TC, R>), grid, dim3(RX, RY), 0, s, (const TS *)u, (TS *)w, b,
                               g->n1, g->n1 * g->n2, stride, c, c.inv2[axis]);
            HSX_LAUNCHED("k_fd_d2_ref");
            return HSX_OK;
        });
    });
}

int hsx_fd_d1(const hsx_grid *g, int axis, const void *u, void *w, void *stream)
{
    int rc = stencil_checks(g, {{u, "u"}, {w, "w"}});
    if (rc) return rc;
    if (axis < 0 || axis >= g->dim) return fail(HSX_ERR_INVALID, "axis %d with dim %d", axis, g->dim);
    hipStream_t s = stream_of(stream);
    if (kernel_family() != HSX_KERNEL_REF) {
        rc = tiled_fd_d1(g, axis, u, w, s);
        if (rc != HSX_ERR_UNSUPPORTED || kernel_family() == HSX_KERNEL_TILED) return rc;
    }
    const Box b = box_of(g->inner);
    dim3 grid;
    if ((rc = ref_grid(b, grid))) return rc;
    const int64_t stride = axis == 0 ? 1 : axis == 1 ? g->n1 : g->n1 * g->n2;
    return dispatch_types(g, [&](auto ts, auto tc) -> int {
        using TS = decltype(ts);
        using TC = decltype(tc);
        const Coefs<TC> c = make_coefs<TC>(g);
        return dispatch_R(g->fd_order / 2, [&](auto rc_) -> int {
            constexpr int R = decltype(rc_)::value;
            hipLaunchKernelGGL((k_fd_d1_ref<TS, TC, R>), grid, dim3(RX, RY), 0, s, (const TS *)u, (TS *)w, b,
                               g->n1, g->n1 * g->n2, stride, c, c.inv[axis]);
            HSX_LAUNCHED("k_fd_d1_ref");
            return HSX_OK;
        });
    });
}

int hsx_laplacian(const hsx_grid *g, const void *u, void *w, void *stream)
{
    int rc = stencil_checks(g, {{u, "u"}, {w, "w"}});
    if (rc) return rc;
    hipStream_t s = stream_of(stream);
    if (kernel_family() != HSX_KERNEL_REF) {
        rc = tiled_laplacian(g, u, w, s);
        if (rc != HSX_ERR_UNSUPPORTED || kernel_family() == HSX_KERNEL_TILED) return rc;
    }
    const Box b = box_of(g->inner);
    dim3 grid;
    if ((rc = ref_grid(b, grid))) return rc;
    return dispatch_types(g, [&](auto ts, auto tc) -> int {
        using TS = decltype(ts);
        using TC = decltype(tc);
        const Coefs<TC> c = make_coefs<TC>(g);
        return dispatch_R(g->fd_order / 2, [&](auto rc_) -> int {
            constexpr int R = decltype(rc_)::value;
            return dispatch_dim(g->dim, [&](auto dc) -> int {
                constexpr int DIM = decltype(dc)::value;
                hipLaunchKernelGGL((k_laplacian_ref<TS, TC, R, DIM>), grid, dim3(RX, RY), 0, s, (const TS *)u,
                                   (TS *)w, b, g->n1, g->n1 * g->n2, c);
                HSX_LAUNCHED("k_laplacian_ref");
                return HSX_OK;
            });
        });
    });
}

int hsx_ac2_step(const hsx_grid *g, void *prn, const void *prc, const void *coef, void *stream)
{
    return hsx_ac2_step_to(g, prn, prn, prc, coef, stream);
}

int hsx_ac2_step_to(const hsx_grid *g, void *out, const void *prn, const void *prc, const void *coef, void *stream)
{
    int rc = stencil_checks(g, {{out, "out"}, {prn, "prn"}, {prc, "prc"}, {coef, "coef"}});
    if (rc) return rc;
    if (prn == prc || out == prc) return fail(HSX_ERR_INVALID, "prn / out and prc must be different grids");
    hipStream_t s = stream_of(stream);
    if (kernel_family() != HSX_KERNEL_REF) {
        rc = tiled_ac2_step(g, out, prn, prc, coef, s);
        if (rc != HSX_ERR_UNSUPPORTED || kernel_family() == HSX_KERNEL_TILED) return rc;
    }
    const Box b = box_of(g->inner);
    dim3 grid;
    if ((rc = ref_grid(b, grid))) return rc;
    return dispatch_types(g, [&](auto ts, auto tc) -> int {
        using TS = decltype(ts);
        using TC = decltype(tc);
        const Coefs<TC> c = make_coefs<TC>(g);
        return dispatch_R(g->fd_order / 2, [&](auto rc_) -> int {
            constexpr int R = decltype(rc_)::value;
            return dispatch_dim(g->dim, [&](auto dc) -> int {
                constexpr int DIM = decltype(dc)::value;
                hipLaunchKernelGGL((k_ac2_ref<TS, TC, R, DIM>), grid, dim3(RX, RY), 0, s, (TS *)out,
                                   (const TS *)prn, (const TS *)prc, (const TS *)coef, b, g->n1, g->n1 * g->n2, c);
                HSX_LAUNCHED("k_ac2_ref");
                return HSX_OK;
            });
        });
    });
}

int hsx_ac2_step_fs(const hsx_grid *g, void *prn, const void *prc, const void *coef, int face_mask, int fold,
                    void *stream)
{
    if (fold & ~(HSX_FOLD_PRC | HSX_FOLD_PRN | HSX_FOLD_HALO)) return fail(HSX_ERR_INVALID, "fold flags 0x%x", fold);
    if (fold == 0) return hsx_ac2_step(g, prn, prc, coef, stream);
    int rc = stencil_checks(g, {{prn, "prn"}, {prc, "prc"}, {coef, "coef"}});
    if (rc) return rc;
    if (prn == prc) return fail(HSX_ERR_INVALID, "prn and prc must be different grids");
    if (g->dim != 3 || kernel_family() == HSX_KERNEL_REF)
        return fail(HSX_ERR_UNSUPPORTED, "folded free surface: 3D grids on the tiled kernels only");
    FaceSet fs;
    bool closed = false;
    int64_t most = 0, most_plane = 0;
    if ((rc = make_faceset(g, face_mask, &fs, &closed, &most, &most_plane))) return rc;
    if (!closed) return fail(HSX_ERR_UNSUPPORTED, "folded free surface: fewer than R+2 inner points on an axis");
    int64_t fb[6];
    for (int k = 0; k < 6; k++) fb[k] = (fs.zmask >> k) & 1 ? fs.zidx[k / 2][k % 2] : -1;
    return tiled_ac2_step_fold(g, prn, prc, coef, fold, fb, stream_of(stream));
}

} // extern "C"
