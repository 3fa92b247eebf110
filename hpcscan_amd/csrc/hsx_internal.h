// Internal helpers shared by the HIP translation units of libhpcscan_mi355x.
#ifndef HSX_INTERNAL_H
#define HSX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "hpcscan_mi355x.h"

namespace hsx {

// ---- errors ---------------------------------------------------------------
int fail(int code, const char *fmt, ...);            // record message, return code
int hip_fail(hipError_t e, const char *what);         // HSX_ERR_HIP with the HIP message

#define HSX_HIP(call)                                                      \
    do {                                                                   \
        hipError_t e_ = (call);                                            \
        if (e_ != hipSuccess) return ::hsx::hip_fail(e_, #call);           \
    } while (0)

// launch-error check after hipLaunchKernelGGL
#define HSX_LAUNCHED(name)                                                 \
    do {                                                                   \
        hipError_t e_ = hipGetLastError();                                 \
        if (e_ != hipSuccess) return ::hsx::hip_fail(e_, name);            \
    } while (0)

// ---- runtime state ----------------------------------------------------------
int ensure_init(void);                 // lazily select device 0 and create the library stream
hipStream_t stream_of(void *s);        // NULL -> library stream
int kernel_family(void);               // HSX_KERNEL_*
int contract_mode(void);               // hsx_set_contract: 1 = contracted tiled step / Laplacian

// device scratch owned by the library (grown on demand, freed at exit)
int scratch_get(int slot, size_t bytes, void **ptr);
enum { SCRATCH_TABLES = 0, SCRATCH_PARTIALS = 1, SCRATCH_RESULT = 2, SCRATCH_SLOTS = 4 };

// ---- descriptor checks ----------------------------------------------------------
int check_grid(const hsx_grid *g);
int check_range(const hsx_grid *g, const int64_t r[6]);
int check_ptr(const void *p, const char *name);

// ---- index helpers ------------------------------------------------------------
struct Box {                    // inclusive index box
    int64_t s1, e1, s2, e2, s3, e3;
    __host__ __device__ int64_t n1() const { return e1 - s1 + 1; }
    __host__ __device__ int64_t n2() const { return e2 - s2 + 1; }
    __host__ __device__ int64_t n3() const { return e3 - s3 + 1; }
};
inline Box box_of(const int64_t r[6]) { return Box{r[0], r[1], r[2], r[3], r[4], r[5]}; }

// ---- free-surface faces (src/grid.cpp:3065-3230) ----------------------------------
struct Face {
    int axis;
    int64_t hs, he, inner;      // halo range and zeroed plane along 'axis'
    int64_t os, oe, bs, be;     // ranges of the outer / inner other axes
    int64_t so, sb, sa;         // strides of outer, inner-other and face axes
};
struct FaceSet {
    Face f[6];
    int n;                       // faces in f
    int64_t zidx[3][2];          // boundary plane index per axis / side
    int zmask;                   // faces of the mask
    int vec1;                    // n2 / n3 faces may use 16-byte groups along n1 (n1 a multiple
                                 // of the vector width, grid 16-byte aligned)
};
// faces of face_mask; closed_form: the two-launch form applies; most / most_plane:
// largest halo slab and boundary plane (points), for the launch sizes
int make_faceset(const hsx_grid *g, int face_mask, FaceSet *fs, bool *closed_form, int64_t *most,
                 int64_t *most_plane);
// free surface of many grids (closed form only): d_fields / d_sets device arrays
int bc_batch_launch(int storage, int nshots, void *const *d_fields, const FaceSet *d_sets, int64_t most,
                    int64_t most_plane, hipStream_t s);
// the n1 faces in one 16-byte pass per row end (k_bc_n1): sides (bit 0 / 1 = low /
// high face, 0 = not applicable) when R is a multiple of the vector width, n1
// a multiple of it, u 16-byte aligned and the halo groups aligned
int n1_vector_sides(const hsx_grid *g, const void *u, const FaceSet &fs, Face *lo, Face *hi);
// whether the n2 / n3 faces of fs may run in 16-byte groups along n1 (FaceSet::vec1)
int n1_groups_ok(const hsx_grid *g, const void *u, const FaceSet &fs);
struct N1Faces {
    Face lo, hi;
    int sides;                   // 0: this grid's n1 faces stay with the generic kernels
};
// the whole free surface of many grids with n1 faces in d_n1 (per grid; nh = R /
// vector width, the same for all; most_rows = the largest number of rows of an
// n1 face) and the other faces in d_sets: n1 rows and mirrors in one launch,
// then the boundary planes
int bc_batch_fs_launch(int storage, int nshots, void *const *d_fields, const N1Faces *d_n1, int nh,
                       int64_t most_rows, const FaceSet *d_sets, int64_t most, int64_t most_plane, hipStream_t s);

// ---- coefficients in the compute type (src/fdm.h:276-507, src/grid.cpp:3251-3257) --
template <typename TC>
struct Coefs {
    TC A[9];        // A0..AR
    TC inv2[3];     // (TC)(1/d) squared in TC
    TC D1[9];       // staggered first-derivative weights, D1[1..R]
    TC inv[3];      // (TC)(1/d)
};
double fd_coef_double(int fd_order, int j);
double fd1_coef_double(int fd_order, int j);
template <typename TC>
inline Coefs<TC> make_coefs(const hsx_grid *g)
{
    Coefs<TC> c;
    const int R = g->fd_order / 2;
    for (int j = 0; j < 9; j++) c.A[j] = j <= R ? (TC)fd_coef_double(g->fd_order, j) : (TC)0;
    for (int j = 0; j < 9; j++) c.D1[j] = j >= 1 && j <= R ? (TC)fd1_coef_double(g->fd_order, j) : (TC)0;
    for (int a = 0; a < 3; a++) {
        const TC inv = (TC)(1.0 / g->d[a]);
        c.inv2[a] = inv * inv;
        c.inv[a] = inv;
    }
    return c;
}

// ---- type / order dispatch ----------------------------------------------------
template <typename F>
inline int dispatch_types(const hsx_grid *g, F &&f)
{
    if (g->storage == HSX_FP32 && g->compute == HSX_FP32) return f(float(), float());
    if (g->storage == HSX_FP64 && g->compute == HSX_FP64) return f(double(), double());
    if (g->storage == HSX_FP32 && g->compute == HSX_FP64) return f(float(), double());
    return fail(HSX_ERR_UNSUPPORTED, "storage/compute precision pair %d/%d not supported", g->storage,
                g->compute);
}

template <int V>
using IC = std::integral_constant<int, V>;

template <typename F>
inline int dispatch_R(int R, F &&f)
{
    switch (R) {
    case 1: return f(IC<1>());
    case 2: return f(IC<2>());
    case 3: return f(IC<3>());
    case 4: return f(IC<4>());
    case 5: return f(IC<5>());
    case 6: return f(IC<6>());
    case 7: return f(IC<7>());
    case 8: return f(IC<8>());
    default: return fail(HSX_ERR_INVALID, "halo width %d not in 1..8", R);
    }
}

template <typename F>
inline int dispatch_dim(int dim, F &&f)
{
    switch (dim) {
    case 1: return f(IC<1>());
    case 2: return f(IC<2>());
    case 3: return f(IC<3>());
    default: return fail(HSX_ERR_INVALID, "dim %d not in 1..3", dim);
    }
}

// tiled kernels (hsx_tiled_*.hip); return HSX_ERR_UNSUPPORTED when the
// grid does not fit their assumptions so the caller uses the reference kernels
int tiled_ac2_step(const hsx_grid *g, void *out, const void *prn, const void *prc, const void *coef, hipStream_t s);
// in-place step with the free surface folded in (flags: FOLD_* of hsx_tiled.h;
// fb: boundary plane of each folded face, -1 for the others)
int tiled_ac2_step_fold(const hsx_grid *g, void *prn, const void *prc, const void *coef, int flags,
                        const int64_t fb[6], hipStream_t s);
int tiled_laplacian(const hsx_grid *g, const void *u, void *w, hipStream_t s);
// the contracted (FMA) instances, in translation units of their own
int tiled_ac2_step_fma(const hsx_grid *g, void *out, const void *prn, const void *prc, const void *coef,
                       hipStream_t s);
int tiled_laplacian_fma(const hsx_grid *g, const void *u, void *w, hipStream_t s);
// the fused step in place on prn with grad += scale * new prn * (q0 - 2 q1 + q2)
// over box (inside the inner points), in double (the adjoint's imaging condition)
int tiled_ac2_step_img(const hsx_grid *g, void *prn, const void *prc, const void *coef, double *grad,
                       const void *q0, const void *q1, const void *q2, const int64_t box[6], double scale,
                       hipStream_t s);
int tiled_fd_d2_x(const hsx_grid *g, const void *u, void *w, hipStream_t s);
int tiled_fd_d2_y(const hsx_grid *g, const void *u, void *w, hipStream_t s);
int tiled_fd_d2_z(const hsx_grid *g, const void *u, void *w, hipStream_t s);
inline int tiled_fd_d2(const hsx_grid *g, int axis, const void *u, void *w, hipStream_t s)
{
    return axis == 0 ? tiled_fd_d2_x(g, u, w, s) : axis == 1 ? tiled_fd_d2_y(g, u, w, s) : tiled_fd_d2_z(g, u, w, s);
}
int tiled_fd_d1_x(const hsx_grid *g, const void *u, void *w, hipStream_t s);
int tiled_fd_d1_y(const hsx_grid *g, const void *u, void *w, hipStream_t s);
int tiled_fd_d1_z(const hsx_grid *g, const void *u, void *w, hipStream_t s);
inline int tiled_fd_d1(const hsx_grid *g, int axis, const void *u, void *w, hipStream_t s)
{
    return axis == 0 ? tiled_fd_d1_x(g, u, w, s) : axis == 1 ? tiled_fd_d1_y(g, u, w, s) : tiled_fd_d1_z(g, u, w, s);
}
int set_tile(int lx, int by, int chunk, int streams);

} // namespace hsx

#endif
