/*
 * hpcscan_mi355x.h -- C ABI of the MI355X (gfx950) stencil library.
 *
 * This is the device boundary of the hpcscan stencil path: every virtual of
 * the reference's Grid plugin interface that touches grid values
 * (src/grid.h:49-245) maps onto one entry point below, called by the C++
 * Grid_Mi355x backend (hpcscan_amd/host/grid_Mi355x.cpp) and, for tests and
 * benchmarks, through ctypes.  Plain pointers and sizes only: device buffers
 * are void*, HIP streams are passed as void* (NULL = the library's stream).
 *
 * Conventions
 *  - Return value: HSX_OK (0) or a negative HSX_ERR_* code; never throws,
 *    never exits.  hsx_last_error() gives the message of the last failure of
 *    the calling thread.
 *  - Grid operations are asynchronous on the given stream unless noted;
 *    reductions synchronise the stream and return their result on the host.
 *  - Arithmetic follows the reference per-point operation order with no
 *    fused multiply-add, so results equal the reference CPU path bit for bit
 *    (reductions: up to the order of a double-precision sum).
 */
#ifndef HPCSCAN_MI355X_H
#define HPCSCAN_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSX_ABI_VERSION 1

/* status codes */
#define HSX_OK               0
#define HSX_ERR_INVALID     -1   /* bad argument (sizes, ranges, order, precision) */
#define HSX_ERR_HIP         -2   /* HIP runtime error */
#define HSX_ERR_NOMEM       -3   /* device allocation failed */
#define HSX_ERR_UNSUPPORTED -4   /* valid request this build does not implement */

/* precisions (reference Myfloat / Myfloat2, src/type_def.h:21-43) */
#define HSX_FP32 0
#define HSX_FP64 1

/* stencil kernel families (hsx_set_kernel) */
#define HSX_KERNEL_AUTO  0   /* tiled kernels where they apply, else reference kernels */
#define HSX_KERNEL_REF   1   /* one thread per point, reference loop semantics */
#define HSX_KERNEL_TILED 2   /* LDS-tiled n3-streaming kernels (3D only) */

/*
 * Grid descriptor: the index layout of the reference Grid (src/grid.cpp:17-38,
 * 252-832).  Flat index i1 + i2*n1 + i3*n1*n2; all ranges inclusive.
 * Inactive axes (dim < 3) have n = 1 and all-zero ranges, as in the reference.
 */
typedef struct {
    int32_t dim;            /* 1, 2 or 3 */
    int32_t fd_order;       /* 2, 4, ..., 16 */
    int32_t storage;        /* HSX_FP32 or HSX_FP64: type of grid values (Myfloat) */
    int32_t compute;        /* HSX_FP32 or HSX_FP64: arithmetic type (Myfloat2) */
    int64_t n1, n2, n3;     /* allocated points per axis */
    int64_t inner[6];       /* i1Start,i1End,i2Start,i2End,i3Start,i3End of INNER_POINTS */
    int64_t halo1[6];       /* per axis: Halo1Start, Halo1End (axes 1,2,3) */
    int64_t halo2[6];       /* per axis: Halo2Start, Halo2End */
    double d[3];            /* grid spacing per axis (inv2 = (Myfloat2)(1/d) squared) */
} hsx_grid;

typedef struct {
    char name[128];
    char arch[64];
    int32_t compute_units;
    int32_t clock_mhz;
    int32_t mem_clock_mhz;
    int32_t bus_width_bits;
    int64_t hbm_bytes;
    int32_t lds_bytes_per_block;
    int32_t l2_bytes;
} hsx_device_props;

/* ---- runtime ---------------------------------------------------------- */
int         hsx_abi_version(void);
const char *hsx_last_error(void);
int         hsx_init(int device);
int         hsx_device_count(int *count);
int         hsx_device_info(int device, hsx_device_props *out);
int         hsx_malloc(void **ptr, size_t bytes);
int         hsx_free(void *ptr);
int         hsx_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int         hsx_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int         hsx_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int         hsx_stream_create(void **stream);
int         hsx_stream_destroy(void *stream);
int         hsx_stream_sync(void *stream);
int         hsx_event_create(void **event);
int         hsx_event_destroy(void *event);
int         hsx_event_record(void *event, void *stream);
int         hsx_event_elapsed_ms(void *start, void *stop, float *ms);
int         hsx_set_kernel(int family);
int         hsx_get_kernel(void);
/* Contracted arithmetic (opt-in; no counterpart in the reference): with on = 1
 * the tiled fused step (hsx_ac2_step, hsx_ac2_step_to) and the tiled Laplacian
 * (hsx_laplacian) evaluate each weighted sum A_j * (u+j + u-j) + acc, and
 * coef * lap + (2 prc - prn), as one fused multiply-add: about a third fewer
 * VALU operations at order 16, where those kernels are bound by VALU issue.
 * Results then differ from the reference CPU path by rounding (one rounding
 * per fused pair instead of two), no longer bit for bit.  The reference-order
 * kernels (grids the tiled kernels do not take) and the folded free surface
 * stay exact.  Process-wide; default 0. */
int         hsx_set_contract(int on);
int         hsx_get_contract(void);
/* Tile shape of the tiled kernels (0 = the library's choice for each):
 *   lanes_n1      16, 32 or 64 lanes of 16 bytes along n1 (default by grid width),
 *                 or 128 (8 rows with the halo ring: the exact FP32 fused step
 *                 and Laplacian up to order 8; the other kernels, precisions and
 *                 orders then keep the library's choice)
 *   rows_n2       8 or 16 rows along n2 (16 needs lanes_n1 >= 32 and its LDS
 *                 planes within the 160 KiB of a workgroup; otherwise 8)
 *   chunk_planes  n3 planes streamed by one workgroup
 *   flags         HSX_TILE_CACHED or HSX_TILE_NT: how prn, coef and the output
 *                 are accessed (nontemporal keeps them out of the caches);
 *                 HSX_TILE_NO_RING: load each plane's halo one plane ahead
 *                 instead of with the queue front (R+1 planes ahead)
 * Process-wide; results do not depend on it. */
#define HSX_TILE_CACHED  1
#define HSX_TILE_NT      2
#define HSX_TILE_NO_RING 4
int         hsx_set_tile(int lanes_n1, int rows_n2, int chunk_planes, int flags);
/* Tile autotuner (no counterpart in the reference): times the candidate
 * shapes of the fused step on scratch grids of g's layout (reps launches each)
 * and keeps the fastest for g's layout, FD order and precisions; the fused
 * step on such grids then uses it unless hsx_set_tile forces a shape.  shape[4] (may be NULL) = lanes, rows, chunk, flags of the choice;
 * *best_ms (may be NULL) its time per launch.  Allocates 3 grids while it runs. */
int         hsx_autotune(const hsx_grid *g, int reps, int shape[4], float *best_ms);
int         hsx_autotune_clear(void);

/* ---- grid filling: Grid::fill (src/grid.cpp:1650-1826) -------------------- */
/* u[range] = val (converted to the storage type) */
int hsx_fill_const(const hsx_grid *g, void *u, const int64_t range[6], double val, void *stream);
/* u[i1,i2,i3] = (Myfloat)(amp * t1[i1-r0] * t2[i2-r2] * t3[i3-r4]) over range;
 * t1/t2/t3 are HOST tables of the 1D factors in double */
int hsx_fill_separable(const hsx_grid *g, void *u, const int64_t range[6], const double *t1,
                       const double *t2, const double *t3, double amp, void *stream);

/* ---- Grid array operations over all npoint values (src/grid.cpp:1830-1920) */
#define HSX_ARRAY_FILL     0   /* w = val       */
#define HSX_ARRAY_COPY     1   /* w = u         */
#define HSX_ARRAY_ADD      2   /* w = u + v     */
#define HSX_ARRAY_MULTIPLY 3   /* w = u * v     */
#define HSX_ARRAY_ADDTO    4   /* w = w + u     */
int hsx_array_op(const hsx_grid *g, int op, void *w, const void *u, const void *v, double val,
                 void *stream);

/* ---- the stencil path --------------------------------------------------- */
/* Grid::applyBoundaryCondition(BOUND_COND_ANTI_MIRROR) (src/grid.cpp:3065-3230):
 * free surface on the faces whose bit is set in face_mask (bit 2a = axis a
 * Halo1 side, bit 2a+1 = Halo2 side; 0x3f = all faces, i.e. no neighbour) */
int hsx_free_surface(const hsx_grid *g, void *u, int face_mask, void *stream);
/* Grid::FD_D2_N1/N2/N3 (src/grid.cpp:3232-4065): w = D2 along axis (0,1,2) of u, inner points */
int hsx_fd_d2(const hsx_grid *g, int axis, const void *u, void *w, void *stream);
/* Grid::FD_D1_N1/N2/N3 (src/grid.cpp:3401-3560, 3735-3894, 4069-4229): w = staggered
 * first derivative along axis (0,1,2) of u, i.e. at i - 1/2, inner points */
int hsx_fd_d1(const hsx_grid *g, int axis, const void *u, void *w, void *stream);
/* Grid::FD_LAPLACIAN (src/grid.cpp:4235-4577): w = Laplacian(u), inner points */
int hsx_laplacian(const hsx_grid *g, const void *u, void *w, void *stream);
/* Grid::updatePressure (src/grid.cpp:4579-4610): prn = 2 prc - prn + coef*lapla */
int hsx_update_pressure(const hsx_grid *g, void *prn, const void *prc, const void *coef,
                        const void *lapla, void *stream);
/* Grid::computePressureWithFD (src/grid.cpp:4612-5112): prn = 2 prc - prn + coef*Lap(prc) */
int hsx_ac2_step(const hsx_grid *g, void *prn, const void *prc, const void *coef, void *stream);
/* the same step out of place: out = 2 prc - prn + coef*Lap(prc) (out may be prn,
 * neither may be prc); used where the old wavefield must be kept (the adjoint's
 * recomputation of a checkpoint segment).  No counterpart in the reference. */
int hsx_ac2_step_to(const hsx_grid *g, void *out, const void *prn, const void *prc, const void *coef,
                    void *stream);
/* Propagator_Ac2::computeWavefieldNextTimeStep's two device calls in one launch
 * (src/propagator_Ac2.cpp:366-393: applyBoundaryCondition(FREESURF) on prc, then
 * computePressureWithFD), with the free surface on the faces of face_mask folded
 * into the step.  fold:
 *   HSX_FOLD_PRC   prc is read as the free surface leaves it: its inner points on
 *                  the boundary planes read as 0.  Its halos must already hold the
 *                  free-surface values of its inner points (written by the previous
 *                  folded step with HSX_FOLD_HALO, or by hsx_free_surface).
 *   HSX_FOLD_PRN   the same for prn (the grid the free surface was applied to one
 *                  step earlier, still only logically).
 *   HSX_FOLD_HALO  after the step, prn's halos on the faces of face_mask hold the
 *                  free-surface values of its new inner points (its boundary planes
 *                  keep the computed values, as after the reference step).
 * Inner results equal hsx_free_surface(prc) [+ hsx_free_surface(prn)] followed by
 * hsx_ac2_step bit for bit; prc is not written.  Returns HSX_ERR_UNSUPPORTED when
 * the grid does not take the folded kernel (not 3D, fewer than R+2 inner points on
 * a folded axis, reference kernels forced): the caller then applies the free
 * surface and steps separately. */
#define HSX_FOLD_PRC 1
#define HSX_FOLD_PRN 2
#define HSX_FOLD_HALO 4
int hsx_ac2_step_fs(const hsx_grid *g, void *prn, const void *prc, const void *coef, int face_mask, int fold,
                    void *stream);

/* ---- reductions: Grid::getSumAbs ... (src/grid.cpp:2006-2370) -------------
 * All accumulate in double on the device; the result is returned on the host. */
int hsx_sum_abs(const hsx_grid *g, const void *u, const int64_t range[6], double *out, void *stream);
int hsx_sum_abs_diff(const hsx_grid *g, const void *u, const void *v, const int64_t range[6],
                     double *out, void *stream);
/* out[0] = sum |u - v|, out[1] = sum |v| (Grid::L1Err before its divide) */
int hsx_l1_sums(const hsx_grid *g, const void *u, const void *v, const int64_t range[6],
                double out[2], void *stream);
/* hsx_l1_sums without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight time loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  out is a DEVICE
 * array of 2 doubles (sum |u - v|, sum |v|), written asynchronously on the
 * stream (no host wait).  The partial sums go to a 16 KiB array of the library
 * (at most 1024 blocks; every range fits, its rows are strided over them), one
 * per device for the life of the process; calls that overlap on one device
 * (other streams, other threads) share it and race. */
int hsx_l1_sums_noalloc(const hsx_grid *g, const void *u, const void *v, const int64_t range[6], double *out,
                        void *stream);
/* max over the range of |u-v|/|v|, or |u-v| where |(float)v| < max_err_float */
int hsx_max_err(const hsx_grid *g, const void *u, const void *v, const int64_t range[6],
                double max_err_float, double *out, void *stream);
/* out[0] = min(FLT_MAX, u...), out[1] = max(-FLT_MAX, u...) in the storage type */
int hsx_min_max(const hsx_grid *g, const void *u, const int64_t range[6], double out[2],
                void *stream);

/* ---- point sources and receivers (DataAcquisition, src/data_Acquisition.cpp:58-305) ----
 * idx, val, w and out are DEVICE arrays; idx holds flat grid indices, distinct
 * within one call; values are in the storage type. */
#define HSX_INJECT_ADD 0   /* u[idx[i]] += scale * val[i] * w[i]   (w may be NULL: 1) */
#define HSX_INJECT_SET 1   /* u[idx[i]]  = scale * val[i] * w[i]   (the reference's source term) */
int hsx_inject(const hsx_grid *g, void *u, int64_t n, const int64_t *idx, const void *val, const void *w,
               double scale, int mode, void *stream);
/* out[i] = u[idx[i]] (DataAcquisition::recordTrace) */
int hsx_gather(const hsx_grid *g, const void *u, int64_t n, const int64_t *idx, void *out, void *stream);

/* ---- adjoint of the Ac2 propagator (no counterpart in the reference: its
 * ADJOINT mode is a stub, src/data_Acquisition.cpp:307-310) ----------------
 * Imaging condition, accumulated in double over range:
 *   grad += scale * mu * (u_next - 2 u + u_prev - f),  f = src_val at src_idx. */
int hsx_imaging(const hsx_grid *g, double *grad, const void *mu, const void *u_next, const void *u,
                const void *u_prev, const int64_t range[6], double scale, int64_t src_idx, double src_val,
                void *stream);

/* ---- low-precision checkpoints of FP32 wavefields ----------------------------
 * A snapshot of n values takes hsx_ckpt_bytes(n) bytes: a 256-byte header
 * (scale, max |value|) and n 16-bit codes, code = round(v / scale) with
 * scale = max|v| / 16384 (FP16) or max|v| (BF16), padded to a multiple of 256
 * bytes so that snapshots stored back to back stay aligned.  packed must be
 * 16-byte aligned (HSX_ERR_INVALID otherwise).  With u_prev non-NULL the
 * snapshot holds u - u_prev; decoding with base non-NULL gives base - value. */
#define HSX_CKPT_FP32 0   /* checkpoints kept in the storage type (bit-exact recomputation) */
#define HSX_CKPT_FP16 1
#define HSX_CKPT_BF16 2
size_t hsx_ckpt_bytes(int64_t n);
int hsx_ckpt_encode(const float *u, const float *u_prev, int64_t n, void *packed, int format, void *stream);
int hsx_ckpt_decode(const void *packed, int64_t n, int format, const float *base, float *out, void *stream);

/* ---- one shot: forward modeling and the misfit gradient -----------------------
 * Step n (0 <= n < nt) is the propagator step (free surface, fused Ac2 step) of
 * q^{n+1} from q^n, q^{n-1} (q^0 = q^-1 = 0), then wavelet[n*nsrc + i] is added
 * at src_idx[i], then traces[n*nrec + r] = q^{n+1}[rec_idx[r]].  All arrays
 * DEVICE, values in the storage type. */
typedef struct {
    int64_t nt;
    double dt;                 /* time step (scales the imaging condition) */
    int64_t nsrc;
    const int64_t *src_idx;
    const void *wavelet;       /* nt x nsrc */
    int64_t nrec;
    const int64_t *rec_idx;
} hsx_shot;

int hsx_forward(const hsx_grid *g, const void *coef, const hsx_shot *shot, void *traces, void *stream);

/* d_obs is the residual itself (the adjoint source, e.g. the gradient of any
 * loss with respect to the traces): res = d_obs instead of traces - d_obs, so
 * that it reaches the adjoint without being rounded against the traces */
#define HSX_ADJ_RESIDUAL 1
/* hsx_adjoint_plan_create tunes the tile shape of the grid's layout
 * (hsx_autotune) before it allocates the work buffers */
#define HSX_ADJ_AUTOTUNE 2

typedef struct {
    int ckpt_every;            /* steps per checkpoint segment; 0 = keep every wavefield */
    int ckpt_format;           /* HSX_CKPT_FP32 / FP16 / BF16 (16-bit: FP32 grids only) */
    int flags;                 /* 0 or HSX_ADJ_RESIDUAL | HSX_ADJ_AUTOTUNE */
} hsx_adjoint_opts;

/* grad (DEVICE, double, npoint values) += dJ/dm for J = 1/2 sum res^2,
 * res = traces - d_obs (or d_obs with HSX_ADJ_RESIDUAL), m = dt^2 / coef the
 * squared slowness, over the inner points not on a boundary plane (source
 * terms outside that box included: they add nothing); *misfit = J (may be
 * NULL).  Forward pass with checkpoints, then the adjoint backwards,
 * recomputing one segment at a time.  Allocates its work buffers (a plan, see
 * below) and returns after the stream has finished with them. */
int hsx_gradient(const hsx_grid *g, const void *coef, const hsx_shot *shot, const void *d_obs,
                 const hsx_adjoint_opts *opts, double *grad, double *misfit, void *stream);

/* The work buffers of hsx_gradient for one grid layout, nt, receiver count
 * and options, allocated once: hsx_gradient_plan then runs gradients of shots
 * of that geometry (any sources, receiver positions, coef, data) without
 * allocating.  A plan serves one call at a time. */
typedef struct hsx_adjoint_plan hsx_adjoint_plan;
int hsx_adjoint_plan_create(const hsx_grid *g, int64_t nt, int64_t nrec, const hsx_adjoint_opts *opts,
                            hsx_adjoint_plan **out);
int hsx_gradient_plan(hsx_adjoint_plan *plan, const hsx_grid *g, const void *coef, const hsx_shot *shot,
                      const void *d_obs, double *grad, double *misfit, void *stream);
int hsx_adjoint_plan_destroy(hsx_adjoint_plan *plan);

/* ---- batched shots (no counterpart in the reference: one grid per process) ----
 * nshots forward modelings (hsx_forward of each shot, same results) advanced
 * together: per time step one launch of the free surface pair, one fused-step
 * launch over the tiles of every grid and one source/receiver launch.  Grids:
 * 3D, one FD order and precision (storage = compute), any sizes, n1 a multiple
 * of 16 bytes; every shot the same nt.  coef[s] and traces[s] are DEVICE
 * buffers kept alive by the caller; the batch owns the wavefields. */
typedef struct hsx_batch hsx_batch;
int hsx_batch_create(int nshots, const hsx_grid *grids, const void *const *coef, const hsx_shot *shots,
                     void *const *traces, hsx_batch **out);
/* all nt steps from zero wavefields; use_graph: capture the time loop into a
 * HIP graph on the first call, replay it on every call */
int hsx_batch_run(hsx_batch *b, int use_graph, void *stream);
/* Tile autotuner of a batch (no counterpart in the reference): times the
 * batched fused step with each candidate shape -- lanes x rows (64 / 32 x 16 / 8),
 * halo ring on / off, n3 chunk (the rule, 16, 32, 64 planes, or -1: one
 * persistent launch, each resident workgroup streaming an equal share of the
 * shots' whole tile columns) -- for reps launches, starting with the shape
 * hsx_batch_create chose, and keeps the fastest for later runs (a captured
 * time loop is captured again).  shape[4] (may be NULL) = lanes, rows, chunk
 * (0: the rule, -1: persistent), ring of the choice; *best_ms (may be NULL)
 * its time per step.  The wavefields are overwritten (every run starts from
 * zero anyway). */
int hsx_batch_autotune(hsx_batch *b, int reps, int shape[4], float *best_ms);
/* The batched step with a given shape (the values hsx_batch_autotune reports):
 * HSX_ERR_UNSUPPORTED when it does not fit this order and precision (LDS,
 * registers); a captured time loop is captured again.  No counterpart in the
 * reference. */
int hsx_batch_set_shape(hsx_batch *b, int lanes, int rows, int chunk, int ring);
int hsx_batch_destroy(hsx_batch *b);

#ifdef __cplusplus
}
#endif

#endif
