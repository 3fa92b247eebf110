/*
 * hpcscan_oracle.h -- CPU restatement of hpcscan's high-order FD stencil path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the MI355X
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (hpcscan_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm in plain C, keeping the
 * per-point operation order of the reference's fdm.h macros so that a build
 * with -ffp-contract=off reproduces the reference CPU path bit for bit.
 * Each function cites the reference file:line it follows.
 *
 * Types: three storage/compute pairs are instantiated, mirroring the
 * reference's Myfloat / Myfloat2 typedefs (src/type_def.h:21-43):
 *   _f32     storage float,  compute float   (default build)
 *   _f64     storage double, compute double  (storage=fp64 compute=fp64)
 *   _f32c64  storage float,  compute double  (compute=fp64 only)
 */
#ifndef HPCSCAN_ORACLE_H
#define HPCSCAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One axis of the reference grid layout (src/grid.cpp:421-468):
 * Offset | Halo1 | Layer1 | Inner | Layer2 | Halo2 | Pad, all inclusive. */
typedef struct {
    int64_t offset_start, offset_end;
    int64_t halo1_start, halo1_end;
    int64_t layer1_start, layer1_end;
    int64_t inner_start, inner_end;
    int64_t layer2_start, layer2_end;
    int64_t halo2_start, halo2_end;
    int64_t pad_start, pad_end;
    int64_t n;          /* total points along the axis */
} orc_axis;

typedef struct {
    int dim;            /* 1, 2 or 3 */
    int fd_order;       /* 2..16 even */
    int halo;           /* fd_order / 2 */
    orc_axis ax[3];
    int64_t n1, n2, n3, npoint;
} orc_grid;

/* Layout options, -1 meaning "unspecified" as in src/config.cpp:57-65. */
typedef struct {
    int dim;
    int64_t n_inner[3];
    int fd_order;
    int storage_fp64;   /* changes the n1 offset / autoPad rules */
    int nlayer;
    int auto_pad;
    int64_t add_pad[3];
    int64_t mul_pad[3];
    int64_t offset[3];
} orc_layout_opts;

/* Grid::initializeGrid + padGridn* + offsetGridn* (src/grid.cpp:252-468, 839-1009). */
int orc_layout(const orc_layout_opts *o, orc_grid *g);

/* Point-type ranges of Grid::getGridIndex (src/grid.cpp:2680-2843).
 * ptype follows the reference Point_type enum order (src/type_def.h:86-88). */
int orc_index_range(const orc_grid *g, int ptype, int64_t r[6]);

/* FD D2 rational coefficients A0..AR in double (src/fdm.h:276-507). */
int orc_d2_coefs(int fd_order, double a[9]);
/* Staggered first-derivative weights A1..AR, exact rationals (src/fdm.h:29-268). */
int orc_d1_coefs(int fd_order, double a[9]);
/* Sum of |coef| over the full stencil (src/fdm.h:771-787). */
double orc_sum_abs_d2_coef(int fd_order);
/* CFL = sqrt(sum|O2| / (dim * sum|Ok|)) rounded to float (src/propagator_Ac2.cpp:410-439). */
float orc_cfl(int fd_order, int dim);

/* 1D function tables used by Grid::fill (src/grid.cpp:1650-1784).
 * ftype follows Func_type (src/type_def.h:83). FUNC_RANDOM is not supported. */
int orc_fill_table(int ftype, int active, double orig, double d,
                   int64_t i_start, int64_t i_end, double param, double *out);

/* inv2_d = (Myfloat2)(1/d) squared in Myfloat2 (src/grid.cpp:3251-3257). */
float  orc_inv2_f32(double d);
double orc_inv2_f64(double d);

/* Loop schedule: 0 = Baseline loops (src/grid.cpp), 1 = CacheBlk blocks
 * (src/grid_CacheBlk.cpp, block sizes cb[3]). Results are identical. */
typedef struct { int mode; int64_t cb1, cb2, cb3; } orc_sched;

#define ORC_DECLARE(SFX, TS, TC)                                                        \
void   orc_fill_const_##SFX(const orc_grid *g, TS *u, const int64_t r[6], TS val);     \
void   orc_fill_sep_##SFX(const orc_grid *g, TS *u, const int64_t r[6],                 \
                          const double *v1, const double *v2, const double *v3,         \
                          double amp);                                                  \
int    orc_free_surface_##SFX(const orc_grid *g, TS *u);                                \
int    orc_free_surface_mask_##SFX(const orc_grid *g, TS *u, int face_mask);            \
int    orc_fd_d2_##SFX(const orc_grid *g, int axis, const TS *u, TS *w,                 \
                       const TC inv2[3], orc_sched s);                                  \
int    orc_fd_d1_##SFX(const orc_grid *g, int axis, const TS *u, TS *w,                 \
                       const TC inv[3], orc_sched s);                                   \
int    orc_laplacian_##SFX(const orc_grid *g, const TS *u, TS *w, const TC inv2[3],     \
                           orc_sched s);                                                \
int    orc_update_pressure_##SFX(const orc_grid *g, TS *prn, const TS *prc,             \
                                 const TS *coef, const TS *lapla);                      \
int    orc_pressure_fd_##SFX(const orc_grid *g, TS *prn, const TS *prc, const TS *coef, \
                             const TC inv2[3], orc_sched s);                            \
double orc_sum_abs_##SFX(const orc_grid *g, const TS *u, const int64_t r[6]);           \
double orc_sum_abs_diff_##SFX(const orc_grid *g, const TS *u, const TS *v,              \
                              const int64_t r[6]);                                      \
double orc_max_err_##SFX(const orc_grid *g, const TS *u, const TS *v,                   \
                         const int64_t r[6]);                                           \
double orc_l1_err_##SFX(const orc_grid *g, const TS *u, const TS *v,                    \
                        const int64_t r[6]);                                            \
double orc_min_##SFX(const orc_grid *g, const TS *u, const int64_t r[6]);               \
double orc_max_##SFX(const orc_grid *g, const TS *u, const int64_t r[6]);              \
int    orc_forward_##SFX(const orc_grid *g, const TS *coef, const TC inv2[3], int64_t nt,   \
                         int64_t nsrc, const int64_t *src_idx, const TS *wavelet,          \
                         int64_t nrec, const int64_t *rec_idx, TS *traces);                \
int    orc_gradient_##SFX(const orc_grid *g, const TS *coef, const TC inv2[3], double dt,   \
                          int64_t nt, int64_t nsrc, const int64_t *src_idx,                \
                          const TS *wavelet, int64_t nrec, const int64_t *rec_idx,         \
                          const TS *d_obs, int residual_given, double *grad, double *misfit);

ORC_DECLARE(f32, float, float)
ORC_DECLARE(f64, double, double)
ORC_DECLARE(f32c64, float, double)

#undef ORC_DECLARE

/* Eigen-mode propagation driver: the loop of src/testCase_Propa.cpp:111-181
 * for one try, with the propagator step of src/propagator_Ac2.cpp:366-393.
 * Returns errTestCase (float, like the reference) and fills the final prn. */
typedef struct {
    int dim, fd_order, nt, snap_inc, split_comp;
    double dt;          /* the Myfloat dt value (float-representable in FP32 builds) */
    int mode[3];        /* ceil(param1..3) */
    double d[3];        /* unit-grid sampling */
    double orig[3];     /* float-truncated unit-grid origins */
} orc_propa_params;

float orc_propa_eigen_f32(const orc_grid *g, const orc_propa_params *p, orc_sched s,
                          float *prn_out, float *prc_out);
float  orc_propa_eigen_f64(const orc_grid *g, const orc_propa_params *p, orc_sched s,
                           double *prn_out, double *prc_out);

#ifdef __cplusplus
}
#endif

#endif
