/*
 * hpcscan_oracle.c -- CPU restatement of hpcscan's FD stencil path.
 *
 * TEST INFRASTRUCTURE ONLY (see hpcscan_oracle.h).  Build with
 * -ffp-contract=off: the reference's own Makefile flags (-O3 -mavx2, no -mfma)
 * never contract a*b+c, and the oracle must not either.
 */
#include "hpcscan_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * FD coefficients (src/fdm.h:276-507), as the rationals the reference writes.
 * ------------------------------------------------------------------------- */
int orc_d2_coefs(int fd_order, double a[9])
{
    for (int j = 0; j < 9; j++) a[j] = 0.0;
    switch (fd_order) {
    case 2:  a[0] = -2.0; a[1] = 1.0; break;
    case 4:  a[0] = -5. / 2.; a[1] = 4. / 3.; a[2] = -1. / 12.; break;
    case 6:  a[0] = -49. / 18.; a[1] = 3. / 2.; a[2] = -3. / 20.; a[3] = 1. / 90.; break;
    case 8:  a[0] = -205. / 72.; a[1] = 8. / 5.; a[2] = -1 / 5.; a[3] = 8. / 315.;
             a[4] = -1 / 560.; break;
    case 10: a[0] = -5269. / 1800.; a[1] = 5. / 3.; a[2] = -5. / 21.; a[3] = 5. / 126.;
             a[4] = -5. / 1008.; a[5] = 1. / 3150.; break;
    case 12: a[0] = -5369. / 1800.; a[1] = 12. / 7.; a[2] = -15. / 56.; a[3] = 10. / 189.;
             a[4] = -1. / 112.; a[5] = 2. / 1925.; a[6] = -1. / 16632.; break;
    case 14: a[0] = -266681. / 88200.; a[1] = 7. / 4.; a[2] = -7. / 24.; a[3] = 7. / 108.;
             a[4] = -7. / 528.; a[5] = 7. / 3300.; a[6] = -7. / 30888.; a[7] = 1. / 84084.;
             break;
    case 16: a[0] = -1077749. / 352800.; a[1] = 16. / 9.; a[2] = -14. / 45.;
             a[3] = 112. / 1485.; a[4] = -7. / 396.; a[5] = 112. / 32175.;
             a[6] = -2. / 3861.; a[7] = 16. / 315315.; a[8] = -1. / 411840.; break;
    default: return -1;
    }
    return 0;
}

int orc_d1_coefs(int fd_order, double a[9])
{
    /* staggered first-derivative weights A1..AR (src/fdm.h:29-268); a[0] unused */
    for (int j = 0; j < 9; j++) a[j] = 0.0;
    switch (fd_order) {
    case 2:  a[1] = 1.0; break;
    case 4:  a[1] = 9. / 8.; a[2] = -1. / 24.; break;
    case 6:  a[1] = 75. / 64.; a[2] = -25. / 384.; a[3] = 3. / 640.; break;
    case 8:  a[1] = 1225. / 1024.; a[2] = -245. / 3072.; a[3] = 49. / 5120.; a[4] = -5. / 7168.; break;
    case 10: a[1] = 19845. / 16384.; a[2] = -735. / 8192.; a[3] = 567. / 40960.; a[4] = -405. / 229376.;
             a[5] = 35. / 294912.; break;
    case 12: a[1] = 160083. / 131072.; a[2] = -12705. / 131072.; a[3] = 22869. / 1310720.;
             a[4] = -5445. / 1835008.; a[5] = 847. / 2359296.; a[6] = -63. / 2883584.; break;
    case 14: a[1] = 1288287. / 1048576.; a[2] = -429429. / 4194304.; a[3] = 429429. / 20971520.;
             a[4] = -61347. / 14680064.; a[5] = 13013. / 18874368.; a[6] = -3549. / 46137344.;
             a[7] = 231. / 54525952.; break;
    case 16: a[1] = 41409225. / 33554432.; a[2] = -3578575. / 33554432.; a[3] = 3864861. / 167772160.;
             a[4] = -1254825. / 234881024.; a[5] = 325325. / 301989888.; a[6] = -61425. / 369098752.;
             a[7] = 7425. / 436207616.; a[8] = -143. / 167772160.; break;
    default: return -1;
    }
    return 0;
}

double orc_sum_abs_d2_coef(int fd_order)
{
    /* getSumAbsFD_D2Coef (src/fdm.h:771-787): sum over the 2R+1 Myfloat2
     * coefficients, each widened to double, in vector order AR..A1 A0 A1..AR. */
    double a[9];
    if (orc_d2_coefs(fd_order, a)) return -1.0;
    const int R = fd_order / 2;
    double s = 0.0;
    for (int j = R; j >= 1; j--) s += fabs((double)(float)a[j]);
    s += fabs((double)(float)a[0]);
    for (int j = 1; j <= R; j++) s += fabs((double)(float)a[j]);
    return s;
}

float orc_cfl(int fd_order, int dim)
{
    /* Propagator_Ac2::getCFL (src/propagator_Ac2.cpp:410-439), FP32 Myfloat2 coefficients */
    double a1 = orc_sum_abs_d2_coef(2);
    double a2 = orc_sum_abs_d2_coef(fd_order) * dim;
    return (float)sqrt(a1 / a2);
}

float orc_inv2_f32(double d)
{
    /* const Myfloat2 inv_d = ONE / d; inv2_d = inv_d * inv_d (src/grid.cpp:3251-3257) */
    float inv = (float)(1.0 / d);
    return inv * inv;
}

double orc_inv2_f64(double d)
{
    double inv = 1.0 / d;
    return inv * inv;
}

/* ---------------------------------------------------------------------------
 * Layout (src/grid.cpp:421-468 per axis; padding :839-947; offset :951-1009)
 * ------------------------------------------------------------------------- */
static void orc_axis_layout(const orc_layout_opts *o, int a, orc_axis *x)
{
    const int64_t h = o->fd_order / 2;
    const int64_t N = o->n_inner[a];
    int64_t off = o->offset[a];
    if (off < 0) off = (a == 0 && (h % 2 == 1) && !o->storage_fp64) ? 1 : 0;
    x->offset_start = 0;
    x->offset_end = off - 1;
    x->halo1_start = x->offset_end + 1;
    x->halo1_end = x->halo1_start + h - 1;
    x->layer1_start = x->halo1_end + 1;
    if (o->nlayer == 0) {
        x->layer1_end = x->layer1_start;
        x->inner_start = x->layer1_start;
    } else {
        x->layer1_end = x->layer1_start + o->nlayer - 1;
        x->inner_start = x->layer1_end + 1;
    }
    x->inner_end = x->inner_start + N - 1;
    if (o->nlayer == 0) {
        x->layer2_start = x->inner_end;
        x->layer2_end = x->inner_end;
    } else {
        x->layer2_start = x->inner_end + 1;
        x->layer2_end = x->layer2_start + o->nlayer - 1;
    }
    x->halo2_start = x->layer2_end + 1;
    x->halo2_end = x->halo2_start + h - 1;

    if (o->auto_pad) {
        x->pad_start = x->pad_end = x->halo2_end;
        if (a == 0 && ((x->halo2_end + 1) % 2) && !o->storage_fp64)
            x->pad_start = x->pad_end = x->halo2_end + 1;
    } else if (o->add_pad[a] >= 0) {
        x->pad_start = x->halo2_end + 1;
        x->pad_end = x->pad_start + o->add_pad[a] - 1;
    } else if (o->mul_pad[a] > 0) {
        const int64_t m = o->mul_pad[a];
        x->pad_start = x->halo2_end + 1;
        int64_t ntmp = (x->pad_start + 1) / m;
        int64_t ntot = (m * ntmp < x->pad_start) ? (ntmp + 1) * m : ntmp * m;
        x->pad_end = ntot - 1;
    } else {
        x->pad_start = x->pad_end = x->halo2_end;
    }
    x->n = x->pad_end + 1;
}

int orc_layout(const orc_layout_opts *o, orc_grid *g)
{
    memset(g, 0, sizeof(*g));
    if (o->dim < 1 || o->dim > 3) return -1;
    if (o->fd_order < 2 || o->fd_order > 16 || (o->fd_order % 2)) return -1;
    g->dim = o->dim;
    g->fd_order = o->fd_order;
    g->halo = o->fd_order / 2;
    for (int a = 0; a < 3; a++) {
        if (a < o->dim) {
            orc_axis_layout(o, a, &g->ax[a]);
        } else {
            /* inactive axes keep the all-zero indices of the Grid constructor
             * (src/grid.cpp:160-249) and n = 1 (src/grid.cpp:516-520, 566-570) */
            memset(&g->ax[a], 0, sizeof(orc_axis));
            g->ax[a].n = 1;
        }
    }
    g->n1 = g->ax[0].n;
    g->n2 = g->ax[1].n;
    g->n3 = g->ax[2].n;
    g->npoint = g->n1 * g->n2 * g->n3;
    return 0;
}

int orc_index_range(const orc_grid *g, int p, int64_t r[6])
{
    /* Grid::getGridIndex (src/grid.cpp:2680-2843) */
    const orc_axis *x = g->ax;
    for (int a = 0; a < 3; a++) {
        r[2 * a] = x[a].inner_start;
        r[2 * a + 1] = x[a].inner_end;
    }
    if (p == 0) return 0;                       /* INNER_POINTS */
    if (p == 1) {                               /* ALL_POINTS: halo1 start .. halo2 end */
        for (int a = 0; a < 3; a++) {
            r[2 * a] = x[a].halo1_start;
            r[2 * a + 1] = x[a].halo2_end;
        }
        return 0;
    }
    if (p == 2) {                               /* MIDDLE_POINT */
        r[0] = r[1] = g->n1 / 2;
        r[2] = r[3] = g->n2 / 2;
        r[4] = r[5] = g->n3 / 2;
        return 0;
    }
    if (p >= 3 && p <= 8) {                     /* IkHALO1 / IkHALO2 */
        int a = (p - 3) / 2, side = (p - 3) % 2;
        r[2 * a] = side == 0 ? x[a].halo1_start : x[a].halo2_start;
        r[2 * a + 1] = side == 0 ? x[a].halo1_end : x[a].halo2_end;
        return 0;
    }
    if (p >= 9 && p <= 14) {                    /* IkINNERHALO1 / IkINNERHALO2 */
        int a = (p - 9) / 2, side = (p - 9) % 2;
        r[2 * a] = side == 0 ? x[a].inner_start : x[a].inner_end - g->halo;
        r[2 * a + 1] = r[2 * a] + g->halo;
        return 0;
    }
    return -1;
}

int orc_fill_table(int ftype, int active, double orig, double d, int64_t i_start,
                   int64_t i_end, double param, double *out)
{
    /* Per-axis table of Grid::fill (src/grid.cpp:1669-1784): coord = Orig + i*d in double */
    for (int64_t i = i_start; i <= i_end; i++) {
        double coord = orig + (double)i * d;
        double v;
        if (!active) v = 1.0;
        else if (ftype == 1) v = sin(coord * param);   /* FUNC_SINE */
        else if (ftype == 2) v = cos(coord * param);   /* FUNC_COSINE */
        else if (ftype == 4) v = coord;                /* FUNC_LINEAR */
        else if (ftype == 6) return -1;                /* FUNC_RANDOM: not restated */
        else v = 1.0;                                  /* FUNC_CONST / NONE / DIRAC */
        out[i - i_start] = v;
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Type-generic kernels
 * ------------------------------------------------------------------------- */
#define TS float
#define TC float
#define SFX f32
#define MAXERRF 1.e-5f
#include "orc_kernels.inc"
#undef TS
#undef TC
#undef SFX
#undef MAXERRF

#define TS double
#define TC double
#define SFX f64
#define MAXERRF 1.e-11f
#include "orc_kernels.inc"
#undef TS
#undef TC
#undef SFX
#undef MAXERRF

#define TS float
#define TC double
#define SFX f32c64
#define MAXERRF 1.e-5f
#include "orc_kernels.inc"
#undef TS
#undef TC
#undef SFX
#undef MAXERRF

/* ---------------------------------------------------------------------------
 * Eigen-mode propagation (src/testCase_Propa.cpp:111-181,
 * src/propagator_Ac2.cpp:366-393 and :441-547)
 * ------------------------------------------------------------------------- */
static const double ORC_PI = 3.1415926535897;  /* src/constant.h:12 */

#define ORC_PROPA(SFX, TS, TC, INV2)                                                        \
float orc_propa_eigen_##SFX(const orc_grid *g, const orc_propa_params *p, orc_sched s,      \
                            TS *prn_out, TS *prc_out)                                       \
{                                                                                           \
    int64_t rall[6], rin[6];                                                                \
    orc_index_range(g, 1, rall);                                                            \
    orc_index_range(g, 0, rin);                                                             \
    double par[3] = {1.0, 1.0, 1.0};                                                        \
    double gamma;                                                                           \
    int m1 = p->mode[0], m2 = p->mode[1], m3 = p->mode[2];                                  \
    if (p->dim == 1) { par[0] = ORC_PI * m1; gamma = m1; }                                  \
    else if (p->dim == 2) { par[0] = ORC_PI * m1; par[1] = ORC_PI * m2;                     \
                            gamma = sqrt((double)(m1 * m1 + m2 * m2)); }                    \
    else { par[0] = ORC_PI * m1; par[1] = ORC_PI * m2; par[2] = ORC_PI * m3;                \
           gamma = sqrt((double)(m1 * m1 + m2 * m2 + m3 * m3)); }                           \
    double *tab[3];                                                                         \
    for (int a = 0; a < 3; a++) {                                                           \
        int64_t len = rall[2 * a + 1] - rall[2 * a] + 1;                                    \
        tab[a] = (double *)malloc(sizeof(double) * (size_t)len);                            \
        orc_fill_table(1, a < p->dim, p->orig[a], p->d[a], rall[2 * a], rall[2 * a + 1],    \
                       par[a], tab[a]);                                                     \
    }                                                                                       \
    TS *prn = prn_out, *prc = prc_out;                                                      \
    TS *coef = (TS *)calloc((size_t)g->npoint, sizeof(TS));                                 \
    TS *ref = (TS *)calloc((size_t)g->npoint, sizeof(TS));                                  \
    TS *lap = p->split_comp ? (TS *)calloc((size_t)g->npoint, sizeof(TS)) : NULL;          \
    TC inv2[3];                                                                             \
    for (int a = 0; a < 3; a++) inv2[a] = INV2(p->d[a]);                                    \
    /* Myfloat scalars: dt, Vp and times are storage-type values */                         \
    const TS dt = (TS)p->dt, vp = (TS)1.0;                                                  \
    /* coef = Vp*Vp*dt*dt in Myfloat (src/propagator_Ac2.cpp:234) */                        \
    orc_fill_const_##SFX(g, coef, rin, vp * vp * dt * dt);                                  \
    /* prn(-2dt), prc(-dt) (src/propagator_Ac2.cpp:270-285) */                              \
    TS t = (TS)(-2.0 * dt);                                                                 \
    orc_fill_sep_##SFX(g, prn, rall, tab[0], tab[1], tab[2], sin(ORC_PI * t * gamma));      \
    t = -dt;                                                                                \
    orc_fill_sep_##SFX(g, prc, rall, tab[0], tab[1], tab[2], sin(ORC_PI * t * gamma));      \
    float sum1 = 0.0f, sum2 = 0.0f;   /* Myfloat32 (src/testCase_Propa.cpp:107) */          \
    for (int it = 0; it < p->nt; it++) {                                                    \
        orc_free_surface_##SFX(g, prc);                                                     \
        if (p->split_comp) {                                                                \
            orc_laplacian_##SFX(g, prc, lap, inv2, s);                                      \
            orc_update_pressure_##SFX(g, prn, prc, coef, lap);                              \
        } else {                                                                            \
            orc_pressure_fd_##SFX(g, prn, prc, coef, inv2, s);                              \
        }                                                                                   \
        if (it % p->snap_inc == 0) {                                                        \
            TS ts = it * dt;                                                                \
            orc_fill_sep_##SFX(g, ref, rall, tab[0], tab[1], tab[2],                        \
                               sin(ORC_PI * ts * gamma));                                   \
            sum1 += (TS)orc_sum_abs_diff_##SFX(g, prn, ref, rin);                           \
            sum2 += (TS)orc_sum_abs_##SFX(g, ref, rin);                                     \
        }                                                                                   \
        TS *tmp = prn; prn = prc; prc = tmp;                                                \
    }                                                                                       \
    /* newest field is now in prc; hand it back in prn_out, the previous in prc_out */      \
    if (prc != prn_out) {                                                                   \
        for (int64_t i = 0; i < g->npoint; i++) {                                           \
            TS x = prn_out[i]; prn_out[i] = prc_out[i]; prc_out[i] = x;                     \
        }                                                                                   \
    }                                                                                       \
    free(coef); free(ref); free(lap);                                                       \
    for (int a = 0; a < 3; a++) free(tab[a]);                                               \
    return sum1 / sum2;                                                                     \
}

ORC_PROPA(f32, float, float, orc_inv2_f32)
ORC_PROPA(f64, double, double, orc_inv2_f64)
