// Grid_Mi355x: MI355X backend of the Grid plugin interface (see grid_Mi355x.h).
//
// Each method validates its arguments the way the reference does
// (sameSize -> KO / -1, src/grid.cpp:2062-2066, 3240-3245), launches the
// library kernels on the library stream and waits for them.  A failing
// library call is reported as a fatal error naming the call and the HIP
// message; nothing falls back to host code.
#include "grid_Mi355x.h"

#include <cmath>
#include <cstdlib>
#include <set>
#include <vector>

#include "config.h"
#include "constant.h"
#include "global.h"
#include "grid_Factory.h"
#include "hardware_Mi355x.h"
#include "output_report.h"

namespace hpcscan {

static const int HSX_STORAGE = sizeof(Myfloat) == 8 ? HSX_FP64 : HSX_FP32;
static const int HSX_COMPUTE = sizeof(Myfloat2) == 8 ? HSX_FP64 : HSX_FP32;

const Mi355xSettings &mi355xSettings(void)
{
    static const Mi355xSettings s = [] {
        Mi355xSettings x;
        auto env = [](const char *k) -> const char * {
            const char *v = getenv(k);
            return v && *v ? v : nullptr;
        };
        if (const char *v = env("HPCSCAN_MI355X_DEVICE")) x.device = atoi(v);
        if (const char *v = env("HPCSCAN_MI355X_KERNEL")) x.kernel = v;
        if (const char *v = env("HPCSCAN_MI355X_FOLD_BC")) x.foldBC = atoi(v) != 0;
        if (const char *v = env("HPCSCAN_MI355X_AUTOTUNE")) x.autotune = atoi(v) != 0;
        if (const char *v = env("HPCSCAN_MI355X_CONTRACT")) x.contract = atoi(v) != 0;
        return x;
    }();
    return s;
}

static bool g_libReady = false;

static int initLibrary(void)
{
    if (g_libReady) return HSX_OK;
    const Mi355xSettings &c = mi355xSettings();
    int rc = hsx_init(c.device);
    if (rc) return rc;
    int fam = HSX_KERNEL_AUTO;
    if (c.kernel == "ref") fam = HSX_KERNEL_REF;
    else if (c.kernel == "tiled") fam = HSX_KERNEL_TILED;
    else if (c.kernel != "auto") {
        printError(" MI355X stencil kernels should be auto, ref or tiled:", c.kernel);
        return HSX_ERR_INVALID;
    }
    rc = hsx_set_kernel(fam);
    if (rc == HSX_OK) rc = hsx_set_contract(c.contract ? 1 : 0);
    if (rc == HSX_OK) g_libReady = true;
    return rc;
}

Grid_Mi355x::Grid_Mi355x(Grid_type t) : Grid(t), d_grid(nullptr) { gridMode = GRID_MODE_MI355X; }

Grid_Mi355x::Grid_Mi355x(Grid_type t, Dim_type d, Myint64 a, Myint64 b, Myint64 c)
    : Grid(t, d, a, b, c), d_grid(nullptr)
{
    gridMode = GRID_MODE_MI355X;
}

Grid_Mi355x::~Grid_Mi355x(void)
{
    if (d_grid) hsx_free(d_grid);
}

bool Grid_Mi355x::check(int rc, const char *what) const
{
    if (rc == HSX_OK) return true;
    string m = string(" In Grid_Mi355x, ") + what + " failed:";
    printError(m.c_str(), hsx_last_error());
    return false;
}

Rtn_code Grid_Mi355x::initializeGrid(void)
{
    if (d_grid != nullptr) {
        printError(" In Grid_Mi355x::initializeGrid, grid already initialized");
        return RTN_CODE_OK;
    }
    Rtn_code rc = Grid::initializeGrid();
    if (rc != RTN_CODE_OK) return rc;
    if (!check(initLibrary(), "hsx_init")) return RTN_CODE_KO;
    if (!check(hsx_malloc(&d_grid, npoint * sizeof(Myfloat)), "hsx_malloc")) return RTN_CODE_KO;
    // points no fill reaches (the halos of a coefficient grid) hold zero, not
    // what the allocation left there: the -writeGrid files are then the same
    // from run to run
    {
        const hsx_grid d = desc();
        if (!check(hsx_array_op(&d, HSX_ARRAY_FILL, d_grid, nullptr, nullptr, 0.0, nullptr), "hsx_array_op") ||
            !check(hsx_stream_sync(nullptr), "hsx_stream_sync"))
            return RTN_CODE_KO;
    }
    if (mi355xSettings().autotune && dim == DIM3) {
        // once per layout in this process; the library keeps the choice for
        // the fused step and the Laplacian of every grid of that layout
        static std::set<std::vector<int64_t>> tuned;
        const hsx_grid d = desc();
        std::vector<int64_t> key = {d.fd_order, d.storage, d.compute, d.n1, d.n2, d.n3};
        key.insert(key.end(), d.inner, d.inner + 6);
        if (tuned.insert(key).second) {
            int shape[4] = {0, 0, 0, 0};
            float ms = 0.f;
            if (!check(hsx_autotune(&d, 3, shape, &ms), "hsx_autotune")) return RTN_CODE_KO;
            const string txt = to_string(shape[0]) + " lanes x " + to_string(shape[1]) + " rows, chunk " +
                               to_string(shape[2]) + ", flags " + to_string(shape[3]) + " (" + to_string(ms) + " ms)";
            printInfo(MASTER, " MI355X autotuned tile", txt);
        }
    }
    return RTN_CODE_OK;
}

// descriptor of this grid for the C ABI; d1..d3 change with defineUnitGrid
hsx_grid Grid_Mi355x::desc(void) const
{
    hsx_grid g;
    g.dim = dim;
    g.fd_order = 2 * haloWidth;
    g.storage = HSX_STORAGE;
    g.compute = HSX_COMPUTE;
    g.n1 = n1;
    g.n2 = n2;
    g.n3 = n3;
    const Myint in[6] = {i1InnerStart, i1InnerEnd, i2InnerStart, i2InnerEnd, i3InnerStart, i3InnerEnd};
    const Myint h1[6] = {i1Halo1Start, i1Halo1End, i2Halo1Start, i2Halo1End, i3Halo1Start, i3Halo1End};
    const Myint h2[6] = {i1Halo2Start, i1Halo2End, i2Halo2Start, i2Halo2End, i3Halo2Start, i3Halo2End};
    for (int k = 0; k < 6; k++) {
        g.inner[k] = in[k];
        g.halo1[k] = h1[k];
        g.halo2[k] = h2[k];
    }
    g.d[0] = d1;
    g.d[1] = d2;
    g.d[2] = d3;
    return g;
}

void Grid_Mi355x::range(Point_type p, int64_t r[6]) const
{
    Myint64 a[6];
    getGridIndex(p, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]);
    for (int k = 0; k < 6; k++) r[k] = a[k];
}

const Grid_Mi355x *Grid_Mi355x::peer(const Grid &g, const char *what) const
{
    const Grid_Mi355x *p = dynamic_cast<const Grid_Mi355x *>(&g);
    if (!p || !p->d_grid) {
        string m = string(" In Grid_Mi355x::") + what + ", argument grid is not an initialized MI355X grid";
        printError(m.c_str());
        return nullptr;
    }
    return p->settle() ? p : nullptr;
}

bool Grid_Mi355x::settle(void) const
{
    if (!fsPending) return true;
    const int mask = fsPending;
    fsPending = 0;
    hsx_grid g = desc();
    return check(hsx_free_surface(&g, d_grid, mask, nullptr), "hsx_free_surface") &&
           check(hsx_stream_sync(nullptr), "hsx_stream_sync");
}

Rtn_code Grid_Mi355x::toHost(void)
{
    if (!settle()) return RTN_CODE_KO;
    return check(hsx_memcpy_d2h(grid_3d, d_grid, npoint * sizeof(Myfloat), nullptr), "hsx_memcpy_d2h")
               ? RTN_CODE_OK
               : RTN_CODE_KO;
}

void Grid_Mi355x::info(void)
{
    Grid::info();
    printInfo(MASTER, " Device buffer (MB)", Myfloat(npoint * sizeof(Myfloat) / 1e6));
}

Rtn_code Grid_Mi355x::write(string fileName)
{
    if (!Config::Instance()->writeGrid) {
        grid_write_time = 0.0;
        return RTN_CODE_OK;
    }
    if (toHost() != RTN_CODE_OK) return RTN_CODE_KO;
    return Grid::write(fileName);
}

Rtn_code Grid_Mi355x::write(Point_type p, string fileName)
{
    if (!Config::Instance()->writeGrid) return RTN_CODE_OK;
    if (toHost() != RTN_CODE_OK) return RTN_CODE_KO;
    return Grid::write(p, fileName);
}

Rtn_code Grid_Mi355x::writeGlobal(Point_type p, string fileName, Myint n_iter)
{
    if (!Config::Instance()->writeGrid) {
        grid_writeGlobal_time = 0.0;
        return RTN_CODE_OK;
    }
    if (toHost() != RTN_CODE_OK) return RTN_CODE_KO;
    return Grid::writeGlobal(p, fileName, n_iter);
}

// the file is read into the host copy over the current device values, then the
// whole grid goes back to the device
Rtn_code Grid_Mi355x::read(string fileName) { return read(ALL_POINTS, fileName); }

Rtn_code Grid_Mi355x::read(Point_type p, string fileName)
{
    if (toHost() != RTN_CODE_OK) return RTN_CODE_KO;
    const Rtn_code rc = p == ALL_POINTS ? Grid::read(fileName) : Grid::read(p, fileName);
    if (rc != RTN_CODE_OK) return rc;
    touched();
    if (!check(hsx_memcpy_h2d(d_grid, grid_3d, npoint * sizeof(Myfloat), nullptr), "hsx_memcpy_h2d"))
        return RTN_CODE_KO;
    return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
}

// ---- fills ------------------------------------------------------------------
void Grid_Mi355x::fill(Point_type p, Myfloat val)
{
    int64_t r[6];
    range(p, r);
    if (!settle()) return;
    touched();
    hsx_grid g = desc();
    if (check(hsx_fill_const(&g, d_grid, r, val, nullptr), "hsx_fill_const"))
        check(hsx_stream_sync(nullptr), "hsx_stream_sync");
}

// separable fill: per-axis factors computed here in double with libm, as
// Grid::fill does (src/grid.cpp:1669-1784), product on the device
void Grid_Mi355x::fill(Point_type p, Func_type t1, Func_type t2, Func_type t3, Myfloat64 p1, Myfloat64 p2,
                       Myfloat64 p3, Myfloat64 amp)
{
    int64_t r[6];
    range(p, r);
    const Func_type ft[3] = {t1, t2, t3};
    const Myfloat64 par[3] = {p1, p2, p3}, orig[3] = {Orig1, Orig2, Orig3}, dd[3] = {d1, d2, d3};
    std::vector<double> tab[3];
    for (int a = 0; a < 3; a++) {
        for (int64_t i = r[2 * a]; i <= r[2 * a + 1]; i++) {
            const Myfloat64 coord = Myfloat64(orig[a] + i * dd[a]);
            double v = 1.0;
            if (a < (int)dim) {
                switch (ft[a]) {
                case FUNC_SINE: v = sin(coord * par[a]); break;
                case FUNC_COSINE: v = cos(coord * par[a]); break;
                case FUNC_LINEAR: v = coord; break;
                case FUNC_RANDOM: v = Myfloat(rand()) / RAND_MAX; break;
                default: v = 1.0; break;
                }
            }
            tab[a].push_back(v);
        }
    }
    if (!settle()) return;
    touched();
    hsx_grid g = desc();
    check(hsx_fill_separable(&g, d_grid, r, tab[0].data(), tab[1].data(), tab[2].data(), amp, nullptr),
          "hsx_fill_separable");
}

void Grid_Mi355x::fillArray(Myfloat val)
{
    if (!settle()) return;
    touched();
    hsx_grid g = desc();
    if (check(hsx_array_op(&g, HSX_ARRAY_FILL, d_grid, nullptr, nullptr, val, nullptr), "hsx_array_op"))
        check(hsx_stream_sync(nullptr), "hsx_stream_sync");
}

void Grid_Mi355x::copyArray(const Grid &u)
{
    const Grid_Mi355x *pu = peer(u, "copyArray");
    if (!pu || !settle()) return;
    touched();
    hsx_grid g = desc();
    if (check(hsx_array_op(&g, HSX_ARRAY_COPY, d_grid, pu->d_grid, nullptr, 0, nullptr), "hsx_array_op"))
        check(hsx_stream_sync(nullptr), "hsx_stream_sync");
}

void Grid_Mi355x::addArray(const Grid &u, const Grid &v)
{
    const Grid_Mi355x *pu = peer(u, "addArray"), *pv = peer(v, "addArray");
    if (!pu || !pv || !settle()) return;
    touched();
    hsx_grid g = desc();
    if (check(hsx_array_op(&g, HSX_ARRAY_ADD, d_grid, pu->d_grid, pv->d_grid, 0, nullptr), "hsx_array_op"))
        check(hsx_stream_sync(nullptr), "hsx_stream_sync");
}

void Grid_Mi355x::multiplyArray(const Grid &u, const Grid &v)
{
    const Grid_Mi355x *pu = peer(u, "multiplyArray"), *pv = peer(v, "multiplyArray");
    if (!pu || !pv || !settle()) return;
    touched();
    hsx_grid g = desc();
    if (check(hsx_array_op(&g, HSX_ARRAY_MULTIPLY, d_grid, pu->d_grid, pv->d_grid, 0, nullptr), "hsx_array_op"))
        check(hsx_stream_sync(nullptr), "hsx_stream_sync");
}

void Grid_Mi355x::addUpdateArray(const Grid &u)
{
    const Grid_Mi355x *pu = peer(u, "addUpdateArray");
    if (!pu || !settle()) return;
    touched();
    hsx_grid g = desc();
    if (check(hsx_array_op(&g, HSX_ARRAY_ADDTO, d_grid, pu->d_grid, nullptr, 0, nullptr), "hsx_array_op"))
        check(hsx_stream_sync(nullptr), "hsx_stream_sync");
}

// ---- stencil path -------------------------------------------------------------
Rtn_code Grid_Mi355x::applyBoundaryCondition(BoundCond_type bc)
{
    if (bc == NO_BOUND_COND) return RTN_CODE_OK;
    if (bc != BOUND_COND_ANTI_MIRROR) {
        printError("IN Grid::applyBoundaryCondition, invalid boundCondType", (Myint)bc);
        return RTN_CODE_KO;
    }
    // a face gets the free surface only where there is no neighbour (src/grid.cpp:3081)
    const Point_type halos[6] = {I1HALO1, I1HALO2, I2HALO1, I2HALO2, I3HALO1, I3HALO2};
    int mask = 0;
    for (int k = 0; k < 2 * (int)dim; k++)
        if (getNeighbourProc(halos[k]) == PROC_NULL) mask |= 1 << k;
    // halos already hold the free surface of the inner points (written by the
    // folded step that produced them): the zeros of the boundary planes are
    // owed to the next step, which reads them as zero
    if (mi355xSettings().foldBC && dim == DIM3 && mask != 0 && fsHalo == mask) {
        fsPending = mask;
        return RTN_CODE_OK;
    }
    if (!settle()) return RTN_CODE_KO;
    hsx_grid g = desc();
    if (!check(hsx_free_surface(&g, d_grid, mask, nullptr), "hsx_free_surface")) return RTN_CODE_KO;
    fsHalo = mask;
    return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
}

Rtn_code Grid_Mi355x::fdD2(int axis, const Grid &W, Myint fdOrder)
{
    static const char *names[3] = {"Grid::FD_D2_N1, grids have not same size",
                                   "Grid::FD_D2_N2, grids have not same size",
                                   "Grid::FD_D2_N3, grids have not same size"};
    if (!sameSize(W)) {
        printError(names[axis]);
        return RTN_CODE_KO;
    }
    if (fdOrder != 2 * haloWidth) {
        printError(" In Grid_Mi355x, fdOrder does not match the grid halo width", fdOrder);
        return RTN_CODE_KO;
    }
    const Grid_Mi355x *pw = peer(W, "FD_D2");
    if (!pw || !settle()) return RTN_CODE_KO;
    pw->touched();
    hsx_grid g = desc();
    if (!check(hsx_fd_d2(&g, axis, d_grid, pw->d_grid, nullptr), "hsx_fd_d2")) return RTN_CODE_KO;
    return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
}

Rtn_code Grid_Mi355x::fdD1(int axis, const Grid &W, Myint fdOrder)
{
    static const char *names[3] = {"Grid::FD_D1_N1, grids have not same size",
                                   "Grid::FD_D1_N2, grids have not same size",
                                   "Grid::FD_D1_N3, grids have not same size"};
    if (!sameSize(W)) {
        printError(names[axis]);
        return RTN_CODE_KO;
    }
    if (fdOrder != 2 * haloWidth) {
        printError(" In Grid_Mi355x, fdOrder does not match the grid halo width", fdOrder);
        return RTN_CODE_KO;
    }
    const Grid_Mi355x *pw = peer(W, "FD_D1");
    if (!pw || !settle()) return RTN_CODE_KO;
    pw->touched();
    hsx_grid g = desc();
    if (!check(hsx_fd_d1(&g, axis, d_grid, pw->d_grid, nullptr), "hsx_fd_d1")) return RTN_CODE_KO;
    return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
}

Rtn_code Grid_Mi355x::FD_D1_N1(Point_type, const Grid &W, Myint o) { return fdD1(0, W, o); }
Rtn_code Grid_Mi355x::FD_D1_N2(Point_type, const Grid &W, Myint o) { return fdD1(1, W, o); }
Rtn_code Grid_Mi355x::FD_D1_N3(Point_type, const Grid &W, Myint o) { return fdD1(2, W, o); }
Rtn_code Grid_Mi355x::FD_D2_N1(Point_type, const Grid &W, Myint o) { return fdD2(0, W, o); }
Rtn_code Grid_Mi355x::FD_D2_N2(Point_type, const Grid &W, Myint o) { return fdD2(1, W, o); }
Rtn_code Grid_Mi355x::FD_D2_N3(Point_type, const Grid &W, Myint o) { return fdD2(2, W, o); }

Rtn_code Grid_Mi355x::FD_LAPLACIAN(Point_type, const Grid &W, Myint fdOrder)
{
    if (!sameSize(W)) {
        printError("Grid::FD_LAPLACIAN, grids have not same size");
        return RTN_CODE_KO;
    }
    if (fdOrder != 2 * haloWidth) {
        printError(" In Grid_Mi355x, fdOrder does not match the grid halo width", fdOrder);
        return RTN_CODE_KO;
    }
    const Grid_Mi355x *pw = peer(W, "FD_LAPLACIAN");
    if (!pw || !settle()) return RTN_CODE_KO;
    pw->touched();
    hsx_grid g = desc();
    if (!check(hsx_laplacian(&g, d_grid, pw->d_grid, nullptr), "hsx_laplacian")) return RTN_CODE_KO;
    return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
}

Rtn_code Grid_Mi355x::updatePressure(Point_type, const Grid &prc, const Grid &coef, const Grid &lapla)
{
    if (!sameSize(prc) || !sameSize(coef) || !sameSize(lapla)) {
        printError("Grid::updatePressure, grids have not same size");
        return RTN_CODE_KO;
    }
    const Grid_Mi355x *pp = peer(prc, "updatePressure"), *pc = peer(coef, "updatePressure"),
                      *pl = peer(lapla, "updatePressure");
    if (!pp || !pc || !pl || !settle()) return RTN_CODE_KO;
    touched();
    hsx_grid g = desc();
    if (!check(hsx_update_pressure(&g, d_grid, pp->d_grid, pc->d_grid, pl->d_grid, nullptr), "hsx_update_pressure"))
        return RTN_CODE_KO;
    return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
}

Rtn_code Grid_Mi355x::computePressureWithFD(Grid &prc, Grid &coef, Myint fdOrder)
{
    if (!sameSize(prc) || !sameSize(coef)) {
        printError("In Grid::computePressureWithFD, grids have not same size");
        return RTN_CODE_KO;
    }
    if (fdOrder != 2 * haloWidth) {
        printError(" In Grid_Mi355x, fdOrder does not match the grid halo width", fdOrder);
        return RTN_CODE_KO;
    }
    const Grid_Mi355x *pp = dynamic_cast<const Grid_Mi355x *>(&prc);
    const Grid_Mi355x *pc = peer(coef, "computePressureWithFD");
    if (!pp || !pp->d_grid || !pc) {
        printError(" In Grid_Mi355x::computePressureWithFD, argument grid is not an initialized MI355X grid");
        return RTN_CODE_KO;
    }
    hsx_grid g = desc();
    // free surface of prc (owed, or applied with its halos) folded into the step
    const int mask = pp->fsPending ? pp->fsPending : pp->fsHalo;
    if (mask && mi355xSettings().foldBC) {
        if (fsPending && fsPending != mask && !settle()) return RTN_CODE_KO;
        const int fold = HSX_FOLD_HALO | (pp->fsPending ? HSX_FOLD_PRC : 0) | (fsPending ? HSX_FOLD_PRN : 0);
        const int rc = hsx_ac2_step_fs(&g, d_grid, pp->d_grid, pc->d_grid, mask, fold, nullptr);
        if (rc == HSX_OK) {
            fsPending = 0;
            fsHalo = mask;
            return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
        }
        if (rc != HSX_ERR_UNSUPPORTED) return check(rc, "hsx_ac2_step_fs") ? RTN_CODE_OK : RTN_CODE_KO;
    }
    // grids the folded kernel does not take: free surface and step apart
    if (!pp->settle() || !settle()) return RTN_CODE_KO;
    touched();
    if (!check(hsx_ac2_step(&g, d_grid, pp->d_grid, pc->d_grid, nullptr), "hsx_ac2_step")) return RTN_CODE_KO;
    return check(hsx_stream_sync(nullptr), "hsx_stream_sync") ? RTN_CODE_OK : RTN_CODE_KO;
}

// ---- reductions (src/grid.cpp:2006-2370) -----------------------------------------
Myfloat Grid_Mi355x::L1Err(Point_type p, const Grid &v) const
{
    if (!sameSize(v)) {
        printError("Grid::L1Err, grids have different size");
        return -1.0;
    }
    const Grid_Mi355x *pv = peer(v, "L1Err");
    if (!pv) return -1.0;
    int64_t r[6];
    range(p, r);
    if (!settle()) return NAN;
    hsx_grid g = desc();
    double s[2];
    if (!check(hsx_l1_sums(&g, d_grid, pv->d_grid, r, s, nullptr), "hsx_l1_sums")) return NAN;
    if (s[1] < MAX_ERR_FLOAT) s[1] = 1.0 * npoint;
    const Myfloat64 err = s[0] / s[1];
    if (std::isnan(err)) printError("In Grid::L1Err, std::isnan(err)");
    return err;
}

Myfloat Grid_Mi355x::maxErr(Point_type p, const Grid &v) const
{
    if (!sameSize(v)) {
        printError("Grid::maxErr, grids have different size");
        return -1.0;
    }
    const Grid_Mi355x *pv = peer(v, "maxErr");
    if (!pv) return -1.0;
    int64_t r[6];
    range(p, r);
    if (!settle()) return NAN;
    hsx_grid g = desc();
    double e;
    if (!check(hsx_max_err(&g, d_grid, pv->d_grid, r, MAX_ERR_FLOAT, &e, nullptr), "hsx_max_err")) return NAN;
    return e;
}

Myfloat Grid_Mi355x::getSumAbs(Point_type p) const
{
    int64_t r[6];
    range(p, r);
    if (!settle()) return NAN;
    hsx_grid g = desc();
    double s;
    if (!check(hsx_sum_abs(&g, d_grid, r, &s, nullptr), "hsx_sum_abs")) return NAN;
    if (std::isnan(s)) printError("In Grid::getSumAbs, std::isnan(sum2)");
    return s;
}

Myfloat Grid_Mi355x::getSumAbsDiff(Point_type p, const Grid &v) const
{
    if (!sameSize(v)) {
        printError("Grid::getSumAbsDiff, grids have different size");
        return -1.0;
    }
    const Grid_Mi355x *pv = peer(v, "getSumAbsDiff");
    if (!pv) return -1.0;
    int64_t r[6];
    range(p, r);
    if (!settle()) return NAN;
    hsx_grid g = desc();
    double s;
    if (!check(hsx_sum_abs_diff(&g, d_grid, pv->d_grid, r, &s, nullptr), "hsx_sum_abs_diff")) return NAN;
    if (std::isnan(s)) printError("In Grid::getSumAbsDiff, std::isnan(sum1)");
    return s;
}

Myfloat Grid_Mi355x::getMin(Point_type p)
{
    int64_t r[6];
    range(p, r);
    if (!settle()) return NAN;
    hsx_grid g = desc();
    double mm[2];
    if (!check(hsx_min_max(&g, d_grid, r, mm, nullptr), "hsx_min_max")) return NAN;
    return mm[0];
}

Myfloat Grid_Mi355x::getMax(Point_type p)
{
    int64_t r[6];
    range(p, r);
    if (!settle()) return NAN;
    hsx_grid g = desc();
    double mm[2];
    if (!check(hsx_min_max(&g, d_grid, r, mm, nullptr), "hsx_min_max")) return NAN;
    return mm[1];
}

// ---- registration: -testMode MI355X ------------------------------------------------
shared_ptr<DataAcquisition> makeDataAcquisition_Mi355x(void);   // data_Acquisition_Mi355x.cpp

static bool registered = Grid_Factory::registerBackend(
    {GRID_MODE_MI355X, [](Grid_type t) -> shared_ptr<Grid> { return make_shared<Grid_Mi355x>(t); },
     [](Grid_type t, Dim_type d, Myint64 a, Myint64 b, Myint64 c) -> shared_ptr<Grid> {
         return make_shared<Grid_Mi355x>(t, d, a, b, c);
     },
     [](string mode) -> shared_ptr<Hardware> { return make_shared<Hardware_Mi355x>(mode); }, nullptr,
     [] { return makeDataAcquisition_Mi355x(); }});

} // namespace hpcscan
