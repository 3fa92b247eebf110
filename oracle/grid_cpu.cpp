// CPU grid backends of the oracle harness build (test infrastructure only).
//
// Registers -testMode Baseline and -testMode CacheBlk in the host harness,
// with every compute method delegated to the C restatement in
// hpcscan_oracle.c: Baseline loops (src/grid.cpp) or CacheBlk blocks
// (src/grid_CacheBlk.cpp, block sizes -cb1/-cb2/-cb3).  The product binary
// never links this file; it is the CPU reference that tests and the bench's
// cpu_baseline leg run beside the MI355X backend.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "config.h"
#include "constant.h"
#include "grid.h"
#include "grid_Factory.h"
#include "hpcscan_oracle.h"
#include "output_report.h"

#if defined(HPCSCAN_STORAGE_FP64)
#define ORC(name) orc_##name##_f64
#elif defined(HPCSCAN_COMPUTE_FP64)
#define ORC(name) orc_##name##_f32c64
#else
#define ORC(name) orc_##name##_f32
#endif

namespace hpcscan {

class Grid_CPU : public Grid
{
public:
    Grid_CPU(Grid_type t, int mode) : Grid(t), blocked(mode) { setMode(); }
    Grid_CPU(Grid_type t, Dim_type d, Myint64 a, Myint64 b, Myint64 c, int mode)
        : Grid(t, d, a, b, c), blocked(mode)
    {
        setMode();
    }

    // points no fill reaches (e.g. the halos of the coefficient grid, which
    // the reference leaves uninitialised) start at zero, so that the files of
    // -writeGrid are the same from run to run and backend to backend
    Rtn_code initializeGrid(void) override
    {
        const Rtn_code rc = Grid::initializeGrid();
        if (rc == RTN_CODE_OK && grid_3d) std::fill(grid_3d, grid_3d + npoint, (Myfloat)0);
        return rc;
    }

    void info(void) override
    {
        Grid::info();
        if (blocked) {
            printInfo(MASTER, " Cache block size axis1", (Myint64)Config::Instance()->cb1);
            printInfo(MASTER, " Cache block size axis2", (Myint64)Config::Instance()->cb2);
            printInfo(MASTER, " Cache block size axis3", (Myint64)Config::Instance()->cb3);
        }
    }

    void fill(Point_type p, Myfloat val) override
    {
        int64_t r[6];
        range(p, r);
        orc_grid g = toOrc();
        ORC(fill_const)(&g, grid_3d, r, val);
    }

    void fill(Point_type p, Func_type t1, Func_type t2, Func_type t3, Myfloat64 p1, Myfloat64 p2,
              Myfloat64 p3, Myfloat64 amp) override
    {
        int64_t r[6];
        range(p, r);
        const Func_type ft[3] = {t1, t2, t3};
        const Myfloat64 par[3] = {p1, p2, p3}, orig[3] = {Orig1, Orig2, Orig3}, dd[3] = {d1, d2, d3};
        std::vector<double> tab[3];
        for (int a = 0; a < 3; a++) {
            tab[a].resize(r[2 * a + 1] - r[2 * a] + 1);
            if (ft[a] == FUNC_RANDOM && a < (int)dim) {
                for (auto &v : tab[a]) v = Myfloat(rand()) / RAND_MAX;
            } else {
                orc_fill_table(ft[a], a < (int)dim, orig[a], dd[a], r[2 * a], r[2 * a + 1], par[a],
                               tab[a].data());
            }
        }
        orc_grid g = toOrc();
        ORC(fill_sep)(&g, grid_3d, r, tab[0].data(), tab[1].data(), tab[2].data(), amp);
    }

    void fillArray(Myfloat v) override
    {
        for (Myint64 i = 0; i < npoint; i++) grid_3d[i] = v;
    }
    void copyArray(const Grid &u) override
    {
        for (Myint64 i = 0; i < npoint; i++) grid_3d[i] = u.grid_3d[i];
    }
    void addArray(const Grid &u, const Grid &v) override
    {
        for (Myint64 i = 0; i < npoint; i++) grid_3d[i] = u.grid_3d[i] + v.grid_3d[i];
    }
    void multiplyArray(const Grid &u, const Grid &v) override
    {
        for (Myint64 i = 0; i < npoint; i++) grid_3d[i] = u.grid_3d[i] * v.grid_3d[i];
    }
    void addUpdateArray(const Grid &u) override
    {
        for (Myint64 i = 0; i < npoint; i++) grid_3d[i] = grid_3d[i] + u.grid_3d[i];
    }

    Rtn_code applyBoundaryCondition(BoundCond_type bc) override
    {
        if (bc == NO_BOUND_COND) return RTN_CODE_OK;
        if (bc != BOUND_COND_ANTI_MIRROR) {
            printError("IN Grid::applyBoundaryCondition, invalid boundCondType", (Myint)bc);
            return RTN_CODE_KO;
        }
        orc_grid g = toOrc();
        return ORC(free_surface)(&g, grid_3d) == 0 ? RTN_CODE_OK : RTN_CODE_KO;
    }

    Rtn_code FD_D2_N1(Point_type, const Grid &W, Myint o) override { return fdD2(0, W, o); }
    Rtn_code FD_D2_N2(Point_type, const Grid &W, Myint o) override { return fdD2(1, W, o); }
    Rtn_code FD_D2_N3(Point_type, const Grid &W, Myint o) override { return fdD2(2, W, o); }
    Rtn_code FD_D1_N1(Point_type, const Grid &W, Myint o) override { return fdD1(0, W, o); }
    Rtn_code FD_D1_N2(Point_type, const Grid &W, Myint o) override { return fdD1(1, W, o); }
    Rtn_code FD_D1_N3(Point_type, const Grid &W, Myint o) override { return fdD1(2, W, o); }

    Rtn_code FD_LAPLACIAN(Point_type, const Grid &W, Myint o) override
    {
        if (!sameSize(W) || o != 2 * haloWidth) {
            printError("Grid::FD_LAPLACIAN, grids have not same size");
            return RTN_CODE_KO;
        }
        orc_grid g = toOrc();
        Myfloat2 iv[3];
        inv2(iv);
        return ORC(laplacian)(&g, grid_3d, W.grid_3d, iv, sched()) == 0 ? RTN_CODE_OK : RTN_CODE_KO;
    }

    Rtn_code updatePressure(Point_type, const Grid &prc, const Grid &coef, const Grid &lapla) override
    {
        orc_grid g = toOrc();
        return ORC(update_pressure)(&g, grid_3d, prc.grid_3d, coef.grid_3d, lapla.grid_3d) == 0
                   ? RTN_CODE_OK
                   : RTN_CODE_KO;
    }

    Rtn_code computePressureWithFD(Grid &prc, Grid &coef, Myint o) override
    {
        if (!sameSize(prc) || o != 2 * haloWidth) {
            printError("In Grid::computePressureWithFD, grids have not same size");
            return RTN_CODE_KO;
        }
        orc_grid g = toOrc();
        Myfloat2 iv[3];
        inv2(iv);
        return ORC(pressure_fd)(&g, grid_3d, prc.grid_3d, coef.grid_3d, iv, sched()) == 0 ? RTN_CODE_OK
                                                                                          : RTN_CODE_KO;
    }

    Myfloat L1Err(Point_type p, const Grid &v) const override
    {
        if (!sameSize(v)) {
            printError("Grid::L1Err, grids have different size");
            return -1.0;
        }
        int64_t r[6];
        range(p, r);
        orc_grid g = toOrc();
        return ORC(l1_err)(&g, grid_3d, v.grid_3d, r);
    }
    Myfloat maxErr(Point_type p, const Grid &v) const override
    {
        if (!sameSize(v)) {
            printError("Grid::maxErr, grids have different size");
            return -1.0;
        }
        int64_t r[6];
        range(p, r);
        orc_grid g = toOrc();
        return ORC(max_err)(&g, grid_3d, v.grid_3d, r);
    }
    Myfloat getSumAbs(Point_type p) const override
    {
        int64_t r[6];
        range(p, r);
        orc_grid g = toOrc();
        return ORC(sum_abs)(&g, grid_3d, r);
    }
    Myfloat getSumAbsDiff(Point_type p, const Grid &v) const override
    {
        if (!sameSize(v)) {
            printError("Grid::getSumAbsDiff, grids have different size");
            return -1.0;
        }
        int64_t r[6];
        range(p, r);
        orc_grid g = toOrc();
        return ORC(sum_abs_diff)(&g, grid_3d, v.grid_3d, r);
    }
    Myfloat getMin(Point_type p) override
    {
        int64_t r[6];
        range(p, r);
        orc_grid g = toOrc();
        return ORC(min)(&g, grid_3d, r);
    }
    Myfloat getMax(Point_type p) override
    {
        int64_t r[6];
        range(p, r);
        orc_grid g = toOrc();
        return ORC(max)(&g, grid_3d, r);
    }

private:
    int blocked;

    void setMode(void) { gridMode = blocked ? GRID_MODE_CACHEBLK : GRID_MODE_BASELINE; }

    orc_sched sched(void) const
    {
        orc_sched s;
        s.mode = blocked;
        s.cb1 = Config::Instance()->cb1;
        s.cb2 = Config::Instance()->cb2;
        s.cb3 = Config::Instance()->cb3;
        return s;
    }

    void range(Point_type p, int64_t r[6]) const
    {
        Myint64 a[6];
        getGridIndex(p, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]);
        for (int i = 0; i < 6; i++) r[i] = a[i];
    }

    void inv2(Myfloat2 iv[3]) const
    {
        // src/grid.cpp:3251-3257 (inactive axes have d = 0 -> inf, never used)
        const Myfloat64 dd[3] = {d1, d2, d3};
        for (int a = 0; a < 3; a++) {
            const Myfloat2 inv = ONE / dd[a];
            iv[a] = inv * inv;
        }
    }

    orc_grid toOrc(void) const
    {
        orc_grid g;
        g.dim = dim;
        g.fd_order = 2 * haloWidth;
        g.halo = haloWidth;
        const Myint src[3][14] = {
            {i1OffsetStart, i1OffsetEnd, i1Halo1Start, i1Halo1End, i1Layer1Start, i1Layer1End, i1InnerStart,
             i1InnerEnd, i1Layer2Start, i1Layer2End, i1Halo2Start, i1Halo2End, i1PadStart, i1PadEnd},
            {i2OffsetStart, i2OffsetEnd, i2Halo1Start, i2Halo1End, i2Layer1Start, i2Layer1End, i2InnerStart,
             i2InnerEnd, i2Layer2Start, i2Layer2End, i2Halo2Start, i2Halo2End, i2PadStart, i2PadEnd},
            {i3OffsetStart, i3OffsetEnd, i3Halo1Start, i3Halo1End, i3Layer1Start, i3Layer1End, i3InnerStart,
             i3InnerEnd, i3Layer2Start, i3Layer2End, i3Halo2Start, i3Halo2End, i3PadStart, i3PadEnd}};
        const Myint nn[3] = {n1, n2, n3};
        for (int a = 0; a < 3; a++) {
            orc_axis &x = g.ax[a];
            int64_t *f = &x.offset_start;
            for (int k = 0; k < 14; k++) f[k] = src[a][k];
            x.n = nn[a];
        }
        g.n1 = n1;
        g.n2 = n2;
        g.n3 = n3;
        g.npoint = npoint;
        return g;
    }

    Rtn_code fdD1(int axis, const Grid &W, Myint o)
    {
        static const char *msg[3] = {"Grid::FD_D1_N1, grids have not same size",
                                     "Grid::FD_D1_N2, grids have not same size",
                                     "Grid::FD_D1_N3, grids have not same size"};
        if (!sameSize(W) || o != 2 * haloWidth) {
            printError(msg[axis]);
            return RTN_CODE_KO;
        }
        orc_grid g = toOrc();
        // const Myfloat2 inv_d1 = ONE / d1 (src/grid.cpp:3415-3417)
        const Myfloat64 dd[3] = {d1, d2, d3};
        Myfloat2 iv[3];
        for (int a = 0; a < 3; a++) iv[a] = ONE / dd[a];
        return ORC(fd_d1)(&g, axis, grid_3d, W.grid_3d, iv, sched()) == 0 ? RTN_CODE_OK : RTN_CODE_KO;
    }

    Rtn_code fdD2(int axis, const Grid &W, Myint o)
    {
        if (!sameSize(W) || o != 2 * haloWidth) {
            printError("Grid::FD_D2, grids have not same size");
            return RTN_CODE_KO;
        }
        orc_grid g = toOrc();
        Myfloat2 iv[3];
        inv2(iv);
        return ORC(fd_d2)(&g, axis, grid_3d, W.grid_3d, iv, sched()) == 0 ? RTN_CODE_OK : RTN_CODE_KO;
    }
};

static bool registeredBaseline = Grid_Factory::registerBackend(
    {GRID_MODE_BASELINE, [](Grid_type t) -> shared_ptr<Grid> { return make_shared<Grid_CPU>(t, 0); },
     [](Grid_type t, Dim_type d, Myint64 a, Myint64 b, Myint64 c) -> shared_ptr<Grid> {
         return make_shared<Grid_CPU>(t, d, a, b, c, 0);
     },
     nullptr, nullptr});

static bool registeredCacheBlk = Grid_Factory::registerBackend(
    {GRID_MODE_CACHEBLK, [](Grid_type t) -> shared_ptr<Grid> { return make_shared<Grid_CPU>(t, 1); },
     [](Grid_type t, Dim_type d, Myint64 a, Myint64 b, Myint64 c) -> shared_ptr<Grid> {
         return make_shared<Grid_CPU>(t, d, a, b, c, 1);
     },
     nullptr, nullptr});

} // namespace hpcscan
