// TestCase FD_D1: staggered first derivative along each axis of a separable
// sine, checked against the analytic derivative taken half a grid step back
// along that axis, with GPoint/s and "apparent" GB/s (reference behaviour:
// src/testCase_FD_D1.cpp:32-395).
#include <cfloat>
#include <functional>

#include "config.h"
#include "constant.h"
#include "fdm.h"
#include "global.h"
#include "grid_Factory.h"
#include "output_report.h"
#include "testCase.h"

namespace hpcscan {

TestCase_FD_D1::TestCase_FD_D1(void)
{
    testCaseName = "FD_D1";
    testCaseVersion = "Standard implementation";
}

namespace {
struct D1Result {
    Myfloat32 gflop = 0, gpointFD = 0, gpointEff = 0, gb = 0;
    double bestTime = 0;
    Myfloat32 err = 0;
};
} // namespace

Rtn_code TestCase_FD_D1::run(void)
{
    if (this->initialize() == RTN_CODE_KO) return RTN_CODE_KO;

    Config *cfg = Config::Instance();
    const string gridMode = cfg->testMode;
    auto U2 = Grid_Factory::create(gridMode, GRID_LOCAL);
    auto W2 = Grid_Factory::create(gridMode, GRID_LOCAL);
    auto R2 = Grid_Factory::create(gridMode, GRID_LOCAL);
    Grid &Ugrid = *U2, &Wgrid = *W2, &Rgrid = *R2;
    Ugrid.initializeGrid();
    Wgrid.initializeGrid();
    Rgrid.initializeGrid();
    Ugrid.defineUnitGrid();
    Wgrid.defineUnitGrid();
    Rgrid.defineUnitGrid();
    Ugrid.info();

    const Myint64 nHaloLoc = Ugrid.getNumberOfGridPointCommHalo(GRID_LOCAL);
    const Myint64 nHaloGlob = Ugrid.getNumberOfGridPointCommHalo(GRID_GLOBAL);
    print_blank();
    printInfo(MASTER, " Halos Com Local (Pts)", nHaloLoc);
    printInfo(MASTER, " Halos Com Global (Pts)", nHaloGlob);
    printInfo(MASTER, " Halos Com Global (MB)", Myfloat(nHaloGlob * sizeof(Myfloat) / 1.e6));

    const Myfloat64 a1 = PI * cfg->param1;
    const Myfloat64 a2 = PI * cfg->param2;
    const Myfloat64 a3 = PI * cfg->param3;
    const Myfloat64 a4 = cfg->param4;

    // U = sin(a1 x1) sin(a2 x2) sin(a3 x3) a4, halos included
    Ugrid.fill(ALL_POINTS, FUNC_SINE, FUNC_SINE, FUNC_SINE, a1, a2, a3, a4);
    Ugrid.write(testCaseName + "U");

    const Myint ntry = cfg->ntry;
    const Myint64 nPtLoc = Ugrid.getNumberOfGridPoint(GRID_LOCAL, INNER_POINTS);
    const Myint64 nPtGlob = Ugrid.getNumberOfGridPoint(GRID_GLOBAL, INNER_POINTS);
    printInfo(MASTER, " Grid Inner Loc. (Pts)", nPtLoc);
    printInfo(MASTER, " Grid Inner Glob. (Pts)", nPtGlob);
    printInfo(MASTER, " Grid Inner Glob. (MB)", Myfloat(nPtGlob * sizeof(Myfloat) / 1.e6));
    print_blank();

    const Myint fdOrder = 2 * Ugrid.haloWidth;
    printInfo(MASTER, " FD space order\t", fdOrder);
    const Myfloat maxErr = 1.e-3;
    printInfo(MASTER, " Max allowed error", maxErr);

    const Myint nOpPerPoint = Ugrid.getFlopPerPtFD_D1(fdOrder);
    const Myint nPtPerStencil = Ugrid.getPtPerStencilFD_D1(fdOrder);

    // one axis: R = ampRef * (sin | cos) on the reference grid shifted half a step
    // back along that axis, ntry timed applications of the operator
    auto runAxis = [&](int axis, std::function<Rtn_code(void)> op) {
        D1Result res;
        print_blank();
        const string caseName = testCaseName + "Axis" + std::to_string(axis + 1);
        printInfo(MASTER, " * Case", caseName);
        const Myfloat64 amp[3] = {a1, a2, a3};
        const Myfloat64 ampRef = amp[axis] * a4;
        // undo the previous axis' shift, then shift this axis by -d/2
        if (axis >= 1) {
            const Myfloat dPrev = Rgrid.getSpaceSampling(axis == 1 ? AXIS1 : AXIS2);
            Rgrid.shiftOriginGrid(axis == 1 ? +dPrev / 2.0 : 0.0, axis == 2 ? +dPrev / 2.0 : 0.0, 0.0);
        }
        const Myfloat d = Rgrid.getSpaceSampling(axis == 0 ? AXIS1 : axis == 1 ? AXIS2 : AXIS3);
        Rgrid.shiftOriginGrid(axis == 0 ? -d / 2.0 : 0.0, axis == 1 ? -d / 2.0 : 0.0, axis == 2 ? -d / 2.0 : 0.0);
        Rgrid.fill(ALL_POINTS, axis == 0 ? FUNC_COSINE : FUNC_SINE, axis == 1 ? FUNC_COSINE : FUNC_SINE,
                   axis == 2 ? FUNC_COSINE : FUNC_SINE, a1, a2, a3, ampRef);
        Rgrid.write(testCaseName + "R");

        double tBest = FLT_MAX, tCom = 0;
        Wgrid.fill(ALL_POINTS, 0.0);
        for (Myint itry = 0; itry < ntry; itry++) {
            const double t0 = hpcWtime();
            Ugrid.exchangeHalos(MPI_COMM_MODE_SENDRECV);
            const double t1 = hpcWtime();
            op();
            const double t2 = hpcWtime();
            if (t2 - t0 < tBest) {
                tBest = t2 - t0;
                tCom = t1 - t0;
            }
            if (itry == 0) {
                Wgrid.write(caseName + "W");
                checkGridL1Err(INNER_POINTS, Wgrid, Rgrid, maxErr);
                res.err = Wgrid.L1Err(INNER_POINTS, Rgrid);
            }
        }
        const Myint nMemOpPerPoint = nPtPerStencil + 1;   // + 1 store
        printInfo(MASTER, " #Flop per point", nOpPerPoint);
        printInfo(MASTER, " #Point in stencil", nPtPerStencil);
        const double tFD = tBest - tCom;
        res.gflop = nPtGlob / tFD / 1.e9 * nOpPerPoint;
        res.gpointEff = nPtGlob / tBest / 1.e9;
        res.gpointFD = nPtGlob / tFD / 1.e9;
        res.gb = nPtGlob / tFD / 1.e9 * nMemOpPerPoint * sizeof(Myfloat);
        res.bestTime = tBest;
        printInfo(MASTER, " Best GFlop/s in FD", res.gflop);
        printInfo(MASTER, " Best Gpoint/s eff.", res.gpointEff);
        printInfo(MASTER, " Best Gpoint/s in FD", res.gpointFD);
        printInfo(MASTER, " Best Apparent BW GB/s", res.gb);
        if (nHaloGlob > 0) {
            printInfo(MASTER, " % MPI Comm\t", Myfloat(tCom / tBest * 100));
            printInfo(MASTER, " MPI Comm. BW GB/s", Myfloat(nHaloGlob / tCom / 1.e9 * sizeof(Myfloat)));
        }
        return res;
    };

    D1Result ax[3];
    if (cfg->dim >= DIM1) ax[0] = runAxis(0, [&] { return Ugrid.FD_D1_N1(INNER_POINTS, Wgrid, fdOrder); });
    if (cfg->dim >= DIM2) ax[1] = runAxis(1, [&] { return Ugrid.FD_D1_N2(INNER_POINTS, Wgrid, fdOrder); });
    if (cfg->dim >= DIM3) ax[2] = runAxis(2, [&] { return Ugrid.FD_D1_N3(INNER_POINTS, Wgrid, fdOrder); });

    // perf log columns 10-34 (src/testCase_FD_D1.cpp:357-385): per axis, then the
    // average of the three axes, then the cache block sizes
    if (myMpiRank == 0) {
        for (const D1Result &r : ax)
            perfLogFile << r.gflop << " " << r.gpointFD << " " << r.gpointEff << " " << r.gb << " " << r.bestTime
                        << " " << r.err << " ";
        perfLogFile << (ax[0].gflop + ax[1].gflop + ax[2].gflop) / 3.0 << " "
                    << (ax[0].gpointFD + ax[1].gpointFD + ax[2].gpointFD) / 3.0 << " "
                    << (ax[0].gpointEff + ax[1].gpointEff + ax[2].gpointEff) / 3.0 << " "
                    << (ax[0].gb + ax[1].gb + ax[2].gb) / 3.0 << " ";
        perfLogFile << cfg->cb1 << " " << cfg->cb2 << " " << cfg->cb3 << "\n";
    }
    this->finalize();
    return RTN_CODE_OK;
}

} // namespace hpcscan
