// TestCase FD_D2: D2 along each axis and the Laplacian of a separable sine,
// checked against the analytic derivative, with GPoint/s and "apparent" GB/s
// (reference behaviour: src/testCase_FD_D2.cpp:32-488).
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

TestCase_FD_D2::TestCase_FD_D2(void)
{
    testCaseName = "FD_D2";
    testCaseVersion = "Standard implementation";
}

namespace {
struct CaseResult {
    Myfloat32 gflop = 0, gpointFD = 0, gpointEff = 0, gb = 0;
    double bestTime = 0;
    Myfloat32 err = 0;
};
} // namespace

Rtn_code TestCase_FD_D2::run(void)
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

    // U = sin(a1 x1) sin(a2 x2) sin(a3 x3) a4, halos included (no boundary condition)
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

    // one case: fill R with ampRef * sines, time ntry applications of op
    auto runCase = [&](const string &caseName, Myfloat64 ampRef, std::function<Rtn_code(void)> op,
                       Myint nOpPerPoint, Myint nPtPerStencil) {
        CaseResult res;
        print_blank();
        printInfo(MASTER, " * Case", caseName);
        Rgrid.fill(ALL_POINTS, FUNC_SINE, FUNC_SINE, FUNC_SINE, a1, a2, a3, ampRef);
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
        const Myint nMemOpPerPoint = nPtPerStencil + 1;
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

    CaseResult ax[3], lap;
    const Myint opD2 = Ugrid.getFlopPerPtFD_D2(fdOrder);
    const Myint ptD2 = Ugrid.getPtPerStencilFD_D2(fdOrder);
    if (cfg->dim >= DIM1)
        ax[0] = runCase(testCaseName + "Axis1", -a1 * a1 * a4,
                        [&] { return Ugrid.FD_D2_N1(INNER_POINTS, Wgrid, fdOrder); }, opD2, ptD2);
    if (cfg->dim >= DIM2)
        ax[1] = runCase(testCaseName + "Axis2", -a2 * a2 * a4,
                        [&] { return Ugrid.FD_D2_N2(INNER_POINTS, Wgrid, fdOrder); }, opD2, ptD2);
    if (cfg->dim >= DIM3)
        ax[2] = runCase(testCaseName + "Axis3", -a3 * a3 * a4,
                        [&] { return Ugrid.FD_D2_N3(INNER_POINTS, Wgrid, fdOrder); }, opD2, ptD2);
    {
        string name;
        Myfloat64 ampRef = 0.0;
        if (cfg->dim == DIM1) {
            name = testCaseName + "Laplacian1D";
            ampRef = -a1 * a1 * a4;
        } else if (cfg->dim == DIM2) {
            name = testCaseName + "Laplacian2D";
            ampRef = -(a1 * a1 + a2 * a2) * a4;
        } else {
            name = testCaseName + "Laplacian3D";
            ampRef = -(a1 * a1 + a2 * a2 + a3 * a3) * a4;
        }
        lap = runCase(name, ampRef, [&] { return Ugrid.FD_LAPLACIAN(INNER_POINTS, Wgrid, fdOrder); },
                      Ugrid.getFlopPerPtFD_LAPLACIAN(fdOrder), Ugrid.getPtPerStencilFD_LAPLACIAN(fdOrder));
    }

    // perf log columns 10-36 (src/testCase_FD_D2.cpp:458-482)
    if (myMpiRank == 0) {
        for (const CaseResult *r : {&ax[0], &ax[1], &ax[2], &lap})
            perfLogFile << r->gflop << " " << r->gpointFD << " " << r->gpointEff << " " << r->gb << " "
                        << r->bestTime << " " << r->err << " ";
        perfLogFile << cfg->cb1 << " " << cfg->cb2 << " " << cfg->cb3 << "\n";
    }
    this->finalize();
    return RTN_CODE_OK;
}

} // namespace hpcscan
