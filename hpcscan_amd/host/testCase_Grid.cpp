// TestCase Grid: the Grid operations around the stencil (fill, error norms,
// reductions, the leapfrog update, the free surface), each timed over ntry
// tries and checked against a closed-form answer on its first try; reports
// GByte/s and GPoint/s per operation (reference behaviour:
// src/testCase_Grid.cpp:32-525; case names, checks, byte counts and perf-log
// columns 10-27 are the reference's).
#include <cfloat>
#include <cmath>
#include <functional>

#include "config.h"
#include "constant.h"
#include "global.h"
#include "grid_Factory.h"
#include "output_report.h"
#include "testCase.h"

namespace hpcscan {

TestCase_Grid::TestCase_Grid(void)
{
    testCaseName = "Grid";
    testCaseVersion = "Standard implementation";
}

Rtn_code TestCase_Grid::run(void)
{
    if (this->initialize() == RTN_CODE_KO) return RTN_CODE_KO;

    Config *cfg = Config::Instance();
    const Myfloat64 a1 = cfg->param1, a2 = cfg->param2, a3 = cfg->param3, a4 = cfg->param4;
    const string gridMode = cfg->testMode;
    auto W2 = Grid_Factory::create(gridMode, GRID_LOCAL);
    auto R2 = Grid_Factory::create(gridMode, GRID_LOCAL);
    Grid &W = *W2, &Rg = *R2;
    W.initializeGrid();
    Rg.initializeGrid();
    W.info();

    const Myint64 nLoc = W.getNumberOfGridPoint(GRID_LOCAL, INNER_POINTS);
    const Myint64 nGlob = W.getNumberOfGridPoint(GRID_GLOBAL, INNER_POINTS);
    printInfo(MASTER, " Inner pts (Loc.)", nLoc);
    printInfo(MASTER, " Inner pts (Glob.)", nGlob);
    const Myint ntry = cfg->ntry;

    // one case: time ntry calls of op, check the first one; bytesPerPoint
    // counts the grid values the operation reads and writes
    struct Perf {
        Myfloat32 gb = 0, gpoint = 0;
    };
    auto runCase = [&](const string &name, Myint bytesPerPoint, std::function<void(void)> op,
                       std::function<void(void)> check, std::function<void(void)> header = nullptr) {
        print_blank();
        printInfo(MASTER, " * Case", testCaseName + name);
        if (header) header();
        double best = FLT_MAX;
        for (Myint itry = 0; itry < ntry; itry++) {
            const double t0 = hpcWtime();
            op();
            const double t = hpcWtime() - t0;
            best = t < best ? t : best;
            if (itry == 0) check();
        }
        Perf p;
        p.gb = Myfloat(bytesPerPoint * nGlob * sizeof(Myfloat) / best / 1.e9);
        p.gpoint = Myfloat(nGlob / best / 1.e9);
        printInfo(MASTER, " Best achieved GByte/s", p.gb);
        printInfo(MASTER, " Best achieved GPoint/s", p.gpoint);
        return p;
    };

    // Fill: W = a1
    Rg.fill(INNER_POINTS, a1);
    const Perf fill = runCase(
        "Fill", 1, [&] { W.fill(INNER_POINTS, a1); },
        [&] { checkAllProcGridL1Err(INNER_POINTS, W, Rg, MAX_ERR_FLOAT); });

    // R = a1 everywhere inside, W = a2 except one point 2 a2
    Rg.fill(INNER_POINTS, a1);
    W.fill(INNER_POINTS, a2);
    if (myMpiRank == 0) W.fill(MIDDLE_POINT, 2 * a2);
    const Myfloat64 sumDiff = (nGlob - 1) * fabs(a2 - a1) + fabs(2 * a2 - a1);
    const Myfloat64 sumRef = nGlob * fabs(a1);

    Myfloat res = 0;
    const Perf maxErr = runCase(
        "MaxErr", 2, [&] { res = W.maxErr(INNER_POINTS, Rg); },
        [&] { checkFloatDiff(res, Myfloat(fabs(2 * a2 - a1) / fabs(a1)), MAX_ERR_FLOAT); });
    const Perf l1Err = runCase(
        "L1Err", 2, [&] { res = W.allProcL1Err(INNER_POINTS, Rg); },
        [&] { checkFloatDiff(res, Myfloat(sumDiff / sumRef), MAX_ERR_FLOAT); });
    const Perf sumAbs = runCase(
        "GetSumAbs", 1, [&] { res = Rg.getSumAbs(INNER_POINTS); },
        [&] { checkFloatDiff(res, Myfloat(sumRef), MAX_ERR_FLOAT); });
    const Perf sumAbsDiff = runCase(
        "GetSumAbsDiff", 2, [&] { res = W.getSumAbsDiff(INNER_POINTS, Rg); },
        [&] { checkFloatDiff(res, Myfloat(sumDiff), MAX_ERR_FLOAT); });
    const Perf getMax = runCase(
        "GetMax", 1, [&] { res = W.getMax(INNER_POINTS); },
        [&] { checkFloatDiff(res, Myfloat(a2 < 0.0 ? a2 : 2 * a2), MAX_ERR_FLOAT); });
    const Perf getMin = runCase(
        "GetMin", 1, [&] { res = W.getMin(INNER_POINTS); },
        [&] { checkFloatDiff(res, Myfloat(a2 < 0.0 ? 2 * a2 : a2), MAX_ERR_FLOAT); });

    // UpdatePressure: W = 2 prc - W + coef * lapla with constant grids
    Perf update;
    {
        auto prc2 = Grid_Factory::create(gridMode, GRID_LOCAL);
        auto coef2 = Grid_Factory::create(gridMode, GRID_LOCAL);
        auto lapla2 = Grid_Factory::create(gridMode, GRID_LOCAL);
        Grid &prc = *prc2, &coef = *coef2, &lapla = *lapla2;
        prc.initializeGrid();
        coef.initializeGrid();
        lapla.initializeGrid();
        W.fill(INNER_POINTS, a1);
        prc.fill(INNER_POINTS, a2);
        coef.fill(INNER_POINTS, a3);
        lapla.fill(INNER_POINTS, a4);
        const Myfloat valRef = 2.0 * a2 - a1 + a3 * a4;
        Rg.fill(INNER_POINTS, valRef);
        update = runCase(
            "UpdatePressure", 5, [&] { W.updatePressure(INNER_POINTS, prc, coef, lapla); },
            [&] { checkAllProcGridL1Err(INNER_POINTS, W, Rg, MAX_ERR_FLOAT); });
    }

    // ApplyBoundaryCondition: a sine of period 2 on the unit grid with halos set
    // to -1; the free surface must restore the sine everywhere
    Perf bc;
    {
        const string caseName = testCaseName + "ApplyBoundaryCondition";
        W.defineUnitGrid();
        Rg.defineUnitGrid();
        W.fill(ALL_POINTS, FUNC_SINE, FUNC_SINE, FUNC_SINE, PI, PI, PI, 1.0);
        const Myfloat fillVal = -1;
        const Point_type faces[6] = {I1HALO1, I1HALO2, I2HALO1, I2HALO2, I3HALO1, I3HALO2};
        for (int f = 0; f < 2 * W.dim; f++)
            if (W.getNeighbourProc(faces[f]) == PROC_NULL) W.fill(faces[f], fillVal);
        W.exchangeHalos(MPI_COMM_MODE_SENDRECV);
        Rg.fill(ALL_POINTS, FUNC_SINE, FUNC_SINE, FUNC_SINE, PI, PI, PI, 1.0);
        Rg.write(caseName + "Ref");
        const Myint64 nBound = W.getNumberOfGridPointBoundaryCondition(BOUND_COND_ANTI_MIRROR);
        bc = runCase(
            "ApplyBoundaryCondition", 2, [&] { W.applyBoundaryCondition(BOUND_COND_ANTI_MIRROR); },
            [&] {
                checkAllProcGridL1Err(ALL_POINTS, W, Rg, MAX_ERR_FLOAT);
                W.write(caseName + "W");
            },
            [&] {
                printInfo(MASTER, " Boundary pts (Loc.)", nBound);
                printInfo(MASTER, " Boundary pts (Glob.)", nBound);   // one rank
                printInfo(MASTER, " % Boundary vs Inner", Myfloat(nBound) / nGlob * 100);
            });
    }

    if (myMpiRank == 0) {
        perfLogFile << fill.gb << " " << fill.gpoint << " " << maxErr.gb << " " << maxErr.gpoint << " "
                    << l1Err.gb << " " << l1Err.gpoint << " " << sumAbs.gb << " " << sumAbs.gpoint << " "
                    << sumAbsDiff.gb << " " << sumAbsDiff.gpoint << " " << getMax.gb << " " << getMax.gpoint << " "
                    << getMin.gb << " " << getMin.gpoint << " " << update.gb << " " << update.gpoint << " "
                    << bc.gb << " " << bc.gpoint << " "
                    << "\n";
    }

    this->finalize();
    return RTN_CODE_OK;
}

} // namespace hpcscan
