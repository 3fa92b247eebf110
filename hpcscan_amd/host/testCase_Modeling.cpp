// TestCase Modeling: forward modeling in a velocity model read from disk, a
// Ricker point source and a receiver line / plane recorded every time step
// (reference behaviour: src/testCase_Modeling.cpp:32-282).
#include <cfloat>

#include "config.h"
#include "constant.h"
#include "data_Acquisition.h"
#include "fdm.h"
#include "global.h"
#include "grid_Factory.h"
#include "output_report.h"
#include "propagator_Ac2.h"
#include "testCase.h"

namespace hpcscan {

TestCase_Modeling::TestCase_Modeling(void)
{
    testCaseName = "Modeling";
    testCaseVersion = "Standard implementation";
}

Rtn_code TestCase_Modeling::run(void)
{
    if (this->initialize() == RTN_CODE_KO) return RTN_CODE_KO;

    Config *cfg = Config::Instance();
    auto propa = Propagator_Factory::create(cfg->propagator);
    if (propa == nullptr) return RTN_CODE_KO;
    if (propa->initialize(MODELING) != RTN_CODE_OK) {
        printError("In TestCase_Modeling::run, propa->initialize() not Ok");
        return RTN_CODE_OK;
    }
    propa->info();

    auto prnGrid = propa->prnGrid;
    auto prcGrid = propa->prcGrid;
    if (myMpiRank == 0) prnGrid->info();

    const Myint64 nHaloLoc = prnGrid->getNumberOfGridPointCommHalo(GRID_LOCAL);
    const Myint64 nHaloGlob = prnGrid->getNumberOfGridPointCommHalo(GRID_GLOBAL);
    print_blank();
    printInfo(MASTER, " Halos Com Local (Pts)", nHaloLoc);
    printInfo(MASTER, " Halos Com Global (Pts)", nHaloGlob);
    printInfo(MASTER, " Halos Com Global (MB)", Myfloat(nHaloGlob * sizeof(Myfloat) / 1.e6));

    double tBest = 0.0, tCom = 0.0;
    const Myint ntry = cfg->ntry;
    const Myint64 nPtLoc = prnGrid->getNumberOfGridPoint(GRID_LOCAL, INNER_POINTS);
    const Myint64 nPtGlob = prnGrid->getNumberOfGridPoint(GRID_GLOBAL, INNER_POINTS);
    printInfo(MASTER, " Grid Inner Loc. (Pts)", nPtLoc);
    printInfo(MASTER, " Grid Inner Glob. (Pts)", nPtGlob);
    printInfo(MASTER, " Grid Inner Glob. (MB)", Myfloat(nPtGlob * sizeof(Myfloat) / 1.e6));
    print_blank();

    const Myint fdOrder = propa->fdOrder;
    const Myint nt = propa->nt;
    const Myfloat dt = propa->dt;
    const Myint snapInc = propa->snapInc;
    const Myfloat maxErr = 1.0;
    printInfo(MASTER, " Max allowed error", maxErr);

    const Myint nPtPerStencil = prcGrid->getPtPerStencilFD_LAPLACIAN(fdOrder);
    const Myint nMemOpPerPoint = nPtPerStencil + 4;                          // + 1 store + 3 load
    const Myint nOpPerPoint = prcGrid->getFlopPerPtFD_LAPLACIAN(fdOrder) + 4;  // + ADD SUB 2 MUL
    Myfloat32 propaGflop = 0, propaGpointEff = 0, propaGpointFD = 0, propaGB = 0;

    auto acquiPtr = DataAcquisition_Factory::create(cfg->testMode);
    DataAcquisition &acqui = *acquiPtr;
    if (acqui.initialize(ACQUI_BUILTIN, *prnGrid, nt) != RTN_CODE_OK) return RTN_CODE_KO;

    print_blank();
    const string caseName = testCaseName + "Forward";
    printInfo(MASTER, " * Case", caseName);

    tBest = FLT_MAX;
    Myfloat32 errTestCase = 0;
    Myint ntCheck = 0;
    Myint nIterSnap = 0;
    double totalComputeTime = 0;

    for (Myint itry = 0; itry < ntry; itry++) {
        ntCheck = 0;
        const double t0 = hpcWtime();
        totalComputeTime = 0;
        for (Myint it = 0; it < nt; it++) {
            const double t0c = hpcWtime();
            propa->computeWavefieldNextTimeStep(*prnGrid, *prcGrid);
            totalComputeTime += hpcWtime() - t0c;

            acqui.appliSourceTerm(*prnGrid, FORWARD, it, dt);

            if (it % snapInc == 0) {
                printInfo(MASTER, " Snapshot at time", Myfloat(it * dt));
                prnGrid->writeGlobal(INNER_POINTS, caseName + "Prn", nIterSnap);
                nIterSnap++;
            }

            acqui.recordTrace(*prnGrid, FORWARD, it, dt);

            auto tmp = prnGrid;
            prnGrid = prcGrid;
            prcGrid = tmp;

            hw->watchTimeAndUpdateHwCounter();
        }
        acqui.writeTrace(caseName + "Prn");

        const double t = hpcWtime() - t0;
        if (t < tBest) tBest = t;
        if (itry == 0) printInfo(MASTER, " ntCheck", ntCheck);
    }

    printInfo(MASTER, " #Flop per point", nOpPerPoint);
    printInfo(MASTER, " #Point in stencil", nPtPerStencil);

    const double timeInFD = propa->compute_pressure_time;
    propaGflop = nt * nPtGlob / timeInFD / 1.e9 * nOpPerPoint;
    propaGpointEff = nt * nPtGlob / tBest / 1.e9;
    propaGpointFD = nt * nPtGlob / timeInFD / 1.e9;
    propaGB = nt * nPtGlob / timeInFD / 1.e9 * nMemOpPerPoint * sizeof(Myfloat);

    printInfo(MASTER, " Total compute time ", totalComputeTime);
    printInfo(MASTER, " Total compute speed ", nt * nPtGlob / totalComputeTime / 1.e9);

    // i/o = grid writes (if asked) + trace write + model read (src/testCase_Modeling.cpp:201-203)
    const double ioTime =
        propa->i_o_read_time + acqui.trace_write_time + prnGrid->grid_writeGlobal_time + propa->i_o_write_coefgrid_time;
    printInfo(MASTER, "\n Total i/o time ", ioTime);
    const double commTime = propa->halo_comm_time + acqui.trace_gather_time;
    printInfo(MASTER, " Total communication time ", commTime);

    printInfo(MASTER, "\n total Read model time ", propa->coef_read_total_time);
    printInfo(MASTER, " i_o Read model time ", propa->i_o_read_time);
    printInfo(MASTER, " i_o Write trace time ", acqui.trace_write_time);
    printInfo(MASTER, " Halo communication time ", propa->halo_comm_time);
    printInfo(MASTER, " Apply boundary condition time ", propa->boundary_condition_time);
    printInfo(MASTER, " Compute pressure time ", propa->compute_pressure_time);
    printInfo(MASTER, " Record trace time ", acqui.record_trace_time);
    if (cfg->writeGrid) {
        printInfo(MASTER, " Total write snapshotgrid time ", prnGrid->grid_writeGlobal_time);
        printInfo(MASTER, " Total write coefGrid time ", propa->i_o_write_coefgrid_time);
    }
    printInfo(MASTER, " Total write trace time ", acqui.total_timeWriteTrace);

    printInfo(MASTER, " \nBest GFlop/s in FD", propaGflop);
    printInfo(MASTER, " Best Gpoint/s eff.", propaGpointEff);
    printInfo(MASTER, " Best Gpoint/s in FD", propaGpointFD);
    printInfo(MASTER, " Best Apparent BW GB/s", propaGB);
    if (nHaloGlob > 0) {
        printInfo(MASTER, " % MPI Comm\t", Myfloat(tCom / tBest * 100));
        printInfo(MASTER, " MPI Comm. BW GB/s", Myfloat(nHaloGlob / tCom / 1.e9 * sizeof(Myfloat)));
    }

    // perf log columns 10-31 (+ 32, 33 with -writeGrid) (src/testCase_Modeling.cpp:231-270)
    if (myMpiRank == 0) {
        perfLogFile << propaGflop << " " << propaGpointFD << " " << propaGpointEff << " " << propaGB << " "
                    << tBest << " " << errTestCase << " " << propa->nt << " " << (Myfloat32)propa->dt << " "
                    << (Myfloat32)propa->stableDt << " " << propa->halo_comm_time << " "
                    << propa->boundary_condition_time << " " << propa->compute_pressure_time << " "
                    << totalComputeTime << " " << nt * nPtGlob / totalComputeTime / 1.e9 << " "
                    << propa->i_o_read_time << " " << acqui.trace_write_time << " " << propa->coef_read_total_time
                    << " " << acqui.record_trace_time << " " << acqui.total_timeWriteTrace << " "
                    << acqui.trace_gather_time << " " << ioTime << " " << commTime << " ";
        if (cfg->writeGrid)
            perfLogFile << prnGrid->grid_writeGlobal_time << " " << propa->i_o_write_coefgrid_time << " ";
        perfLogFile << "\n";
    }
    this->finalize();
    return RTN_CODE_OK;
}

} // namespace hpcscan
