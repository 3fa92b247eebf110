// Acoustic 2nd-order propagator: initialisation (CFL, dt, nt, eigen mode),
// time step and the SplitComp variant.
//
// Scalar arithmetic keeps the reference's Myfloat typing because it changes nt
// and dt by one ulp / one step (src/propagator_Ac2.cpp:84-304, 410-547):
// dt, stableDt, tmax and the check times are Myfloat, the CFL is Myfloat.
#include "propagator_Ac2.h"

#include <cmath>

#include "config.h"
#include "constant.h"
#include "fdm.h"
#include "global.h"
#include "grid_Factory.h"
#include "output_report.h"

namespace hpcscan {

Propagator_Ac2::Propagator_Ac2(void)
{
    propaWaveEq = "Acoustic Wave Eq. 2nd order";
    propaKernelType = "Standard implementation";
    boundCondType = BOUND_COND_ANTI_MIRROR;
    fdOrder = Config::Instance()->fdOrder;
    snapInc = 0;
    tmax = dt = stableDt = freqMax = minVelocity = maxVelocity = 0.0;
    nt = 0;
    halo_comm_time = boundary_condition_time = compute_pressure_time = 0.0;
    coef_read_total_time = i_o_read_time = i_o_write_coefgrid_time = 0.0;
}

Rtn_code Propagator_Ac2::initialize(PropaInit_type propaInitType)
{
    Config *cfg = Config::Instance();
    const string gridMode = cfg->testMode;
    coefGrid = Grid_Factory::create(gridMode, GRID_LOCAL);
    prnGrid = Grid_Factory::create(gridMode, GRID_LOCAL);
    prcGrid = Grid_Factory::create(gridMode, GRID_LOCAL);
    if (!coefGrid || !prnGrid || !prcGrid) return RTN_CODE_KO;
    if (coefGrid->initializeGrid() != RTN_CODE_OK || prnGrid->initializeGrid() != RTN_CODE_OK ||
        prcGrid->initializeGrid() != RTN_CODE_OK)
        return RTN_CODE_KO;

    if (cfg->boundary == "FreeSurf") {
        boundCondType = BOUND_COND_ANTI_MIRROR;
    } else if (cfg->boundary == "None") {
        boundCondType = NO_BOUND_COND;
    } else {
        printError("In Propagator_Ac2::initialize, invalid boundary", cfg->boundary);
        return RTN_CODE_KO;
    }

    if (propaInitType == EIGEN_MODE) {
        coefGrid->defineUnitGrid();
        prnGrid->defineUnitGrid();
        prcGrid->defineUnitGrid();
        const Myfloat Vp = 1.0;
        minVelocity = maxVelocity = Vp;
    } else if (propaInitType == MODELING) {
        // Vp model read from disk into coefGrid (src/propagator_Ac2.cpp:138-160)
        if (cfg->modelVpFile == "UNSPECIFIED") {
            printError("In Propagator_Ac2::initialize, modelVpFile has not been specified");
            return RTN_CODE_KO;
        }
        coefGrid->fill(ALL_POINTS, 0.0);
        const double t0 = hpcWtime();
        if (coefGrid->read(INNER_POINTS, cfg->modelVpFile) != RTN_CODE_OK) return RTN_CODE_KO;
        coef_read_total_time = hpcWtime() - t0;
        i_o_read_time = coefGrid->i_o_pread_time;
        minVelocity = coefGrid->allProcGetMin(INNER_POINTS);
        maxVelocity = coefGrid->allProcGetMax(INNER_POINTS);
    } else {
        printError("In Propagator_Ac2::initialize, propaInitType not supported", (Myint)propaInitType);
        return RTN_CODE_KO;
    }
    if (minVelocity <= 0.0) {
        printError("In Propagator_Ac2::initialize, minVelocity <= 0.0");
        return RTN_CODE_KO;
    }

    // stable dt = CFL * min(h) / vmax, all Myfloat (src/propagator_Ac2.cpp:181)
    stableDt = getCFL() * coefGrid->getMinSpaceSampling() / maxVelocity;
    if (cfg->dt > 0.0)
        dt = cfg->dt;
    else
        dt = stableDt * cfg->ratioCFL;

    // snapshot increment (src/propagator_Ac2.cpp:197-212)
    if (cfg->snapDt > 0.0) {
        const Myint incTmp = ceil(cfg->snapDt / dt);
        dt = cfg->snapDt / incTmp;
        snapInc = cfg->snapDt / dt;
        if ((cfg->snapDt - (snapInc * dt)) > dt / 2.0) snapInc++;
    } else {
        snapInc = cfg->snapInc;
    }

    // tmax and nt (src/propagator_Ac2.cpp:218-226)
    if (cfg->tmax != UNSPECIFIED) {
        tmax = cfg->tmax;
        nt = (tmax / dt) + 1;
    } else {
        nt = cfg->nt;
        tmax = (nt - 1) * dt;
    }

    if (propaInitType == EIGEN_MODE) {
        coefGrid->fill(INNER_POINTS, minVelocity * minVelocity * dt * dt);
    } else {
        // coef = vp * vp * dt * dt, left to right in Myfloat (src/propagator_Ac2.cpp:236-258),
        // as grid operations so that it runs where the grid lives (the halos stay 0)
        auto dtGrid = Grid_Factory::create(gridMode, GRID_LOCAL);
        if (!dtGrid || dtGrid->initializeGrid() != RTN_CODE_OK) return RTN_CODE_KO;
        dtGrid->fillArray(dt);
        coefGrid->multiplyArray(*coefGrid, *coefGrid);
        coefGrid->multiplyArray(*coefGrid, *dtGrid);
        coefGrid->multiplyArray(*coefGrid, *dtGrid);
    }
    coefGrid->write("coefGrid");
    i_o_write_coefgrid_time = coefGrid->grid_write_time;

    if (propaInitType == EIGEN_MODE) {
        // prn at t = -2 dt, prc at t = -dt (src/propagator_Ac2.cpp:270-285)
        Myfloat timeSec = -2.0 * dt;
        if (initializeGrid(*prnGrid, propaInitType, timeSec) != RTN_CODE_OK) {
            printError("In Propagator_Ac2::initialize, initializeGrid not Ok");
            return RTN_CODE_KO;
        }
        timeSec = -dt;
        if (initializeGrid(*prcGrid, propaInitType, timeSec) != RTN_CODE_OK) {
            printError("In Propagator_Ac2::initialize, initializeGrid not Ok");
            return RTN_CODE_KO;
        }
    } else {
        // zero wavefields, maximum frequency from the command line (src/propagator_Ac2.cpp:287-296)
        prnGrid->fill(ALL_POINTS, 0.0);
        prcGrid->fill(ALL_POINTS, 0.0);
        freqMax = cfg->freqMax;
    }

    halo_comm_time = compute_pressure_time = boundary_condition_time = 0.0;
    return RTN_CODE_OK;
}

void Propagator_Ac2::info(void)
{
    print_blank();
    printInfo(MASTER, " * Propagator parameters *");
    printInfo(MASTER, " Wave equation\t", propaWaveEq);
    printInfo(MASTER, " Kernel type\t", propaKernelType);
    printInfo(MASTER, " Boundary type\t", boundCondType == NO_BOUND_COND ? "None" : "Free Surface");
    printInfo(MASTER, " FD space order\t", fdOrder);
    printInfo(MASTER, " CFL \t\t", getCFL());
    printInfo(MASTER, " Stable time step (s)", stableDt);
    printInfo(MASTER, " No. time steps\t", nt);
    printInfo(MASTER, " Time step (s)\t", dt);
    printInfo(MASTER, " Ratio Eff. / Stable dt", dt / stableDt);
    printInfo(MASTER, " Snapshot inc. (sec.)", (snapInc - 1) * dt);
    printInfo(MASTER, " Snapshot inc. (steps)", snapInc);
    printInfo(MASTER, "");
    printInfo(MASTER, " Max. frequency (Hz)", freqMax);
    printInfo(MASTER, " Min. period T (s)", 1.0 / freqMax);
    printInfo(MASTER, " Max. time (s)\t", tmax);
    printInfo(MASTER, " No. of periods\t", tmax * freqMax);
    printInfo(MASTER, " No. of steps / period", (1.0 / freqMax) / dt);
    printInfo(MASTER, "");
    printInfo(MASTER, " Min. velocity (m/s)", minVelocity);
    printInfo(MASTER, " Max. velocity (m/s)", maxVelocity);
    const Myfloat lambdaMin = minVelocity / freqMax;
    printInfo(MASTER, " Min. lambda (m)", lambdaMin);
    printInfo(MASTER, " Min. grid spacing (m)", prnGrid->getMinSpaceSampling());
    const Axis_type ax[3] = {AXIS1, AXIS2, AXIS3};
    const char *lbl[3] = {" No. of lambda axis1", " No. of lambda axis2", " No. of lambda axis3"};
    for (int a = 0; a < (int)prnGrid->dim; a++) {
        const Myfloat size = prnGrid->getMaxCoord(ax[a]) - prnGrid->getMinCoord(ax[a]);
        printInfo(MASTER, lbl[a], size / lambdaMin);
    }
    printInfo(MASTER, " No. grid pts / lambda", (minVelocity / freqMax) / prnGrid->getMinSpaceSampling());
}

// src/propagator_Ac2.cpp:366-393: halos, boundary, FD; each phase timed separately
Rtn_code Propagator_Ac2::computeWavefieldNextTimeStep(Grid &prnGridIn, Grid &prcGridIn)
{
    const double t0 = hpcWtime();
    prcGridIn.myExchangeAll_halos();
    const double t1 = hpcWtime();
    halo_comm_time += t1 - t0;

    const double t2 = hpcWtime();
    prcGridIn.applyBoundaryCondition(boundCondType);
    const double t3 = hpcWtime();
    boundary_condition_time += t3 - t2;

    const double t4 = hpcWtime();
    Rtn_code rc = computePressureWithFD(prnGridIn, prcGridIn);
    const double t5 = hpcWtime();
    compute_pressure_time += t5 - t4;
    return rc;
}

Rtn_code Propagator_Ac2::computePressureWithFD(Grid &prnGridIn, Grid &prcGridIn)
{
    return prnGridIn.computePressureWithFD(prcGridIn, *coefGrid, fdOrder);
}

// CFL = sqrt(sum|O2 coef| / (dim * sum|O_k coef|)) (src/propagator_Ac2.cpp:410-439)
Myfloat Propagator_Ac2::getCFL(void)
{
    if (coefGrid == nullptr) {
        printInfo(MASTER, "In Propagator_Ac2::getCFL, coefGrid == nullptr");
        return -1.0;
    }
    const Myfloat64 a1 = getSumAbsFD_D2Coef(2);
    const Myfloat64 a2 = getSumAbsFD_D2Coef(fdOrder) * coefGrid->dim;
    const Myfloat CFL = sqrt(a1 / a2);
    return CFL;
}

// Eigen mode p = prod_i sin(pi m_i x_i) sin(pi gamma t), gamma = |m|
// (src/propagator_Ac2.cpp:441-547); m_i = ceil(param_i)
Rtn_code Propagator_Ac2::initializeGrid(Grid &gridIn, PropaInit_type propaInitType, Myfloat timeSec)
{
    if (propaInitType != EIGEN_MODE) {
        printError("In Propagator_Ac2::initializeGrid, propaInitType not supported", (Myint)propaInitType);
        return RTN_CODE_KO;
    }
    Config *cfg = Config::Instance();
    const Myint m[3] = {(Myint)ceil(cfg->param1), (Myint)ceil(cfg->param2), (Myint)ceil(cfg->param3)};
    const char *name[3] = {"nMode1 < 1", "nMode2 < 1", "nMode3 < 1"};
    Myfloat64 val[3] = {1.0, 1.0, 1.0};
    Myint sq = 0;
    for (int a = 0; a < (int)cfg->dim; a++) {
        if (m[a] < 1) {
            printError(name[a]);
            return RTN_CODE_KO;
        }
        val[a] = PI * m[a];
        sq += m[a] * m[a];
    }
    const Myfloat64 gamma = cfg->dim == DIM1 ? (Myfloat64)m[0] : sqrt(sq);
    const Myfloat64 amp = sin(PI * timeSec * gamma);
    gridIn.fill(ALL_POINTS, FUNC_SINE, FUNC_SINE, FUNC_SINE, val[0], val[1], val[2], amp);
    freqMax = gamma / 2.0;
    return RTN_CODE_OK;
}

// ---------------------------------------------------------------------------
// SplitComp: Laplacian into a fourth grid, then the leapfrog update
// (src/propagator_Ac2SplitComp.cpp:26-85)
// ---------------------------------------------------------------------------
Propagator_Ac2SplitComp::Propagator_Ac2SplitComp(void) : Propagator_Ac2()
{
    propaKernelType = "With Laplacian operator";
    laplaGrid = nullptr;
}

Rtn_code Propagator_Ac2SplitComp::initialize(PropaInit_type propaInitType)
{
    Rtn_code rc = Propagator_Ac2::initialize(propaInitType);
    if (rc != RTN_CODE_OK) return rc;
    laplaGrid = Grid_Factory::create(Config::Instance()->testMode, GRID_LOCAL);
    if (!laplaGrid || laplaGrid->initializeGrid() != RTN_CODE_OK) return RTN_CODE_KO;
    if (propaInitType == EIGEN_MODE) laplaGrid->defineUnitGrid();
    return RTN_CODE_OK;
}

Rtn_code Propagator_Ac2SplitComp::computePressureWithFD(Grid &prnGridIn, Grid &prcGridIn)
{
    Rtn_code rc = prcGridIn.FD_LAPLACIAN(INNER_POINTS, *laplaGrid, fdOrder);
    if (rc != RTN_CODE_OK) return rc;
    return prnGridIn.updatePressure(INNER_POINTS, prcGridIn, *coefGrid, *laplaGrid);
}

shared_ptr<Propagator_Ac2> Propagator_Factory::create(string propagator)
{
    if (propagator == PROPA_TYPE_AC2STANDARD) return make_shared<Propagator_Ac2>();
    if (propagator == PROPA_TYPE_AC2SPLITCOMP) return make_shared<Propagator_Ac2SplitComp>();
    printError("IN Propagator_Factory::create, invalid propagator", propagator);
    return nullptr;
}

} // namespace hpcscan
