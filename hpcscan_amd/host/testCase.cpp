// TestCase base: selection by name, hardware + grid probe, perf log header,
// PASS/FAIL checks (reference behaviour: src/testCase.cpp:25-320).
#include "testCase.h"

#include <cmath>

#include <omp.h>

#include "config.h"
#include "constant.h"
#include "global.h"
#include "grid_Factory.h"
#include "output_report.h"

namespace hpcscan {

Rtn_code TestCase::initialize(void)
{
    Config *cfg = Config::Instance();
    if (cfg->testCaseName != testCaseName && cfg->testCaseName != "All") return RTN_CODE_KO;

    print_blank();
    print_line2();
    printInfo(MASTER, " TestCase name\t", testCaseName);
    printInfo(MASTER, " TestCase version", testCaseVersion);
    printInfo(MASTER, " TestCase mode\t", cfg->testMode);

    hw = Hardware_Factory::create(cfg->testMode);
    if (hw == nullptr) {
        printError("In TestCase::initialize, can not initialize hardware");
        return RTN_CODE_KO;
    }
    hw->info();

    auto gridTest = Grid_Factory::create(cfg->testMode, GRID_LOCAL);
    if (gridTest == nullptr) {
        printError("In TestCase::initialize, Not supported or invalid testMode");
        return RTN_CODE_KO;
    }
    if (gridTest->initializeGrid() != RTN_CODE_OK) {
        printError("In TestCase::initialize, Can not initialize grid with testMode");
        return RTN_CODE_KO;
    }

    if (myMpiRank == 0) {
        perfLogFile.open("hpcscan.perf." + testCaseName + ".log", ios::app);
        perfLogFile << cfg->hostName << " " << cfg->testCaseName << " " << cfg->testMode << " "
                    << cfg->propagator << " ";
        perfLogFile << nMpiProc << " " << cfg->nsub1 << " " << cfg->nsub2 << " " << cfg->nsub3 << " "
                    << omp_get_max_threads() << " " << cfg->n1 << " " << cfg->n2 << " " << cfg->n3 << " "
                    << cfg->fdOrder << " ";
    }
    testCaseStart = hpcWtime();
    hw->updateHwCounter();
    return RTN_CODE_OK;
}

void TestCase::finalize(void)
{
    testCaseEnd = hpcWtime();
    hw->updateHwCounter();
    hw->displayCounterStat();
    if (myMpiRank == 0) perfLogFile.close();
    print_blank();
    printInfo(MASTER, " Case run time (s)", testCaseEnd - testCaseStart);
    print_line2();
}

static void passFail(bool ok, const char *what, Myfloat64 v)
{
    string pass = string(" Test PASSED / ") + what;
    string fail = string(">> Test FAILED! / ") + what;
    printInfo(MASTER, ok ? pass.c_str() : fail.c_str(), v);
}

Rtn_code TestCase::checkGridL1Err(Point_type p, const Grid &g, const Grid &r, Myfloat maxAllowed)
{
    const Myfloat64 e = g.L1Err(p, r);
    passFail(!std::isnan(e) && e >= 0.0 && e < maxAllowed, "L1 err.", e);
    return RTN_CODE_OK;
}

Rtn_code TestCase::checkAllProcGridL1Err(Point_type p, const Grid &g, const Grid &r, Myfloat maxAllowed)
{
    return checkGridL1Err(p, g, r, maxAllowed);
}

Rtn_code TestCase::checkGridMaxErr(Point_type p, const Grid &g, const Grid &r, Myfloat maxAllowed)
{
    const Myfloat64 e = g.maxErr(p, r);
    passFail(!std::isnan(e) && e >= 0.0 && e < maxAllowed, "Max err.", e);
    return RTN_CODE_OK;
}

Rtn_code TestCase::checkAllProcGridMaxErr(Point_type p, const Grid &g, const Grid &r, Myfloat maxAllowed)
{
    return checkGridMaxErr(p, g, r, maxAllowed);
}

Rtn_code TestCase::checkIntegerDiff(Myint a, Myint b)
{
    printInfo(MASTER, a == b ? " Test PASSED" : ">> Test FAILED!");
    return a == b ? RTN_CODE_OK : RTN_CODE_KO;
}

Rtn_code TestCase::checkBoolDiff(bool a, bool b)
{
    printInfo(MASTER, a == b ? " Test PASSED" : ">> Test FAILED!");
    return a == b ? RTN_CODE_OK : RTN_CODE_KO;
}

Rtn_code TestCase::checkFloatDiff(Myfloat a, Myfloat b, Myfloat maxAllowed)
{
    const Myfloat e = relErr(a, b);
    if (e <= maxAllowed) {
        printInfo(MASTER, " Test PASSED / Rel err.", e);
        return RTN_CODE_OK;
    }
    printInfo(MASTER, ">> Test FAILED! / Rel err.", e);
    return RTN_CODE_KO;
}

// src/testCase.cpp:302-320
Myfloat TestCase::relErr(Myfloat a, Myfloat b)
{
    if (fabs((Myfloat64)b) > MAX_ERR_FLOAT) return fabs((Myfloat64)a - b) / fabs((Myfloat64)b);
    return fabs((Myfloat64)a - b);
}

} // namespace hpcscan
