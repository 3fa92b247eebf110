// hpcscan host driver for the stencil path (reference behaviour: src/main.cpp:32-159).
// Runs the FD_D1, FD_D2, Grid, Modeling and Propa test cases selected with -testCase on the grid mode
// selected with -testMode.  One process, no MPI.
#include "config.h"
#include "global.h"
#include "output_report.h"
#include "testCase.h"

int main(int argc, char *argv[])
{
    using namespace hpcscan;
    const double tStart = hpcWtime();

    Rtn_code rc = Config::Instance()->parse_argument(argc, argv);
    if (rc != RTN_CODE_OK) {
        Config::Instance()->finalize(rc);
        return rc == RTN_CODE_EXIT ? 0 : 1;
    }
    Config::Instance()->initialize();
    print_header_of_output_report();
    rc = Config::Instance()->info();
    if (rc != RTN_CODE_OK) {
        Config::Instance()->finalize(rc);
        return 1;
    }

    Myint numTestCase = 0;
    {
        TestCase_FD_D1 t;
        if (t.run() == RTN_CODE_OK) numTestCase++;
    }
    {
        TestCase_FD_D2 t;
        if (t.run() == RTN_CODE_OK) numTestCase++;
    }
    {
        TestCase_Grid t;
        if (t.run() == RTN_CODE_OK) numTestCase++;
    }
    {
        TestCase_Modeling t;
        if (t.run() == RTN_CODE_OK) numTestCase++;
    }
    {
        TestCase_Propa t;
        if (t.run() == RTN_CODE_OK) numTestCase++;
    }

    if (numTestCase == 0)
        printWarning(" No test case was executed");
    else
        printInfo(MASTER, " Number of test cases executed", numTestCase);
    printInfo(MASTER, " Total run time (s)", hpcWtime() - tStart);
    Config::Instance()->finalize(RTN_CODE_OK);
    return 0;
}
