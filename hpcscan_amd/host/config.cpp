// Configuration: defaults, command-line parsing and consistency checks.
// Behaviour follows src/config.cpp:32-1208 (flag names, defaults, validation
// rules and exit codes); the parser itself is table driven.
#include "config.h"

#include <cstdlib>
#include <cstring>
#include <functional>
#include <unistd.h>
#include <vector>

#include <omp.h>

#include "constant.h"
#include "fdm.h"
#include "global.h"
#include "grid_Factory.h"
#include "output_report.h"

namespace hpcscan {

Config *Config::m_pInstance = nullptr;

Config *Config::Instance()
{
    if (!m_pInstance) m_pInstance = new Config;
    return m_pInstance;
}

// defaults: src/config.cpp:32-76
Config::Config(void)
{
    autoPad = false;
    boundary = "FreeSurf";
    cb1 = 9999; cb2 = 4; cb3 = 16;
    d1 = d2 = d3 = PI / 30;
    dim = DIM3;
    dt = 0.0;
    fdOrder = 8;
    freqMax = 25.0;
    gpuBlkSize = 256; gpuBlkSize1 = 32; gpuBlkSize2 = 2; gpuBlkSize3 = 16;
    gpuGridSize = 512;
    gpuMpiAware = false;
    hwCounterDt = 0;
    modelVpFile = "UNSPECIFIED";
    n1 = n2 = n3 = 61;
    n1AddPad = n2AddPad = n3AddPad = UNSPECIFIED;
    n1MulPad = n2MulPad = n3MulPad = UNSPECIFIED;
    n1Offset = n2Offset = n3Offset = UNSPECIFIED;
    nlayer = 0;
    nt = 20;
    ntry = 10;
    nsub1 = nsub2 = nsub3 = 1;
    param1 = 0.8; param2 = 0.9; param3 = 1.0; param4 = 1.1;
    propagator = PROPA_TYPE_AC2STANDARD;
    ratioCFL = 1.0;
    snapDt = 0.0;
    snapInc = 1;
    testCaseName = "UNSPECIFIED";
    testMode = GRID_MODE_BASELINE;
    tmax = UNSPECIFIED;
    writeGrid = false;

    memset(hostName, 0, sizeof(hostName));
    memset(userName, 0, sizeof(userName));
    gethostname(hostName, HOST_NAME_MAX);
    if (getlogin_r(userName, LOGIN_NAME_MAX) != 0) {
        const char *u = getenv("USER");
        strncpy(userName, u ? u : "unknown", LOGIN_NAME_MAX);
    }
}

namespace {

// EnvSet / EnvOn / EnvOff: backend settings, passed on as environment
// variables (target = the variable's name) rather than Config members, so that
// a backend reads them without fields the reference's Config lacks
enum class Kind { On, Off, Int, Float, Str, DebugLvl, Dim, FdOrder, EnvSet, EnvOn, EnvOff };
enum class Check { None, Pos, NonNeg };

struct Flag {
    const char *name;
    Kind kind;
    void *target;
    const char *label;
    Check check;
    const char *help;
};

std::vector<Flag> flagTable(Config &c)
{
    using K = Kind;
    using C = Check;
    return {
        {"-autoPad", K::On, &c.autoPad, " Automatic padding", C::None, "automatic grid padding on all axis"},
        {"-boundary", K::Str, &c.boundary, " Boundary\t", C::None, "boundary condition type FreeSurf (DEFAULT) | None"},
        {"-cb1", K::Int, &c.cb1, " cb1\t\t", C::Pos, "cache block size axis 1 [grid pts]"},
        {"-cb2", K::Int, &c.cb2, " cb2\t\t", C::Pos, "cache block size axis 2 [grid pts]"},
        {"-cb3", K::Int, &c.cb3, " cb3\t\t", C::Pos, "cache block size axis 3 [grid pts]"},
        {"-d1", K::Float, &c.d1, " d1 (m)\t\t", C::None, "grid spacing (m) axis 1"},
        {"-d2", K::Float, &c.d2, " d2 (m)\t\t", C::None, "grid spacing (m) axis 2"},
        {"-d3", K::Float, &c.d3, " d3 (m)\t\t", C::None, "grid spacing (m) axis 3"},
        {"-debug", K::DebugLvl, nullptr, " Debug info level", C::None, "debug trace none|light|mid|full"},
        {"-dim", K::Dim, &c.dim, " Space dimension", C::None, "space dimension [1,2 or 3]"},
        {"-dt", K::Float, &c.dt, " Time step (s)\t", C::None, "time step (s) for propagator"},
        {"-fdOrder", K::FdOrder, &c.fdOrder, " FD space Order\t", C::None, "spatial FD order [2,4,6,8,10,12,14 or 16]"},
        {"-freqMax", K::Float, &c.freqMax, " max. freq. (Hz)", C::None, "maximum frequency (Hz) for modeling"},
        {"-gpuBlkSize", K::Int, &c.gpuBlkSize, " gpuBlkSize\t", C::Pos, "GPU, no. of threads / 1D block"},
        {"-gpuBlkSize1", K::Int, &c.gpuBlkSize1, " gpuBlkSize1\t", C::Pos, "GPU, no. of threads / 3D block axis 1"},
        {"-gpuBlkSize2", K::Int, &c.gpuBlkSize2, " gpuBlkSize2\t", C::Pos, "GPU, no. of threads / 3D block axis 2"},
        {"-gpuBlkSize3", K::Int, &c.gpuBlkSize3, " gpuBlkSize3\t", C::Pos, "GPU, no. of threads / 3D block axis 3"},
        {"-gpuGridSize", K::Int, &c.gpuGridSize, " gpuGridSize\t", C::Pos, "GPU, no. of 1D blocks per grid"},
        {"-gpuMpiAware", K::On, &c.gpuMpiAware, " gpuMpiAware\t", C::None, "use MPI GPU-aware library"},
        {"-hwCounterDt", K::Float, &c.hwCounterDt, " Dt hw counter (s)", C::None, "time (s) between hardware counter update"},
        {"-modelVpFile", K::Str, &c.modelVpFile, " model Vp file\t", C::None, "path/name of Vp file (P-wave velocity m/s)"},
        {"-n1", K::Int, &c.n1, " n1\t\t", C::Pos, "inner domain size axis 1 [grid pts]"},
        {"-n2", K::Int, &c.n2, " n2\t\t", C::Pos, "inner domain size axis 2 [grid pts]"},
        {"-n3", K::Int, &c.n3, " n3\t\t", C::Pos, "inner domain size axis 3 [grid pts]"},
        {"-n1AddPad", K::Int, &c.n1AddPad, " n1AddPad\t", C::NonNeg, "add N points to grid at the end of axis 1"},
        {"-n2AddPad", K::Int, &c.n2AddPad, " n2AddPad\t", C::NonNeg, "add N points to grid at the end of axis 2"},
        {"-n3AddPad", K::Int, &c.n3AddPad, " n3AddPad\t", C::NonNeg, "add N points to grid at the end of axis 3"},
        {"-n1MulPad", K::Int, &c.n1MulPad, " n1MulPad\t", C::Pos, "pad axis 1 to a multiple of N"},
        {"-n2MulPad", K::Int, &c.n2MulPad, " n2MulPad\t", C::Pos, "pad axis 2 to a multiple of N"},
        {"-n3MulPad", K::Int, &c.n3MulPad, " n3MulPad\t", C::Pos, "pad axis 3 to a multiple of N"},
        {"-n1Offset", K::Int, &c.n1Offset, " n1Offset\t", C::NonNeg, "add N points to grid at the beginning of axis 1"},
        {"-n2Offset", K::Int, &c.n2Offset, " n2Offset\t", C::NonNeg, "add N points to grid at the beginning of axis 2"},
        {"-n3Offset", K::Int, &c.n3Offset, " n3Offset\t", C::NonNeg, "add N points to grid at the beginning of axis 3"},
        {"-nsub1", K::Int, &c.nsub1, " No. of nsub1\t", C::None, "no. of subdomains axis 1"},
        {"-nsub2", K::Int, &c.nsub2, " No. of nsub2\t", C::None, "no. of subdomains axis 2"},
        {"-nsub3", K::Int, &c.nsub3, " No. of nsub3\t", C::None, "no. of subdomains axis 3"},
        {"-nt", K::Int, &c.nt, " No. of time steps", C::Pos, "no. of time steps for propagator"},
        {"-ntry", K::Int, &c.ntry, " No. of tries\t", C::None, "no. of tries for each testCase"},
        {"-param1", K::Float, &c.param1, " Parameter 1\t", C::None, "parameter 1 used in testCases"},
        {"-param2", K::Float, &c.param2, " Parameter 2\t", C::None, "parameter 2 used in testCases"},
        {"-param3", K::Float, &c.param3, " Parameter 3\t", C::None, "parameter 3 used in testCases"},
        {"-param4", K::Float, &c.param4, " Parameter 4\t", C::None, "parameter 4 used in testCases"},
        {"-propagator", K::Str, &c.propagator, " Propagator\t", C::None, "Ac2Standard (DEFAULT) | Ac2SplitComp"},
        {"-ratioCFL", K::Float, &c.ratioCFL, " Ratio CFL\t", C::None, "ratio of stability dt for propagator"},
        {"-snapInc", K::Int, &c.snapInc, " Snapshots inc. (steps)", C::Pos, "snaphots increment (no. of time steps)"},
        {"-snapDt", K::Float, &c.snapDt, " Snapshots inc. (sec.)", C::Pos, "snaphots increment (time in sec.), overrides snapInc"},
        {"-testCase", K::Str, &c.testCaseName, " TestCase name\t", C::None, "run specific testCase by name: FD_D1 | FD_D2 | Grid | Modeling | Propa | All"},
        {"-testMode", K::Str, &c.testMode, " Test mode\t", C::None, "test mode (see enabled modes in the header)"},
        {"-tmax", K::Float, &c.tmax, " Tmax\t\t", C::Pos, "max. time (s) for propagator, overrides nt"},
        {"-writeGrid", K::On, &c.writeGrid, " Write grids on disk", C::None, "write grids on disk"},
        // MI355X backend settings (environment variables HPCSCAN_MI355X_*)
        {"-gpuDevice", K::EnvSet, (void *)"HPCSCAN_MI355X_DEVICE", " GPU device\t", C::NonNeg, "MI355X: HIP device ordinal"},
        {"-hipKernel", K::EnvSet, (void *)"HPCSCAN_MI355X_KERNEL", " HIP kernel\t", C::None, "MI355X: stencil kernels auto (DEFAULT) | ref | tiled"},
        {"-hipFoldBC", K::EnvOn, (void *)"HPCSCAN_MI355X_FOLD_BC", " HIP fold BC\t", C::None, "MI355X: fold the free surface into the fused step"},
        {"-hipNoFoldBC", K::EnvOff, (void *)"HPCSCAN_MI355X_FOLD_BC", " HIP fold BC\t", C::None, "MI355X: apply the free surface as a separate pass (DEFAULT)"},
        {"-hipAutotune", K::EnvOn, (void *)"HPCSCAN_MI355X_AUTOTUNE", " HIP autotune\t", C::None, "MI355X: time the tile shapes once per grid layout, keep the fastest"},
        {"-hipContract", K::EnvOn, (void *)"HPCSCAN_MI355X_CONTRACT", " HIP contract\t", C::None, "MI355X: fused multiply-adds in the fused step and Laplacian (not bit-exact)"},
    };
}

void printHelp(const std::vector<Flag> &table)
{
    printInfo(MASTER, " List of command line parameters:");
    for (const Flag &f : table) {
        string arg;
        switch (f.kind) {
        case Kind::Int: arg = " <int>"; break;
        case Kind::Float: arg = " <float>"; break;
        case Kind::Str: case Kind::EnvSet: arg = " <string>"; break;
        case Kind::DebugLvl: arg = " <OPT>"; break;
        case Kind::Dim: case Kind::FdOrder: arg = " <int>"; break;
        default: break;
        }
        string left = string(" ") + f.name + arg;
        if (left.size() < 23) left.resize(23, ' ');
        printInfo(MASTER, (left + "= " + f.help).c_str());
    }
    printInfo(MASTER, " -help or -h           = list of command line parameters");
    printInfo(MASTER, " -version or -v        = print version information");
}

} // namespace

Rtn_code Config::parse_argument(int argc, char *argv[])
{
    printDebug(MID_DEBUG, "IN Config::parse_argument");
    std::vector<Flag> table = flagTable(*this);

    for (int ii = 0; ii < argc; ii++) {
        string a = argv[ii];
        if (a == "-version" || a == "-v") {
            print_header_of_output_report();
            return RTN_CODE_EXIT;
        }
        if (argc == 1 || a == "-help" || a == "-h") {
            printHelp(table);
            return RTN_CODE_EXIT;
        }
    }

    print_line2();
    printInfo(MASTER, " COMMAND LINE ARGUMENTS\n");
    printInfo(MASTER, " Executable\t ", argv[0]);

    for (int ii = 1; ii < argc; ii++) {
        string a = argv[ii];
        const Flag *f = nullptr;
        for (const Flag &x : table)
            if (a == x.name) { f = &x; break; }
        if (!f) {
            printError(" Unknown argument ", argv[ii]);
            return RTN_CODE_KO;
        }
        if (f->kind == Kind::On || f->kind == Kind::Off) {
            *(bool *)f->target = (f->kind == Kind::On);
            printInfo(MASTER, f->label, f->kind == Kind::On ? "YES" : "NO");
            continue;
        }
        if (f->kind == Kind::EnvOn || f->kind == Kind::EnvOff) {
            setenv((const char *)f->target, f->kind == Kind::EnvOn ? "1" : "0", 1);
            printInfo(MASTER, f->label, f->kind == Kind::EnvOn ? "YES" : "NO");
            continue;
        }
        if (++ii >= argc) {
            printError((string(" parameter is needed after ") + f->name).c_str());
            return RTN_CODE_KO;
        }
        const char *v = argv[ii];
        switch (f->kind) {
        case Kind::Int: {
            Myint x = atoi(v);
            *(Myint *)f->target = x;
            printInfo(MASTER, f->label, x);
            if ((f->check == Check::Pos && x <= 0) || (f->check == Check::NonNeg && x < 0)) {
                string m = string(" ") + (f->name + 1) + (f->check == Check::Pos ? " should be > 0" : " should be >= 0");
                printError(m.c_str());
                return RTN_CODE_KO;
            }
            break;
        }
        case Kind::Float: {
            Myfloat64 x = atof(v);
            *(Myfloat64 *)f->target = x;
            printInfo(MASTER, f->label, x);
            if (f->check == Check::Pos && x <= 0.0) {
                printError((string(" ") + (f->name + 1) + " should be > 0").c_str());
                return RTN_CODE_KO;
            }
            break;
        }
        case Kind::Str:
            *(string *)f->target = v;
            printInfo(MASTER, f->label, string(v));
            break;
        case Kind::EnvSet:
            if (f->check == Check::NonNeg && atoi(v) < 0) {
                printError((string(" ") + (f->name + 1) + " should be >= 0").c_str());
                return RTN_CODE_KO;
            }
            setenv((const char *)f->target, v, 1);
            printInfo(MASTER, f->label, string(v));
            break;
        case Kind::DebugLvl: {
            string s = v;
            if (s == "none") { debug = NO_DEBUG; printInfo(MASTER, f->label, "NONE"); }
            else if (s == "light") { debug = LIGHT_DEBUG; printInfo(MASTER, f->label, "LIGHT"); }
            else if (s == "mid") { debug = MID_DEBUG; printInfo(MASTER, f->label, "MEDIUM"); }
            else if (s == "full") { debug = FULL_DEBUG; printInfo(MASTER, f->label, "FULL"); }
            else { printError(" Invalid debug level", v); return RTN_CODE_KO; }
            break;
        }
        case Kind::Dim: {
            dim = (Dim_type)atoi(v);
            printInfo(MASTER, f->label, (Myint)dim);
            if (dim != DIM1 && dim != DIM2 && dim != DIM3) {
                printError(" Space dimension should be 1, 2 or 3");
                return RTN_CODE_KO;
            }
            break;
        }
        case Kind::FdOrder: {
            fdOrder = atoi(v);
            printInfo(MASTER, f->label, fdOrder);
            if (!validFdOrder(fdOrder)) {
                printError(" FD space order should be 2,4,8,10,12,14 or 16");
                return RTN_CODE_KO;
            }
            break;
        }
        default:
            break;
        }
    }
    printDebug(MID_DEBUG, "OUT Config::parse_argument");
    return RTN_CODE_OK;
}

// src/config.cpp:1069-1160
Rtn_code Config::info(void)
{
    print_blank();
    print_line5();
    printInfo(MASTER, " Configuration parameters");
    printInfo(MASTER, " TestCase name\t", testCaseName);
    if (testCaseName == "UNSPECIFIED") {
        printWarning(" Please specify test case name with -testCase");
        return RTN_CODE_KO;
    }
    printInfo(MASTER, " No. of MPI process", nMpiProc);
    printInfo(MASTER, " No. subdomains axis1", nsub1);
    printInfo(MASTER, " No. subdomains axis2", nsub2);
    printInfo(MASTER, " No. subdomains axis3", nsub3);
    printInfo(MASTER, " OpenMP threads / MPI", omp_get_max_threads());
    printInfo(MASTER, " Number of tries", ntry);
    printInfo(MASTER, " FD space order\t", fdOrder);
    print_line5();

    if (nMpiProc != nsub1 * nsub2 * nsub3) {
        printError("Number subdomains inconsistent with nproc");
        return RTN_CODE_KO;
    }
    const Myint MIN_NPOINT_NSUB = 20;  // src/config.cpp:81
    if (dim == DIM1 && (nsub2 > 1 || nsub3 > 1)) {
        printError(nsub2 > 1 ? "nsub2 > 1 not allowed for 1D" : "nsub3 > 1 not allowed for 1D");
        return RTN_CODE_KO;
    }
    if (dim == DIM2 && nsub3 > 1) {
        printError("nsub3 > 1 not allowed for 2D");
        return RTN_CODE_KO;
    }
    const Myint nn[3] = {n1, n2, n3}, ns[3] = {nsub1, nsub2, nsub3};
    for (int a = 0; a < (int)dim; a++) {
        if (nn[a] / ns[a] < MIN_NPOINT_NSUB) {
            string m = "(n" + to_string(a + 1) + " / nsub" + to_string(a + 1) + ") < MIN_NPOINT_NSUB, enlarge n" + to_string(a + 1);
            printError(m.c_str());
            printInfo(MASTER, "MIN_NPOINT_NSUB", MIN_NPOINT_NSUB);
            return RTN_CODE_KO;
        }
    }
    return RTN_CODE_OK;
}

Rtn_code Config::initialize(void)
{
    if (debug > NO_DEBUG) {
        string file_name = "hpcscan.debug.proc" + to_string(myMpiRank) + ".log";
        debugLogFile.open(file_name);
    }
    return RTN_CODE_OK;
}

Rtn_code Config::finalize(Rtn_code rtnCode)
{
    print_end_of_output_report(rtnCode);
    if (debugLogFile.is_open()) debugLogFile.close();
    Grid_Factory::shutdown();
    return RTN_CODE_OK;
}

} // namespace hpcscan
