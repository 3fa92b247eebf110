// Grid base class: backend-independent layout, bookkeeping and file output.
//
// Layout per axis (src/grid.cpp:17-38, 421-468):
//   Offset | Halo1 | Layer1 | Inner | Layer2 | Halo2 | Pad
// flat index i1 + i2*n1 + i3*n1*n2, n1 contiguous.
#include "grid.h"

#include <algorithm>
#include <cassert>
#include <cfloat>
#include <cmath>
#include <fstream>

#include "config.h"
#include "constant.h"
#include "fdm.h"
#include "global.h"
#include "output_report.h"

namespace hpcscan {

void Grid::resetIndexes(void)
{
    i1OffsetStart = i1OffsetEnd = i1Halo1Start = i1Halo1End = i1Layer1Start = i1Layer1End = 0;
    i1InnerStart = i1InnerEnd = i1Layer2Start = i1Layer2End = i1Halo2Start = i1Halo2End = 0;
    i1PadStart = i1PadEnd = 0;
    i2OffsetStart = i2OffsetEnd = i2Halo1Start = i2Halo1End = i2Layer1Start = i2Layer1End = 0;
    i2InnerStart = i2InnerEnd = i2Layer2Start = i2Layer2End = i2Halo2Start = i2Halo2End = 0;
    i2PadStart = i2PadEnd = 0;
    i3OffsetStart = i3OffsetEnd = i3Halo1Start = i3Halo1End = i3Layer1Start = i3Layer1End = 0;
    i3InnerStart = i3InnerEnd = i3Layer2Start = i3Layer2End = i3Halo2Start = i3Halo2End = 0;
    i3PadStart = i3PadEnd = 0;
    i1ProcIdStart = i1ProcIdEnd = i2ProcIdStart = i2ProcIdEnd = i3ProcIdStart = i3ProcIdEnd = PROC_NULL;
    d1 = d2 = d3 = 0;
    i1OffsetGlobInner = i2OffsetGlobInner = i3OffsetGlobInner = 0;
    Orig1 = Orig2 = Orig3 = 0;
    n1 = n2 = n3 = 0;
    npoint = 0;
    haloWidth = 0;
    grid_3d = nullptr;
    grid_writeGlobal_time = grid_write_time = i_o_pread_time = 0.0;
}

// src/grid.cpp:62-143
Grid::Grid(Grid_type gridTypeIn)
{
    resetIndexes();
    gridMode = GRID_MODE_BASELINE;
    n1Inner = Config::Instance()->n1;
    n2Inner = Config::Instance()->n2;
    n3Inner = Config::Instance()->n3;
    gridType = gridTypeIn;
    dim = Config::Instance()->dim;
}

// src/grid.cpp:147-249
Grid::Grid(Grid_type gridTypeIn, Dim_type dimTypeIn, Myint64 n1InnerIn, Myint64 n2InnerIn,
           Myint64 n3InnerIn)
{
    resetIndexes();
    gridMode = GRID_MODE_BASELINE;
    n1Inner = n1InnerIn;
    n2Inner = dimTypeIn >= DIM2 ? n2InnerIn : 1;
    n3Inner = dimTypeIn >= DIM3 ? n3InnerIn : 1;
    gridType = gridTypeIn;
    dim = dimTypeIn;
}

Grid::~Grid(void)
{
    delete[] grid_3d;
}

// One axis of the layout (src/grid.cpp:421-468); 'offset' and 'pad' hooks are virtual.
#define HPCSCAN_AXIS_LAYOUT(I, NLOC)                                                 \
    I##Halo1Start = I##OffsetEnd + 1;                                                \
    I##Halo1End = I##Halo1Start + haloWidth - 1;                                     \
    I##Layer1Start = I##Halo1End + 1;                                                \
    I##Layer1End = nlayer == 0 ? I##Layer1Start : I##Layer1Start + nlayer - 1;       \
    I##InnerStart = nlayer == 0 ? I##Layer1Start : I##Layer1End + 1;                 \
    I##InnerEnd = I##InnerStart + (NLOC) - 1;                                        \
    I##Layer2Start = nlayer == 0 ? I##InnerEnd : I##InnerEnd + 1;                    \
    I##Layer2End = nlayer == 0 ? I##InnerEnd : I##Layer2Start + nlayer - 1;          \
    I##Halo2Start = I##Layer2End + 1;                                                \
    I##Halo2End = I##Halo2Start + haloWidth - 1;

Rtn_code Grid::initializeGrid(void)
{
    printDebug(MID_DEBUG, "IN Grid::initializeGrid");
    if (grid_3d != nullptr) {
        printError(" In Grid::initializeGrid, grid_3d != NULL");
        return RTN_CODE_OK;
    }
    Config *cfg = Config::Instance();
    const Myint nlayer = cfg->nlayer;
    haloWidth = getFD_haloWidth(cfg->fdOrder);
    if (haloWidth == UNSPECIFIED) {
        printError(" In Grid::initializeGrid, haloWidth = UNSPECIFIED");
        return RTN_CODE_OK;
    }

    // sub-domain of this rank (src/grid.cpp:289-404); at one rank it is the whole grid
    Myint nLoc[3] = {(Myint)n1Inner, (Myint)n2Inner, (Myint)n3Inner};
    Myint subIdx[3] = {0, 0, 0};
    const Myint nsub[3] = {cfg->nsub1, cfg->nsub2, cfg->nsub3};
    if (gridType == GRID_LOCAL) {
        subIdx[2] = myMpiRank / (nsub[0] * nsub[1]);
        subIdx[1] = (myMpiRank - subIdx[2] * nsub[0] * nsub[1]) / nsub[0];
        subIdx[0] = myMpiRank - subIdx[2] * nsub[0] * nsub[1] - subIdx[1] * nsub[0];
        for (int a = 0; a < 3; a++) {
            const Myint q = nLoc[a] / nsub[a];
            nLoc[a] = subIdx[a] < nsub[a] - 1 ? q : nLoc[a] - (nsub[a] - 1) * q;
        }
        const Myint stride[3] = {1, nsub[0], nsub[0] * nsub[1]};
        Myint *start[3] = {&i1ProcIdStart, &i2ProcIdStart, &i3ProcIdStart};
        Myint *end[3] = {&i1ProcIdEnd, &i2ProcIdEnd, &i3ProcIdEnd};
        for (int a = 0; a < 3; a++) {
            *start[a] = (nsub[a] == 1 || subIdx[a] == 0) ? PROC_NULL : myMpiRank - stride[a];
            *end[a] = (nsub[a] == 1 || subIdx[a] == nsub[a] - 1) ? PROC_NULL : myMpiRank + stride[a];
        }
    }

    offsetGridn1();
    HPCSCAN_AXIS_LAYOUT(i1, nLoc[0])
    padGridn1();
    n1 = i1PadEnd + 1;
    d1 = cfg->d1;
    i1OffsetGlobInner = subIdx[0] * (cfg->n1 / nsub[0]);
    Orig1 = (i1OffsetGlobInner - i1InnerStart) * d1;

    if (dim >= DIM2) {
        offsetGridn2();
        HPCSCAN_AXIS_LAYOUT(i2, nLoc[1])
        padGridn2();
        n2 = i2PadEnd + 1;
        d2 = cfg->d2;
        i2OffsetGlobInner = subIdx[1] * (cfg->n2 / nsub[1]);
        Orig2 = (i2OffsetGlobInner - i2InnerStart) * d2;
    } else {
        n2 = 1;
    }

    if (dim >= DIM3) {
        offsetGridn3();
        HPCSCAN_AXIS_LAYOUT(i3, nLoc[2])
        padGridn3();
        n3 = i3PadEnd + 1;
        d3 = cfg->d3;
        i3OffsetGlobInner = subIdx[2] * (cfg->n3 / nsub[2]);
        Orig3 = (i3OffsetGlobInner - i3InnerStart) * d3;
    } else {
        n3 = 1;
    }

    npoint = (Myint64)n1 * n2 * n3;
    // host copy; pages are only committed when a backend actually uses them
    grid_3d = new Myfloat[npoint];
    printDebug(MID_DEBUG, "OUT Grid::initializeGrid");
    return RTN_CODE_OK;
}
#undef HPCSCAN_AXIS_LAYOUT

// padding rules (src/grid.cpp:839-947)
static void padAxis(Myint halo2End, bool autoPadThisAxis, bool autoPad, Myint addPad, Myint mulPad,
                    Myint &padStart, Myint &padEnd)
{
    if (autoPad) {
        padStart = padEnd = halo2End;
#ifndef HPCSCAN_STORAGE_FP64
        if (autoPadThisAxis && ((halo2End + 1) % 2)) padStart = padEnd = halo2End + 1;
#else
        (void)autoPadThisAxis;
#endif
    } else if (addPad != UNSPECIFIED) {
        padStart = halo2End + 1;
        padEnd = padStart + addPad - 1;
    } else if (mulPad != UNSPECIFIED) {
        padStart = halo2End + 1;
        const Myint nTmp = (padStart + 1) / mulPad;
        const Myint nTot = (mulPad * nTmp < padStart) ? (nTmp + 1) * mulPad : nTmp * mulPad;
        padEnd = nTot - 1;
    } else {
        padStart = padEnd = halo2End;
    }
}

void Grid::padGridn1(void)
{
    Config *c = Config::Instance();
    padAxis(i1Halo2End, true, c->autoPad, c->n1AddPad, c->n1MulPad, i1PadStart, i1PadEnd);
}
void Grid::padGridn2(void)
{
    Config *c = Config::Instance();
    padAxis(i2Halo2End, false, c->autoPad, c->n2AddPad, c->n2MulPad, i2PadStart, i2PadEnd);
}
void Grid::padGridn3(void)
{
    Config *c = Config::Instance();
    padAxis(i3Halo2End, false, c->autoPad, c->n3AddPad, c->n3MulPad, i3PadStart, i3PadEnd);
}

// offset rules (src/grid.cpp:951-1009): one point on n1 when the halo is odd in FP32
void Grid::offsetGridn1(void)
{
    i1OffsetStart = 0;
    Myint offset = Config::Instance()->n1Offset;
    if (offset == UNSPECIFIED) {
#ifndef HPCSCAN_STORAGE_FP64
        offset = (haloWidth % 2 == 1) ? 1 : 0;
#else
        offset = 0;
#endif
    }
    i1OffsetEnd = i1OffsetStart + offset - 1;
}
void Grid::offsetGridn2(void)
{
    i2OffsetStart = 0;
    Myint offset = Config::Instance()->n2Offset;
    i2OffsetEnd = i2OffsetStart + (offset == UNSPECIFIED ? 0 : offset) - 1;
}
void Grid::offsetGridn3(void)
{
    i3OffsetStart = 0;
    Myint offset = Config::Instance()->n3Offset;
    i3OffsetEnd = i3OffsetStart + (offset == UNSPECIFIED ? 0 : offset) - 1;
}

// src/grid.cpp:1078-1241
void Grid::info(void)
{
    print_blank();
    printInfo(MASTER, " * Grid parameters *");
    printInfo(MASTER, " Grid type\t", gridType == GRID_LOCAL ? "LOCAL" : "GLOBAL");
    printInfo(MASTER, " Grid mode\t", gridMode);
    if (dim == DIM1) {
        printInfo(MASTER, " Grid dim.\t", "1D");
        printInfo(MASTER, " Inner n1\t", i1InnerEnd - i1InnerStart + 1);
        printInfo(MASTER, " Total n1\t", n1);
    } else if (dim == DIM2) {
        printInfo(MASTER, " Grid dim.\t", "2D");
        printInfo(MASTER, " Total n1\t", n1);
        printInfo(MASTER, " Total n2\t", n2);
        printInfo(MASTER, " Inner n1\t", i1InnerEnd - i1InnerStart + 1);
        printInfo(MASTER, " Inner n2\t", i2InnerEnd - i2InnerStart + 1);
    } else {
        printInfo(MASTER, " Grid dim.\t", "3D");
        printInfo(MASTER, " Total n1\t", n1);
        printInfo(MASTER, " Total n2\t", n2);
        printInfo(MASTER, " Total n3\t", n3);
        printInfo(MASTER, " Inner n1\t", i1InnerEnd - i1InnerStart + 1);
        printInfo(MASTER, " Inner n2\t", i2InnerEnd - i2InnerStart + 1);
        printInfo(MASTER, " Inner n3\t", i3InnerEnd - i3InnerStart + 1);
    }
    if (i1Halo1Start - i1OffsetStart > 0) printInfo(MASTER, " n1 offset\t", i1Halo1Start - i1OffsetStart);
    if (i2Halo1Start - i2OffsetStart > 0) printInfo(MASTER, " n2 offset\t", i2Halo1Start - i2OffsetStart);
    if (i3Halo1Start - i3OffsetStart > 0) printInfo(MASTER, " n3 offset\t", i3Halo1Start - i3OffsetStart);
    if (i1PadEnd - i1Halo2End > 0) printInfo(MASTER, " n1 padding\t", i1PadEnd - i1Halo2End);
    if (i2PadEnd - i2Halo2End > 0) printInfo(MASTER, " n2 padding\t", i2PadEnd - i2Halo2End);
    if (i3PadEnd - i3Halo2End > 0) printInfo(MASTER, " n3 padding\t", i3PadEnd - i3Halo2End);

    const Myint64 gridSize = npoint * sizeof(Myfloat);
    if (gridSize < 1e3) printInfo(MASTER, " Grid size (Byte)", gridSize);
    else if (gridSize < 1e6) printInfo(MASTER, " Grid size (KB)\t", Myfloat(gridSize / 1e3));
    else if (gridSize < 1e9) printInfo(MASTER, " Grid size (MB)\t", Myfloat(gridSize / 1e6));
    else printInfo(MASTER, " Grid size (GB)\t", Myfloat(gridSize / 1e9));

    // alignment of the first inner point, in elements from the start of the array
    const Myint64 first = i1InnerStart + (Myint64)i2InnerStart * n1 + (Myint64)i3InnerStart * n1 * n2;
    printInfo(MASTER, " 64 bits align (inner)", (first * sizeof(Myfloat)) % 8 == 0 ? "YES" : "NO");
}

// src/grid.cpp:1245-1287
Rtn_code Grid::write(string fileName)
{
    if (!Config::Instance()->writeGrid) {
        grid_write_time = 0.0;
        return RTN_CODE_OK;
    }
    const double t0 = hpcWtime();
    string fbin = fileName + ".proc" + to_string(myMpiRank) + ".grid.bin";
    printInfo(MASTER, "* write on disk\t", fbin.c_str());
    ofstream out(fbin.c_str(), ios::binary | ios::app);
    if (!out.is_open()) {
        printError(" In Grid::write, cannot open", fbin);
        return RTN_CODE_KO;
    }
    out.write((const char *)grid_3d, npoint * sizeof(Myfloat));
    out.close();
    ofstream info(fileName + ".proc" + to_string(myMpiRank) + ".grid.info");
    info << n1 << "\n" << n2 << "\n" << n3 << "\n";
    const double t1 = hpcWtime();
    grid_write_time = t1 - t0;
    printInfo(MASTER, "* time to write (s)", t1 - t0);
    return RTN_CODE_OK;
}

// src/grid.cpp:1291-1351
Rtn_code Grid::write(Point_type pointType, string fileName)
{
    if (!Config::Instance()->writeGrid) return RTN_CODE_OK;
    if (pointType == ALL_POINTS) return this->write(fileName);
    const double t0 = hpcWtime();
    string fbin = fileName + ".proc" + to_string(myMpiRank) + ".grid.bin";
    printInfo(MASTER, "* write on disk\t", fbin.c_str());
    ofstream out(fbin.c_str(), ios::binary | ios::app);
    if (!out.is_open()) {
        printError(" In Grid::write, cannot open", fbin);
        return RTN_CODE_KO;
    }
    Myint64 i1s, i1e, i2s, i2e, i3s, i3e;
    getGridIndex(pointType, &i1s, &i1e, &i2s, &i2e, &i3s, &i3e);
    for (Myint64 i3 = i3s; i3 <= i3e; i3++)
        for (Myint64 i2 = i2s; i2 <= i2e; i2++)
            out.write((const char *)&grid_3d[i1s + i2 * n1 + i3 * n2 * n1], (i1e - i1s + 1) * sizeof(Myfloat));
    out.close();
    ofstream info(fileName + ".proc" + to_string(myMpiRank) + ".grid.info");
    info << (i1e - i1s + 1) << "\n" << (i2e - i2s + 1) << "\n" << (i3e - i3s + 1) << "\n";
    printInfo(MASTER, "* time to write (s)", hpcWtime() - t0);
    return RTN_CODE_OK;
}

// snapshot num_iter of the inner points appended to <file>.global.grid.bin, one
// rank (src/grid.cpp:1358-1441: the MPI-IO view of the global inner grid)
Rtn_code Grid::writeGlobal(Point_type pointType, string fileName, Myint num_iter)
{
    if (!Config::Instance()->writeGrid) {
        grid_writeGlobal_time = 0.0;
        return RTN_CODE_OK;
    }
    const double t0 = hpcWtime();
    const string fbin = fileName + ".global.grid.bin";
    printInfo(MASTER, "* write on disk\t", fbin.c_str());
    Myint64 i1s, i1e, i2s, i2e, i3s, i3e;
    getGridIndex(pointType, &i1s, &i1e, &i2s, &i2e, &i3s, &i3e);
    const Myint64 nRow = i1e - i1s + 1;
    const Myint64 snap = nRow * (i2e - i2s + 1) * (i3e - i3s + 1);
    fstream out(fbin.c_str(), ios::binary | ios::in | ios::out);
    if (!out.is_open()) out.open(fbin.c_str(), ios::binary | ios::out);
    if (!out.is_open()) {
        printError(" In Grid::writeGlobal, cannot open", fbin);
        return RTN_CODE_KO;
    }
    out.seekp((std::streamoff)(sizeof(Myfloat) * snap * num_iter), ios::beg);
    for (Myint64 i3 = i3s; i3 <= i3e; i3++)
        for (Myint64 i2 = i2s; i2 <= i2e; i2++)
            out.write((const char *)&grid_3d[i1s + i2 * n1 + i3 * n2 * n1], nRow * sizeof(Myfloat));
    out.close();
    ofstream info(fileName + ".global.grid.info");
    info << Config::Instance()->n1 << "\n" << Config::Instance()->n2 << "\n" << Config::Instance()->n3 << "\n";
    grid_writeGlobal_time = hpcWtime() - t0;
    printInfo(MASTER, "* time to write (s)", grid_writeGlobal_time);
    return RTN_CODE_OK;
}

// the whole array, as Grid::write stores it (src/grid.cpp:1560-1583)
Rtn_code Grid::read(string fileName)
{
    const double t0 = hpcWtime();
    printInfo(MASTER, "* read on disk\t", fileName.c_str());
    ifstream in(fileName.c_str(), ios::binary);
    if (!in.is_open()) {
        printError(" In Grid::read, cannot open", fileName);
        return RTN_CODE_KO;
    }
    in.read((char *)grid_3d, npoint * sizeof(Myfloat));
    if (!in) {
        printError(" In Grid::read, file too short", fileName);
        return RTN_CODE_KO;
    }
    const double t = hpcWtime() - t0;
    i_o_pread_time = t;
    printInfo(MASTER, "* read time (s)\t", t);
    printInfo(MASTER, "* read bwth (GB/s)", npoint * sizeof(Myfloat) / t / 1e9);
    return RTN_CODE_OK;
}

// the points of pointType from a file holding the global inner grid (n1 x n2 x n3
// of the command line, n1 fastest), row by row (src/grid.cpp:1587-1643)
Rtn_code Grid::read(Point_type pointType, string fileName)
{
    if (pointType == ALL_POINTS) return this->read(fileName);
    const double t0 = hpcWtime();
    printInfo(MASTER, "* read on disk\t", fileName.c_str());
    ifstream in(fileName.c_str(), ios::binary);
    if (!in.is_open()) {
        printError(" In Grid::read, cannot open", fileName);
        return RTN_CODE_KO;
    }
    Myint64 i1s, i1e, i2s, i2e, i3s, i3e;
    getGridIndex(pointType, &i1s, &i1e, &i2s, &i2e, &i3s, &i3e);
    const Myint64 nRead = i1e - i1s + 1;
    const Myint64 g1 = Config::Instance()->n1, g2 = Config::Instance()->n2;
    for (Myint64 i3 = i3s; i3 <= i3e; i3++) {
        for (Myint64 i2 = i2s; i2 <= i2e; i2++) {
            const Myint64 off = (Myint64)sizeof(Myfloat) * (i1OffsetGlobInner + g1 * (i2OffsetGlobInner + i2 - i2s) +
                                                             g1 * g2 * (i3OffsetGlobInner + i3 - i3s));
            in.seekg((std::streamoff)off, ios::beg);
            in.read((char *)&grid_3d[i1s + i2 * n1 + i3 * n2 * n1], nRead * sizeof(Myfloat));
            if (!in) {
                printError(" In Grid::read, file too short", fileName);
                return RTN_CODE_KO;
            }
        }
    }
    const double t = hpcWtime() - t0;
    i_o_pread_time = t;
    printInfo(MASTER, "* read time (s)\t", t);
    printInfo(MASTER, "* read bwth (GB/s)", npoint * sizeof(Myfloat) / t / 1e9);
    return RTN_CODE_OK;
}

void Grid::notProvided(const char *method) const
{
    string m = string(" Grid mode ") + gridMode + " does not provide";
    printError(m.c_str(), method);
}

Myfloat Grid::L1Err(Point_type, const Grid &) const { notProvided("L1Err"); return -1; }
Myfloat Grid::allProcL1Err(Point_type p, const Grid &g) const { return L1Err(p, g); }
Myfloat Grid::maxErr(Point_type, const Grid &) const { notProvided("maxErr"); return -1; }
Myfloat Grid::getMin(Point_type) { notProvided("getMin"); return NAN; }
Myfloat Grid::getMax(Point_type) { notProvided("getMax"); return NAN; }
Myfloat Grid::allProcGetMin(Point_type p) { return getMin(p); }
Myfloat Grid::allProcGetMax(Point_type p) { return getMax(p); }
Myfloat Grid::getSumAbs(Point_type) const { notProvided("getSumAbs"); return -1; }
Myfloat Grid::getSumAbsDiff(Point_type, const Grid &) const { notProvided("getSumAbsDiff"); return -1; }
void Grid::fill(Point_type, Myfloat) { notProvided("fill"); }
void Grid::fill(Point_type, Func_type, Func_type, Func_type, Myfloat64, Myfloat64, Myfloat64, Myfloat64)
{
    notProvided("fill");
}
void Grid::fillArray(Myfloat) { notProvided("fillArray"); }
void Grid::copyArray(const Grid &) { notProvided("copyArray"); }
void Grid::addArray(const Grid &, const Grid &) { notProvided("addArray"); }
void Grid::multiplyArray(const Grid &, const Grid &) { notProvided("multiplyArray"); }
void Grid::addUpdateArray(const Grid &) { notProvided("addUpdateArray"); }
Rtn_code Grid::applyBoundaryCondition(BoundCond_type) { notProvided("applyBoundaryCondition"); return RTN_CODE_KO; }
Rtn_code Grid::FD_D1_N1(Point_type, const Grid &, Myint) { notProvided("FD_D1_N1"); return RTN_CODE_KO; }
Rtn_code Grid::FD_D1_N2(Point_type, const Grid &, Myint) { notProvided("FD_D1_N2"); return RTN_CODE_KO; }
Rtn_code Grid::FD_D1_N3(Point_type, const Grid &, Myint) { notProvided("FD_D1_N3"); return RTN_CODE_KO; }
Rtn_code Grid::FD_D2_N1(Point_type, const Grid &, Myint) { notProvided("FD_D2_N1"); return RTN_CODE_KO; }
Rtn_code Grid::FD_D2_N2(Point_type, const Grid &, Myint) { notProvided("FD_D2_N2"); return RTN_CODE_KO; }
Rtn_code Grid::FD_D2_N3(Point_type, const Grid &, Myint) { notProvided("FD_D2_N3"); return RTN_CODE_KO; }
Rtn_code Grid::FD_LAPLACIAN(Point_type, const Grid &, Myint) { notProvided("FD_LAPLACIAN"); return RTN_CODE_KO; }
Rtn_code Grid::updatePressure(Point_type, const Grid &, const Grid &, const Grid &)
{
    notProvided("updatePressure");
    return RTN_CODE_KO;
}
Rtn_code Grid::computePressureWithFD(Grid &, Grid &, Myint)
{
    notProvided("computePressureWithFD");
    return RTN_CODE_KO;
}

// src/grid.cpp:1924-2000
void Grid::defineUnitGrid(void)
{
    const Myfloat64 d1Old = d1, d2Old = d2, d3Old = d3;
    const Myfloat64 extent = 1.0;
    d1 = extent / (Config::Instance()->n1 - 1);
    if (dim >= DIM2) d2 = extent / (Config::Instance()->n2 - 1);
    if (dim >= DIM3) d3 = extent / (Config::Instance()->n3 - 1);
    // new origins are rounded to Myfloat on purpose (resetOriginGrid takes Myfloat)
    Myfloat o1 = 0.0, o2 = 0.0, o3 = 0.0;
    o1 = Orig1 * d1 / d1Old;
    if (dim >= DIM2) o2 = Orig2 * d2 / d2Old;
    if (dim >= DIM3) o3 = Orig3 * d3 / d3Old;
    resetOriginGrid(o1, o2, o3);
}

void Grid::resetOriginGrid(Myfloat o1, Myfloat o2, Myfloat o3)
{
    Orig1 = o1;
    if (dim >= DIM2) Orig2 = o2;
    if (dim >= DIM3) Orig3 = o3;
}

void Grid::shiftOriginGrid(Myfloat s1, Myfloat s2, Myfloat s3)
{
    Orig1 = Orig1 + s1;
    Orig2 = Orig2 + s2;
    Orig3 = Orig3 + s3;
}

Myfloat Grid::getMinCoord(Axis_type a)
{
    if (a == AXIS1) return Orig1 + d1 * i1InnerStart;
    if (a == AXIS2) return Orig2 + d2 * i2InnerStart;
    return Orig3 + d3 * i3InnerStart;
}

Myfloat Grid::getMaxCoord(Axis_type a)
{
    if (a == AXIS1) return Orig1 + d1 * i1InnerEnd;
    if (a == AXIS2) return Orig2 + d2 * i2InnerEnd;
    return Orig3 + d3 * i3InnerEnd;
}

Myfloat64 Grid::getSpaceSampling(Axis_type a)
{
    return a == AXIS1 ? d1 : a == AXIS2 ? d2 : d3;
}

// src/grid.cpp:2533-2546: the minimum is returned as Myfloat
Myfloat Grid::getMinSpaceSampling(void)
{
    Myfloat m = d1;
    if (dim >= DIM2) m = min((double)m, d2);
    if (dim >= DIM3) m = min((double)m, d3);
    return m;
}

Myint Grid::getNeighbourProc(Point_type p)
{
    switch (p) {
    case I1HALO1: return i1ProcIdStart;
    case I1HALO2: return i1ProcIdEnd;
    case I2HALO1: return i2ProcIdStart;
    case I2HALO2: return i2ProcIdEnd;
    case I3HALO1: return i3ProcIdStart;
    case I3HALO2: return i3ProcIdEnd;
    default: printError("IN Grid::getNeighbourProc, invalid pointType"); return PROC_NULL;
    }
}

Myint Grid::getNumberOfNeighbour(void)
{
    Myint v = 0;
    for (Myint p : {i1ProcIdStart, i1ProcIdEnd, i2ProcIdStart, i2ProcIdEnd, i3ProcIdStart, i3ProcIdEnd})
        if (p != PROC_NULL) v++;
    return v;
}

// src/grid.cpp:2680-2843
Rtn_code Grid::getGridIndex(Point_type p, Myint64 *s1, Myint64 *e1, Myint64 *s2, Myint64 *e2,
                            Myint64 *s3, Myint64 *e3) const
{
    *s1 = i1InnerStart; *e1 = i1InnerEnd;
    *s2 = i2InnerStart; *e2 = i2InnerEnd;
    *s3 = i3InnerStart; *e3 = i3InnerEnd;
    switch (p) {
    case INNER_POINTS: break;
    case ALL_POINTS:
        *s1 = i1Halo1Start; *e1 = i1Halo2End;
        *s2 = i2Halo1Start; *e2 = i2Halo2End;
        *s3 = i3Halo1Start; *e3 = i3Halo2End;
        break;
    case MIDDLE_POINT:
        *s1 = *e1 = n1 / 2;
        *s2 = *e2 = n2 / 2;
        *s3 = *e3 = n3 / 2;
        break;
    case I1HALO1: *s1 = i1Halo1Start; *e1 = i1Halo1End; break;
    case I1HALO2: *s1 = i1Halo2Start; *e1 = i1Halo2End; break;
    case I2HALO1: *s2 = i2Halo1Start; *e2 = i2Halo1End; break;
    case I2HALO2: *s2 = i2Halo2Start; *e2 = i2Halo2End; break;
    case I3HALO1: *s3 = i3Halo1Start; *e3 = i3Halo1End; break;
    case I3HALO2: *s3 = i3Halo2Start; *e3 = i3Halo2End; break;
    case I1INNERHALO1: *e1 = *s1 + haloWidth; break;
    case I1INNERHALO2: *s1 = i1InnerEnd - haloWidth; *e1 = *s1 + haloWidth; break;
    case I2INNERHALO1: *e2 = *s2 + haloWidth; break;
    case I2INNERHALO2: *s2 = i2InnerEnd - haloWidth; *e2 = *s2 + haloWidth; break;
    case I3INNERHALO1: *e3 = *s3 + haloWidth; break;
    case I3INNERHALO2: *s3 = i3InnerEnd - haloWidth; *e3 = *s3 + haloWidth; break;
    default:
        printError("IN Grid::getGridIndex, Invalid pointType");
        return RTN_CODE_KO;
    }
    return RTN_CODE_OK;
}

Myint64 Grid::getNumberOfGridPoint(Grid_type, Point_type p)
{
    Myint64 s1, e1, s2, e2, s3, e3;
    getGridIndex(p, &s1, &e1, &s2, &e2, &s3, &e3);
    return (e1 - s1 + 1) * (e2 - s2 + 1) * (e3 - s3 + 1);
}

Myint64 Grid::getNumberOfGridPointCommHalo(Grid_type)
{
    Myint64 n = 0;
    const Point_type halos[6] = {I1HALO1, I1HALO2, I2HALO1, I2HALO2, I3HALO1, I3HALO2};
    for (Point_type h : halos)
        if (getNeighbourProc(h) != PROC_NULL) n += getNumberOfGridPoint(GRID_LOCAL, h);
    return n;
}

// src/grid.cpp:2588-2676: halo points plus the boundary plane of every free face
Myint64 Grid::getNumberOfGridPointBoundaryCondition(BoundCond_type bc)
{
    if (bc == NO_BOUND_COND) return 0;
    if (bc != BOUND_COND_ANTI_MIRROR) {
        printError("IN Grid::getNumberOfGridPointBoundaryCondition", (Myint)bc);
        return -1;
    }
    Myint64 n = 0;
    const Point_type halos[6] = {I1HALO1, I1HALO2, I2HALO1, I2HALO2, I3HALO1, I3HALO2};
    for (int k = 0; k < 2 * (int)dim; k++) {
        if (getNeighbourProc(halos[k]) != PROC_NULL) continue;
        Myint64 s[3], e[3];
        getGridIndex(halos[k], &s[0], &e[0], &s[1], &e[1], &s[2], &e[2]);
        const int a = k / 2;
        Myint64 len = e[a] - s[a] + 2;  // halo width plus the zeroed plane
        Myint64 prod = len;
        for (int b = 0; b < 3; b++)
            if (b != a) prod *= (e[b] - s[b] + 1);
        n += prod;
    }
    return n;
}

bool Grid::sameSize(const Grid &g) const
{
    return npoint == g.npoint && n1 == g.n1 && n2 == g.n2 && n3 == g.n3;
}

Rtn_code Grid::exchangeHalos(MPI_comm_mode_type mode)
{
    if (mode != MPI_COMM_MODE_SENDRECV) return RTN_CODE_KO;
    const Point_type halos[6] = {I1HALO1, I1HALO2, I2HALO1, I2HALO2, I3HALO1, I3HALO2};
    for (Point_type h : halos)
        if (exchangeHalo(mode, h) != RTN_CODE_OK) return RTN_CODE_KO;
    return RTN_CODE_OK;
}

Rtn_code Grid::exchangeHalo(MPI_comm_mode_type, Point_type p)
{
    if (getNeighbourProc(p) == PROC_NULL) return RTN_CODE_OK;
    printError(" In Grid::exchangeHalo, halo exchange needs more than one process");
    return RTN_CODE_KO;
}

Rtn_code Grid::myExchangeAll_halos()
{
    return getNumberOfNeighbour() == 0 ? RTN_CODE_OK : exchangeHalos(MPI_COMM_MODE_SENDRECV);
}

Myint Grid::getFlopPerPtFD_D1(Myint fdOrder) { return getFD_D1nMathOp(fdOrder); }
Myint Grid::getFlopPerPtFD_D2(Myint fdOrder) { return getFD_D2nMathOp(fdOrder); }
Myint Grid::getFlopPerPtFD_LAPLACIAN(Myint fdOrder)
{
    const Myint op = getFlopPerPtFD_D2(fdOrder);
    if (dim == DIM1) return op;
    if (dim == DIM2) return 2 * op + 1;
    return 3 * op + 2;
}
Myint Grid::getPtPerStencilFD_D1(Myint fdOrder) { return (fdOrder / 2) * 2; }
Myint Grid::getPtPerStencilFD_D2(Myint fdOrder) { return (fdOrder / 2) * 2 + 1; }
Myint Grid::getPtPerStencilFD_LAPLACIAN(Myint fdOrder)
{
    const Myint h = fdOrder / 2;
    return dim == DIM1 ? 2 * h + 1 : dim == DIM2 ? 4 * h + 1 : 6 * h + 1;
}

Myint Grid::getOffsetGlobal(Axis_type a)
{
    return a == AXIS1 ? i1OffsetGlobInner : a == AXIS2 ? i2OffsetGlobInner : i3OffsetGlobInner;
}

Myint64 Grid::getGlobalnPoints(Axis_type a)
{
    return a == AXIS1 ? n1Inner : a == AXIS2 ? n2Inner : n3Inner;
}

} // namespace hpcscan
