// Data acquisition of the Modeling test case (reference behaviour:
// src/data_Acquisition.cpp:58-481), one rank: the built-in geometry, the
// Ricker source term, trace recording and the trace files.
#include "data_Acquisition.h"

#include <cerrno>
#include <cmath>
#include <cstring>
#include <fcntl.h>
#include <fstream>
#include <unistd.h>

#include "config.h"
#include "constant.h"
#include "global.h"
#include "grid_Factory.h"
#include "output_report.h"

namespace hpcscan {

DataAcquisition::DataAcquisition(void)
{
    nt = 0;
    nSrc = 0;
    nRec = nRecLoc = 0;
    nSampleInTraceLoc = total_nSampleInTrace = 0;
    record_trace_time = trace_write_time = trace_gather_time = total_timeWriteTrace = 0.0;
}

DataAcquisition::~DataAcquisition(void)
{
    delete[] srcIdx;
    delete[] tab_recIdxLoc;
    delete[] tab_idxglob;
    delete[] trace;
}

// Built-in geometry (src/data_Acquisition.cpp:72-279): one source one point
// below the top n1 plane, in the middle of n2 (x n3); receivers on that plane:
// one point in 1D, every n2 point in 2D, every n2 point of every 10th n3 plane in 3D.
Rtn_code DataAcquisition::initialize(DataAcqui_type acquiType, Grid &coefGrid, Myint ntIn)
{
    print_blank();
    printInfo(MASTER, " * Initialize Data Acquisition *");
    nt = ntIn;
    if (acquiType != ACQUI_BUILTIN) {
        printError("In DataAcquisition::initialize, unsupported acquiType", (Myint)acquiType);
        return RTN_CODE_KO;
    }
    printInfo(MASTER, " Acquisition type", "BUILTIN");
    Myint64 i1s, i1e, i2s, i2e, i3s, i3e;
    coefGrid.getGridIndex(INNER_POINTS, &i1s, &i1e, &i2s, &i2e, &i3s, &i3e);
    const Myint64 n1 = coefGrid.n1, n2 = coefGrid.n2;
    const Myint64 N2 = i2e - i2s + 1, N3 = i3e - i3s + 1;
    auto flat = [&](Myint64 g1, Myint64 g2, Myint64 g3) {
        return (i1s + g1) + (i2s + g2) * n1 + (i3s + g3) * n2 * n1;
    };

    // source
    delete[] srcIdx;
    srcIdx = new Myint64[1];
    srcIdx[0] = flat(1, N2 / 2, N3 / 2);
    nSrc = 1;
    printInfo(MASTER, " Number of sources", nSrc);

    // receivers
    std::vector<Myint64> rec;
    if (coefGrid.dim == DIM1) {
        rec.push_back(flat(1, 0, 0));
    } else if (coefGrid.dim == DIM2) {
        for (Myint64 r2 = 0; r2 < N2; r2++) rec.push_back(flat(1, r2, 0));
    } else {
        for (Myint64 r3 = 0; r3 < N3; r3 += 10)
            for (Myint64 r2 = 0; r2 < N2; r2++) rec.push_back(flat(1, r2, r3));
    }
    nRec = nRecLoc = (Myint)rec.size();
    delete[] tab_recIdxLoc;
    delete[] tab_idxglob;
    tab_recIdxLoc = new Myint64[rec.size() ? rec.size() : 1];
    tab_idxglob = new Myint[rec.size() ? rec.size() : 1];
    for (size_t i = 0; i < rec.size(); i++) {
        tab_recIdxLoc[i] = rec[i];
        tab_idxglob[i] = (Myint)i;
    }
    printInfo(MASTER, " Number of receivers", nRec);
    printInfo(ALL, " Number of receivers locally", nRecLoc);
    nSampleInTraceLoc = total_nSampleInTrace = (Myint64)nRecLoc * nt;
    printInfo(MASTER, " Trace file size (MB)", total_nSampleInTrace * sizeof(Myfloat) / 1e6);
    delete[] trace;
    trace = new Myfloat[nSampleInTraceLoc ? nSampleInTraceLoc : 1];
    for (Myint64 i = 0; i < nSampleInTraceLoc; i++) trace[i] = ZERO;
    return RTN_CODE_OK;
}

// Ricker wavelet of dominant frequency freqMax / 2.5, delayed by 1.5 sqrt(6) / (pi f)
// (src/data_Acquisition.cpp:292-301)
static Myfloat rickerTerm(Myint it, Myfloat dt)
{
    const Myfloat64 rickerFreq = Config::Instance()->freqMax / 2.5;
    const Myfloat64 da = M_PI * rickerFreq;
    const Myfloat64 t00 = 1.5 * sqrt(6.) / da;
    const Myfloat64 tt = it * dt;
    const Myfloat64 aa = M_PI * rickerFreq * (tt - t00);
    const Myfloat64 a2 = pow(aa, 2.);
    return (1. - 2. * a2) * exp(-a2);
}

Rtn_code DataAcquisition::appliSourceTerm(Grid &grid, Modeling_type modType, Myint it, Myfloat dt)
{
    if (modType == FORWARD) {
        if (nSrc != 0) grid.grid_3d[srcIdx[0]] = rickerTerm(it, dt);
    } else if (modType != ADJOINT) {
        printError("In DataAcquisition::appliSourceTerm, unsupported modType", (Myint)modType);
        return RTN_CODE_KO;
    }
    return RTN_CODE_OK;
}

Rtn_code DataAcquisition::recordTrace(Grid &grid, Modeling_type modType, Myint it, Myfloat)
{
    const double t0 = hpcWtime();
    if (modType == FORWARD) {
        for (Myint iRec = 0; iRec < nRecLoc; iRec++) trace[it + (size_t)nt * iRec] = grid.grid_3d[tab_recIdxLoc[iRec]];
    } else if (modType != ADJOINT) {
        printError("In DataAcquisition::recordTrace, unsupported modType", (Myint)modType);
        return RTN_CODE_KO;
    }
    record_trace_time = hpcWtime() - t0;
    return RTN_CODE_OK;
}

// traces of all receivers in global receiver order, appended to <fileName>.trace.bin;
// <fileName>.trace.info holds nt, nRec, 1 (src/data_Acquisition.cpp:355-478)
Rtn_code DataAcquisition::writeTrace(string fileName)
{
    const double t0 = hpcWtime();
    const string fbin = fileName + ".trace.bin";
    printInfo(MASTER, "* write on disk\t", fbin.c_str());
    std::vector<Myfloat> all((size_t)nRec * nt, 0.0);
    for (Myint i = 0; i < nRecLoc; i++)
        memcpy(&all[(size_t)tab_idxglob[i] * nt], &trace[(size_t)i * nt], nt * sizeof(Myfloat));
    const double t0w = hpcWtime();
    const int fd = open(fbin.c_str(), O_WRONLY | O_APPEND | O_CREAT, S_IRUSR | S_IWUSR);
    if (fd == -1) {
        printError(" In DataAcquisition::writeTrace, cannot open", fbin + ": " + strerror(errno));
        return RTN_CODE_KO;
    }
    const char *p = reinterpret_cast<const char *>(all.data());
    size_t left = all.size() * sizeof(Myfloat);
    while (left > 0) {
        const ssize_t w = ::write(fd, p, left);
        if (w < 0) {
            close(fd);
            printError(" In DataAcquisition::writeTrace, write failed", fbin);
            return RTN_CODE_KO;
        }
        p += w;
        left -= (size_t)w;
    }
    close(fd);
    std::ofstream info(fileName + ".trace.info");
    info << nt << "\n" << nRec << "\n" << 1 << "\n";
    info.close();
    trace_write_time = hpcWtime() - t0w;
    total_timeWriteTrace = hpcWtime() - t0;
    printInfo(MASTER, "* write time (s)", total_timeWriteTrace);
    printInfo(MASTER, "* write bwth (GB/s)", all.size() * sizeof(Myfloat) / total_timeWriteTrace / 1e9);
    return RTN_CODE_OK;
}

shared_ptr<DataAcquisition> DataAcquisition_Factory::create(string gridMode)
{
    const GridBackend *b = Grid_Factory::find(gridMode);
    if (b && b->acquisition) return b->acquisition();
    return make_shared<DataAcquisition>();
}

} // namespace hpcscan
