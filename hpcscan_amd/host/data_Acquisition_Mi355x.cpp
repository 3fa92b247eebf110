// Data acquisition on the MI355X backend: the source term and the receivers
// work on the device grid (hsx_inject / hsx_gather); traces stay in HBM,
// time-major, until writeTrace copies them back once. Both are queued on the
// library stream without waiting: the next Grid call on that stream (which
// completes before it returns) or writeTrace's copy orders them, so the time
// loop pays no host round trip for them (400^3 order 8 modeling: the time loop
// ran at 163.9 GPoint/s when each of them waited for the stream).
#include <vector>

#include "data_Acquisition.h"
#include "global.h"
#include "grid_Mi355x.h"
#include "output_report.h"

namespace hpcscan {

class DataAcquisition_Mi355x : public DataAcquisition
{
public:
    ~DataAcquisition_Mi355x(void) override
    {
        for (void *p : {d_src, d_wavelet, d_rec, d_traces})
            if (p) hsx_free(p);
    }

    Rtn_code appliSourceTerm(Grid &grid, Modeling_type modType, Myint it, Myfloat dt) override
    {
        if (modType != FORWARD) return DataAcquisition::appliSourceTerm(grid, modType, it, dt);
        if (nSrc == 0) return RTN_CODE_OK;
        Grid_Mi355x *g = device(grid);
        if (!g || !ready(dt)) return RTN_CODE_KO;
        if (it < 0 || it >= nt) {
            printError(" In DataAcquisition_Mi355x::appliSourceTerm, time step out of range", it);
            return RTN_CODE_KO;
        }
        hsx_grid d = g->desc();
        const char *w = static_cast<const char *>(d_wavelet) + (size_t)it * sizeof(Myfloat);
        if (!ok(hsx_inject(&d, g->d_grid, 1, static_cast<const int64_t *>(d_src), w, nullptr, 1.0, HSX_INJECT_SET,
                           nullptr),
                "hsx_inject"))
            return RTN_CODE_KO;
        g->touched();
        return RTN_CODE_OK;
    }

    Rtn_code recordTrace(Grid &grid, Modeling_type modType, Myint it, Myfloat dt) override
    {
        if (modType != FORWARD) return DataAcquisition::recordTrace(grid, modType, it, dt);
        const double t0 = hpcWtime();
        Grid_Mi355x *g = device(grid);
        if (!g || !ready(dt)) return RTN_CODE_KO;
        if (it < 0 || it >= nt) {
            printError(" In DataAcquisition_Mi355x::recordTrace, time step out of range", it);
            return RTN_CODE_KO;
        }
        hsx_grid d = g->desc();
        char *out = static_cast<char *>(d_traces) + (size_t)it * nRecLoc * sizeof(Myfloat);
        if (!ok(hsx_gather(&d, g->d_grid, nRecLoc, static_cast<const int64_t *>(d_rec), out, nullptr), "hsx_gather"))
            return RTN_CODE_KO;
        record_trace_time = hpcWtime() - t0;
        return RTN_CODE_OK;
    }

    Rtn_code writeTrace(string fileName) override
    {
        if (d_traces) {
            std::vector<Myfloat> tm((size_t)nt * nRecLoc);
            if (!ok(hsx_memcpy_d2h(tm.data(), d_traces, tm.size() * sizeof(Myfloat), nullptr), "hsx_memcpy_d2h"))
                return RTN_CODE_KO;
            for (Myint it = 0; it < nt; it++)
                for (Myint r = 0; r < nRecLoc; r++) trace[it + (size_t)nt * r] = tm[(size_t)it * nRecLoc + r];
        }
        return DataAcquisition::writeTrace(fileName);
    }

private:
    void *d_src = nullptr, *d_wavelet = nullptr, *d_rec = nullptr, *d_traces = nullptr;
    Myfloat dtReady = 0;

    static bool ok(int rc, const char *what)
    {
        if (rc == HSX_OK) return true;
        printError((string(" In DataAcquisition_Mi355x, ") + what + " failed:").c_str(), hsx_last_error());
        return false;
    }

    Grid_Mi355x *device(Grid &grid)
    {
        Grid_Mi355x *g = dynamic_cast<Grid_Mi355x *>(&grid);
        if (!g || !g->d_grid) printError(" In DataAcquisition_Mi355x, the grid is not an initialized MI355X grid");
        return g && g->d_grid ? g : nullptr;
    }

    // device copies of the geometry, the wavelet of every step and the trace buffer
    bool ready(Myfloat dt)
    {
        if (d_traces && dt == dtReady) return true;
        for (void **p : {&d_src, &d_wavelet, &d_rec, &d_traces})
            if (*p) {
                hsx_free(*p);
                *p = nullptr;
            }
        std::vector<Myfloat> wav(nt);
        for (Myint it = 0; it < nt; it++) wav[it] = rickerTerm(it, dt);
        const size_t nr = (size_t)(nRecLoc > 0 ? nRecLoc : 1);
        if (!ok(hsx_malloc(&d_src, sizeof(int64_t)), "hsx_malloc") ||
            !ok(hsx_malloc(&d_wavelet, (size_t)(nt > 0 ? nt : 1) * sizeof(Myfloat)), "hsx_malloc") ||
            !ok(hsx_malloc(&d_rec, nr * sizeof(int64_t)), "hsx_malloc") ||
            !ok(hsx_malloc(&d_traces, (size_t)(nt > 0 ? nt : 1) * nr * sizeof(Myfloat)), "hsx_malloc"))
            return false;
        const int64_t src = nSrc ? srcIdx[0] : 0;
        if (!ok(hsx_memcpy_h2d(d_src, &src, sizeof(int64_t), nullptr), "hsx_memcpy_h2d") ||
            !ok(hsx_memcpy_h2d(d_wavelet, wav.data(), wav.size() * sizeof(Myfloat), nullptr), "hsx_memcpy_h2d") ||
            (nRecLoc > 0 &&
             !ok(hsx_memcpy_h2d(d_rec, recIdx.data(), recIdx.size() * sizeof(int64_t), nullptr), "hsx_memcpy_h2d")))
            return false;
        dtReady = dt;
        return true;
    }
};

DataAcquisition *newDataAcquisition_Mi355x(void) { return new DataAcquisition_Mi355x(); }

} // namespace hpcscan
