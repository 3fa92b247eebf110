// Data acquisition on the MI355X backend: the source term and the receivers
// work on the device grid (hsx_inject / hsx_gather); traces stay in HBM,
// time-major, until writeTrace copies them back once. Both are queued on the
// library stream without waiting: the next Grid call on that stream (which
// completes before it returns) or writeTrace's copy orders them, so the time
// loop pays no host round trip for them (400^3 order 8 modeling: the time loop
// ran at 163.9 GPoint/s when each of them waited for the stream).
//
// The class builds against the reference's DataAcquisition
// (src/data_Acquisition.h:20-80: its member names, its virtuals).  There
// appliSourceTerm and the destructor are not virtual: a call through a
// DataAcquisition& would reach the host version, which writes the source into
// the host array while the wavefield lives in HBM.  initialize() therefore
// checks that appliSourceTerm reaches this class through a base reference and
// fails with a message naming the one-word fix otherwise; the factory helper
// makes the shared pointer from this class, so the destructor runs.
#include <cmath>
#include <memory>
#include <vector>

#include "config.h"
#include "data_Acquisition.h"
#include "global.h"
#include "grid_Mi355x.h"
#include "output_report.h"

namespace hpcscan {

class DataAcquisition_Mi355x : public DataAcquisition
{
public:
    ~DataAcquisition_Mi355x(void)
    {
        for (void *p : {d_src, d_wavelet, d_rec, d_traces})
            if (p) hsx_free(p);
    }

    Rtn_code initialize(DataAcqui_type acquiType, Grid &coefGrid, Myint ntIn) override
    {
        const Rtn_code rc = DataAcquisition::initialize(acquiType, coefGrid, ntIn);
        if (rc != RTN_CODE_OK) return rc;
        // does appliSourceTerm dispatch to this class through a base reference?
        // (an ADJOINT call with a step number no caller uses: the base version
        // does nothing for ADJOINT)
        probed = false;
        DataAcquisition &base = *this;
        base.appliSourceTerm(coefGrid, ADJOINT, PROBE_STEP, 0);
        if (!probed) {
            printError(" In DataAcquisition_Mi355x::initialize, DataAcquisition::appliSourceTerm is not virtual: "
                       "the source term would go to the host array; declare it virtual (src/data_Acquisition.h:33)");
            return RTN_CODE_KO;
        }
        return RTN_CODE_OK;
    }

    // not marked override: the reference declares it non-virtual (see above)
    Rtn_code appliSourceTerm(Grid &grid, Modeling_type modType, Myint it, Myfloat dt)
    {
        if (modType == ADJOINT && it == PROBE_STEP) {
            probed = true;
            return RTN_CODE_OK;
        }
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
    static constexpr Myint PROBE_STEP = -0x5eed;
    bool probed = false;
    void *d_src = nullptr, *d_wavelet = nullptr, *d_rec = nullptr, *d_traces = nullptr;
    Myfloat dtReady = 0;

    // the reference's source term of step it: Ricker wavelet of dominant
    // frequency freqMax / 2.5 delayed by 1.5 sqrt(6) / (pi f), in double, then
    // Myfloat (src/data_Acquisition.cpp:290-301)
    static Myfloat ricker(Myint it, Myfloat dt)
    {
        const Myfloat64 f = Config::Instance()->freqMax / 2.5;
        const Myfloat64 t00 = 1.5 * sqrt(6.) / (M_PI * f);
        const Myfloat64 aa = M_PI * f * (it * dt - t00);
        const Myfloat64 a2 = pow(aa, 2.);
        return (1. - 2. * a2) * exp(-a2);
    }

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
        for (Myint it = 0; it < nt; it++) wav[it] = ricker(it, dt);
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
             !ok(hsx_memcpy_h2d(d_rec, tab_recIdxLoc, (size_t)nRecLoc * sizeof(int64_t), nullptr), "hsx_memcpy_h2d")))
            return false;
        dtReady = dt;
        return true;
    }
};

// for the factory: the shared pointer deletes a DataAcquisition_Mi355x
shared_ptr<DataAcquisition> makeDataAcquisition_Mi355x(void) { return make_shared<DataAcquisition_Mi355x>(); }

} // namespace hpcscan
