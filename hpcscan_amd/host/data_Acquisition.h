// Data acquisition of the Modeling test case: one source, a line / plane of
// receivers, traces written to disk (reference API: src/data_Acquisition.h,
// src/data_Acquisition_Factory.h).
//
// The base class works on the host array grid_3d (the CPU backends); a backend
// whose grids live elsewhere registers its own subclass with Grid_Factory.
#ifndef HPCSCAN_DATA_ACQUISITION_H_
#define HPCSCAN_DATA_ACQUISITION_H_

#include <memory>
#include <string>
#include <vector>

#include "grid.h"
#include "type_def.h"

namespace hpcscan {

class DataAcquisition
{
public:
    DataAcquisition(void);
    virtual ~DataAcquisition(void) {}

    Rtn_code initialize(DataAcqui_type acquiType, Grid &coefGrid, Myint nt);
    // sets the source point of grid to the Ricker wavelet at it * dt
    virtual Rtn_code appliSourceTerm(Grid &grid, Modeling_type modType, Myint it, Myfloat dt);
    // trace[it + nt * iRec] = grid at receiver iRec
    virtual Rtn_code recordTrace(Grid &grid, Modeling_type modType, Myint it, Myfloat dt);
    // <fileName>.trace.bin (receiver-major, appended) and <fileName>.trace.info
    virtual Rtn_code writeTrace(string fileName);

    // source term of time step it (src/data_Acquisition.cpp:292-301)
    static Myfloat rickerTerm(Myint it, Myfloat dt);

    Myint nt;
    Myint nSrc;
    std::vector<Myint64> srcIdx;        // flat grid index of each source
    Myint nRec, nRecLoc;
    std::vector<Myint64> recIdx;        // flat grid index of each local receiver
    std::vector<Myint> recIdxGlob;      // global receiver number of each local receiver
    std::vector<Myfloat> trace;         // nRecLoc x nt, receiver-major

    double record_trace_time;
    double trace_write_time;
    double trace_gather_time;
    double total_timeWriteTrace;
};

class DataAcquisition_Factory
{
public:
    // the acquisition of the grid mode's backend, else the host one
    static shared_ptr<DataAcquisition> create(string gridMode);
};

} // namespace hpcscan

#endif
