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
    // virtual here; not in the reference (src/data_Acquisition.h:27), whose
    // factory deletes through the base pointer
    virtual ~DataAcquisition(void);

    virtual Rtn_code initialize(DataAcqui_type acquiType, Grid &coefGrid, Myint nt);
    // sets the source point of grid to the Ricker wavelet at it * dt; virtual
    // here, not in the reference (src/data_Acquisition.h:33): a backend whose
    // grids live on a device overrides it
    virtual Rtn_code appliSourceTerm(Grid &grid, Modeling_type modType, Myint it, Myfloat dt);
    // trace[it + nt * iRec] = grid at receiver iRec
    virtual Rtn_code recordTrace(Grid &grid, Modeling_type modType, Myint it, Myfloat dt);
    // <fileName>.trace.bin (receiver-major, appended) and <fileName>.trace.info
    virtual Rtn_code writeTrace(string fileName);

    double trace_gather_time;
    double trace_write_time;
    double record_trace_time;
    double total_timeWriteTrace;

protected:
    // members of the reference class (src/data_Acquisition.h:48-77), one rank
    Myint nSrc;
    Myint nt;
    Myint64 *srcIdx = nullptr;            // flat grid index of each source
    Myint nRec;                           // receivers in all
    Myint nRecLoc;                        // receivers of this process
    Myint64 *tab_recIdxLoc = nullptr;     // flat grid index of each local receiver
    Myint *tab_idxglob = nullptr;         // global receiver number of each local receiver
    Myfloat *trace = nullptr;             // nRecLoc x nt, receiver-major
    Myint64 nSampleInTraceLoc;
    Myint64 total_nSampleInTrace;
};

class DataAcquisition_Factory
{
public:
    // the acquisition of the grid mode's backend, else the host one
    static shared_ptr<DataAcquisition> create(string gridMode);
};

} // namespace hpcscan

#endif
