// Grid_Factory: maps the -testMode string to a Grid backend
// (reference: src/grid_Factory.cpp:38-207 and src/hardware_Factory.cpp:33-125).
//
// The reference hard-codes every backend in both factories behind #ifdefs.
// Here each backend registers itself once, with its Grid creators and the
// Hardware object that describes its target; the binary then offers exactly
// the modes it was linked with.
#ifndef HPCSCAN_GRID_FACTORY_H_
#define HPCSCAN_GRID_FACTORY_H_

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "grid.h"
#include "hardware.h"
#include "type_def.h"

namespace hpcscan {

class DataAcquisition;

// The creators return shared pointers made from the derived type, so that the
// backend's destructor runs although ~Grid / ~Hardware / ~DataAcquisition are
// not virtual in the reference (src/grid.h:47, src/hardware.h:26,
// src/data_Acquisition.h:27).
struct GridBackend {
    string mode;
    std::function<shared_ptr<Grid>(Grid_type)> create;
    std::function<shared_ptr<Grid>(Grid_type, Dim_type, Myint64, Myint64, Myint64)> createSized;
    std::function<shared_ptr<Hardware>(string)> hardware;   // may be empty: the host description
    std::function<void(void)> shutdown;   // release process-wide resources, may be empty
    std::function<shared_ptr<DataAcquisition>(void)> acquisition;   // may be empty: host acquisition
};

class Grid_Factory
{
public:
    static shared_ptr<Grid> create(string gridMode, Grid_type gridType);
    static shared_ptr<Grid> create(string gridMode, Grid_type gridType, Dim_type dim, Myint64 n1,
                                   Myint64 n2, Myint64 n3);
    static bool registerBackend(const GridBackend &b);
    static const GridBackend *find(const string &mode);
    static vector<string> modes(void);
    static void shutdown(void);
};

class Hardware_Factory
{
public:
    static shared_ptr<Hardware> create(string gridMode);
};

} // namespace hpcscan

#endif
