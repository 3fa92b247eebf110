// Backend registry (reference behaviour: src/grid_Factory.cpp:38-207,
// src/hardware_Factory.cpp:33-125: unknown mode -> error message and nullptr).
#include "grid_Factory.h"

#include "output_report.h"

namespace hpcscan {

static vector<GridBackend> &registry(void)
{
    static vector<GridBackend> r;
    return r;
}

bool Grid_Factory::registerBackend(const GridBackend &b)
{
    registry().push_back(b);
    return true;
}

const GridBackend *Grid_Factory::find(const string &mode)
{
    for (const GridBackend &b : registry())
        if (b.mode == mode) return &b;
    return nullptr;
}

vector<string> Grid_Factory::modes(void)
{
    vector<string> m;
    for (const GridBackend &b : registry()) m.push_back(b.mode);
    return m;
}

void Grid_Factory::shutdown(void)
{
    for (const GridBackend &b : registry())
        if (b.shutdown) b.shutdown();
}

shared_ptr<Grid> Grid_Factory::create(string gridMode, Grid_type gridType)
{
    const GridBackend *b = find(gridMode);
    if (!b) {
        printError("IN Grid_Factory::create, not supported or invalid gridMode", gridMode);
        printError("IN Grid_Factory::create, return nullptr", gridMode);
        return nullptr;
    }
    return b->create(gridType);
}

shared_ptr<Grid> Grid_Factory::create(string gridMode, Grid_type gridType, Dim_type dim, Myint64 n1,
                                      Myint64 n2, Myint64 n3)
{
    const GridBackend *b = find(gridMode);
    if (!b) {
        printError("IN Grid_Factory::create, not supported or invalid gridMode", gridMode);
        printError("IN Grid_Factory::create, return nullptr", gridMode);
        return nullptr;
    }
    return b->createSized(gridType, dim, n1, n2, n3);
}

shared_ptr<Hardware> Hardware_Factory::create(string gridMode)
{
    const GridBackend *b = Grid_Factory::find(gridMode);
    if (!b) {
        printError("IN Hardware_Factory::create, not supported or invalid gridMode", gridMode);
        printError("IN Hardware_Factory::create, return nullptr", gridMode);
        return nullptr;
    }
    return b->hardware ? b->hardware(gridMode) : make_shared<Hardware>(gridMode);
}

} // namespace hpcscan
