// Grid_Mi355x: the MI355X backend of the Grid plugin interface.
//
// Selected with -testMode MI355X.  Grid values live in HBM; every compute
// virtual of Grid (src/grid.h:49-245) runs as hand-written gfx950 kernels
// behind the C ABI of include/hpcscan_mi355x.h and is complete on return, so
// the callers' wall-clock timers (src/propagator_Ac2.cpp:371-389) measure it.
// The host array grid_3d is only filled on demand (write()).
#ifndef HPCSCAN_GRID_MI355X_H_
#define HPCSCAN_GRID_MI355X_H_

#include "grid.h"
#include "hpcscan_mi355x.h"

namespace hpcscan {

// Settings of the MI355X backend, read once from the environment, so that the
// backend needs no member beyond the reference's Config (src/config.h):
//   HPCSCAN_MI355X_DEVICE    HIP device ordinal (default 0)
//   HPCSCAN_MI355X_KERNEL    stencil kernels: auto (default) | ref | tiled
//   HPCSCAN_MI355X_FOLD_BC   1: fold the free surface into the fused step (default 0)
//   HPCSCAN_MI355X_AUTOTUNE  1: time the tile shapes once per 3D grid layout and
//                            keep the fastest (hsx_autotune; default 0)
//   HPCSCAN_MI355X_CONTRACT  1: fused multiply-adds in the tiled fused step and
//                            Laplacian (hsx_set_contract; results no longer the
//                            reference's bits; default 0)
// The harness's -gpuDevice / -hipKernel / -hipFoldBC / -hipAutotune /
// -hipContract flags set them.
struct Mi355xSettings {
    int device = 0;
    string kernel = "auto";
    bool foldBC = false;
    bool autotune = false;
    bool contract = false;
};
const Mi355xSettings &mi355xSettings(void);

class Grid_Mi355x : public Grid
{
public:
    Grid_Mi355x(Grid_type);
    Grid_Mi355x(Grid_type, Dim_type, Myint64, Myint64, Myint64);
    // not virtual in the reference (src/grid.h:47): the factory makes the
    // shared pointer from this class, so this runs
    ~Grid_Mi355x(void);

    Rtn_code initializeGrid(void) override;
    void info(void) override;
    Rtn_code write(string) override;
    Rtn_code write(Point_type, string) override;
    Rtn_code writeGlobal(Point_type, string, Myint) override;
    Rtn_code read(string) override;
    Rtn_code read(Point_type, string) override;

    Myfloat L1Err(Point_type, const Grid &) const override;
    Myfloat maxErr(Point_type, const Grid &) const override;
    Myfloat getMin(Point_type) override;
    Myfloat getMax(Point_type) override;
    Myfloat getSumAbs(Point_type) const override;
    Myfloat getSumAbsDiff(Point_type, const Grid &) const override;

    void fill(Point_type, Myfloat) override;
    void fill(Point_type, Func_type, Func_type, Func_type, Myfloat64, Myfloat64, Myfloat64, Myfloat64) override;
    void fillArray(Myfloat) override;
    void copyArray(const Grid &) override;
    void addArray(const Grid &, const Grid &) override;
    void multiplyArray(const Grid &, const Grid &) override;
    void addUpdateArray(const Grid &) override;

    Rtn_code applyBoundaryCondition(BoundCond_type) override;
    Rtn_code FD_D1_N1(Point_type, const Grid &, Myint) override;
    Rtn_code FD_D1_N2(Point_type, const Grid &, Myint) override;
    Rtn_code FD_D1_N3(Point_type, const Grid &, Myint) override;
    Rtn_code FD_D2_N1(Point_type, const Grid &, Myint) override;
    Rtn_code FD_D2_N2(Point_type, const Grid &, Myint) override;
    Rtn_code FD_D2_N3(Point_type, const Grid &, Myint) override;
    Rtn_code FD_LAPLACIAN(Point_type, const Grid &, Myint) override;
    Rtn_code updatePressure(Point_type, const Grid &, const Grid &, const Grid &) override;
    Rtn_code computePressureWithFD(Grid &, Grid &, Myint) override;

    // device copy of the grid values (npoint * sizeof(Myfloat) bytes)
    void *d_grid;
    // C-ABI descriptor of this grid
    hsx_grid desc(void) const;
    // values changed behind the backend's back (source injection): any halo
    // values held for the folded free surface are stale
    void touched(void) const { fsHalo = 0; }

protected:
    // Free surface folded into the next step.  applyBoundaryCondition on a grid
    // whose halos already hold the free-surface values of its inner points only
    // records the zeros it owes on the boundary planes (fsPending = face mask);
    // computePressureWithFD reads those planes as zero (hsx_ac2_step_fs) and
    // writes the new grid's halos (fsHalo).  Any other use of a grid first
    // applies what it owes (settle).
    mutable int fsPending = 0;
    mutable int fsHalo = 0;
    bool settle(void) const;
    void range(Point_type, int64_t r[6]) const;
    bool check(int rc, const char *what) const;
    const Grid_Mi355x *peer(const Grid &g, const char *what) const;
    Rtn_code toHost(void);
    Rtn_code fdD2(int axis, const Grid &W, Myint fdOrder);
    Rtn_code fdD1(int axis, const Grid &W, Myint fdOrder);
};

} // namespace hpcscan

#endif
