// Grid: the plugin interface of the stencil path.
//
// The public and protected members keep the names and signatures of the
// reference class (src/grid.h:35-337) so that a backend written for this
// harness (e.g. Grid_Mi355x) drops into the reference tree unchanged.
//
// Unlike the reference, whose base class is also the CPU "Baseline"
// implementation, this base class implements only what is backend
// independent: the index layout (src/grid.cpp:252-1009), host-side
// bookkeeping, unit grids and file output.  Every compute method fails loudly
// unless a backend overrides it, so a missing override can never fall back to
// some other implementation silently.
#ifndef HPCSCAN_GRID_H_
#define HPCSCAN_GRID_H_

#include <cstdint>
#include <string>

#include "type_def.h"

namespace hpcscan {

class Grid
{
public:
    Grid(Grid_type);
    Grid(Grid_type gridTypeIn, Dim_type dimTypeIn, Myint64 n1InnerIn, Myint64 n2InnerIn,
         Myint64 n3InnerIn);
    // virtual here: the factories hand out shared_ptr<Grid> and backends own device memory
    virtual ~Grid(void);

    virtual Rtn_code initializeGrid(void);
    virtual void info(void);

    // file output (src/grid.cpp:1245-1351); backends bring data to grid_3d first
    virtual Rtn_code write(string);
    virtual Rtn_code write(Point_type pointType, string);
    virtual Rtn_code writeGlobal(Point_type pointType, string file_name, Myint n_iter);
    virtual Rtn_code read(string);
    virtual Rtn_code read(Point_type pointType, string);

    // error metrics and reductions (src/grid.cpp:2006-2370)
    virtual Myfloat L1Err(Point_type pointType, const Grid &gridIn) const;
    virtual Myfloat allProcL1Err(Point_type, const Grid &) const;
    virtual Myfloat maxErr(Point_type, const Grid &) const;
    virtual Myfloat getMin(Point_type);
    virtual Myfloat getMax(Point_type);
    virtual Myfloat allProcGetMin(Point_type);
    virtual Myfloat allProcGetMax(Point_type);
    virtual Myfloat getSumAbs(Point_type) const;
    virtual Myfloat getSumAbsDiff(Point_type, const Grid &) const;

    Rtn_code getGridIndex(Point_type, Myint64 *i1Start, Myint64 *i1End, Myint64 *i2Start,
                          Myint64 *i2End, Myint64 *i3Start, Myint64 *i3End) const;
    Myint getNumberOfNeighbour();
    Myint64 getNumberOfGridPoint(Grid_type gridType, Point_type pointType);
    Myint64 getNumberOfGridPointCommHalo(Grid_type gridType);
    Myint64 getNumberOfGridPointBoundaryCondition(BoundCond_type boundCondType);
    Myfloat getMinCoord(Axis_type);
    Myfloat getMaxCoord(Axis_type);
    Myfloat64 getSpaceSampling(Axis_type axisType);
    Myfloat getMinSpaceSampling(void);
    Myint getNeighbourProc(Point_type);
    bool sameSize(const Grid &) const;

    // grid filling (src/grid.cpp:1650-1920)
    virtual void fill(Point_type pointType, Myfloat val);
    virtual void fill(Point_type pType, Func_type t1, Func_type t2, Func_type t3,
                      Myfloat64 param1, Myfloat64 param2, Myfloat64 param3, Myfloat64 amp);
    virtual void fillArray(Myfloat val);
    virtual void copyArray(const Grid &gridIn);
    virtual void addArray(const Grid &gridIn1, const Grid &gridIn2);
    virtual void multiplyArray(const Grid &gridIn1, const Grid &gridIn2);
    virtual void addUpdateArray(const Grid &gridIn);

    void defineUnitGrid();
    void resetOriginGrid(Myfloat Orig1In, Myfloat Orig2In, Myfloat Orig3In);
    void shiftOriginGrid(Myfloat Shift1, Myfloat Shift2, Myfloat Shift3);

    // halo exchanges: one rank, no neighbour, nothing to exchange
    virtual Rtn_code exchangeHalos(MPI_comm_mode_type);
    virtual Rtn_code exchangeHalo(MPI_comm_mode_type, Point_type pointType);
    virtual Rtn_code myExchangeAll_halos();

    // the stencil path (src/grid.cpp:3065-5112)
    virtual Rtn_code applyBoundaryCondition(BoundCond_type boundCondType);
    virtual Rtn_code FD_D1_N1(Point_type pType, const Grid &Wgrid, Myint fdOrder);
    virtual Rtn_code FD_D1_N2(Point_type pType, const Grid &Wgrid, Myint fdOrder);
    virtual Rtn_code FD_D1_N3(Point_type pType, const Grid &Wgrid, Myint fdOrder);
    virtual Rtn_code FD_D2_N1(Point_type pType, const Grid &Wgrid, Myint fdOrder);
    virtual Rtn_code FD_D2_N2(Point_type pType, const Grid &Wgrid, Myint fdOrder);
    virtual Rtn_code FD_D2_N3(Point_type pType, const Grid &Wgrid, Myint fdOrder);
    virtual Rtn_code FD_LAPLACIAN(Point_type pType, const Grid &Wgrid, Myint fdOrder);
    virtual Rtn_code updatePressure(Point_type pType, const Grid &prcGrid, const Grid &coefGrid,
                                    const Grid &laplaGrid);
    virtual Rtn_code computePressureWithFD(Grid &prcGridIn, Grid &coefGridIn, Myint fdOrder);

    // operation counts used by the reports (src/grid.cpp:2374-2466)
    Myint getFlopPerPtFD_D1(Myint fdOrder);
    Myint getFlopPerPtFD_D2(Myint fdOrder);
    Myint getFlopPerPtFD_LAPLACIAN(Myint fdOrder);
    Myint getPtPerStencilFD_D1(Myint fdOrder);
    Myint getPtPerStencilFD_D2(Myint fdOrder);
    Myint getPtPerStencilFD_LAPLACIAN(Myint fdOrder);

    Myint getOffsetGlobal(Axis_type axis);
    Myint64 getGlobalnPoints(Axis_type axis);

    Dim_type dim;
    Grid_type gridType;
    Myfloat *grid_3d;         // host copy of the grid (allocated by initializeGrid)
    Myint n1, n2, n3;
    Myfloat64 d1, d2, d3;
    Myint haloWidth;
    string gridMode;
    Myint64 n1Inner, n2Inner, n3Inner;
    Myint64 npoint;
    double grid_writeGlobal_time;
    double grid_write_time;
    double i_o_pread_time;

protected:
    Myint i1OffsetStart, i1OffsetEnd, i2OffsetStart, i2OffsetEnd, i3OffsetStart, i3OffsetEnd;
    Myint i1Halo1Start, i1Halo1End, i2Halo1Start, i2Halo1End, i3Halo1Start, i3Halo1End;
    Myint i1Halo2Start, i1Halo2End, i2Halo2Start, i2Halo2End, i3Halo2Start, i3Halo2End;
    Myint i1Layer1Start, i1Layer1End, i2Layer1Start, i2Layer1End, i3Layer1Start, i3Layer1End;
    Myint i1Layer2Start, i1Layer2End, i2Layer2Start, i2Layer2End, i3Layer2Start, i3Layer2End;
    Myint i1InnerStart, i1InnerEnd, i2InnerStart, i2InnerEnd, i3InnerStart, i3InnerEnd;
    Myint i1PadStart, i1PadEnd, i2PadStart, i2PadEnd, i3PadStart, i3PadEnd;

    Myfloat64 Orig1, Orig2, Orig3;
    Myint i1OffsetGlobInner, i2OffsetGlobInner, i3OffsetGlobInner;
    Myint i1ProcIdStart, i1ProcIdEnd, i2ProcIdStart, i2ProcIdEnd, i3ProcIdStart, i3ProcIdEnd;

    virtual void padGridn1(void);
    virtual void padGridn2(void);
    virtual void padGridn3(void);
    virtual void offsetGridn1(void);
    virtual void offsetGridn2(void);
    virtual void offsetGridn3(void);

    // report a compute method that this grid mode does not provide
    void notProvided(const char *method) const;

private:
    void resetIndexes(void);
};

} // namespace hpcscan

#endif
