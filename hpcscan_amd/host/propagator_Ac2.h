// Acoustic 2nd-order propagator (reference API: src/propagator_Ac2.h,
// src/propagator_Ac2SplitComp.h, src/propagator_Factory.h).
//
//   pn = coef * LAP(pc) + 2 pc - pn,   coef = vp^2 dt^2,   pn/pc swapped each step
#ifndef HPCSCAN_PROPAGATOR_AC2_H_
#define HPCSCAN_PROPAGATOR_AC2_H_

#include <memory>

#include "grid.h"
#include "type_def.h"

namespace hpcscan {

class Propagator_Ac2
{
public:
    Propagator_Ac2(void);
    virtual ~Propagator_Ac2(void) {}

    Rtn_code computeWavefieldNextTimeStep(Grid &prnGridIn, Grid &prcGridIn);
    virtual Rtn_code computePressureWithFD(Grid &prnGridIn, Grid &prcGridIn);
    Myfloat getCFL(void);
    virtual Rtn_code initialize(PropaInit_type propaInitType);
    virtual Rtn_code initializeGrid(Grid &gridIn, PropaInit_type propaInitType, Myfloat timeSec);
    void info(void);

    shared_ptr<Grid> coefGrid;
    shared_ptr<Grid> prnGrid;
    shared_ptr<Grid> prcGrid;

    BoundCond_type boundCondType;
    Myint fdOrder;
    Myfloat freqMax;
    Myint nt;
    Myfloat dt, stableDt;
    Myfloat tmax;
    Myint snapInc;
    Myfloat minVelocity, maxVelocity;
    string propaWaveEq;
    string propaKernelType;

    double halo_comm_time;
    double boundary_condition_time;
    double compute_pressure_time;
    double coef_read_total_time;
    double i_o_read_time;
    double i_o_write_coefgrid_time;
};

class Propagator_Ac2SplitComp : public Propagator_Ac2
{
public:
    Propagator_Ac2SplitComp(void);
    virtual Rtn_code computePressureWithFD(Grid &prnGridIn, Grid &prcGridIn);
    virtual Rtn_code initialize(PropaInit_type propaInitType);

private:
    shared_ptr<Grid> laplaGrid;
};

class Propagator_Factory
{
public:
    static shared_ptr<Propagator_Ac2> create(string propagator);
};

} // namespace hpcscan

#endif
