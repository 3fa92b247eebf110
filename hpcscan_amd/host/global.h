// Process-wide state of the host harness (reference: src/global.h, src/main.cpp:20-30).
//
// The harness runs one process per GPU with no MPI: nMpiProc is 1 and
// myMpiRank 0.  hpcWtime() replaces MPI_Wtime for the timers that bracket
// every Grid call (src/propagator_Ac2.cpp:371-389).
#ifndef HPCSCAN_GLOBAL_H_
#define HPCSCAN_GLOBAL_H_

#include "type_def.h"

namespace hpcscan {

extern Debug_level debug;
extern int nMpiProc;
extern int myMpiRank;

// wall clock in seconds (monotonic)
double hpcWtime(void);

} // namespace hpcscan

#endif
