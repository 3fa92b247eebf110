// Constants of the hpcscan host API (reference: src/constant.h:1-53).
#ifndef HPCSCAN_CONSTANT_H_
#define HPCSCAN_CONSTANT_H_

#include "type_def.h"

namespace hpcscan {

// truncated on purpose: every published error value depends on it (src/constant.h:12)
const Myfloat64 PI = 3.1415926535897;

#ifdef HPCSCAN_STORAGE_FP64
const Myfloat32 MAX_ERR_FLOAT = 1.e-11;
#else
const Myfloat32 MAX_ERR_FLOAT = 1.e-5;
#endif

const Myfloat2 ZERO = 0.0;
const Myfloat2 ONE  = 1.0;
const Myfloat2 TWO  = 2.0;

// grid modes (names of the reference, src/constant.h:28-38, plus this backend)
const string GRID_MODE_BASELINE = "Baseline";
const string GRID_MODE_CACHEBLK = "CacheBlk";
const string GRID_MODE_MI355X   = "MI355X";

// propagators (src/constant.h:43-44)
const string PROPA_TYPE_AC2STANDARD  = "Ac2Standard";
const string PROPA_TYPE_AC2SPLITCOMP = "Ac2SplitComp";

// rank id of "no neighbour" in a single-process run
const Myint PROC_NULL = -1;

} // namespace hpcscan

#endif
