// Configuration singleton (reference API: src/config.h, defaults src/config.cpp:32-83).
// Every reference flag keeps its name, default and validation; flags that only
// exist for the MI355X backend are add-only and listed after them.
#ifndef HPCSCAN_CONFIG_H_
#define HPCSCAN_CONFIG_H_

#include <climits>
#include <fstream>
#include <string>

#include "type_def.h"

namespace hpcscan {

class Config
{
public:
    static Config *Instance();

    Rtn_code parse_argument(int argc, char *argv[]);
    Rtn_code info(void);
    Rtn_code initialize(void);
    Rtn_code finalize(Rtn_code);

    bool autoPad;
    string boundary;
    Myint cb1, cb2, cb3;
    Myfloat64 d1, d2, d3;
    ofstream debugLogFile;
    Dim_type dim;
    Myfloat64 dt;
    Myint fdOrder;
    Myfloat64 freqMax;
    Myint gpuBlkSize, gpuBlkSize1, gpuBlkSize2, gpuBlkSize3;
    bool gpuMpiAware;
    Myint gpuGridSize;
    Myfloat64 hwCounterDt;
    string modelVpFile;
    Myint n1, n2, n3;
    Myint n1AddPad, n2AddPad, n3AddPad;
    Myint n1MulPad, n2MulPad, n3MulPad;
    Myint n1Offset, n2Offset, n3Offset;
    Myint nlayer;
    Myint nt;
    Myint ntry;
    Myint nsub1, nsub2, nsub3;
    Myfloat64 param1, param2, param3, param4;
    string propagator;
    Myfloat64 ratioCFL;
    Myfloat64 snapDt;
    Myint snapInc;
    string testCaseName;
    string testMode;
    Myfloat64 tmax;
    bool writeGrid;
    char hostName[HOST_NAME_MAX + 1];
    char userName[LOGIN_NAME_MAX + 1];

    // MI355X backend options (no reference counterpart)

private:
    Config(void);
    Config(Config const &) = delete;
    Config &operator=(Config const &) = delete;
    static Config *m_pInstance;
};

} // namespace hpcscan

#endif
