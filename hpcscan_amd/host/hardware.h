// Hardware: target description and power counters (reference API: src/hardware.h).
#ifndef HPCSCAN_HARDWARE_H_
#define HPCSCAN_HARDWARE_H_

#include <string>
#include <vector>

#include "type_def.h"

namespace hpcscan {

class Hardware
{
public:
    Hardware(string gridMode);
    virtual ~Hardware(void);

    virtual void info(void);
    void hostInfo(void);
    void updateHwCounter(void);
    void watchTimeAndUpdateHwCounter(void);
    virtual Myfloat measureCurrentPower(void);
    void displayCounterStat(void);

    bool supportGetPowerUsage;

protected:
    virtual bool checkSupportGetPowerUsage(void);
    vector<hwCounter_struct_type> hwCounterVec;
    string mode;
};

} // namespace hpcscan

#endif
