// Hardware_Mi355x: describes the GPU behind -testMode MI355X and reads its
// board power from the amdgpu hwmon interface when the node exposes it.
#ifndef HPCSCAN_HARDWARE_MI355X_H_
#define HPCSCAN_HARDWARE_MI355X_H_

#include <string>

#include "hardware.h"

namespace hpcscan {

class Hardware_Mi355x : public Hardware
{
public:
    Hardware_Mi355x(string gridMode);
    void info(void) override;
    Myfloat measureCurrentPower(void) override;

protected:
    bool checkSupportGetPowerUsage(void) override;
    string powerFile;
};

} // namespace hpcscan

#endif
