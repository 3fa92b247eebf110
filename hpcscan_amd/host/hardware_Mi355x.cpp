// Hardware_Mi355x: GPU description through the C ABI, power through hwmon.
#include "hardware_Mi355x.h"

#include <dirent.h>
#include <fstream>

#include "grid_Mi355x.h"
#include "hpcscan_mi355x.h"
#include "output_report.h"

namespace hpcscan {

Hardware_Mi355x::Hardware_Mi355x(string gridMode) : Hardware(gridMode) {}

void Hardware_Mi355x::info(void)
{
    print_blank();
    printInfo(MASTER, " Hardware information");
    printInfo(MASTER, " Target hardware:");
    hostInfo();
    hsx_device_props p;
    const int dev = mi355xSettings().device;
    if (hsx_device_info(dev, &p) == HSX_OK) {
        printInfo(MASTER, " GPU device\t", dev);
        printInfo(MASTER, " GPU name\t", p.name);
        printInfo(MASTER, " GPU arch\t", p.arch);
        printInfo(MASTER, " GPU compute units", p.compute_units);
        printInfo(MASTER, " GPU clock (MHz)", p.clock_mhz);
        printInfo(MASTER, " GPU memory (GB)", Myfloat(p.hbm_bytes / 1e9));
        printInfo(MASTER, " GPU L2 cache (KB)", (Myint)(p.l2_bytes / 1024));
    } else {
        printError(" In Hardware_Mi355x::info, hsx_device_info failed:", hsx_last_error());
    }
    supportGetPowerUsage = checkSupportGetPowerUsage();
    printInfo(MASTER, " Read power usage", supportGetPowerUsage ? "SUPPORTED" : "NOT SUPPORTED");
    print_line2();
}

// first amdgpu hwmon node with a power sensor (micro-watts)
bool Hardware_Mi355x::checkSupportGetPowerUsage(void)
{
    DIR *d = opendir("/sys/class/drm");
    if (!d) return false;
    struct dirent *e;
    while ((e = readdir(d)) != nullptr && powerFile.empty()) {
        string card = e->d_name;
        if (card.rfind("card", 0) != 0 || card.find('-') != string::npos) continue;
        string hw = "/sys/class/drm/" + card + "/device/hwmon";
        DIR *h = opendir(hw.c_str());
        if (!h) continue;
        struct dirent *f;
        while ((f = readdir(h)) != nullptr && powerFile.empty()) {
            if (string(f->d_name).rfind("hwmon", 0) != 0) continue;
            for (const char *s : {"power1_average", "power1_input"}) {
                string p = hw + "/" + f->d_name + "/" + s;
                ifstream t(p);
                long v;
                if (t >> v) {
                    powerFile = p;
                    break;
                }
            }
        }
        closedir(h);
    }
    closedir(d);
    return !powerFile.empty();
}

Myfloat Hardware_Mi355x::measureCurrentPower(void)
{
    if (powerFile.empty()) return UNSPECIFIED;
    ifstream t(powerFile);
    long uw = 0;
    if (!(t >> uw)) return UNSPECIFIED;
    return Myfloat(uw / 1e6);
}

} // namespace hpcscan
