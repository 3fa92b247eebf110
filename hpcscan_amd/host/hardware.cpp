// Hardware base: host CPU description and power-counter statistics
// (reference behaviour: src/hardware.cpp:33-285).
#include "hardware.h"

#include <cfloat>
#include <cpuid.h>
#include <cstring>
#include <fstream>

#include <omp.h>

#include "config.h"
#include "global.h"
#include "output_report.h"

namespace hpcscan {

Hardware::Hardware(string gridMode) : supportGetPowerUsage(false), mode(gridMode) {}

Hardware::~Hardware(void) {}

void Hardware::info(void)
{
    print_blank();
    printInfo(MASTER, " Hardware information");
    printInfo(MASTER, " Target hardware:");
    hostInfo();
    supportGetPowerUsage = checkSupportGetPowerUsage();
    printInfo(MASTER, " Read power usage", supportGetPowerUsage ? "SUPPORTED" : "NOT SUPPORTED");
    print_line2();
}

// CPU brand string from cpuid leaves 0x80000002..4 (src/hardware.cpp:76-105)
void Hardware::hostInfo(void)
{
    unsigned int regs[12] = {0};
    char brand[49] = {0};
    unsigned int maxExt = __get_cpuid_max(0x80000000, nullptr);
    if (maxExt >= 0x80000004) {
        for (unsigned int i = 0; i < 3; i++)
            __get_cpuid(0x80000002 + i, &regs[4 * i], &regs[4 * i + 1], &regs[4 * i + 2], &regs[4 * i + 3]);
        memcpy(brand, regs, 48);
    }
    printInfo(MASTER, " CPU (Host) type", brand[0] ? brand : "unknown");
}

bool Hardware::checkSupportGetPowerUsage(void) { return false; }

Myfloat Hardware::measureCurrentPower(void) { return UNSPECIFIED; }

void Hardware::updateHwCounter(void)
{
    hwCounter_struct_type s;
    s.timeOfMeasure = hpcWtime();
    s.powerWatt = measureCurrentPower();
    hwCounterVec.push_back(s);
}

void Hardware::watchTimeAndUpdateHwCounter(void)
{
    if (Config::Instance()->hwCounterDt <= 0.) return;
    if (hwCounterVec.empty()) {
        printWarning("IN Hardware::watchTimeAndUpdateHwCounter, hwCounterVec.size() == 0");
        return;
    }
    if (hpcWtime() - hwCounterVec.back().timeOfMeasure >= Config::Instance()->hwCounterDt) updateHwCounter();
}

void Hardware::displayCounterStat(void)
{
    print_blank();
    print_line5();
    printInfo(MASTER, " Hardware counters statistics");
    if (supportGetPowerUsage && !hwCounterVec.empty()) {
        const Myint n = hwCounterVec.size();
        printInfo(MASTER, " Number of measures", n);
        const double statTime = hwCounterVec.back().timeOfMeasure - hwCounterVec.front().timeOfMeasure;
        printInfo(MASTER, " Stat. over period (s)", statTime);
        Myfloat32 pmin = FLT_MAX, pmax = -FLT_MAX, pavg = 0;
        for (const auto &s : hwCounterVec) {
            pmin = min(pmin, (Myfloat32)s.powerWatt);
            pmax = max(pmax, (Myfloat32)s.powerWatt);
            pavg += s.powerWatt;
        }
        pavg /= n;
        const Myfloat32 wh = pavg * statTime / 3600.;
        printInfo(MASTER, " Min. power (Watt)", pmin);
        printInfo(MASTER, " Max. power (Watt)", pmax);
        printInfo(MASTER, " Aver. power (Watt)", pavg);
        printInfo(MASTER, " Aver. consump. (W.h)", wh);
        if (myMpiRank == 0) {
            Config *c = Config::Instance();
            ofstream log("hpcscan.hwCounter." + c->testCaseName + ".log", ios::app);
            log << c->hostName << " " << c->testCaseName << " " << c->testMode << " " << c->propagator << " "
                << nMpiProc << " " << c->nsub1 << " " << c->nsub2 << " " << c->nsub3 << " "
                << omp_get_max_threads() << " " << c->n1 << " " << c->n2 << " " << c->n3 << " " << c->fdOrder
                << " " << statTime << " " << hwCounterVec.size() << " " << pmin << " " << pmax << " " << pavg
                << " " << wh << " \n";
        }
    } else {
        printInfo(MASTER, " No statistic available");
    }
    print_line5();
}

} // namespace hpcscan
