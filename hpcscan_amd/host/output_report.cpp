// Text report (reference behaviour: src/output_report.cpp:35-456).
#include "output_report.h"

#include <chrono>
#include <iostream>

#include "config.h"
#include "global.h"
#include "grid_Factory.h"

#ifndef HPCSCAN_GIT_COMMIT
#define HPCSCAN_GIT_COMMIT "unknown"
#endif

namespace hpcscan {

static const char *VERSION_STRING = "1.2-mi355x";

static const string LINE_T1(80, '=');
static const string LINE_T2(80, '-');
static const string LINE_T3(80, '.');
static string repeatPattern(const char *p)
{
    string s;
    while (s.size() < 80) s += p;
    return s.substr(0, 80);
}
static const string LINE_T4 = repeatPattern("Xx");
static const string LINE_T5 = repeatPattern("-*");

Debug_level debug = NO_DEBUG;
int nMpiProc = 1;
int myMpiRank = 0;

double hpcWtime(void)
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static inline bool master(Display_type d) { return d == ALL || myMpiRank == 0; }

Rtn_code print_header_of_output_report(void)
{
    if (myMpiRank != 0) return RTN_CODE_OK;
    print_line1();
    cout << "\t\t\t H P C S C A N - ver " << VERSION_STRING << "\n\n";
    printInfo(MASTER, " Git version", HPCSCAN_GIT_COMMIT);
    printInfo(MASTER, "");
    printInfo(MASTER, " Host name", Config::Instance()->hostName);
    printInfo(MASTER, " User name", Config::Instance()->userName);
    string st = (sizeof(Myfloat) == 8 ? "FP64" : "FP32");
    st += " [" + to_string(sizeof(Myfloat)) + " Bytes]";
    printInfo(MASTER, " Storage", st);
    string ct = (sizeof(Myfloat2) == 8 ? "FP64" : "FP32");
    ct += " [" + to_string(sizeof(Myfloat2)) + " Bytes]";
    printInfo(MASTER, " Compute", ct);
    for (const string &m : Grid_Factory::modes()) {
        string label = " Mode " + m;
        printInfo(MASTER, label.c_str(), "ENABLED");
    }
    print_line1();
    return RTN_CODE_OK;
}

Rtn_code print_end_of_output_report(Rtn_code rtnCode)
{
    if (myMpiRank != 0) return RTN_CODE_OK;
    cout << "\n";
    if (rtnCode == RTN_CODE_OK) {
        print_line1();
        cout << "\t\t\t HPCSCAN TERMINATED SUCCESSFULLY\n";
        print_line1();
    } else if (rtnCode == RTN_CODE_KO) {
        print_line1();
        cout << "\t\t\t HPCSCAN TERMINATED WITH ERROR(S)\n";
        print_line1();
    }
    return RTN_CODE_OK;
}

void print_line1(void) { if (myMpiRank == 0) cout << LINE_T1 << "\n" << flush; }
void print_line2(void) { if (myMpiRank == 0) cout << LINE_T2 << "\n" << flush; }
void print_line3(void) { if (myMpiRank == 0) cout << LINE_T3 << "\n" << flush; }
void print_line4(void) { if (myMpiRank == 0) cout << LINE_T4 << "\n" << flush; }
void print_line5(void) { if (myMpiRank == 0) cout << LINE_T5 << "\n" << flush; }
void print_blank(void) { if (myMpiRank == 0) cout << "\n" << flush; }

void printInfo(Display_type d, const char *t) { if (master(d)) cout << t << "\n" << flush; }
void printInfo(Display_type d, const char *t, Myint v) { if (master(d)) cout << t << "\t" << v << "\n" << flush; }
void printInfo(Display_type d, const char *t, Myint v, Myint w)
{
    if (master(d)) cout << t << "\t" << v << " " << w << "\n" << flush;
}
void printInfo(Display_type d, const char *t, Myint64 v) { if (master(d)) cout << t << "\t" << v << "\n" << flush; }
void printInfo(Display_type d, const char *t, Myfloat32 v) { if (master(d)) cout << t << "\t" << v << "\n" << flush; }
void printInfo(Display_type d, const char *t, Myfloat64 v) { if (master(d)) cout << t << "\t" << v << "\n" << flush; }
void printInfo(Display_type d, const char *t, string v) { if (master(d)) cout << t << "\t" << v << "\n" << flush; }
void printInfo(Display_type d, const char *t, const char *v) { if (master(d)) cout << t << "\t" << v << "\n" << flush; }
void printInfo(Display_type d, string t, const char *v) { if (master(d)) cout << t << "\t" << v << "\n" << flush; }

static void warningBox(const string &body)
{
    cout << "\n" << LINE_T3 << "\n\t\t\t W A R N I N G \n" << body << "\n\n" << LINE_T3 << "\n" << flush;
}
void printWarning(const char *t) { warningBox(t); }
void printWarning(const char *t, Myint v) { warningBox(string(t) + "\t" + to_string(v)); }

static void errorBox(const string &body)
{
    cout << "\n" << LINE_T4 << "\n\t\t\t F A T A L    E R R O R\n\n" << body << "\n\n" << LINE_T4 << "\n" << flush;
}
void printError(const char *t) { errorBox(t); }
void printError(const char *t, Myint v) { errorBox(string(t) + "\t" + to_string(v)); }
void printError(const char *t, Myint v, Myint w) { errorBox(string(t) + "\t" + to_string(v) + "\t" + to_string(w)); }
void printError(const char *t, string v) { errorBox(string(t) + "\t" + v); }
void printError(const char *t, const char *v) { errorBox(string(t) + "\t" + v); }

template <typename T>
static void dbg(Debug_level l, const char *t, const T &v, const char *sep = " ")
{
    if (l <= debug && Config::Instance()->debugLogFile.is_open())
        Config::Instance()->debugLogFile << t << sep << v << "\n" << flush;
}
void printDebug(Debug_level l, const char *t)
{
    if (l <= debug && Config::Instance()->debugLogFile.is_open())
        Config::Instance()->debugLogFile << t << "\n" << flush;
}
void printDebug(Debug_level l, const char *t, string v) { dbg(l, t, v); }
void printDebug(Debug_level l, const char *t, Myint v) { dbg(l, t, v); }
void printDebug(Debug_level l, const char *t, Myint v, Myint w) { dbg(l, t, to_string(v) + " " + to_string(w)); }
void printDebug(Debug_level l, const char *t, Myint64 v) { dbg(l, t, v); }
void printDebug(Debug_level l, const char *t, Myfloat32 v) { dbg(l, t, v); }
void printDebug(Debug_level l, const char *t, Myfloat64 v) { dbg(l, t, v); }
void printDebug(Debug_level l, const char *t, const char *v) { dbg(l, t, v, "\t"); }

} // namespace hpcscan
