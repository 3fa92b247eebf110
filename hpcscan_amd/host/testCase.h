// Test cases of the stencil path (reference API: src/testCase.h,
// src/testCase_FD_D2.h, src/testCase_Propa.h).
#ifndef HPCSCAN_TESTCASE_H_
#define HPCSCAN_TESTCASE_H_

#include <fstream>
#include <memory>
#include <string>

#include "grid.h"
#include "hardware.h"
#include "type_def.h"

namespace hpcscan {

class TestCase
{
public:
    virtual ~TestCase(void) {}
    virtual Rtn_code run(void) = 0;

protected:
    virtual Rtn_code initialize(void);
    virtual void finalize(void);

    Rtn_code checkGridL1Err(Point_type, const Grid &, const Grid &, Myfloat);
    Rtn_code checkAllProcGridL1Err(Point_type, const Grid &, const Grid &, Myfloat);
    Rtn_code checkGridMaxErr(Point_type, const Grid &, const Grid &, Myfloat);
    Rtn_code checkAllProcGridMaxErr(Point_type, const Grid &, const Grid &, Myfloat);
    Rtn_code checkFloatDiff(Myfloat, Myfloat, Myfloat);
    Rtn_code checkIntegerDiff(Myint, Myint);
    Rtn_code checkBoolDiff(bool, bool);
    Myfloat relErr(Myfloat float1, Myfloat float2);

    ofstream perfLogFile;
    double testCaseStart, testCaseEnd;
    string testCaseName;
    string testCaseVersion;
    shared_ptr<Hardware> hw;
};

class TestCase_Modeling : protected TestCase
{
public:
    TestCase_Modeling(void);
    virtual Rtn_code run(void);
};

class TestCase_FD_D1 : protected TestCase
{
public:
    TestCase_FD_D1(void);
    virtual Rtn_code run(void);
};

class TestCase_FD_D2 : protected TestCase
{
public:
    TestCase_FD_D2(void);
    virtual Rtn_code run(void);
};

class TestCase_Propa : protected TestCase
{
public:
    TestCase_Propa(void);
    virtual Rtn_code run(void);
};

class TestCase_Grid : protected TestCase
{
public:
    TestCase_Grid(void);
    virtual Rtn_code run(void);
};

} // namespace hpcscan

#endif
