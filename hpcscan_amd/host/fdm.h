// Finite-difference D2 coefficients and helpers (reference: src/fdm.h:276-922).
//
// The weights are the exact centred second-derivative rationals of the
// reference, rounded to Myfloat2 when stored.  The reference expands them in
// per-order macros; here they live in one table indexed by half-width R and
// the per-point evaluation order is implemented once by each backend.
#ifndef HPCSCAN_FDM_H_
#define HPCSCAN_FDM_H_

#include <cmath>
#include <vector>

#include "constant.h"
#include "type_def.h"

namespace hpcscan {

// numerator / denominator of A0..AR for R = 1..8 (src/fdm.h:276-507)
struct FD_Rational { double num, den; };
static const FD_Rational FD_D2_TABLE[8][9] = {
    {{-2, 1}, {1, 1}},
    {{-5, 2}, {4, 3}, {-1, 12}},
    {{-49, 18}, {3, 2}, {-3, 20}, {1, 90}},
    {{-205, 72}, {8, 5}, {-1, 5}, {8, 315}, {-1, 560}},
    {{-5269, 1800}, {5, 3}, {-5, 21}, {5, 126}, {-5, 1008}, {1, 3150}},
    {{-5369, 1800}, {12, 7}, {-15, 56}, {10, 189}, {-1, 112}, {2, 1925}, {-1, 16632}},
    {{-266681, 88200}, {7, 4}, {-7, 24}, {7, 108}, {-7, 528}, {7, 3300}, {-7, 30888},
     {1, 84084}},
    {{-1077749, 352800}, {16, 9}, {-14, 45}, {112, 1485}, {-7, 396}, {112, 32175},
     {-2, 3861}, {16, 315315}, {-1, 411840}},
};

// number of math ops of one D2 along one axis (src/fdm.h:880-921)
static const Myint FD_D2_NOP[8] = {4, 8, 11, 14, 17, 20, 23, 26};

// staggered first-derivative weights A1..AR for R = 1..8 (src/fdm.h:29-268; index 0 unused)
static const FD_Rational FD_D1_TABLE[8][9] = {
    {{0, 1}, {1, 1}},
    {{0, 1}, {9, 8}, {-1, 24}},
    {{0, 1}, {75, 64}, {-25, 384}, {3, 640}},
    {{0, 1}, {1225, 1024}, {-245, 3072}, {49, 5120}, {-5, 7168}},
    {{0, 1}, {19845, 16384}, {-735, 8192}, {567, 40960}, {-405, 229376}, {35, 294912}},
    {{0, 1}, {160083, 131072}, {-12705, 131072}, {22869, 1310720}, {-5445, 1835008}, {847, 2359296},
     {-63, 2883584}},
    {{0, 1}, {1288287, 1048576}, {-429429, 4194304}, {429429, 20971520}, {-61347, 14680064},
     {13013, 18874368}, {-3549, 46137344}, {231, 54525952}},
    {{0, 1}, {41409225, 33554432}, {-3578575, 33554432}, {3864861, 167772160}, {-1254825, 234881024},
     {325325, 301989888}, {-61425, 369098752}, {7425, 436207616}, {-143, 167772160}},
};

// number of math ops of one D1 along one axis (src/fdm.h:35-235, getFD_D1nMathOp :836-875)
static const Myint FD_D1_NOP[8] = {2, 6, 9, 12, 15, 18, 21, 24};

static inline bool validFdOrder(Myint fdOrder)
{
    return fdOrder >= 2 && fdOrder <= 16 && fdOrder % 2 == 0;
}

// A_j in double, j = 0..R (the value the reference writes as a C double literal)
static inline double getFD_D2coefDouble(Myint fdOrder, Myint j)
{
    const FD_Rational &q = FD_D2_TABLE[fdOrder / 2 - 1][j];
    return q.num / q.den;
}

// full stencil AR..A1 A0 A1..AR as Myfloat2 (src/fdm.h:655-769)
static inline vector<Myfloat2> getFD_D2coefVector(Myint fdOrder)
{
    vector<Myfloat2> v;
    if (!validFdOrder(fdOrder)) return v;
    const Myint R = fdOrder / 2;
    for (Myint j = R; j >= 1; j--) v.push_back((Myfloat2)getFD_D2coefDouble(fdOrder, j));
    v.push_back((Myfloat2)getFD_D2coefDouble(fdOrder, 0));
    for (Myint j = 1; j <= R; j++) v.push_back((Myfloat2)getFD_D2coefDouble(fdOrder, j));
    return v;
}

// sum of |coef| used by the CFL (src/fdm.h:771-787)
static inline Myfloat64 getSumAbsFD_D2Coef(Myint fdOrder)
{
    Myfloat64 s = 0.0;
    for (Myfloat2 c : getFD_D2coefVector(fdOrder)) s += fabs((Myfloat64)c);
    return s;
}

// sum of coef, expected to be 0 (src/fdm.h:808-822)
static inline Myfloat64 getSumFD_D2Coef(Myint fdOrder)
{
    Myfloat64 s = 0.0;
    for (Myfloat2 c : getFD_D2coefVector(fdOrder)) s += c;
    return s;
}

static inline Myint getFD_haloWidth(Myint fdOrder)
{
    return validFdOrder(fdOrder) ? fdOrder / 2 : UNSPECIFIED;
}

static inline Myint getFD_D1nMathOp(Myint fdOrder)
{
    return validFdOrder(fdOrder) ? FD_D1_NOP[fdOrder / 2 - 1] : UNSPECIFIED;
}

static inline Myint getFD_D2nMathOp(Myint fdOrder)
{
    return validFdOrder(fdOrder) ? FD_D2_NOP[fdOrder / 2 - 1] : UNSPECIFIED;
}

} // namespace hpcscan

#endif
