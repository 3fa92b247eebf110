// Per-point stencil arithmetic shared by every kernel family.
//
// The operation order is the reference's (src/fdm.h:281-545,
// src/grid.cpp:4476-5112); the library is built with -ffp-contract=off so
// no a*b+c below becomes an fma and every kernel family returns the same
// bits as the reference CPU path.  The contracted variants (FMA = true,
// leapfrog_fma) are the opt-in exception: explicit fused multiply-adds.
#ifndef HSX_STENCIL_POINT_H
#define HSX_STENCIL_POINT_H

#include "hsx_internal.h"

namespace hsx {

// D2 of u at flat index c along stride s, times inv2 (src/fdm.h:281-545)
template <typename TS, typename TC, int R>
__device__ __forceinline__ TC d2_point(const TS *__restrict__ u, int64_t c, int64_t s, const TC *A, TC inv2)
{
    TC acc = A[0] * (TC)u[c];
    if (R == 1) {
        // FD_D2_O2: the unit weight A1 is not multiplied
        acc = acc + (TC)u[c + s];
        acc = acc + (TC)u[c - s];
    } else {
#pragma unroll
        for (int j = 1; j <= R; j++) {
            const TS pair = u[c + j * s] + u[c - j * s];   // Myfloat + Myfloat
            acc = acc + A[j] * (TC)pair;
        }
    }
    return acc * inv2;
}

// a * b + c rounded once (the opt-in contracted kernels only)
template <typename TC>
__device__ __forceinline__ TC fma_tc(TC a, TC b, TC c)
{
    if constexpr (sizeof(TC) == 4)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}

// same arithmetic from values already gathered: v0 centre, vp[j-1]/vm[j-1] at +-j;
// FMA: each A_j * pair + acc in one fused multiply-add (not the reference's bits)
template <typename TS, typename TC, int R, bool FMA = false>
__device__ __forceinline__ TC d2_values(TS v0, const TS *vp, const TS *vm, const TC *A, TC inv2)
{
    TC acc = A[0] * (TC)v0;
    if (R == 1) {
        acc = acc + (TC)vp[0];
        acc = acc + (TC)vm[0];
    } else {
#pragma unroll
        for (int j = 1; j <= R; j++) {
            const TS pair = vp[j - 1] + vm[j - 1];
            acc = FMA ? fma_tc(A[j], (TC)pair, acc) : acc + A[j] * (TC)pair;
        }
    }
    return acc * inv2;
}

// staggered D1 from gathered values: vp[j-1] at +(j-1), vm[j-1] at -j, i.e. the
// derivative at i - 1/2: (A1*(u0 - u-1) + A2*(u+1 - u-2) + ...) * inv, each
// difference in the storage type, left to right (src/fdm.h:29-268); order 2
// has no weight: (u0 - u-1) * inv
template <typename TS, typename TC, int R>
__device__ __forceinline__ TC d1_values(const TS *vp, const TS *vm, const TC *D1, TC inv)
{
    if (R == 1) return (TC)(TS)(vp[0] - vm[0]) * inv;
    TC acc = D1[1] * (TC)(TS)(vp[0] - vm[0]);
#pragma unroll
    for (int j = 2; j <= R; j++) {
        const TS diff = vp[j - 1] - vm[j - 1];
        acc = acc + D1[j] * (TC)diff;
    }
    return acc * inv;
}

template <typename TS, typename TC, int R>
__device__ __forceinline__ TC d1_point(const TS *__restrict__ u, int64_t c, int64_t s, const TC *D1, TC inv)
{
    TS vp[R], vm[R];
#pragma unroll
    for (int j = 1; j <= R; j++) {
        vp[j - 1] = u[c + (j - 1) * s];
        vm[j - 1] = u[c - j * s];
    }
    return d1_values<TS, TC, R>(vp, vm, D1, inv);
}

// Laplacian: N1 (+ N2 (+ N3)) summed left to right in the compute type
template <typename TS, typename TC, int R, int DIM>
__device__ __forceinline__ TC laplacian_point(const TS *__restrict__ u, int64_t ii, int64_t n1, int64_t n12,
                                              const Coefs<TC> &c)
{
    TC lap = d2_point<TS, TC, R>(u, ii, 1, c.A, c.inv2[0]);
    if (DIM >= 2) lap = lap + d2_point<TS, TC, R>(u, ii, n1, c.A, c.inv2[1]);
    if (DIM >= 3) lap = lap + d2_point<TS, TC, R>(u, ii, n12, c.A, c.inv2[2]);
    return lap;
}

// prn = TWO*prc - prn + coef*lap (src/grid.cpp:5017-5020)
template <typename TS, typename TC>
__device__ __forceinline__ TS leapfrog(TS prc, TS prn, TS coef, TC lap)
{
    const TC two = (TC)2.0;
    return (TS)(two * (TC)prc - (TC)prn + (TC)coef * lap);
}

// the same with coef * lap + (2 prc - prn) in one fused multiply-add
template <typename TS, typename TC>
__device__ __forceinline__ TS leapfrog_fma(TS prc, TS prn, TS coef, TC lap)
{
    const TC two = (TC)2.0;
    return (TS)fma_tc((TC)coef, lap, two * (TC)prc - (TC)prn);
}

} // namespace hsx

#endif
