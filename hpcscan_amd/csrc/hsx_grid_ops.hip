// Grid operations of the C ABI other than the stencils: fills, array
// operations, the free-surface boundary, the leapfrog update and the
// double-precision reductions.
//
// These are HBM-streaming passes: one thread per value, rows of the index box
// mapped to waves so that n1 (contiguous) runs across the 64 lanes.
#include <algorithm>
#include <cfloat>
#include <mutex>

#include "hsx_internal.h"

namespace hsx {

constexpr int ROW_THREADS = 64;   // one wave per row
constexpr int ROWS_PER_BLOCK = 4; // 256 threads

static dim3 row_grid(const Box &b)
{
    const int64_t rows = b.n2() * b.n3();
    int64_t nb = (rows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
    return dim3((unsigned)(nb > 65535 * 16 ? 65535 * 16 : nb));
}

// Visit every point of box b: f(flat index, i1, i2, i3)
template <typename F>
__device__ inline void for_box(const Box &b, int64_t n1, int64_t n12, F &&f)
{
    const int64_t N1 = b.n1(), N2 = b.n2(), rows = N2 * b.n3();
    for (int64_t row = (int64_t)blockIdx.x * ROWS_PER_BLOCK + threadIdx.y; row < rows;
         row += (int64_t)gridDim.x * ROWS_PER_BLOCK) {
        const int64_t i2 = b.s2 + row % N2, i3 = b.s3 + row / N2;
        const int64_t base = i2 * n1 + i3 * n12;
        for (int64_t k = threadIdx.x; k < N1; k += ROW_THREADS) {
            const int64_t i1 = b.s1 + k;
            f(base + i1, i1, i2, i3);
        }
    }
}

// Visit box b row by row in 16-byte groups of VW points (flat indices aligned
// to VW, so the buffer must be 16-byte aligned): group(ii, i1, i2, i3) for each
// whole group inside the row's [s1, e1] span (points ii .. ii + VW - 1, i1 the
// first one's n1 index), point(ii, i1, i2, i3) for the unaligned head and tail
// points of the row.  One wave per row, as for_box.
template <int VW, typename FP, typename FG>
__device__ inline void for_box_groups(const Box &b, int64_t n1, int64_t n12, FP &&point, FG &&group)
{
    const int64_t N2 = b.n2(), rows = N2 * b.n3();
    const int lane = threadIdx.x;
    for (int64_t row = (int64_t)blockIdx.x * ROWS_PER_BLOCK + threadIdx.y; row < rows;
         row += (int64_t)gridDim.x * ROWS_PER_BLOCK) {
        const int64_t i2 = b.s2 + row % N2, i3 = b.s3 + row / N2;
        const int64_t base = i2 * n1 + i3 * n12;
        const int64_t a0 = base + b.s1, a1 = base + b.e1 + 1;       // [a0, a1) flat
        const int64_t g0 = (a0 + VW - 1) / VW, g1 = a1 / VW;          // whole groups
        if (g0 >= g1) {
            for (int64_t i = a0 + lane; i < a1; i += ROW_THREADS) point(i, i - base, i2, i3);
            continue;
        }
        const int64_t h = g0 * VW - a0, t = a1 - g1 * VW;
        if (lane < h) point(a0 + lane, a0 + lane - base, i2, i3);
        if (lane < t) point(g1 * VW + lane, g1 * VW + lane - base, i2, i3);
        for (int64_t gi = g0 + lane; gi < g1; gi += ROW_THREADS) group(gi * VW, gi * VW - base, i2, i3);
    }
}

// ---- fills (src/grid.cpp:1650-1826) -----------------------------------------
// VEC: 16-byte stores (u 16-byte aligned)
template <typename TS, bool VEC>
__global__ void __launch_bounds__(256) k_fill_const(TS *u, Box b, int64_t n1, int64_t n12, TS val)
{
    if constexpr (!VEC) {
        for_box(b, n1, n12, [&](int64_t ii, int64_t, int64_t, int64_t) { u[ii] = val; });
    } else {
        constexpr int VW = 16 / sizeof(TS);
        using NV = TS __attribute__((ext_vector_type(VW)));
        const NV vv = (NV)val;
        for_box_groups<VW>(
            b, n1, n12, [&](int64_t ii, int64_t, int64_t, int64_t) { u[ii] = val; },
            [&](int64_t ii, int64_t, int64_t, int64_t) { *reinterpret_cast<NV *>(u + ii) = vv; });
    }
}

// amp * v1 * v2 * v3 left to right in double, then rounded to the storage type
template <typename TS, bool VEC>
__global__ void __launch_bounds__(256) k_fill_sep(TS *u, Box b, int64_t n1, int64_t n12, const double *t1,
                                                  const double *t2, const double *t3, double amp)
{
    auto one = [&](int64_t i1, int64_t i2, int64_t i3) -> TS {
        return (TS)(amp * t1[i1 - b.s1] * t2[i2 - b.s2] * t3[i3 - b.s3]);
    };
    if constexpr (!VEC) {
        for_box(b, n1, n12, [&](int64_t ii, int64_t i1, int64_t i2, int64_t i3) { u[ii] = one(i1, i2, i3); });
    } else {
        constexpr int VW = 16 / sizeof(TS);
        using NV = TS __attribute__((ext_vector_type(VW)));
        for_box_groups<VW>(
            b, n1, n12, [&](int64_t ii, int64_t i1, int64_t i2, int64_t i3) { u[ii] = one(i1, i2, i3); },
            [&](int64_t ii, int64_t i1, int64_t i2, int64_t i3) {
                NV r;
#pragma unroll
                for (int e = 0; e < VW; e++) r[e] = one(i1 + e, i2, i3);
                *reinterpret_cast<NV *>(u + ii) = r;
            });
    }
}

// ---- array operations (src/grid.cpp:1830-1920) --------------------------------
template <typename TS>
__device__ __forceinline__ void array_elem(int op, TS *w, const TS *u, const TS *v, TS val, int64_t i)
{
    switch (op) {
    case HSX_ARRAY_FILL: w[i] = val; break;
    case HSX_ARRAY_COPY: w[i] = u[i]; break;
    case HSX_ARRAY_ADD: w[i] = u[i] + v[i]; break;
    case HSX_ARRAY_MULTIPLY: w[i] = u[i] * v[i]; break;
    default: w[i] = w[i] + u[i]; break;
    }
}

// 16-byte vectors per lane when every pointer is 16-byte aligned (the usual
// case: whole grids from hsx_malloc), element loop for the tail
template <typename TS>
__global__ void __launch_bounds__(256) k_array16(int op, TS *w, const TS *u, const TS *v, TS val, int64_t n)
{
    constexpr int VW = 16 / sizeof(TS);
    struct alignas(16) V { TS x[VW]; };
    const int64_t nv = n / VW;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        V r, a, b;
        if (op != HSX_ARRAY_FILL) a = reinterpret_cast<const V *>(u)[i];
        if (op == HSX_ARRAY_ADD || op == HSX_ARRAY_MULTIPLY) b = reinterpret_cast<const V *>(v)[i];
        if (op == HSX_ARRAY_ADDTO) r = reinterpret_cast<const V *>(w)[i];
#pragma unroll
        for (int k = 0; k < VW; k++) {
            switch (op) {
            case HSX_ARRAY_FILL: r.x[k] = val; break;
            case HSX_ARRAY_COPY: r.x[k] = a.x[k]; break;
            case HSX_ARRAY_ADD: r.x[k] = a.x[k] + b.x[k]; break;
            case HSX_ARRAY_MULTIPLY: r.x[k] = a.x[k] * b.x[k]; break;
            default: r.x[k] = r.x[k] + a.x[k]; break;
            }
        }
        reinterpret_cast<V *>(w)[i] = r;
    }
    for (int64_t i = nv * VW + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        array_elem(op, w, u, v, val, i);
}

template <typename TS>
__global__ void __launch_bounds__(256) k_array(int op, TS *w, const TS *u, const TS *v, TS val, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        array_elem(op, w, u, v, val, i);
}

// ---- free surface, one face (src/grid.cpp:3065-3230) ---------------------------
// For every row of the two other axes: u[inner] = 0 and u[i] = -u[2*inner - i]
// for i in the halo range.  Sources are inner points at distance >= 1 from the
// zeroed plane and no halo point is a source, so one face is a parallel map;
// faces run one after the other, in the reference order.
template <typename TS>
__global__ void __launch_bounds__(256) k_free_surface(TS *u, Face f)
{
    const int64_t H = f.he - f.hs + 2;           // zero + halo points per row
    const int64_t NB = f.be - f.bs + 1, NO = f.oe - f.os + 1;
    const int64_t total = H * NB * NO;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t k, ib, io;
        if (f.axis == 0) {                        // face axis contiguous: k fastest
            k = t % H;
            ib = (t / H) % NB;
            io = t / (H * NB);
        } else {                                  // i1 contiguous: i1 fastest
            ib = t % NB;
            k = (t / NB) % H;
            io = t / (NB * H);
        }
        const int64_t base = (f.os + io) * f.so + (f.bs + ib) * f.sb;
        if (k == 0) {
            u[base + f.inner * f.sa] = (TS)0.0;
        } else {
            const int64_t i = f.hs + k - 1;
            u[base + i * f.sa] = -u[base + (2 * f.inner - i) * f.sa];
        }
    }
}

// ---- free surface, all faces in two launches -------------------------------------
// The six sequential phases of the reference (I1HALO1, I1HALO2, I2HALO1, ... each
// zeroing its boundary plane then mirroring) have a closed form when every
// active axis has at least R+2 inner points, so that a mirror source never
// reaches the far boundary plane or a halo:
//  * halo point p of face (a, side), mirror m:  u[p] = -0 if some face of an
//    axis a' < a in the mask has its boundary plane through p (p and m share
//    that coordinate; the plane was zeroed before phase a), else -u_old[m];
//  * then every boundary plane of the mask is zero.
// A halo point is written once and mirrors read only inner points that the
// first launch does not write, so the first launch has no ordering inside it,
// and the second launch (the planes) only writes inner points.
// The face loops index with IT = uint32_t when the face fits (always, below
// 65536^2 points per face): the 64-bit divisions otherwise dominate these
// kernels, which move only a few bytes per index.
// bx / nbx: this block's index and the number of blocks sharing the face (a
// launch may give other blocks other work).
template <typename TS, typename IT>
__device__ __forceinline__ void bc_mirror_loop(TS *u, const FaceSet &fs, const Face &f, IT H, IT NB, IT total,
                                               uint32_t bx, uint32_t nbx)
{
    const int a = f.axis;
    const int oa = a == 2 ? 1 : 2, ob = a == 0 ? 1 : 0;
    for (IT t = (IT)bx * blockDim.x + threadIdx.x; t < total; t += (IT)nbx * blockDim.x) {
        IT k, ib, io;
        if (a == 0) {
            const IT r = t / H;
            k = t - r * H;
            io = r / NB;
            ib = r - io * NB;
        } else {
            const IT r = t / NB;
            ib = t - r * NB;
            io = r / H;
            k = r - io * H;
        }
        int64_t c[3];
        c[a] = f.hs + (int64_t)k;
        c[oa] = f.os + (int64_t)io;
        c[ob] = f.bs + (int64_t)ib;
        bool zero = false;
        for (int ap = 0; ap < a; ap++)
            for (int sp = 0; sp < 2; sp++)
                if ((fs.zmask >> (2 * ap + sp)) & 1) zero |= c[ap] == fs.zidx[ap][sp];
        const int64_t base = c[oa] * f.so + c[ob] * f.sb;
        const int64_t i = c[a];
        u[base + i * f.sa] = zero ? -(TS)0.0 : -u[base + (2 * f.inner - i) * f.sa];
    }
}

// n2 / n3 faces with 16-byte groups along n1 (fs.vec1): one thread per group of a
// halo row; the elements of an edge group outside the face's n1 range are left
// alone (scalar stores there), the -0 rule is applied per element for the n1
// boundary planes and per row for the n2 ones. 1000^3 FP32 order 8, mirror and
// zero: n2 faces 0.020 ms, n3 faces 0.020 ms, against 0.033 / 0.035 ms with one
// point per thread.
template <typename TS>
__device__ __forceinline__ void bc_mirror_vec(TS *u, const FaceSet &fs, const Face &f, uint32_t bx, uint32_t nbx)
{
    constexpr int VW = 16 / sizeof(TS);
    using NV = TS __attribute__((ext_vector_type(VW)));
    const int a = f.axis;                          // 1 or 2; n1 is the inner-other axis (stride 1)
    const int64_t g0 = f.bs / VW;
    const uint32_t NG = (uint32_t)(f.be / VW - g0 + 1), H = (uint32_t)(f.he - f.hs + 1);
    const uint32_t total = NG * H * (uint32_t)(f.oe - f.os + 1);
    const bool z1lo = fs.zmask & 1, z1hi = fs.zmask & 2;
    for (uint32_t t = bx * blockDim.x + threadIdx.x; t < total; t += nbx * blockDim.x) {
        const uint32_t r = t / NG, g = t - r * NG;
        const uint32_t io = r / H, k = r - io * H;
        const int64_t i = f.hs + (int64_t)k;       // halo index along the face axis
        const int64_t o = f.os + (int64_t)io;      // index along the outer axis
        // -0 for the whole row: its n2 coordinate on a folded n2 boundary plane (n3 faces)
        bool zrow = false;
        if (a == 2)
            for (int sp = 0; sp < 2; sp++)
                if ((fs.zmask >> (2 + sp)) & 1) zrow |= o == fs.zidx[1][sp];
        const int64_t i1 = (g0 + (int64_t)g) * VW;
        const int64_t row = o * f.so + i1;
        const NV m = *reinterpret_cast<const NV *>(u + row + (2 * f.inner - i) * f.sa);
        NV w;
#pragma unroll
        for (int e = 0; e < VW; e++) {
            const bool z = zrow || (z1lo && i1 + e == fs.zidx[0][0]) || (z1hi && i1 + e == fs.zidx[0][1]);
            w[e] = z ? -(TS)0.0 : -m[e];
        }
        TS *p = u + row + i * f.sa;
        if (i1 >= f.bs && i1 + VW - 1 <= f.be) {
            *reinterpret_cast<NV *>(p) = w;
        } else {
#pragma unroll
            for (int e = 0; e < VW; e++)
                if (i1 + e >= f.bs && i1 + e <= f.be) p[e] = w[e];
        }
    }
}

template <typename TS>
__device__ __forceinline__ void bc_mirror_face(TS *u, const FaceSet &fs, int face, uint32_t bx, uint32_t nbx)
{
    const Face &f = fs.f[face];
    if (fs.vec1 && f.axis != 0) {
        bc_mirror_vec(u, fs, f, bx, nbx);
        return;
    }
    const int64_t H = f.he - f.hs + 1;
    const int64_t NB = f.be - f.bs + 1, NO = f.oe - f.os + 1;
    const int64_t total = H * NB * NO;
    if (total <= (int64_t)UINT32_MAX - (int64_t)nbx * blockDim.x)
        bc_mirror_loop<TS, uint32_t>(u, fs, f, (uint32_t)H, (uint32_t)NB, (uint32_t)total, bx, nbx);
    else
        bc_mirror_loop<TS, uint64_t>(u, fs, f, (uint64_t)H, (uint64_t)NB, (uint64_t)total, bx, nbx);
}

template <typename TS, typename IT>
__device__ __forceinline__ void bc_zero_loop(TS *u, const Face &f, IT NB, IT total)
{
    for (IT t = (IT)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (IT)gridDim.x * blockDim.x) {
        const IT io = t / NB, ib = t - io * NB;
        u[(f.os + (int64_t)io) * f.so + (f.bs + (int64_t)ib) * f.sb + f.inner * f.sa] = (TS)0.0;
    }
}

template <typename TS>
__device__ __forceinline__ void bc_zero_face(TS *u, const FaceSet &fs, int face)
{
    const Face &f = fs.f[face];
    const int64_t NB = f.be - f.bs + 1, NO = f.oe - f.os + 1;
    const int64_t total = NB * NO;
    if (total <= (int64_t)UINT32_MAX - (int64_t)gridDim.x * blockDim.x)
        bc_zero_loop<TS, uint32_t>(u, f, (uint32_t)NB, (uint32_t)total);
    else
        bc_zero_loop<TS, uint64_t>(u, f, (uint64_t)NB, (uint64_t)total);
}

// The n1 faces with 16-byte accesses, one thread per row end, when R is a
// multiple of the vector width and the halo groups are aligned: the halo
// groups get -u of their mirrors and the boundary group its zero in the same
// pass. Closed form: an n1 halo point never lies on a boundary plane of an
// earlier axis, and no mirror of another face reads the n1 boundary plane
// (those halo points are -0 by the same rule), so the zero may be written
// here; the mirrors read old values of the other faces' boundary planes,
// which k_bc_zero clears after this launch. The rest of each written group
// is stored back as read. 1000^3 FP32 order 8: 0.075 ms for both n1 faces,
// against 0.111 ms through k_bc_mirror + k_bc_zero (a second pass over the
// same 64-byte sectors of 2 x 10^6 rows).
template <typename TS, int NH>
__device__ __forceinline__ void bc_n1_rows(TS *u, const Face &lo, const Face &hi, int sides, uint32_t t)
{
    constexpr int VW = 16 / sizeof(TS), R = NH * VW;
    using NV = TS __attribute__((ext_vector_type(VW)));
    const Face &f = (sides & 1) ? lo : hi;           // both faces span the same rows
    const uint32_t NB = (uint32_t)(f.be - f.bs + 1), rows = NB * (uint32_t)(f.oe - f.os + 1);
    const uint32_t ns = (uint32_t)(sides & 1) + (uint32_t)(sides >> 1);
    if (t >= rows * ns) return;
    const uint32_t row = ns == 2 ? t >> 1 : t;
    const int side = ns == 2 ? (int)(t & 1) : ((sides & 1) ? 0 : 1);
    const uint32_t io = row / NB, ib = row - io * NB;
    NV *r = reinterpret_cast<NV *>(u + (f.os + (int64_t)io) * f.so + (f.bs + (int64_t)ib) * f.sb);
    NV w[NH + 1];
    if (side == 0) {
        // boundary b = lo.inner (first of its group), halo b-R .. b-1, mirrors b+1 .. b+R
        const int64_t g = lo.inner / VW;
#pragma unroll
        for (int j = 0; j <= NH; j++) w[j] = r[g + j];
#pragma unroll
        for (int h = 0; h < NH; h++) {
            NV o;
#pragma unroll
            for (int e = 0; e < VW; e++) {
                const int k = R - h * VW - e;            // halo point b - k, mirror b + k
                o[e] = -w[k / VW][k % VW];
            }
            r[g - NH + h] = o;
        }
        w[0][0] = (TS)0.0;
        r[g] = w[0];
    } else {
        // boundary b = hi.inner (last of its group), halo b+1 .. b+R, mirrors b-1 .. b-R
        const int64_t g = hi.inner / VW;
#pragma unroll
        for (int j = 0; j <= NH; j++) w[j] = r[g - NH + j];
#pragma unroll
        for (int h = 0; h < NH; h++) {
            NV o;
#pragma unroll
            for (int e = 0; e < VW; e++) {
                const int k = 1 + h * VW + e;            // halo point b + k, mirror b - k
                const int p = (NH + 1) * VW - 1 - k;     // position of b - k in w
                o[e] = -w[p / VW][p % VW];
            }
            r[g + 1 + h] = o;
        }
        w[NH][VW - 1] = (TS)0.0;
        r[g] = w[NH];
    }
}

template <typename TS, int NH>
__global__ void __launch_bounds__(256) k_bc_n1(TS *u, Face lo, Face hi, int sides)
{
    bc_n1_rows<TS, NH>(u, lo, hi, sides, blockIdx.x * 256u + threadIdx.x);
}

// The n1 faces (bc_n1_rows) and the mirrors of the other faces in one launch:
// they write disjoint halo points and neither reads what the other writes
// (bc_n1_rows rewrites the n1 boundary group with its own values and a zero
// that the other faces' mirrors take as -0 without reading it), so the two
// need no order; blocks [0, nb_n1) take the n1 rows, the next nb1 * rest.n the
// mirrors (nb1 blocks per face).  The boundary planes follow in k_bc_zero.
template <typename TS, int NH>
__global__ void __launch_bounds__(256) k_bc_n1_mirror(TS *u, Face lo, Face hi, int sides, FaceSet rest,
                                                     uint32_t nb_n1, uint32_t nb1)
{
    if (blockIdx.x < nb_n1) {
        bc_n1_rows<TS, NH>(u, lo, hi, sides, blockIdx.x * 256u + threadIdx.x);
        return;
    }
    const uint32_t m = blockIdx.x - nb_n1;
    bc_mirror_face(u, rest, (int)(m / nb1), m % nb1, nb1);
}

// the same for many grids: blockIdx.y = grid; blocks past nb_n1 take the
// mirrors of faces (blockIdx.x - nb_n1) / nb1 < fs[grid].n
template <typename TS, int NH>
__global__ void __launch_bounds__(256) k_bc_n1_mirror_batch(TS *const *u, const N1Faces *nf, const FaceSet *fs,
                                                           uint32_t nb_n1, uint32_t nb1)
{
    const int s = blockIdx.y;
    if (blockIdx.x < nb_n1) {
        const N1Faces &x = nf[s];
        if (x.sides) bc_n1_rows<TS, NH>(u[s], x.lo, x.hi, x.sides, blockIdx.x * 256u + threadIdx.x);
        return;
    }
    const uint32_t m = blockIdx.x - nb_n1;
    const int face = (int)(m / nb1);
    if (face < fs[s].n) bc_mirror_face(u[s], fs[s], face, m % nb1, nb1);
}

int n1_groups_ok(const hsx_grid *g, const void *u, const FaceSet &fs)
{
    const int VW = g->storage == HSX_FP64 ? 2 : 4;
    if (g->n1 % VW || (uintptr_t)u % 16) return 0;
    for (int k = 0; k < fs.n; k++) {
        const Face &f = fs.f[k];
        if (f.axis == 0) continue;
        const int64_t n = (f.be / VW - f.bs / VW + 1) * (f.he - f.hs + 1) * (f.oe - f.os + 1);
        if (n >= ((int64_t)1 << 31)) return 0;
    }
    return 1;
}

int n1_vector_sides(const hsx_grid *g, const void *u, const FaceSet &fs, Face *lo, Face *hi)
{
    const int VW = g->storage == HSX_FP64 ? 2 : 4;
    if (g->n1 % VW || (uintptr_t)u % 16) return 0;
    int sides = 0;
    for (int k = 0; k < fs.n; k++) {
        const Face &f = fs.f[k];
        if (f.axis != 0) continue;
        const int64_t R = f.he - f.hs + 1;
        if (R % VW || R / VW < 1 || R / VW > 4 || f.hs % VW) return 0;
        const int64_t rows = (f.be - f.bs + 1) * (f.oe - f.os + 1);
        if (rows <= 0 || 2 * rows >= ((int64_t)1 << 31)) return 0;
        if (f.inner > f.he) {
            *lo = f;
            sides |= 1;
        } else {
            *hi = f;
            sides |= 2;
        }
    }
    if (sides == 3 && lo->he - lo->hs != hi->he - hi->hs) return 0;
    return sides;
}

// one grid: blockIdx.y = face
template <typename TS>
__global__ void __launch_bounds__(256) k_bc_mirror(TS *u, FaceSet fs)
{
    bc_mirror_face(u, fs, blockIdx.y, blockIdx.x, gridDim.x);
}

template <typename TS>
__global__ void __launch_bounds__(256) k_bc_zero(TS *u, FaceSet fs)
{
    bc_zero_face(u, fs, blockIdx.y);
}

// many grids: blockIdx.y = 6 * grid + face
template <typename TS>
__global__ void __launch_bounds__(256) k_bc_mirror_batch(TS *const *u, const FaceSet *fs)
{
    const int s = blockIdx.y / 6, face = blockIdx.y % 6;
    if (face < fs[s].n) bc_mirror_face(u[s], fs[s], face, blockIdx.x, gridDim.x);
}

template <typename TS>
__global__ void __launch_bounds__(256) k_bc_zero_batch(TS *const *u, const FaceSet *fs)
{
    const int s = blockIdx.y / 6, face = blockIdx.y % 6;
    if (face < fs[s].n) bc_zero_face(u[s], fs[s], face);
}

int make_faceset(const hsx_grid *g, int face_mask, FaceSet *fs, bool *closed_form, int64_t *most,
                 int64_t *most_plane)
{
    const int64_t st[3] = {1, g->n1, g->n1 * g->n2};
    fs->n = 0;
    fs->zmask = 0;
    fs->vec1 = 0;
    for (int a = 0; a < 3; a++) fs->zidx[a][0] = fs->zidx[a][1] = -1;
    *closed_form = true;
    *most = *most_plane = 0;
    for (int axis = 0; axis < g->dim; axis++) {
        for (int side = 0; side < 2; side++) {
            if (!(face_mask & (1 << (2 * axis + side)))) continue;
            const int64_t *hr = side == 0 ? g->halo1 : g->halo2;
            Face f;
            f.axis = axis;
            f.hs = hr[2 * axis];
            f.he = hr[2 * axis + 1];
            f.inner = side == 0 ? f.he + 1 : f.hs - 1;
            // other axes: outer = the higher one, inner-other = the lower one
            const int oa = axis == 2 ? 1 : 2, ob = axis == 0 ? 1 : 0;
            f.os = g->inner[2 * oa];
            f.oe = g->inner[2 * oa + 1];
            f.bs = g->inner[2 * ob];
            f.be = g->inner[2 * ob + 1];
            f.so = st[oa];
            f.sb = st[ob];
            f.sa = st[axis];
            // mirror sources must stay inside the allocation
            const int64_t n_ax = axis == 0 ? g->n1 : axis == 1 ? g->n2 : g->n3;
            if (2 * f.inner - f.hs >= n_ax || 2 * f.inner - f.he < 0 || f.inner < 0 || f.inner >= n_ax)
                return fail(HSX_ERR_INVALID, "free surface: mirror of axis %d face %d outside the grid", axis + 1,
                            side + 1);
            fs->f[fs->n++] = f;
            fs->zidx[axis][side] = f.inner;
            fs->zmask |= 1 << (2 * axis + side);
            const int64_t plane = (f.be - f.bs + 1) * (f.oe - f.os + 1);
            *most = std::max(*most, plane * (f.he - f.hs + 1));
            *most_plane = std::max(*most_plane, plane);
        }
        // closed form: the mirrors of each face stay strictly between the two
        // boundary planes of the axis (both computed as the sequential phases do);
        // an axis with no face in the mask adds no phase and no condition
        const int64_t z0 = g->halo1[2 * axis + 1] + 1, z1 = g->halo2[2 * axis] - 1;
        const int64_t h1 = g->halo1[2 * axis + 1] - g->halo1[2 * axis] + 1;
        const int64_t h2 = g->halo2[2 * axis + 1] - g->halo2[2 * axis] + 1;
        if (((fs->zmask >> (2 * axis)) & 3) && z1 - z0 <= (h1 > h2 ? h1 : h2)) *closed_form = false;
    }
    return HSX_OK;
}

int bc_batch_launch(int storage, int nshots, void *const *d_fields, const FaceSet *d_sets, int64_t most,
                    int64_t most_plane, hipStream_t s)
{
    const unsigned nb1 = (unsigned)std::max<int64_t>(1, std::min<int64_t>((most + 255) / 256, 4096));
    const unsigned nb2 = (unsigned)std::max<int64_t>(1, std::min<int64_t>((most_plane + 255) / 256, 4096));
    if (storage == HSX_FP64) {
        hipLaunchKernelGGL(k_bc_mirror_batch<double>, dim3(nb1, 6 * nshots), dim3(256), 0, s,
                           (double *const *)d_fields, d_sets);
        HSX_LAUNCHED("k_bc_mirror_batch");
        hipLaunchKernelGGL(k_bc_zero_batch<double>, dim3(nb2, 6 * nshots), dim3(256), 0, s,
                           (double *const *)d_fields, d_sets);
    } else {
        hipLaunchKernelGGL(k_bc_mirror_batch<float>, dim3(nb1, 6 * nshots), dim3(256), 0, s,
                           (float *const *)d_fields, d_sets);
        HSX_LAUNCHED("k_bc_mirror_batch");
        hipLaunchKernelGGL(k_bc_zero_batch<float>, dim3(nb2, 6 * nshots), dim3(256), 0, s,
                           (float *const *)d_fields, d_sets);
    }
    HSX_LAUNCHED("k_bc_zero_batch");
    return HSX_OK;
}

int bc_batch_fs_launch(int storage, int nshots, void *const *d_fields, const N1Faces *d_n1, int nh,
                       int64_t most_rows, const FaceSet *d_sets, int64_t most, int64_t most_plane, hipStream_t s)
{
    if (nh < 1 || nh > 4) return fail(HSX_ERR_INVALID, "batch free surface: %d n1 halo groups", nh);
    const uint32_t nbn = (uint32_t)std::max<int64_t>(1, (2 * most_rows + 255) / 256);
    const uint32_t nb1 = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((most + 255) / 256, 4096));
    const uint32_t nb2 = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((most_plane + 255) / 256, 4096));
    const dim3 both(nbn + 6 * nb1, (unsigned)nshots);
    auto launch = [&](auto ts) {
        using TS = decltype(ts);
        TS *const *u = (TS *const *)d_fields;
        switch (nh) {
        case 1: hipLaunchKernelGGL((k_bc_n1_mirror_batch<TS, 1>), both, dim3(256), 0, s, u, d_n1, d_sets, nbn, nb1); break;
        case 2: hipLaunchKernelGGL((k_bc_n1_mirror_batch<TS, 2>), both, dim3(256), 0, s, u, d_n1, d_sets, nbn, nb1); break;
        case 3: hipLaunchKernelGGL((k_bc_n1_mirror_batch<TS, 3>), both, dim3(256), 0, s, u, d_n1, d_sets, nbn, nb1); break;
        default: hipLaunchKernelGGL((k_bc_n1_mirror_batch<TS, 4>), both, dim3(256), 0, s, u, d_n1, d_sets, nbn, nb1); break;
        }
        HSX_LAUNCHED("k_bc_n1_mirror_batch");
        hipLaunchKernelGGL(k_bc_zero_batch<TS>, dim3(nb2, 6 * nshots), dim3(256), 0, s, u, d_sets);
        HSX_LAUNCHED("k_bc_zero_batch");
    };
    if (storage == HSX_FP64)
        launch(0.0);
    else
        launch(0.0f);
    return HSX_OK;
}

// ---- leapfrog update of the split propagator (src/grid.cpp:4579-4610) -----------
// prn = TWO*prc - prn + coef*lapla; coef*lapla is a product of two storage values
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_update_pressure(TS *prn, const TS *prc, const TS *coef, const TS *lapla,
                                                         Box b, int64_t n1, int64_t n12)
{
    const TC two = (TC)2.0;
    for_box(b, n1, n12, [&](int64_t ii, int64_t, int64_t, int64_t) {
        const TS cl = coef[ii] * lapla[ii];
        prn[ii] = (TS)(two * (TC)prc[ii] - (TC)prn[ii] + (TC)cl);
    });
}

// The same update with 16-byte lanes (n1 a multiple of the vector width, 16-byte
// aligned grids): one thread per aligned group of a row's [s1, e1] span, 256
// threads per block over the (row, group) space in address order; points of an
// edge group outside the box are written back with the prn value they were
// read with. Every grid is read once and prn written once, so the accesses are
// nontemporal. At 1000^3 FP32 this measured 3.64-3.70 ms against 3.95 ms for
// the row kernel above (wave-per-row 16-byte lanes: 3.86-3.91 ms cached,
// 4.16-4.20 ms nontemporal).
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_update_pressure16(TS *prn, const TS *prc, const TS *coef, const TS *lapla,
                                                           Box b, int64_t n1, int64_t n12, int64_t ng, int64_t total)
{
    constexpr int VW = 16 / sizeof(TS);
    using NV = TS __attribute__((ext_vector_type(VW)));
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int64_t row = t / ng, k = t - row * ng, N2 = b.n2();
    const int64_t i2 = b.s2 + row % N2, i3 = b.s3 + row / N2;
    const int64_t i1 = (b.s1 / VW + k) * VW, ii = i2 * n1 + i3 * n12 + i1;
    const TC two = (TC)2.0;
    const NV p = __builtin_nontemporal_load(reinterpret_cast<const NV *>(prn + ii));
    const NV c = __builtin_nontemporal_load(reinterpret_cast<const NV *>(prc + ii));
    const NV co = __builtin_nontemporal_load(reinterpret_cast<const NV *>(coef + ii));
    const NV la = __builtin_nontemporal_load(reinterpret_cast<const NV *>(lapla + ii));
    NV r;
#pragma unroll
    for (int v = 0; v < VW; v++) {
        const TS cl = co[v] * la[v];
        const TS nv = (TS)(two * (TC)c[v] - (TC)p[v] + (TC)cl);
        r[v] = (i1 + v >= b.s1 && i1 + v <= b.e1) ? nv : p[v];
    }
    __builtin_nontemporal_store(r, reinterpret_cast<NV *>(prn + ii));
}

// ---- reductions (src/grid.cpp:2006-2370) ----------------------------------------
enum RedOp { RED_SUM_ABS, RED_SUM_ABS_DIFF, RED_L1, RED_MAX_ERR, RED_MIN_MAX };

template <int OP>
struct RedInit {
    static __device__ double a() { return OP == RED_MAX_ERR ? -FLT_MAX : OP == RED_MIN_MAX ? FLT_MAX : 0.0; }
    static __device__ double b() { return OP == RED_MIN_MAX ? -FLT_MAX : 0.0; }
};

template <int OP>
__device__ inline void red_combine(double &a, double &b, double a2, double b2)
{
    if (OP == RED_MAX_ERR) {
        a = a2 > a ? a2 : a;
    } else if (OP == RED_MIN_MAX) {
        a = a2 < a ? a2 : a;
        b = b2 > b ? b2 : b;
    } else {
        a += a2;
        b += b2;
    }
}

template <int OP>
__device__ inline void block_reduce(double &a, double &b)
{
    __shared__ double sa[8], sb[8];
    for (int off = 32; off > 0; off >>= 1) {
        double a2 = __shfl_down(a, off, 64), b2 = __shfl_down(b, off, 64);
        red_combine<OP>(a, b, a2, b2);
    }
    const int lane = (threadIdx.y * blockDim.x + threadIdx.x) & 63;
    const int wave = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
    const int nwave = (blockDim.x * blockDim.y + 63) >> 6;
    if (lane == 0) {
        sa[wave] = a;
        sb[wave] = b;
    }
    __syncthreads();
    if (wave == 0) {
        a = lane < nwave ? sa[lane] : RedInit<OP>::a();
        b = lane < nwave ? sb[lane] : RedInit<OP>::b();
        for (int off = 4; off > 0; off >>= 1) {
            double a2 = __shfl_down(a, off, 64), b2 = __shfl_down(b, off, 64);
            red_combine<OP>(a, b, a2, b2);
        }
    }
}

// one point of a reduction: the sums in double, min / max in the storage type
// as the reference's Myfloat accumulator (src/grid.cpp:2006-2370)
template <typename TS, int OP>
__device__ __forceinline__ void red_point(TS x, TS y, double &a, double &c, TS &mn, TS &mx, float maxerr)
{
    if (OP == RED_SUM_ABS) {
        a += fabs((double)x);
    } else if (OP == RED_SUM_ABS_DIFF) {
        a += fabs((double)x - (double)y);
    } else if (OP == RED_L1) {
        a += fabs((double)x - (double)y);
        c += fabs((double)y);
    } else if (OP == RED_MAX_ERR) {
        const double d = fabs((double)x - (double)y);
        const double e = fabsf((float)y) < maxerr ? d : d / fabs((double)y);
        if (e > a) a = e;
    } else {
        if (x < mn) mn = x;
        if (x > mx) mx = x;
    }
}

// VEC: 16-byte loads over the aligned groups of each row (u and v 16-byte
// aligned), four loads in flight per operand, the unaligned head and tail of
// the row point by point; one wave per row either way.  1000^3 FP32 inner
// box: sum |u| 0.68 ms (1.86 ms point by point), sum |u - v| 1.40 ms (2.26 ms).
template <typename TS, int OP, bool VEC>
__global__ void __launch_bounds__(256) k_reduce(const TS *u, const TS *v, Box b, int64_t n1, int64_t n12,
                                                float maxerr, double *part)
{
    double a = RedInit<OP>::a(), c = RedInit<OP>::b();
    TS mn = (TS)FLT_MAX, mx = (TS)-FLT_MAX;
    constexpr bool TWO = OP != RED_SUM_ABS && OP != RED_MIN_MAX;
    if constexpr (!VEC) {
        for_box(b, n1, n12, [&](int64_t ii, int64_t, int64_t, int64_t) {
            red_point<TS, OP>(u[ii], TWO ? v[ii] : (TS)0, a, c, mn, mx, maxerr);
        });
    } else {
        constexpr int VW = 16 / sizeof(TS), UN = 4;                    // 4 loads in flight per operand
        using NV = TS __attribute__((ext_vector_type(VW)));
        const int64_t N2 = b.n2(), rows = N2 * b.n3();
        const int lane = threadIdx.x;
        for (int64_t row = (int64_t)blockIdx.x * ROWS_PER_BLOCK + threadIdx.y; row < rows;
             row += (int64_t)gridDim.x * ROWS_PER_BLOCK) {
            const int64_t base = (b.s2 + row % N2) * n1 + (b.s3 + row / N2) * n12;
            const int64_t a0 = base + b.s1, a1 = base + b.e1 + 1;       // [a0, a1) flat
            const int64_t g0 = (a0 + VW - 1) / VW, g1 = a1 / VW;          // whole groups
            if (g0 >= g1) {                                               // short row
                for (int64_t i = a0 + lane; i < a1; i += ROW_THREADS)
                    red_point<TS, OP>(u[i], TWO ? v[i] : (TS)0, a, c, mn, mx, maxerr);
                continue;
            }
            const int64_t h = g0 * VW - a0, t = a1 - g1 * VW;             // head / tail points
            if (lane < h) red_point<TS, OP>(u[a0 + lane], TWO ? v[a0 + lane] : (TS)0, a, c, mn, mx, maxerr);
            if (lane < t) {
                const int64_t i = g1 * VW + lane;
                red_point<TS, OP>(u[i], TWO ? v[i] : (TS)0, a, c, mn, mx, maxerr);
            }
            for (int64_t gi = g0 + lane; gi < g1; gi += UN * ROW_THREADS) {
                NV x[UN], y[UN];
#pragma unroll
                for (int k = 0; k < UN; k++) {
                    const int64_t gk = gi + k * ROW_THREADS;
                    if (gk < g1) {
                        x[k] = *reinterpret_cast<const NV *>(u + gk * VW);
                        if constexpr (TWO) y[k] = *reinterpret_cast<const NV *>(v + gk * VW);
                    }
                }
#pragma unroll
                for (int k = 0; k < UN; k++) {
                    if (gi + k * ROW_THREADS >= g1) break;
#pragma unroll
                    for (int e = 0; e < VW; e++)
                        red_point<TS, OP>(x[k][e], TWO ? y[k][e] : (TS)0, a, c, mn, mx, maxerr);
                }
            }
        }
    }
    if (OP == RED_MIN_MAX) {
        a = (double)mn;
        c = (double)mx;
    }
    block_reduce<OP>(a, c);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = c;
    }
}

// the 16-byte path where both operands are 16-byte aligned
static bool red_vec(const void *u, const void *v)
{
    return ((uintptr_t)u % 16) == 0 && ((uintptr_t)v % 16) == 0;
}

template <int OP>
__global__ void __launch_bounds__(256) k_reduce_final(const double *part, int nb, double *out)
{
    double a = RedInit<OP>::a(), c = RedInit<OP>::b();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) red_combine<OP>(a, c, part[i], part[nb + i]);
    block_reduce<OP>(a, c);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = c;
    }
}

static std::mutex g_red_mu;   // the partial-sum scratch is shared by all callers

// partial results per reduction: one block per ROWS_PER_BLOCK rows, at most
// RED_BLOCKS blocks (the rows of larger ranges are strided over them)
constexpr int RED_BLOCKS = 1024;
static int red_blocks(const Box &b)
{
    const int64_t rows = b.n2() * b.n3();
    const int64_t nb = (rows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
    return (int)(nb < RED_BLOCKS ? nb : RED_BLOCKS);
}

// Scratch of hsx_l1_sums_noalloc: the two partial sums of each of at most
// RED_BLOCKS blocks (16 KiB), one copy per device for the life of the process.
// Host code never uses this name as an address: it takes the device address
// of each device's copy from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) double scratch[2 * RED_BLOCKS];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static double *g_noalloc_ptr[NOALLOC_MAX_DEVICES];   // device address of scratch, per device

// the current device's copy of scratch, looked up once per device (with that
// device current)
static int noalloc_scratch(double **out)
{
    int dev = -1;
    HSX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(HSX_ERR_UNSUPPORTED, "device %d: the allocation-free entry points serve devices 0..%d", dev,
                    NOALLOC_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        HSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HSX_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_noalloc_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_noalloc_ptr[dev];
    return HSX_OK;
}

template <int OP>
static int run_reduce(const hsx_grid *g, const void *u, const void *v, const int64_t r[6], float maxerr,
                      double out[2], void *stream)
{
    int rc = check_grid(g);
    if (!rc) rc = check_range(g, r);
    if (!rc) rc = check_ptr(u, "u");
    if (!rc && OP != RED_SUM_ABS && OP != RED_MIN_MAX) rc = check_ptr(v, "v");
    if (!rc && !out) rc = fail(HSX_ERR_INVALID, "null out");
    if (!rc) rc = ensure_init();
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_red_mu);
    const Box b = box_of(r);
    const int nb = red_blocks(b);
    double *part = nullptr, *res = nullptr;
    if ((rc = scratch_get(SCRATCH_PARTIALS, sizeof(double) * 2 * nb, (void **)&part))) return rc;
    if ((rc = scratch_get(SCRATCH_RESULT, sizeof(double) * 2, (void **)&res))) return rc;
    hipStream_t s = stream_of(stream);
    rc = dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        if (red_vec(u, v))
            hipLaunchKernelGGL((k_reduce<TS, OP, true>), dim3(nb), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (const TS *)u, (const TS *)v, b, g->n1, g->n1 * g->n2, maxerr, part);
        else
            hipLaunchKernelGGL((k_reduce<TS, OP, false>), dim3(nb), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (const TS *)u, (const TS *)v, b, g->n1, g->n1 * g->n2, maxerr, part);
        HSX_LAUNCHED("k_reduce");
        return HSX_OK;
    });
    if (rc) return rc;
    hipLaunchKernelGGL((k_reduce_final<OP>), dim3(1), dim3(256), 0, s, part, nb, res);
    HSX_LAUNCHED("k_reduce_final");
    HSX_HIP(hipMemcpyAsync(out, res, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    HSX_HIP(hipStreamSynchronize(s));
    return HSX_OK;
}

static int common_checks(const hsx_grid *g, std::initializer_list<std::pair<const void *, const char *>> ptrs)
{
    int rc = check_grid(g);
    for (auto &p : ptrs)
        if (!rc) rc = check_ptr(p.first, p.second);
    if (!rc) rc = ensure_init();
    return rc;
}

} // namespace hsx

using namespace hsx;

extern "C" {

int hsx_fill_const(const hsx_grid *g, void *u, const int64_t range[6], double val, void *stream)
{
    int rc = common_checks(g, {{u, "u"}});
    if (!rc) rc = check_range(g, range);
    if (rc) return rc;
    const Box b = box_of(range);
    hipStream_t s = stream_of(stream);
    return dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        if (((uintptr_t)u % 16) == 0)
            hipLaunchKernelGGL((k_fill_const<TS, true>), row_grid(b), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (TS *)u, b, g->n1, g->n1 * g->n2, (TS)val);
        else
            hipLaunchKernelGGL((k_fill_const<TS, false>), row_grid(b), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (TS *)u, b, g->n1, g->n1 * g->n2, (TS)val);
        HSX_LAUNCHED("k_fill_const");
        return HSX_OK;
    });
}

int hsx_fill_separable(const hsx_grid *g, void *u, const int64_t range[6], const double *t1,
                       const double *t2, const double *t3, double amp, void *stream)
{
    int rc = common_checks(g, {{u, "u"}, {t1, "t1"}, {t2, "t2"}, {t3, "t3"}});
    if (!rc) rc = check_range(g, range);
    if (rc) return rc;
    const Box b = box_of(range);
    const size_t l1 = b.n1(), l2 = b.n2(), l3 = b.n3();
    double *tab = nullptr;
    static std::mutex mu;   // the table staging buffer is shared
    std::lock_guard<std::mutex> lk(mu);
    if ((rc = scratch_get(SCRATCH_TABLES, sizeof(double) * (l1 + l2 + l3), (void **)&tab))) return rc;
    hipStream_t s = stream_of(stream);
    HSX_HIP(hipMemcpyAsync(tab, t1, sizeof(double) * l1, hipMemcpyHostToDevice, s));
    HSX_HIP(hipMemcpyAsync(tab + l1, t2, sizeof(double) * l2, hipMemcpyHostToDevice, s));
    HSX_HIP(hipMemcpyAsync(tab + l1 + l2, t3, sizeof(double) * l3, hipMemcpyHostToDevice, s));
    rc = dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        if (((uintptr_t)u % 16) == 0)
            hipLaunchKernelGGL((k_fill_sep<TS, true>), row_grid(b), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (TS *)u, b, g->n1, g->n1 * g->n2, tab, tab + l1, tab + l1 + l2, amp);
        else
            hipLaunchKernelGGL((k_fill_sep<TS, false>), row_grid(b), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (TS *)u, b, g->n1, g->n1 * g->n2, tab, tab + l1, tab + l1 + l2, amp);
        HSX_LAUNCHED("k_fill_sep");
        return HSX_OK;
    });
    if (rc) return rc;
    // the staging buffer is reused by the next call: finish with it here
    HSX_HIP(hipStreamSynchronize(s));
    return HSX_OK;
}

int hsx_array_op(const hsx_grid *g, int op, void *w, const void *u, const void *v, double val, void *stream)
{
    int rc = common_checks(g, {{w, "w"}});
    if (rc) return rc;
    if (op < HSX_ARRAY_FILL || op > HSX_ARRAY_ADDTO) return fail(HSX_ERR_INVALID, "array op %d", op);
    if (op != HSX_ARRAY_FILL && (rc = check_ptr(u, "u"))) return rc;
    if ((op == HSX_ARRAY_ADD || op == HSX_ARRAY_MULTIPLY) && (rc = check_ptr(v, "v"))) return rc;
    const int64_t n = g->n1 * g->n2 * g->n3;
    hipStream_t s = stream_of(stream);
    auto a16 = [](const void *p) { return p == nullptr || ((uintptr_t)p & 15) == 0; };
    const bool vec = a16(w) && a16(u) && a16(v);
    return dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        const int64_t work = vec ? n / (16 / (int64_t)sizeof(TS)) + 1 : n;
        int64_t nb = (work + 255) / 256;
        if (nb > 8192) nb = 8192;
        if (vec)
            hipLaunchKernelGGL(k_array16<TS>, dim3((unsigned)nb), dim3(256), 0, s, op, (TS *)w, (const TS *)u,
                               (const TS *)v, (TS)val, n);
        else
            hipLaunchKernelGGL(k_array<TS>, dim3((unsigned)nb), dim3(256), 0, s, op, (TS *)w, (const TS *)u,
                               (const TS *)v, (TS)val, n);
        HSX_LAUNCHED("k_array");
        return HSX_OK;
    });
}

int hsx_free_surface(const hsx_grid *g, void *u, int face_mask, void *stream)
{
    int rc = common_checks(g, {{u, "u"}});
    if (rc) return rc;
    hipStream_t s = stream_of(stream);
    FaceSet fs;
    bool closed_form;
    int64_t most, most_plane;
    if ((rc = make_faceset(g, face_mask, &fs, &closed_form, &most, &most_plane))) return rc;
    if (fs.n == 0) return HSX_OK;
    return dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        if (closed_form) {
            const unsigned nb1 = (unsigned)std::min<int64_t>((most + 255) / 256, 4096);
            const unsigned nb2 = (unsigned)std::min<int64_t>((most_plane + 255) / 256, 4096);
            FaceSet rest = fs;                   // the faces left to the two generic launches
            rest.vec1 = n1_groups_ok(g, u, fs);
            Face lo{}, hi{};
            const int sides = n1_vector_sides(g, u, fs, &lo, &hi);
            if (sides) {
                const Face &f = (sides & 1) ? lo : hi;
                const int64_t threads = (f.be - f.bs + 1) * (f.oe - f.os + 1) * (sides == 3 ? 2 : 1);
                const dim3 grid((unsigned)((threads + 255) / 256));
                rest.n = 0;
                for (int k = 0; k < fs.n; k++)
                    if (fs.f[k].axis != 0) rest.f[rest.n++] = fs.f[k];
                if (rest.n == 0) {
                    switch ((f.he - f.hs + 1) / (16 / (int64_t)sizeof(TS))) {
                    case 1: hipLaunchKernelGGL((k_bc_n1<TS, 1>), grid, dim3(256), 0, s, (TS *)u, lo, hi, sides); break;
                    case 2: hipLaunchKernelGGL((k_bc_n1<TS, 2>), grid, dim3(256), 0, s, (TS *)u, lo, hi, sides); break;
                    case 3: hipLaunchKernelGGL((k_bc_n1<TS, 3>), grid, dim3(256), 0, s, (TS *)u, lo, hi, sides); break;
                    default: hipLaunchKernelGGL((k_bc_n1<TS, 4>), grid, dim3(256), 0, s, (TS *)u, lo, hi, sides); break;
                    }
                    HSX_LAUNCHED("k_bc_n1");
                    return HSX_OK;
                }
                const uint32_t nbn = grid.x;
                const dim3 both(nbn + nb1 * (uint32_t)rest.n);
                switch ((f.he - f.hs + 1) / (16 / (int64_t)sizeof(TS))) {
                case 1: hipLaunchKernelGGL((k_bc_n1_mirror<TS, 1>), both, dim3(256), 0, s, (TS *)u, lo, hi, sides, rest, nbn, nb1); break;
                case 2: hipLaunchKernelGGL((k_bc_n1_mirror<TS, 2>), both, dim3(256), 0, s, (TS *)u, lo, hi, sides, rest, nbn, nb1); break;
                case 3: hipLaunchKernelGGL((k_bc_n1_mirror<TS, 3>), both, dim3(256), 0, s, (TS *)u, lo, hi, sides, rest, nbn, nb1); break;
                default: hipLaunchKernelGGL((k_bc_n1_mirror<TS, 4>), both, dim3(256), 0, s, (TS *)u, lo, hi, sides, rest, nbn, nb1); break;
                }
                HSX_LAUNCHED("k_bc_n1_mirror");
                hipLaunchKernelGGL(k_bc_zero<TS>, dim3(nb2, rest.n), dim3(256), 0, s, (TS *)u, rest);
                HSX_LAUNCHED("k_bc_zero");
                return HSX_OK;
            }
            hipLaunchKernelGGL(k_bc_mirror<TS>, dim3(nb1, rest.n), dim3(256), 0, s, (TS *)u, rest);
            HSX_LAUNCHED("k_bc_mirror");
            hipLaunchKernelGGL(k_bc_zero<TS>, dim3(nb2, rest.n), dim3(256), 0, s, (TS *)u, rest);
            HSX_LAUNCHED("k_bc_zero");
            return HSX_OK;
        }
        // small grids: the reference's six phases one after the other
        for (int k = 0; k < fs.n; k++) {
            const Face &f = fs.f[k];
            const int64_t total = (f.he - f.hs + 2) * (f.be - f.bs + 1) * (f.oe - f.os + 1);
            const unsigned nb = (unsigned)std::min<int64_t>((total + 255) / 256, 16384);
            hipLaunchKernelGGL(k_free_surface<TS>, dim3(nb), dim3(256), 0, s, (TS *)u, f);
            HSX_LAUNCHED("k_free_surface");
        }
        return HSX_OK;
    });
}

int hsx_update_pressure(const hsx_grid *g, void *prn, const void *prc, const void *coef, const void *lapla,
                        void *stream)
{
    int rc = common_checks(g, {{prn, "prn"}, {prc, "prc"}, {coef, "coef"}, {lapla, "lapla"}});
    if (rc) return rc;
    const Box b = box_of(g->inner);
    hipStream_t s = stream_of(stream);
    return dispatch_types(g, [&](auto ts, auto tc) -> int {
        using TS = decltype(ts);
        using TC = decltype(tc);
        constexpr int VW = 16 / sizeof(TS);
        const bool wide = g->n1 % VW == 0 &&
                          ((uintptr_t)prn | (uintptr_t)prc | (uintptr_t)coef | (uintptr_t)lapla) % 16 == 0;
        const int64_t ng = b.e1 / VW - b.s1 / VW + 1, total = ng * b.n2() * b.n3();
        if (wide && total > 0 && (total + 255) / 256 <= INT32_MAX) {
            hipLaunchKernelGGL((k_update_pressure16<TS, TC>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                               (TS *)prn, (const TS *)prc, (const TS *)coef, (const TS *)lapla, b, g->n1,
                               g->n1 * g->n2, ng, total);
            HSX_LAUNCHED("k_update_pressure16");
            return HSX_OK;
        }
        hipLaunchKernelGGL((k_update_pressure<TS, TC>), row_grid(b), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                           (TS *)prn, (const TS *)prc, (const TS *)coef, (const TS *)lapla, b, g->n1,
                           g->n1 * g->n2);
        HSX_LAUNCHED("k_update_pressure");
        return HSX_OK;
    });
}

int hsx_sum_abs(const hsx_grid *g, const void *u, const int64_t range[6], double *out, void *stream)
{
    double r[2];
    int rc = run_reduce<RED_SUM_ABS>(g, u, nullptr, range, 0.f, r, stream);
    if (!rc) *out = r[0];
    return rc;
}

int hsx_sum_abs_diff(const hsx_grid *g, const void *u, const void *v, const int64_t range[6], double *out,
                     void *stream)
{
    double r[2];
    int rc = run_reduce<RED_SUM_ABS_DIFF>(g, u, v, range, 0.f, r, stream);
    if (!rc) *out = r[0];
    return rc;
}

int hsx_l1_sums(const hsx_grid *g, const void *u, const void *v, const int64_t range[6], double out[2],
                void *stream)
{
    return run_reduce<RED_L1>(g, u, v, range, 0.f, out, stream);
}

int hsx_l1_sums_noalloc(const hsx_grid *g, const void *u, const void *v, const int64_t range[6], double *out,
                        void *stream)
{
    int rc = check_grid(g);
    if (!rc) rc = check_range(g, range);
    if (!rc) rc = check_ptr(u, "u");
    if (!rc) rc = check_ptr(v, "v");
    if (!rc) rc = check_ptr(out, "out");
    if (!rc) rc = ensure_init();
    if (rc) return rc;
    double *part = nullptr;
    if ((rc = noalloc_scratch(&part))) return rc;
    const Box b = box_of(range);
    const int nb = red_blocks(b);
    if (nb < 1 || nb > RED_BLOCKS)
        return fail(HSX_ERR_INVALID, "l1_sums_noalloc: %d partial blocks, the scratch holds %d", nb, RED_BLOCKS);
    hipStream_t s = stream_of(stream);
    // every block writes its own partials before k_reduce_final reads them:
    // nothing in scratch has to start at zero
    rc = dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        if (red_vec(u, v))
            hipLaunchKernelGGL((k_reduce<TS, RED_L1, true>), dim3(nb), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (const TS *)u, (const TS *)v, b, g->n1, g->n1 * g->n2, 0.f, part);
        else
            hipLaunchKernelGGL((k_reduce<TS, RED_L1, false>), dim3(nb), dim3(ROW_THREADS, ROWS_PER_BLOCK), 0, s,
                               (const TS *)u, (const TS *)v, b, g->n1, g->n1 * g->n2, 0.f, part);
        HSX_LAUNCHED("k_reduce");
        return HSX_OK;
    });
    if (rc) return rc;
    hipLaunchKernelGGL((k_reduce_final<RED_L1>), dim3(1), dim3(256), 0, s, part, nb, out);
    HSX_LAUNCHED("k_reduce_final");
    return HSX_OK;
}

int hsx_max_err(const hsx_grid *g, const void *u, const void *v, const int64_t range[6], double max_err_float,
                double *out, void *stream)
{
    double r[2];
    int rc = run_reduce<RED_MAX_ERR>(g, u, v, range, (float)max_err_float, r, stream);
    if (!rc) *out = r[0];
    return rc;
}

int hsx_min_max(const hsx_grid *g, const void *u, const int64_t range[6], double out[2], void *stream)
{
    return run_reduce<RED_MIN_MAX>(g, u, nullptr, range, 0.f, out, stream);
}

} // extern "C"
