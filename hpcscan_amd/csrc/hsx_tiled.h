// LDS-tiled, n3-streaming stencil kernels for gfx950 (3D grids).
//
// One workgroup owns a tile of LX*VW points along n1 by BY rows along n2 and
// marches along n3 over a chunk of planes:
//  * n3 neighbours come from a per-thread register queue of the 2R+1 planes
//    z-R..z+R (each slot holds VW consecutive n1 values);
//  * n1 / n2 neighbours come from the current plane staged in LDS with its
//    +-R halo (double buffered, one barrier per plane);
//  * loads and stores are VW-wide (16 bytes per lane: float4 / double2) and a
//    wave covers contiguous n1, so every global access is coalesced;
//  * the next plane (queue front, halo rows/columns, prn and coef) is loaded
//    while the current one is computed.
// The queue is never shifted: the plane loop is unrolled by 2R+1 so that the
// slot of plane z+k is a compile-time index.
//
// Per-point arithmetic is d2_values / leapfrog of hsx_stencil_point.h, i.e.
// the reference operation order; results equal the reference kernels bit for
// bit.  Launch: 1D grid of tiles x chunks, block ids remapped so that the
// tiles handled by one XCD (one L2) are neighbours and share their halos.
#ifndef HSX_TILED_H
#define HSX_TILED_H

#include "hsx_internal.h"
#include "hsx_stencil_point.h"

#include <algorithm>
#include <utility>

namespace hsx {

enum TiledMode {
    MODE_AC2 = 0, MODE_LAP = 1, MODE_D2X = 2, MODE_D2Y = 3, MODE_D2Z = 4,
    MODE_D1X = 5, MODE_D1Y = 6, MODE_D1Z = 7,       // staggered first derivatives
    // the fused step and the Laplacian with each weighted sum contracted into
    // fused multiply-adds (opt-in, hsx_set_contract): not the reference's bits
    MODE_AC2_FMA = 8, MODE_LAP_FMA = 9,
    // the fused step with the adjoint's imaging condition accumulated from its
    // output (hsx_gradient): grad += scale * out * (q0 - 2 q1 + q2) in double
    MODE_AC2_IMG = 10
};
constexpr bool mode_ac2(int m) { return m == MODE_AC2 || m == MODE_AC2_FMA || m == MODE_AC2_IMG; }
constexpr bool mode_lap(int m) { return m == MODE_LAP || m == MODE_LAP_FMA; }
constexpr bool mode_fma(int m) { return m == MODE_AC2_FMA || m == MODE_LAP_FMA; }
constexpr int mode_base(int m)
{
    return m == MODE_AC2_FMA || m == MODE_AC2_IMG ? MODE_AC2 : m == MODE_LAP_FMA ? MODE_LAP : m;
}

template <typename T, int VW>
struct alignas(sizeof(T) * VW) Pack {
    T v[VW];
};

template <typename TC>
struct TiledArgs {
    void *out;                // new prn (MODE_AC2) or w
    const void *prn_in;       // MODE_AC2: old prn (may be out)
    const void *u;            // prc / u
    const void *coef;         // MODE_AC2 only
    int64_t n1, n2, n12;      // row length, rows per plane, plane stride
    int64_t s1, e1, s2, e2, s3, e3;
    int64_t t1s;              // n1 origin of the first tile (multiple of VW)
    int ntx, nty, nchunk, chunk;
    int nt;                   // 1: prn / coef / output streamed with nontemporal accesses
    // MODE_AC2: the free surface folded into the step (0 = none):
    //  FOLD_PRC  prc is read as if its boundary planes were zero,
    //  FOLD_PRN  prn likewise,
    //  FOLD_HALO the output's halos of the folded faces are written (-mirror).
    int fold;
    int fb[6];                // boundary plane of each folded face (axis-major, low side first), -1 if not folded
    // MODE_AC2_IMG: grad[i] += img_scale * out[i] * ((q0[i] - 2 q1[i]) + q2[i]) over img_box
    const void *img_q[3];
    double *img_grad;
    double img_scale;
    int64_t img_box[6];
    Coefs<TC> c;
};

constexpr int FOLD_PRC = HSX_FOLD_PRC, FOLD_PRN = HSX_FOLD_PRN, FOLD_HALO = HSX_FOLD_HALO;

constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }
constexpr int round_up(int a, int b) { return ceil_div(a, b) * b; }

template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING_ = false>
struct Tiled {
    static constexpr int P = 2 * R + 1;               // queue length
    static constexpr int HL = round_up(R, VW);        // halo slots left / right (multiple of VW)
    static constexpr int W = HL + LX * VW + HL;       // LDS row pitch in elements
    static constexpr int ROWS = BY + 2 * R;
    static constexpr int PLANE = ROWS * W;            // elements of one LDS plane
    static constexpr int NHG = HL / VW;               // halo groups per side
    static constexpr int NHR = ceil_div(2 * R, BY);   // n2 halo rows loaded per thread
    static constexpr bool NEED_X = mode_ac2(MODE) || mode_lap(MODE) || MODE == MODE_D2X || MODE == MODE_D1X;
    static constexpr bool NEED_Y = mode_ac2(MODE) || mode_lap(MODE) || MODE == MODE_D2Y || MODE == MODE_D1Y;
    static constexpr bool NEED_Z = mode_ac2(MODE) || mode_lap(MODE) || MODE == MODE_D2Z || MODE == MODE_D1Z;
    static constexpr bool D1 = MODE == MODE_D1X || MODE == MODE_D1Y || MODE == MODE_D1Z;
    static constexpr bool NEED_LDS = NEED_X || NEED_Y;
    using V = Pack<TS, VW>;
    // Halo ring: the n1 / n2 halo of plane z+R is loaded together with the
    // queue front (plane z+R), i.e. while the neighbouring tiles load the same
    // rows as their own queue front, so those loads meet in L2; it waits in an
    // LDS ring of R+1 planes, private to the thread that loaded it.
    static constexpr bool RING = RING_ && NEED_X && NEED_Y && NEED_Z;
    static constexpr int NR = R + 1;
    static constexpr size_t PLANES_BYTES = NEED_LDS ? 2 * PLANE * sizeof(TS) : 0;   // two plane buffers
    static constexpr size_t RING_H = RING ? (size_t)NR * 2 * R * LX * sizeof(V) : 0;
    static constexpr size_t RING_C = RING ? (size_t)NR * BY * 2 * NHG * sizeof(V) : 0;
    static constexpr size_t LDS = PLANES_BYTES + RING_H + RING_C;
};

template <typename TS, int VW>
__device__ __forceinline__ Pack<TS, VW> vload(const TS *p)
{
    return *reinterpret_cast<const Pack<TS, VW> *>(p);
}

// the same 16 bytes as a clang vector, for the nontemporal builtins
template <typename TS, int VW>
using NVec = TS __attribute__((ext_vector_type(VW)));

template <typename TS, int VW>
__device__ __forceinline__ Pack<TS, VW> vload_nt(const TS *p)
{
    const NVec<TS, VW> t = __builtin_nontemporal_load(reinterpret_cast<const NVec<TS, VW> *>(p));
    Pack<TS, VW> r;
#pragma unroll
    for (int v = 0; v < VW; v++) r.v[v] = t[v];
    return r;
}

template <typename TS, int VW>
__device__ __forceinline__ void vstore_nt(TS *p, const Pack<TS, VW> &x)
{
    NVec<TS, VW> t;
#pragma unroll
    for (int v = 0; v < VW; v++) t[v] = x.v[v];
    __builtin_nontemporal_store(t, reinterpret_cast<NVec<TS, VW> *>(p));
}

template <typename TS, int VW>
__device__ __forceinline__ Pack<TS, VW> vzero()
{
    Pack<TS, VW> z;
#pragma unroll
    for (int v = 0; v < VW; v++) z.v[v] = (TS)0;
    return z;
}

// compile-time loop: f(integral_constant<0>), ..., f(integral_constant<N-1>)
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>()), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>());
}

// XCD-aware block order: the hardware deals workgroups to the 8 XCDs round
// robin; logical = (blockIdx % 8) * per_xcd + blockIdx / 8 gives each XCD (one
// L2) a contiguous run of logical blocks, i.e. neighbouring tiles
__device__ __forceinline__ int xcd_logical(int total)
{
    const int per_xcd = (total + 7) / 8;
    return (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
}

// one workgroup: planes [z0, z1) of tile 'tile' of the grid described by a
template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING_, bool FOLD = false>
__device__ __forceinline__ void tiled_body_tz(const TiledArgs<TC> &a, int tile, int64_t z0, int64_t z1)
{
    using K = Tiled<TS, TC, R, VW, LX, BY, MODE, RING_>;
    using V = typename K::V;
    constexpr int P = K::P, HL = K::HL, W = K::W, NHG = K::NHG, NHR = K::NHR, NR = K::NR;
    constexpr bool RING = K::RING;
    constexpr int64_t ES = sizeof(TS);
    extern __shared__ __align__(16) unsigned char hsx_lds_raw[];
    TS *lds = reinterpret_cast<TS *>(hsx_lds_raw);
    V *ring_h = reinterpret_cast<V *>(hsx_lds_raw + K::PLANES_BYTES);
    V *ring_c = reinterpret_cast<V *>(hsx_lds_raw + K::PLANES_BYTES + K::RING_H);

    const int tx = threadIdx.x, ty = threadIdx.y;

    const int64_t t1 = a.t1s + (int64_t)(tile % a.ntx) * (LX * VW);            // first n1 index of the tile
    const int64_t i1 = t1 + tx * VW;                                           // first n1 index of this lane
    const int64_t t2s = a.s2 + (int64_t)(tile / a.ntx) * BY;                   // first n2 row of the tile
    const int64_t i2 = t2s + ty;

    const bool lane_in_row = i1 + VW - 1 < a.n1;        // vector fully inside the row
    const bool row_ok = i2 < a.n2;
    const bool col_ok = lane_in_row && row_ok;
    const int64_t n1 = a.n1;

    // Addressing: plane base (uniform, scalar registers) + 32-bit byte offset of
    // the lane inside the plane (the host checks that a plane is below 4 GiB).
    const int64_t pstride = a.n12 * ES;
    const char *ub = reinterpret_cast<const char *>(a.u);
    const char *pb = reinterpret_cast<const char *>(a.prn_in);
    const char *cb = reinterpret_cast<const char *>(a.coef);
    char *ob = reinterpret_cast<char *>(a.out);
    auto at = [&](const char *base, int64_t z, uint32_t off) -> const TS * {
        return reinterpret_cast<const TS *>((base + z * pstride) + (uint64_t)off);
    };
    const uint32_t col = (uint32_t)((i1 + i2 * n1) * ES);

    // ---- folded free surface (FOLD instances; the closed form of
    // Grid::applyBoundaryCondition ANTI_MIRROR, src/grid.cpp:3065-3230): an inner
    // point on a boundary plane of a folded face reads as +0; halos are read as
    // stored (the previous folded step wrote them)
    constexpr bool FK = FOLD && mode_ac2(MODE);
    constexpr bool FMA = mode_fma(MODE);
    const int fold = FK ? a.fold : 0;
    const int64_t tw = (int64_t)LX * VW;
    auto in_range = [](int64_t x, int64_t lo, int64_t hi) { return x >= lo && x < hi; };
    // the tile (with its halo) holds a boundary column / row of a folded face (uniform)
    const bool edge_t = FK && (in_range(a.fb[0], t1 - HL, t1 + tw + HL) || in_range(a.fb[1], t1 - HL, t1 + tw + HL) ||
                               in_range(a.fb[2], t2s - R, t2s + BY + R) || in_range(a.fb[3], t2s - R, t2s + BY + R));
    auto bplane = [&](int64_t z) { return (int)z == a.fb[4] || (int)z == a.fb[5]; };
    // per-lane flags (VW <= 4 elements):
    //   bits 0-3 / 4-7     element column inner / on a folded boundary column
    //   bits 8 / 9         row inner / on a folded boundary row
    //   bits 16-19 / 20-23 the same as 0-7 for the n1 halo group of this lane
    //   bits 24+2k / 25+2k n2 halo row k inner / on a folded boundary row
    uint32_t fl = 0;
    if constexpr (FK) {
#pragma unroll
        for (int v = 0; v < VW; v++) {
            const int64_t c = i1 + v;
            if (c >= a.s1 && c <= a.e1) fl |= 1u << v;
            if (c == a.fb[0] || c == a.fb[1]) fl |= 1u << (4 + v);
        }
        if (i2 >= a.s2 && i2 <= a.e2) fl |= 1u << 8;
        if (i2 == a.fb[2] || i2 == a.fb[3]) fl |= 1u << 9;
    }
    // element v (column bits cb: inner | boundary << 4; row bits rb: inner |
    // boundary << 1) of plane z becomes +0 when the point is inner and lies on a
    // folded boundary plane; the uniform test skips most tiles and planes
    // The flags are made opaque to the compiler at each use (an empty asm
    // statement), so that what is derived from them is recomputed in place
    // instead of being hoisted out of the plane loop and kept live; the zeroing
    // is a bit mask on the value (VALU only, no lane masks).
    using UB = std::conditional_t<sizeof(TS) == 8, uint64_t, uint32_t>;
    auto zero_fold = [&](V &x, uint32_t cb, uint32_t rb, int64_t z) {
        if (!(z >= a.s3 && z <= a.e3 && (edge_t || bplane(z)))) return;
        const uint32_t pz = bplane(z) ? 1u : 0u;
        asm volatile("" : "+v"(cb), "+v"(rb));
#pragma unroll
        for (int v = 0; v < VW; v++) {
            const uint32_t zb = (cb >> v) & rb & ((cb >> (4 + v)) | (rb >> 1) | pz) & 1u;
            x.v[v] = __builtin_bit_cast(TS, __builtin_bit_cast(UB, x.v[v]) & ((UB)zb - 1));
        }
    };

    // ---- queue: slot (z - z0 + R) % P holds plane z
    V q[P];
#pragma unroll
    for (int k = 0; k < P; k++) q[k] = vzero<TS, VW>();
#pragma unroll
    for (int k = 0; k < 2 * R; k++)        // planes z0-R .. z0+R-1
        if (col_ok) q[k] = vload<TS, VW>(at(ub, z0 - R + k, col));
    if (fold & FOLD_PRC) {
#pragma unroll
        for (int k = 0; k < 2 * R; k++) zero_fold(q[k], fl & 0xff, (fl >> 8) & 3, z0 - R + k);
    }

    // prefetch registers for the first plane
    V front = vzero<TS, VW>();
    if (col_ok && K::NEED_Z) front = vload<TS, VW>(at(ub, z0 + R, col));
    if (col_ok && !K::NEED_Z) q[R] = vload<TS, VW>(at(ub, z0, col));
    V hrow[NHR];          // n2 halo rows of the next plane
    V hcol = vzero<TS, VW>();   // n1 halo group of the next plane (edge lanes)
    V nprn = vzero<TS, VW>(), ncoef = vzero<TS, VW>();

    // which halo loads this thread owns
    uint32_t hrow_off[NHR];
    int hrow_lds[NHR];
    bool hrow_ok[NHR];
#pragma unroll
    for (int k = 0; k < NHR; k++) {
        const int h = ty + k * BY;                         // 0 .. 2R-1 for the halo rows
        const int row = h < R ? h : BY + h;                // LDS row
        const int64_t hi2 = t2s + row - R;
        hrow_lds[k] = row;
        hrow_off[k] = (uint32_t)((i1 + hi2 * n1) * ES);
        hrow_ok[k] = K::NEED_Y && h < 2 * R && hi2 < a.n2 && lane_in_row;
        hrow[k] = vzero<TS, VW>();
    }
    const bool left = K::NEED_X && tx < NHG;
    const bool right = K::NEED_X && tx >= LX - NHG;
    const int64_t hc_i1 = left ? t1 - HL + tx * VW : t1 + LX * VW + (tx - (LX - NHG)) * VW;
    const int hc_lds = left ? tx * VW : HL + LX * VW + (tx - (LX - NHG)) * VW;
    const bool hc_ok = (left || right) && row_ok && hc_i1 >= 0 && hc_i1 + VW - 1 < a.n1;
    const uint32_t hc_off = (uint32_t)((hc_i1 + i2 * n1) * ES);

    // ring slots owned by this thread
    // (threads that own no halo get in-range addresses too, so that the ring is
    // read without branches; what they read is not used)
    const int rc_idx = left ? tx : right ? NHG + (tx - (LX - NHG)) : 0;
    auto ring_h_at = [&](int slot, int k) -> V * {
        return ring_h + ((slot * 2 * R + (ty + k * BY) % (2 * R)) * LX + tx);
    };
    auto ring_c_at = [&](int slot) -> V * { return ring_c + ((slot * BY + ty) * 2 * NHG + rc_idx); };

    auto load_halo = [&](int64_t z) {
#pragma unroll
        for (int k = 0; k < NHR; k++)
            if (hrow_ok[k]) hrow[k] = vload<TS, VW>(at(ub, z, hrow_off[k]));
        if (hc_ok) hcol = vload<TS, VW>(at(ub, z, hc_off));
    };
    // the free-surface zeros in the halo just loaded for plane z
    if constexpr (FK) {
#pragma unroll
        for (int k = 0; k < NHR; k++) {
            const int64_t hi2 = t2s + hrow_lds[k] - R;
            if (hi2 >= a.s2 && hi2 <= a.e2) fl |= 1u << (24 + 2 * k);
            if (hi2 == a.fb[2] || hi2 == a.fb[3]) fl |= 1u << (25 + 2 * k);
        }
#pragma unroll
        for (int v = 0; v < VW; v++) {
            const int64_t c = hc_i1 + v;
            if (c >= a.s1 && c <= a.e1) fl |= 1u << (16 + v);
            if (c == a.fb[0] || c == a.fb[1]) fl |= 1u << (20 + v);
        }
    }
    auto fold_halo = [&](int64_t z) {
        if (!(fold & FOLD_PRC)) return;
#pragma unroll
        for (int k = 0; k < NHR; k++) zero_fold(hrow[k], fl & 0xff, (fl >> (24 + 2 * k)) & 3, z);
        zero_fold(hcol, (fl >> 16) & 0xff, (fl >> 8) & 3, z);
    };
    auto store_halo = [&](int slot) {
#pragma unroll
        for (int k = 0; k < NHR; k++)
            if (hrow_ok[k]) *ring_h_at(slot, k) = hrow[k];
        if (hc_ok) *ring_c_at(slot) = hcol;
    };
    // FP32 fused step of orders 6-10: prn and coef are loaded two planes ahead
    // (registers to spare at 1024 threads; more bytes in flight per CU)
    constexpr bool DEEP = mode_ac2(MODE) && sizeof(TS) == 4 && sizeof(TC) == 4 && R >= 3 && R <= 5;
    V nprn2 = vzero<TS, VW>(), ncoef2 = vzero<TS, VW>();   // plane z+2 (DEEP)
    auto load_streams_to = [&](int64_t z, V &vp, V &vc) {
        if (mode_ac2(MODE) && col_ok) {
            if (a.nt) {
                vp = vload_nt<TS, VW>(at(pb, z, col));
                vc = vload_nt<TS, VW>(at(cb, z, col));
            } else {
                vp = vload<TS, VW>(at(pb, z, col));
                vc = vload<TS, VW>(at(cb, z, col));
            }
        }
    };
    auto load_streams = [&](int64_t z) { load_streams_to(z, nprn, ncoef); };
    if (RING) {
        // halos of planes z0 .. z0+R-1 into the ring, plane z0+R in registers
        for (int j = 0; j < R; j++) {
            load_halo(z0 + j);
            fold_halo(z0 + j);
            store_halo((int)((z0 + j) % NR));
        }
        load_halo(z0 + R);
    } else {
        load_halo(z0);
    }
    load_streams(z0);
    if constexpr (DEEP) {
        if (z0 + 1 < z1) load_streams_to(z0 + 1, nprn2, ncoef2);
    }

    // compute mask: which of the VW points of this lane are inner points
    bool inner[VW];
#pragma unroll
    for (int v = 0; v < VW; v++) inner[v] = i1 + v >= a.s1 && i1 + v <= a.e1 && i2 <= a.e2;
    bool any_inner = false, all_inner = true;
#pragma unroll
    for (int v = 0; v < VW; v++) {
        any_inner |= inner[v];
        all_inner &= inner[v];
    }

    // FOLD_HALO: each output point within R of a folded face is also written,
    // negated, to its mirror in the halo of that face; -0 instead where the
    // point lies on a boundary plane of an earlier axis (the closed form of the
    // reference's six face phases)
    auto near = [&](int64_t b, int64_t lo, int64_t hi, int side) {   // [lo, hi) meets the R points inside face b
        return b >= 0 && (side == 0 ? lo <= b + R && hi > b + 1 : lo <= b - 1 && hi > b - R);
    };
    const bool near1 = FK && (near(a.fb[0], t1, t1 + tw, 0) || near(a.fb[1], t1, t1 + tw, 1));
    const bool near2 = FK && (near(a.fb[2], t2s, t2s + BY, 0) || near(a.fb[3], t2s, t2s + BY, 1));
    // (flags and thread indices pass through empty asm statements: what is
    // derived from them is computed here, when a face is near, instead of
    // being hoisted out of the plane loop into registers)
    auto mirror_out = [&](const V &res, int64_t z) {
        uint32_t f = fl;
        asm volatile("" : "+v"(f));
        for (int side = 0; side < 2; side++) {
            const int64_t b = a.fb[4 + side];
            const int64_t d = side == 0 ? z - b : b - z;
            if (b < 0 || d < 1 || d > R) continue;                        // uniform
            TS *o = reinterpret_cast<TS *>((ob + (2 * b - z) * pstride) + (uint64_t)col);
#pragma unroll
            for (int v = 0; v < VW; v++)
                if ((f >> v) & (f >> 8) & 1) o[v] = (((f >> (4 + v)) | (f >> 9)) & 1) ? -(TS)0 : -res.v[v];
        }
        if (near2) {
            uint32_t y = threadIdx.y;
            asm volatile("" : "+v"(y));
            const int r = (int)t2s + (int)y;
            for (int side = 0; side < 2; side++) {
                const int b = a.fb[2 + side];
                const int d = side == 0 ? r - b : b - r;
                if (b < 0 || d < 1 || d > R) continue;
                TS *o = reinterpret_cast<TS *>((ob + z * pstride) + (uint64_t)col +
                                               (int64_t)(side == 0 ? -2 * d : 2 * d) * n1 * ES);
#pragma unroll
                for (int v = 0; v < VW; v++)
                    if ((f >> v) & (f >> 8) & 1) o[v] = ((f >> (4 + v)) & 1) ? -(TS)0 : -res.v[v];
            }
        }
        if (near1) {
            uint32_t x = threadIdx.x;
            asm volatile("" : "+v"(x));
            const int c0 = (int)t1 + (int)x * VW;
            TS *o = reinterpret_cast<TS *>((ob + z * pstride) + (uint64_t)col);
#pragma unroll
            for (int v = 0; v < VW; v++) {
                if (!((f >> v) & (f >> 8) & 1)) continue;
                const int dl = c0 + v - a.fb[0], dh = a.fb[1] - (c0 + v);
                if (a.fb[0] >= 0 && dl >= 1 && dl <= R) o[v - 2 * dl] = -res.v[v];
                if (a.fb[1] >= 0 && dh >= 1 && dh <= R) o[v + 2 * dh] = -res.v[v];
            }
        }
    };

    const TC *A = a.c.A;
    int buf = 0;
    int rslot = (int)(z0 % NR);              // ring slot of plane z
    for (int64_t zb = z0; zb < z1; zb += P) {
        // P planes per trip, each with compile-time queue slots
        static_for<P>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            const int64_t z = zb + s;
            if (z >= z1) return;                 // uniform over the workgroup
            constexpr int cz = (s + R) % P;      // slot of plane z
            // the front loaded during the previous plane is plane z+R
            if (K::NEED_Z) {
                q[(s + 2 * R) % P] = front;
                if (fold & FOLD_PRC) zero_fold(q[(s + 2 * R) % P], fl & 0xff, (fl >> 8) & 3, z + R);
            }
            const V cur = q[cz];
            V prn_v = nprn;
            const V coef_v = ncoef;
            if constexpr (DEEP) {
                nprn = nprn2;
                ncoef = ncoef2;
            }
            if (fold & FOLD_PRN) zero_fold(prn_v, fl & 0xff, (fl >> 8) & 3, z);

            // stage plane z in LDS: centre rows from the queue, halo from the
            // ring (loaded R planes ago) or from the registers loaded last plane
            TS *pl = lds + buf * K::PLANE;
            if (K::NEED_LDS) {
                *reinterpret_cast<V *>(pl + (ty + R) * W + HL + tx * VW) = cur;
                if constexpr (RING) {
#pragma unroll
                    for (int k = 0; k < NHR; k++) {
                        const V t = *ring_h_at(rslot, k);
                        if (hrow_ok[k]) *reinterpret_cast<V *>(pl + hrow_lds[k] * W + HL + tx * VW) = t;
                    }
                    const V t = *ring_c_at(rslot);
                    if (hc_ok) *reinterpret_cast<V *>(pl + (ty + R) * W + hc_lds) = t;
                    // the halo of plane z+R (registers) goes to the slot of plane z-1
                    fold_halo(z + R);
                    store_halo(rslot == 0 ? NR - 1 : rslot - 1);
                    rslot = rslot == NR - 1 ? 0 : rslot + 1;
                } else {
                    fold_halo(z);
#pragma unroll
                    for (int k = 0; k < NHR; k++)
                        if (hrow_ok[k]) *reinterpret_cast<V *>(pl + hrow_lds[k] * W + HL + tx * VW) = hrow[k];
                    if (hc_ok) *reinterpret_cast<V *>(pl + (ty + R) * W + hc_lds) = hcol;
                }
            }

            // prefetch plane z+1: queue front z+1+R, halos (of z+1+R with the
            // ring, else of z+1), prn and coef
            if (z + 1 < z1) {
                if (K::NEED_Z && col_ok) front = vload<TS, VW>(at(ub, z + 1 + R, col));
                if (!K::NEED_Z && col_ok) q[(s + 1 + R) % P] = vload<TS, VW>(at(ub, z + 1, col));
                load_halo(RING ? z + 1 + R : z + 1);
                if constexpr (DEEP) {
                    if (z + 2 < z1) load_streams_to(z + 2, nprn2, ncoef2);
                } else {
                    load_streams(z + 1);
                }
            }

            // n3-only kernels (no LDS) meet here too: it keeps the
            // workgroup's waves on one plane (FD_D2 along n3, 1000^3 FP32:
            // order 8 1.85 ms against 1.97 ms without, order 16 1.68 / 1.85)
            __syncthreads();

            if (any_inner) {
                // n1 neighbours: the row segment [tx*VW, tx*VW + VW + 2 HL) of the LDS row
                TS xr[2 * HL + VW];
                // n2: accumulated row pair by row pair for all VW points at once
                // (the order of d2_values for each point), so that only one pair
                // of LDS rows is live at a time
                TC ay[VW];
                if (K::NEED_X) {
                    const TS *rowp = pl + (ty + R) * W + tx * VW;
#pragma unroll
                    for (int g = 0; g < 2 * NHG + 1; g++) {
                        const V t = g == NHG ? cur : *reinterpret_cast<const V *>(rowp + g * VW);
#pragma unroll
                        for (int v = 0; v < VW; v++) xr[g * VW + v] = t.v[v];
                    }
                }
                if (K::NEED_Y && K::D1) {
                    // staggered: A1*(u0 - u-1) + Aj*(u+(j-1) - u-j), j = 2..R (order 2: no weight)
                    const V td = *reinterpret_cast<const V *>(pl + (ty + R - 1) * W + HL + tx * VW);
#pragma unroll
                    for (int v = 0; v < VW; v++) {
                        const TS diff = cur.v[v] - td.v[v];
                        ay[v] = R == 1 ? (TC)diff : a.c.D1[1] * (TC)diff;
                    }
#pragma unroll
                    for (int j = 2; j <= R; j++) {
                        const V tu = *reinterpret_cast<const V *>(pl + (ty + R + j - 1) * W + HL + tx * VW);
                        const V tdj = *reinterpret_cast<const V *>(pl + (ty + R - j) * W + HL + tx * VW);
#pragma unroll
                        for (int v = 0; v < VW; v++) {
                            const TS diff = tu.v[v] - tdj.v[v];
                            ay[v] = ay[v] + a.c.D1[j] * (TC)diff;
                        }
                    }
                } else if (K::NEED_Y) {
#pragma unroll
                    for (int v = 0; v < VW; v++) ay[v] = A[0] * (TC)cur.v[v];
#pragma unroll
                    for (int j = 1; j <= R; j++) {
                        const V tu = *reinterpret_cast<const V *>(pl + (ty + R + j) * W + HL + tx * VW);
                        const V td = *reinterpret_cast<const V *>(pl + (ty + R - j) * W + HL + tx * VW);
#pragma unroll
                        for (int v = 0; v < VW; v++) {
                            if (R == 1) {
                                ay[v] = ay[v] + (TC)tu.v[v];
                                ay[v] = ay[v] + (TC)td.v[v];
                            } else {
                                const TS pair = tu.v[v] + td.v[v];
                                ay[v] = FMA ? fma_tc(A[j], (TC)pair, ay[v]) : ay[v] + A[j] * (TC)pair;
                            }
                        }
                    }
                }
                V res;
#pragma unroll
                for (int v = 0; v < VW; v++) {
                    const TS c0 = cur.v[v];
                    TS p[R], m[R];
                    TC d1 = 0, d2 = 0, d3 = 0;
                    if (K::D1) {
                        // staggered first derivative at i - 1/2: values at +(j-1) and -j
                        if (K::NEED_X) {
#pragma unroll
                            for (int j = 1; j <= R; j++) {
                                p[j - 1] = xr[HL + v + j - 1];
                                m[j - 1] = xr[HL + v - j];
                            }
                            d1 = d1_values<TS, TC, R>(p, m, a.c.D1, a.c.inv[0]);
                        }
                        if (K::NEED_Y) d2 = ay[v] * a.c.inv[1];
                        if (K::NEED_Z) {
#pragma unroll
                            for (int j = 1; j <= R; j++) {
                                p[j - 1] = q[(s + R + j - 1) % P].v[v];
                                m[j - 1] = q[(s + R - j + P) % P].v[v];
                            }
                            d3 = d1_values<TS, TC, R>(p, m, a.c.D1, a.c.inv[2]);
                        }
                    } else {
                        if (K::NEED_X) {
#pragma unroll
                            for (int j = 1; j <= R; j++) {
                                p[j - 1] = xr[HL + v + j];
                                m[j - 1] = xr[HL + v - j];
                            }
                            d1 = d2_values<TS, TC, R, FMA>(c0, p, m, A, a.c.inv2[0]);
                        }
                        if (K::NEED_Y) d2 = ay[v] * a.c.inv2[1];
                        if (K::NEED_Z) {
#pragma unroll
                            for (int j = 1; j <= R; j++) {
                                p[j - 1] = q[(s + R + j) % P].v[v];
                                m[j - 1] = q[(s + R - j + P) % P].v[v];
                            }
                            d3 = d2_values<TS, TC, R, FMA>(c0, p, m, A, a.c.inv2[2]);
                        }
                    }
                    TS o;
                    if (mode_ac2(MODE)) {
                        const TC lap = d1 + d2 + d3;
                        o = FMA ? leapfrog_fma<TS, TC>(c0, prn_v.v[v], coef_v.v[v], lap)
                                : leapfrog<TS, TC>(c0, prn_v.v[v], coef_v.v[v], lap);
                    } else if (mode_lap(MODE)) {
                        o = (TS)(d1 + d2 + d3);
                    } else if (MODE == MODE_D2X || MODE == MODE_D1X) {
                        o = (TS)d1;
                    } else if (MODE == MODE_D2Y || MODE == MODE_D1Y) {
                        o = (TS)d2;
                    } else {
                        o = (TS)d3;
                    }
                    res.v[v] = o;
                }
                TS *outp = reinterpret_cast<TS *>((ob + z * pstride) + (uint64_t)col);
                if (all_inner) {
                    if (a.nt)
                        vstore_nt<TS, VW>(outp, res);
                    else
                        *reinterpret_cast<V *>(outp) = res;
                } else {
#pragma unroll
                    for (int v = 0; v < VW; v++)
                        if (inner[v]) outp[v] = res.v[v];
                }
                if constexpr (FK) {
                    if (fold & FOLD_HALO) mirror_out(res, z);
                }
                if constexpr (MODE == MODE_AC2_IMG) {
                    // the imaging condition from the values just computed, in
                    // the order of k_imaging (the box lies inside the inner points)
                    const int64_t *ib = a.img_box;
                    if (z >= ib[4] && z <= ib[5] && i2 >= ib[2] && i2 <= ib[3] && i1 + VW - 1 >= ib[0] &&
                        i1 <= ib[1]) {
                        const V q0 = vload<TS, VW>(at(reinterpret_cast<const char *>(a.img_q[0]), z, col));
                        const V q1 = vload<TS, VW>(at(reinterpret_cast<const char *>(a.img_q[1]), z, col));
                        const V q2 = vload<TS, VW>(at(reinterpret_cast<const char *>(a.img_q[2]), z, col));
                        double *gp = a.img_grad + (z * a.n12 + i2 * n1 + i1);
#pragma unroll
                        for (int v = 0; v < VW; v++) {
                            if (i1 + v < ib[0] || i1 + v > ib[1]) continue;
                            const double d2 = ((double)q0.v[v] - 2.0 * (double)q1.v[v]) + (double)q2.v[v];
                            gp[v] += a.img_scale * (double)res.v[v] * d2;
                        }
                    }
                }
            }
            buf ^= 1;
        });
    }
}

// one workgroup: tile / chunk number 'logical' of the grid described by a
template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING_, bool FOLD = false>
__device__ __forceinline__ void tiled_body(const TiledArgs<TC> &a, int logical)
{
    const int tile = logical % (a.ntx * a.nty);
    const int chunk = logical / (a.ntx * a.nty);
    const int64_t z0 = a.s3 + (int64_t)chunk * a.chunk;
    const int64_t z1 = min(z0 + a.chunk, a.e3 + 1);    // planes [z0, z1)
    tiled_body_tz<TS, TC, R, VW, LX, BY, MODE, RING_, FOLD>(a, tile, z0, z1);
}

template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING_, bool FOLD = false>
__global__ void __launch_bounds__(LX *BY) k_tiled(TiledArgs<TC> a)
{
    const int nblk = a.ntx * a.nty * a.nchunk;
    const int logical = xcd_logical(nblk);
    if (logical >= nblk) return;                       // whole block leaves, before any barrier
    tiled_body<TS, TC, R, VW, LX, BY, MODE, RING_, FOLD>(a, logical);
}

// Many grids in one launch (the batched shots): off[s] = first logical block
// of grid s (off[nshots] = total); the workgroup finds its grid, then runs the
// same tile code on that grid's arguments.
template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING_>
__global__ void __launch_bounds__(LX *BY) k_tiled_batch(const TiledArgs<TC> *__restrict__ grids,
                                                      const int *__restrict__ off, int nshots)
{
    const int total = off[nshots];
    const int logical = xcd_logical(total);
    if (logical >= total) return;
    int s = 0;
    while (s + 1 < nshots && off[s + 1] <= logical) s++;
    const TiledArgs<TC> a = grids[s];
    tiled_body<TS, TC, R, VW, LX, BY, MODE, RING_>(a, logical - off[s]);
}

// The batched step as one persistent launch: as many workgroups as the device
// holds at once, each running its own list of column segments seg[so[g]] ..
// seg[so[g + 1] - 1] (shot, tile, planes [z0, z1)) one after the other.  The
// host cuts the columns of all shots, laid end to end, into equal numbers of
// planes per workgroup, so that a workgroup streams one or two long segments
// (2R planes of queue priming each) instead of many short chunks, and the
// launch has no second round.  Consecutive parts go to workgroups of one XCD
// (xcd_logical), so that tiles sharing halo rows share an L2.
struct BatchSeg {
    int shot, tile;
    int z0, z1;
};

template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING_>
__global__ void __launch_bounds__(LX *BY) k_tiled_batch_seg(const TiledArgs<TC> *__restrict__ grids,
                                                          const BatchSeg *__restrict__ seg, const int *__restrict__ so,
                                                          int parts)
{
    const int part = xcd_logical(parts);
    if (part >= parts) return;                          // whole block leaves, before any barrier
    for (int k = so[part]; k < so[part + 1]; k++) {
        const BatchSeg sg = seg[k];
        if (k > so[part]) __syncthreads();              // the previous segment's last plane is read
        tiled_body_tz<TS, TC, R, VW, LX, BY, MODE, RING_>(grids[sg.shot], sg.tile, sg.z0, sg.z1);
    }
}

// ---- launch configuration -------------------------------------------------------------------
// Tile shape: LX lanes along n1 (64 = one wave per row, 16/32 for narrow grids),
// BY rows along n2, chunk = planes per workgroup along n3, streams = cached or
// nontemporal prn / coef / output, halo ring on or off.  hsx_set_tile()
// overrides each of them (0 = the choice made here):
//  * BY = 16 with the halo ring where its LDS fits (LX >= 32), else BY = 8;
//  * 16-row tiles: one chunk per tile when the tiles fill the device once
//    (every workgroup starts together, so tiles that share halo rows stream
//    through L2 in step), nontemporal streams;
//  * 8-row tiles: short chunks, cached streams.
struct TileCfg {
    int lx, by, chunk, flags;
};
// the shape forced by hsx_set_tile (a copy, read under its lock); every launch
// takes its shape as a value, so that threads stepping different grids or
// tuning never see each other's choice
TileCfg tile_forced(void);            // defined in hsx_tiled_ac2.hip

// free surface folded into the next MODE_AC2 launch of this thread (hsx_ac2_step_fs)
struct FoldReq {
    int flags;           // FOLD_* bits, 0 = none
    int64_t fb[6];
};
extern thread_local FoldReq t_fold;   // defined in hsx_tiled_ac2.hip

// imaging condition fused into the next MODE_AC2_IMG launch of this thread
struct ImgReq {
    const void *q[3];
    double *grad;
    double scale;
    int64_t box[6];
};
extern thread_local ImgReq t_img;     // defined in hsx_tiled_ac2_img.hip
constexpr int TILE_LDS_MAX = 163840;  // gfx950: 160 KiB of LDS per workgroup

// 64 lanes unless the tiles of 64 lanes span more than 5/4 of the n1 width
// that tiles of 32 lanes span (their last tile mostly idle): 328^3 FP32 order
// 8 ran 0.125 ms with 32 x 16 tiles against 0.150 ms with 64 x 16 (512 lanes'
// width for 320 points); 608^3 and 808^3 stay faster with 64 lanes (0.75 /
// 1.81 ms against 0.93 / 2.22 ms) although a fifth of their lanes idle
inline int pick_lx(const hsx_grid *g, int vw, const TileCfg &cfg)
{
    if (cfg.lx) return cfg.lx;
    const int64_t N1 = g->inner[1] - g->inner[0] + 1;
    if (N1 <= 16 * vw) return 16;
    if (N1 <= 32 * vw) return 32;
    const int64_t w64 = (N1 + 64 * vw - 1) / (64 * vw) * (64 * vw);
    const int64_t w32 = (N1 + 32 * vw - 1) / (32 * vw) * (32 * vw);
    return 4 * w64 > 5 * w32 ? 32 : 64;
}

// 8-row tiles: chunks of ~48 planes (at least 4R), at least 2048 workgroups
inline int64_t chunk_short(int64_t N3, int R, int64_t tiles)
{
    int64_t len = 4 * R > 48 ? 4 * R : 48;
    int64_t nch = (N3 + len - 1) / len;
    if (tiles * nch < 2048) {
        nch = (2048 + tiles - 1) / tiles;
        len = (N3 + nch - 1) / nch;
    }
    if (len < 4 * R) len = 4 * R;
    return len < N3 ? len : N3;
}

// 16-row tiles and tiles with the halo ring: the chunk length c (at least 4R
// planes) that minimises rounds(c) x (c + 2R), rounds = the waves of resident
// workgroups that tiles x ceil(N3 / c) chunks need and 2R the planes a chunk
// loads before its first output; ties go to the longer chunk. When the tiles
// fill the device once this is one whole column per tile (1000^3 FP32 order 8:
// 252 tiles on 256 slots, 3.13 ms against 3.25 ms with 64-plane chunks); 700^3
// (132 tiles) gets 78-plane chunks, 1188 workgroups in five rounds: 1.22 ms,
// against 1.62 ms for two chunks per tile (264 workgroups: a second round of 8)
inline int64_t chunk_one_round(int64_t N3, int R, int64_t tiles, int64_t slots)
{
    const int64_t minc = 4 * R < N3 ? 4 * R : N3;
    int64_t best_c = N3, best_t = INT64_MAX;
    for (int64_t k = 1; k <= N3; k++) {
        const int64_t c = (N3 + k - 1) / k;
        if (c < minc) break;
        const int64_t chunks = (N3 + c - 1) / c;
        if (chunks != k) continue;                 // the same c as a smaller k
        const int64_t rounds = (tiles * chunks + slots - 1) / slots;
        const int64_t t = rounds * (c + 2 * R);
        if (t < best_t) {
            best_t = t;
            best_c = c;
        }
    }
    return best_c;
}

int device_cus(void);   // compute units of the current device (hsx_runtime.hip)

// resident workgroups per CU of a kernel instance (once per instantiation;
// also lifts the 64 KiB default cap on dynamic LDS where the tile needs more)
template <typename KERN>
inline int tiled_per_cu(KERN kern, int threads, size_t lds)
{
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, lds) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

// whether a kernel instance keeps more than a few per-lane values in scratch
// memory (register spills): at 1000^3 FP32 the order-4 64x16 instance spills
// 32-36 bytes per lane and is still the fastest shape (3.25 ms against 3.46 ms for
// 64x8), while spills of 80+ bytes (order 16) cost several times the launch
constexpr size_t TILED_SPILL_OK = 48;
template <typename KERN>
inline bool tiled_spills(KERN kern)
{
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(kern)) != hipSuccess) return false;
    return attr.localSizeBytes > TILED_SPILL_OK;
}

// n3-only kernels (D2 and D1 along n3: register queue, no LDS): the chunk
// length that gives about 1024 threads per CU in one round of workgroups
// (tiles x chunks <= CUs x 1024 / threads, as many chunks as fit), or 0
// when the tiles alone fill less than 3/4 of that round (then the general
// rule).  FD_D2 along n3, nontemporal streams, against the general rule:
// FP32 order 8 (64 x 16 tiles) 1000^3 1.71 / 1.84 ms, 500^3 0.201 / 0.233
// ms, 300^3 0.040 / 0.050 ms; FP64 order 8 (64 x 8) 1000^3 3.82 / 3.87 ms,
// 700^3 1.31 / 1.38 ms, 500^3 0.419 / 0.488 ms, 300^3 0.069 / 0.108 ms;
// FP32 order 16 1000^3 and 500^3 unchanged, 300^3 0.056 / 0.048 ms
inline int64_t chunk_stream(int64_t N3, int R, int64_t tiles, int threads)
{
    const int64_t want = (int64_t)device_cus() * 1024 / threads;
    const int64_t k = std::max<int64_t>(1, want / tiles);
    if (k == 1 && 4 * tiles < 3 * want) return 0;
    const int64_t c = (N3 + k - 1) / k;
    return std::min<int64_t>(std::max<int64_t>(c, 4 * R), N3);
}

// kernel arguments of one grid; slots = resident workgroups available to it
// (chunk > 0: n3 planes per workgroup, at least 4R, unless hsx_set_tile forces one)
template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING = false>
inline int fill_tiled_args(const hsx_grid *g, void *out, const void *prn_in, const void *u, const void *coef,
                           int64_t slots, const TileCfg &cfg, TiledArgs<TC> &a, int64_t chunk = 0)
{
    a.out = out;
    a.prn_in = prn_in ? prn_in : out;
    a.u = u;
    a.coef = coef;
    a.n1 = g->n1;
    a.n2 = g->n2;
    a.n12 = g->n1 * g->n2;
    if (a.n12 * (int64_t)sizeof(TS) >= ((int64_t)1 << 32))
        return fail(HSX_ERR_UNSUPPORTED, "tiled kernels: one n3 plane must stay below 4 GiB");
    a.s1 = g->inner[0];
    a.e1 = g->inner[1];
    a.s2 = g->inner[2];
    a.e2 = g->inner[3];
    a.s3 = g->inner[4];
    a.e3 = g->inner[5];
    a.t1s = (a.s1 / VW) * VW;
    const int64_t span1 = a.e1 - a.t1s + 1;
    a.ntx = (int)((span1 + LX * VW - 1) / (LX * VW));
    a.nty = (int)((a.e2 - a.s2 + 1 + BY - 1) / BY);
    const int64_t N3 = a.e3 - a.s3 + 1;
    const int64_t tiles = (int64_t)a.ntx * a.nty;
    constexpr bool STREAM = !Tiled<TS, TC, R, VW, LX, BY, MODE, RING>::NEED_LDS;
    const int64_t cs = STREAM && cfg.chunk <= 0 && chunk <= 0 ? chunk_stream(N3, R, tiles, LX * BY) : 0;
    const int64_t len = cfg.chunk > 0 ? std::min<int64_t>(cfg.chunk, N3)
                        : chunk > 0           ? std::min<int64_t>(std::max<int64_t>(chunk, 4 * R), N3)
                        : cs > 0              ? cs
                        : (BY == 16 || RING) ? chunk_one_round(N3, R, tiles, slots)
                                          : chunk_short(N3, R, tiles);
    a.chunk = (int)len;
    a.nchunk = (int)((N3 + len - 1) / len);
    const int streams = cfg.flags & (HSX_TILE_CACHED | HSX_TILE_NT);
    a.nt = streams == HSX_TILE_NT || (streams == 0 && (BY == 16 || RING || STREAM));
    a.fold = MODE == MODE_AC2 ? t_fold.flags : 0;   // the folded free surface: exact step only
    for (int k = 0; k < 6; k++) a.fb[k] = a.fold ? (int)t_fold.fb[k] : -1;
    if constexpr (MODE == MODE_AC2_IMG) {
        for (int k = 0; k < 3; k++) a.img_q[k] = t_img.q[k];
        a.img_grad = t_img.grad;
        a.img_scale = t_img.scale;
        for (int k = 0; k < 6; k++) a.img_box[k] = t_img.box[k];
    } else {
        for (int k = 0; k < 3; k++) a.img_q[k] = nullptr;
        a.img_grad = nullptr;
        a.img_scale = 0.0;
        for (int k = 0; k < 6; k++) a.img_box[k] = 0;
    }
    a.c = make_coefs<TC>(g);
    if (tiles * a.nchunk > (int64_t)1 << 30) return fail(HSX_ERR_UNSUPPORTED, "too many tiles");
    return HSX_OK;
}

template <typename TS, typename TC, int R, int VW, int LX, int BY, int MODE, bool RING, bool FOLD = false>
inline int launch_tiled(const hsx_grid *g, void *out, const void *prn_in, const void *u, const void *coef, hipStream_t s,
                        const TileCfg &cfg)
{
    using K = Tiled<TS, TC, R, VW, LX, BY, MODE, RING>;
    static_assert(K::LDS <= TILE_LDS_MAX, "LDS tile above the workgroup limit");
    auto kern = k_tiled<TS, TC, R, VW, LX, BY, MODE, RING, FOLD>;
    static const int per_cu = tiled_per_cu(kern, LX * BY, K::LDS);
    TiledArgs<TC> a;
    int rc = fill_tiled_args<TS, TC, R, VW, LX, BY, MODE, RING>(g, out, prn_in, u, coef, (int64_t)per_cu * device_cus(),
                                                                cfg, a);
    if (rc) return rc;
    const int64_t nblk = (int64_t)a.ntx * a.nty * a.nchunk;
    const unsigned grid = (unsigned)(((nblk + 7) / 8) * 8);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(LX, BY), K::LDS, s, a);
    HSX_LAUNCHED("k_tiled");
    return HSX_OK;
}

// 16-row instance to use at LX lanes: 1 with the halo ring, 2 without, 0 none.
// 1024-thread tiles leave 128 VGPRs per lane, too few for FP64 values with an
// n3 register queue (the kernels without one, D2 and D1 along n1 and n2, keep
// few); by default an instance is used only if it keeps its values in
// registers (no per-lane scratch); rows forced to 16 are used as asked
template <typename TS, typename TC, int R, int VW, int LX, int MODE>
struct TallFits {
    static constexpr bool fits_regs = (sizeof(TS) == 4 && sizeof(TC) == 4) || LX * 16 <= 512 ||
                                      !Tiled<TS, TC, R, VW, LX, 16, MODE, false>::NEED_Z;
    static constexpr bool ring = LX >= 32 && fits_regs && Tiled<TS, TC, R, VW, LX, 16, MODE, true>::RING &&
                                 Tiled<TS, TC, R, VW, LX, 16, MODE, true>::LDS <= TILE_LDS_MAX;
    static constexpr bool tall = LX >= 32 && fits_regs &&
                                 Tiled<TS, TC, R, VW, LX, 16, MODE, false>::LDS <= TILE_LDS_MAX;
};

template <typename TS, typename TC, int R, int VW, int LX, int MODE, bool FOLD = false>
inline int tall_variant(const TileCfg &cfg)
{
    constexpr bool ring_ok = TallFits<TS, TC, R, VW, LX, MODE>::ring;
    constexpr bool tall_ok = TallFits<TS, TC, R, VW, LX, MODE>::tall;
    const bool forced = cfg.by == 16;
    if (cfg.by == 8) return 0;
    if constexpr (ring_ok) {
        static const bool ring_spills = tiled_spills(k_tiled<TS, TC, R, VW, LX, 16, MODE, true, FOLD>);
        if (!(cfg.flags & HSX_TILE_NO_RING) && (forced || !ring_spills)) return 1;
    }
    if constexpr (tall_ok) {
        static const bool tall_spills = tiled_spills(k_tiled<TS, TC, R, VW, LX, 16, MODE, false, FOLD>);
        if (forced || !tall_spills) return 2;
    }
    return 0;
}

// BY = 16 (with the halo ring where it fits) for tiles of at least 32 lanes, else 8
template <typename TS, typename TC, int R, int VW, int LX, int MODE, bool FOLD = false>
inline int launch_tiled_by(const hsx_grid *g, void *out, const void *prn_in, const void *u, const void *coef, hipStream_t s,
                           const TileCfg &cfg)
{
    const int tall = tall_variant<TS, TC, R, VW, LX, MODE, FOLD>(cfg);
    if constexpr (TallFits<TS, TC, R, VW, LX, MODE>::ring) {
        if (tall == 1) return launch_tiled<TS, TC, R, VW, LX, 16, MODE, true, FOLD>(g, out, prn_in, u, coef, s, cfg);
    }
    if constexpr (TallFits<TS, TC, R, VW, LX, MODE>::tall) {
        if (tall == 2) return launch_tiled<TS, TC, R, VW, LX, 16, MODE, false, FOLD>(g, out, prn_in, u, coef, s, cfg);
    }
    // 8 rows: the ring by default for FP64 storage only (1000^3 fused step,
    // 64 x 8: FP64 6.88 / 7.26 / 7.48 ms at orders 4 / 8 / 12 with the ring and
    // nontemporal streams, 6.99 / 7.36 / 7.78 without; FP32 order 8 3.83 ms
    // with against 3.66 ms without, FP32 storage with FP64 compute 3.89 against
    // 3.70: there its LDS halves the resident workgroups while the 8-row halo
    // stays large); otherwise only when 8 rows are asked for
    constexpr bool ring8_ok = Tiled<TS, TC, R, VW, LX, 8, MODE, true>::RING &&
                              Tiled<TS, TC, R, VW, LX, 8, MODE, true>::LDS <= TILE_LDS_MAX;
    if constexpr (ring8_ok && !FOLD) {
        const bool ring8 = cfg.by == 8 || (cfg.by == 0 && sizeof(TS) == 8);
        if (ring8 && !(cfg.flags & HSX_TILE_NO_RING))
            return launch_tiled<TS, TC, R, VW, LX, 8, MODE, true>(g, out, prn_in, u, coef, s, cfg);
    }
    return launch_tiled<TS, TC, R, VW, LX, 8, MODE, false, FOLD>(g, out, prn_in, u, coef, s, cfg);
}

// alignment and size conditions of the tiled kernels
inline int tiled_fits(const hsx_grid *g, int vw_elems, size_t elem)
{
    if (g->dim != 3) return fail(HSX_ERR_UNSUPPORTED, "tiled kernels need a 3D grid");
    if (g->n1 % vw_elems) return fail(HSX_ERR_UNSUPPORTED, "n1 %% %d != 0", vw_elems);
    (void)elem;
    return HSX_OK;
}

// shape chosen by hsx_autotune for this grid and operation, if any
bool tuned_shape(const hsx_grid *g, int mode, TileCfg *out);

// whether the 128 x 8 tiles exist for g (FP32, the ring's LDS fits, spills
// within TILED_SPILL_OK): elsewhere a request for them runs the library's
// choice (hsx_tiled_ac2.hip, which holds their instances)
bool wide_tiles_fit(const hsx_grid *g);

// The fused step's own choice of 128 x 8 tiles (with the ring) over 64 x 16:
// where both span the same width and either one 512-point tile covers the row
// or the row fills at least 93 % of the 128-lane tiles. Order 8 FP32, median
// of 48 launches each on one device (128 x 8 / 64 x 16): 400^3 0.958, 460^3
// 0.894, 500^3 0.911, 512^3 0.967, 960^3 0.992, 1000^3 0.966, 1020^3 0.988;
// not taken at 600^3 1.202, 800^3 1.032, 900^3 1.016, 1100^3 1.053.
inline bool wide_by_default(const hsx_grid *g)
{
    const int64_t N1 = g->inner[1] - g->inner[0] + 1;
    const int64_t w64 = (N1 + 255) / 256 * 256, w128 = (N1 + 511) / 512 * 512;
    return w64 == w128 && (w128 == 512 || 100 * N1 >= 93 * w128);
}

template <int MODE, bool FOLD = false>
inline int run_tiled_shape(const hsx_grid *g, void *out, const void *prn_in, const void *u, const void *coef,
                           hipStream_t s, const TileCfg &cfg_in)
{
    const int R = g->fd_order / 2;
    return dispatch_types(g, [&](auto ts, auto tc) -> int {
        using TS = decltype(ts);
        using TC = decltype(tc);
        constexpr int VW = 16 / sizeof(TS);
        int rc = tiled_fits(g, VW, sizeof(TS));
        if (rc) return rc;
        if (((uintptr_t)out | (uintptr_t)prn_in | (uintptr_t)u | (uintptr_t)coef) % 16)
            return fail(HSX_ERR_UNSUPPORTED, "buffers not 16-byte aligned");
        TileCfg cfg = cfg_in;
        int lx = pick_lx(g, VW, cfg);
        // FP32 defaults measured per size (HIP events, kernel alone): the fused
        // step and the order-4 Laplacian take 128 x 8 tiles where
        // wide_by_default holds (Laplacian order 4: 1000^3 1.54 / 1.61 ms,
        // 500^3 0.189 / 0.196, 400^3 0.119 / 0.122; 700^3, not taken, 0.671 /
        // 0.587); the order-8 Laplacian drops the halo ring, whose halo rows
        // still meet in L2 with two streams instead of four (1000^3 1.84 /
        // 1.88 ms, 800^3 1.11 / 1.24, 700^3 0.70 / 0.79, 500^3 0.237 / 0.258,
        // 400^3 0.131 / 0.141)
        constexpr bool fp32_vals = sizeof(TS) == 4 && sizeof(TC) == 4;
        const bool dflt = cfg.lx == 0 && cfg.by == 0;
        if (fp32_vals && dflt && wide_by_default(g) &&
            (MODE == MODE_AC2 ? wide_tiles_fit(g) : MODE == MODE_LAP && R == 2))
            lx = 128;
        if (fp32_vals && dflt && cfg.flags == 0 && MODE == MODE_LAP && R == 4) cfg.flags = HSX_TILE_NO_RING;
        // FP64 Laplacian: order 12 on 32 x 16 tiles (1000^3 4.54 / 4.71 ms,
        // 800^3 2.45 / 2.71, 700^3 1.65 / 1.93, 500^3 0.535 / 0.622, 400^3
        // 0.363 / 0.425 against 64 x 8), order 16 on 16 x 8 (1000^3 5.55 /
        // 6.49 ms, 800^3 3.04 / 3.47, 700^3 2.08 / 2.36, 500^3 0.802 / 0.864,
        // 400^3 0.426 / 0.486 against 32 x 16)
        if (sizeof(TS) == 8 && dflt && MODE == MODE_LAP && lx == 64) {
            if (R == 6) {
                lx = 32;
                cfg.by = 16;
            } else if (R == 8) {
                lx = 16;
            }
        }
        return dispatch_R(R, [&](auto rc_) -> int {
            constexpr int RR = decltype(rc_)::value;
            if (lx == 128) {
                // 128 lanes x 8 rows with the halo ring (rows of 2 KiB, two
                // waves each): the exact FP32 fused step and Laplacian up to
                // order 8, where the ring's LDS fits and spills stay within
                // TILED_SPILL_OK; the fused step's default where
                // wide_by_default holds, otherwise on request (hsx_set_tile,
                // the autotuner). 1000^3 order 8: 3.04-3.08 ms against
                // 3.12-3.19 ms for 64 x 16; 500^3 0.363 against 0.394 ms;
                // 700^3 and 800^3 6 % and 2 % slower.
                using K = Tiled<TS, TC, RR, VW, 128, 8, MODE, true>;
                if constexpr (sizeof(TS) == 4 && sizeof(TC) == 4 && !FOLD && (MODE == MODE_AC2 || MODE == MODE_LAP) &&
                              K::RING && K::LDS <= TILE_LDS_MAX) {
                    static const bool spills = tiled_spills(k_tiled<TS, TC, RR, VW, 128, 8, MODE, true>);
                    if (!spills && !(cfg.flags & HSX_TILE_NO_RING))
                        return launch_tiled<TS, TC, RR, VW, 128, 8, MODE, true>(g, out, prn_in, u, coef, s, cfg);
                }
                cfg.lx = 0;     // elsewhere: the library's choice of lanes and rows
                cfg.by = 0;
                lx = pick_lx(g, VW, cfg);
            }
            if (lx == 16)
                return launch_tiled<TS, TC, RR, VW, 16, 8, MODE, false, FOLD>(g, out, prn_in, u, coef, s, cfg);
            // 64 lanes without a 16-row instance that keeps its values in
            // registers: FP64 values up to order 12 take 64 x 8 tiles, the rest
            // (order 16 FP32, orders 14-16 in FP64) 32 lanes x 16 rows.
            // 1000^3 FP64 order 8: 64 x 8 7.26 ms, 32 x 16 7.96 ms; order 16:
            // 7.73 against 7.47 ms. FP32 storage with FP64 compute, order 8:
            // 3.70 against 4.27 ms; FP32 order 16: 4.09 against 3.90 ms (and
            // 4.50 ms for 64 x 16 with 8-byte lanes, which fits in registers
            // but issues twice the loads and LDS reads per point)
            // (the kernels without an n3 queue, D2 and D1 along n1 and n2,
            // take 64 x 16 in FP64 where it keeps its values in registers,
            // else 32 x 16: 1000^3 FP64 D2 along n2 3.41 ms at order 8 and
            // 4.27 ms at order 12 against 4.05 and 4.70 ms for 64 x 8)
            constexpr bool fp64_vals = !(sizeof(TS) == 4 && sizeof(TC) == 4);
            constexpr bool queue = Tiled<TS, TC, RR, VW, 64, 16, MODE, false>::NEED_Z;
            const bool narrow = lx == 64 && cfg.lx == 0 && !(fp64_vals && RR <= 6 && queue) &&
                                tall_variant<TS, TC, RR, VW, 64, MODE, FOLD>(cfg) == 0 &&
                                tall_variant<TS, TC, RR, VW, 32, MODE, FOLD>(cfg) != 0;
            if (lx == 32 || narrow)
                return launch_tiled_by<TS, TC, RR, VW, 32, MODE, FOLD>(g, out, prn_in, u, coef, s, cfg);
            return launch_tiled_by<TS, TC, RR, VW, 64, MODE, FOLD>(g, out, prn_in, u, coef, s, cfg);
        });
    });
}

// the forced shape (hsx_set_tile) if any, else the tuned one, else the defaults
template <int MODE, bool FOLD = false>
inline int run_tiled(const hsx_grid *g, void *out, const void *prn_in, const void *u, const void *coef, hipStream_t s)
{
    const TileCfg forced = tile_forced();
    TileCfg tuned;
    const bool use_tuned = forced.lx == 0 && forced.by == 0 && forced.chunk == 0 && forced.flags == 0 &&
                           tuned_shape(g, mode_base(MODE), &tuned);
    return run_tiled_shape<MODE, FOLD>(g, out, prn_in, u, coef, s, use_tuned ? tuned : forced);
}

} // namespace hsx

#endif
