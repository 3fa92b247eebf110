// Modeling kernels: point sources and receivers (the reference's
// DataAcquisition, src/data_Acquisition.cpp:58-305), the imaging condition of
// the adjoint, and the low-precision checkpoint codec.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>

#include "hsx_internal.h"

namespace hsx {

// ---- sources / receivers ------------------------------------------------------------
// u[idx[i]] (+)= scale * val[i] * (w ? w[i] : 1); indices are distinct within one call;
// an index outside [0, npoint) is skipped (the host checks them where it builds them)
template <typename TS>
__global__ void __launch_bounds__(256) k_inject(TS *u, int64_t npoint, int64_t n, const int64_t *idx, const TS *val,
                                                const TS *w, double scale, int mode)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if ((uint64_t)idx[i] >= (uint64_t)npoint) continue;
        double v = scale * (double)val[i];
        if (w) v *= (double)w[i];
        TS *p = u + idx[i];
        *p = mode == HSX_INJECT_SET ? (TS)v : (TS)((double)*p + v);
    }
}

// out[i] = u[idx[i]]; 0 for an index outside [0, npoint)
template <typename TS>
__global__ void __launch_bounds__(256) k_gather(const TS *u, int64_t npoint, int64_t n, const int64_t *idx, TS *out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx[i];
        out[i] = (uint64_t)j < (uint64_t)npoint ? u[j] : (TS)0;
    }
}

// ---- imaging condition -----------------------------------------------------------
// grad[ii] += scale * mu[ii] * (un[ii] - 2 u[ii] + up[ii] - f(ii)) in double over
// the box, f = src_val at src_idx (the source term of that step), 0 elsewhere.
template <typename TS>
__global__ void __launch_bounds__(256) k_imaging(double *grad, const TS *mu, const TS *un, const TS *u, const TS *up,
                                                 Box b, int64_t n1, int64_t n12, double scale, int64_t src_idx,
                                                 double src_val)
{
    const int64_t N1 = b.n1(), N2 = b.n2(), rows = N2 * b.n3();
    for (int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.y; row < rows; row += (int64_t)gridDim.x * 4) {
        const int64_t base = (b.s2 + row % N2) * n1 + (b.s3 + row / N2) * n12 + b.s1;
        for (int64_t k = threadIdx.x; k < N1; k += 64) {
            const int64_t ii = base + k;
            double d2 = ((double)un[ii] - 2.0 * (double)u[ii]) + (double)up[ii];
            if (ii == src_idx) d2 -= src_val;
            grad[ii] += scale * (double)mu[ii] * d2;
        }
    }
}

// ---- checkpoint codec ------------------------------------------------------------
// A snapshot of n values is stored as n 16-bit codes after a 256-byte header
// holding the scale (float) and the max |value| (float):
//   code = round_to_format(v / scale), v = code * scale,
//   scale = maxabs / FORMAT_TARGET (1 if the snapshot is zero).
// Pair snapshots hold u_n and d = u_n - u_{n-1} (each with its own scale), so
// that the small step-to-step difference keeps its relative precision.
constexpr size_t CKPT_HEADER = 256;

__device__ __forceinline__ float fmt_target(int fmt) { return fmt == HSX_CKPT_FP16 ? 16384.f : 1.f; }

__device__ __forceinline__ uint16_t enc16(float x, int fmt)
{
    if (fmt == HSX_CKPT_FP16) return __half_as_ushort(__float2half_rn(x));
    return __bfloat16_as_ushort(__float2bfloat16(x));     // round to nearest even
}

__device__ __forceinline__ float dec16(uint16_t c, int fmt)
{
    if (fmt == HSX_CKPT_FP16) return __half2float(__ushort_as_half(c));
    return __bfloat162float(__ushort_as_bfloat16(c));
}

// max |a - (b ? b : 0)| into hdr[1] (as float bits; values are >= 0 so the
// unsigned order of the bits is the float order)
__global__ void __launch_bounds__(256) k_ckpt_maxabs(const float *a, const float *b, int64_t n, unsigned *hdr_bits)
{
    float m = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = b ? a[i] - b[i] : a[i];
        m = fmaxf(m, fabsf(v));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    __shared__ float wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) m = fmaxf(m, wm[k]);
        atomicMax(hdr_bits + 1, __float_as_uint(m));
    }
}

__global__ void __launch_bounds__(256) k_ckpt_encode(const float *a, const float *b, int64_t n, char *packed, int fmt)
{
    float *hdr = reinterpret_cast<float *>(packed);
    const float maxabs = hdr[1];
    const float scale = maxabs > 0.f ? maxabs / fmt_target(fmt) : 1.f;
    const float inv = 1.f / scale;
    uint16_t *code = reinterpret_cast<uint16_t *>(packed + CKPT_HEADER);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = b ? a[i] - b[i] : a[i];
        code[i] = enc16(v * inv, fmt);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) hdr[0] = scale;
}

// out = decoded (mode 0) or out = base - decoded (mode 1)
__global__ void __launch_bounds__(256) k_ckpt_decode(const char *packed, int64_t n, int fmt, const float *base,
                                                     float *out)
{
    const float scale = reinterpret_cast<const float *>(packed)[0];
    const uint16_t *code = reinterpret_cast<const uint16_t *>(packed + CKPT_HEADER);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = dec16(code[i], fmt) * scale;
        out[i] = base ? base[i] - v : v;
    }
}

__global__ void k_ckpt_clear(unsigned *hdr_bits)
{
    if (threadIdx.x < 2) hdr_bits[threadIdx.x] = 0u;
}

static unsigned blocks_for(int64_t n)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

// k_ckpt_maxabs ends each block with one atomic on the same word: 512 blocks
// keep those few (8192 of them took ~97 us on a 208^3 grid, the pass itself ~10)
static unsigned blocks_for_max(int64_t n)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 512));
}

} // namespace hsx

using namespace hsx;

extern "C" {

// header + codes, rounded up to whole 256-byte units: snapshots stored back to
// back keep every header 256-byte aligned (its max |value| is updated with a
// 32-bit atomic)
size_t hsx_ckpt_bytes(int64_t n) { return CKPT_HEADER + ((size_t)n * 2 + CKPT_HEADER - 1) / CKPT_HEADER * CKPT_HEADER; }

int hsx_inject(const hsx_grid *g, void *u, int64_t n, const int64_t *idx, const void *val, const void *w,
               double scale, int mode, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if ((rc = check_ptr(u, "u")) || (rc = check_ptr(idx, "idx")) || (rc = check_ptr(val, "val"))) return rc;
    if (n < 0) return fail(HSX_ERR_INVALID, "n = %lld", (long long)n);
    if (mode != HSX_INJECT_ADD && mode != HSX_INJECT_SET) return fail(HSX_ERR_INVALID, "inject mode %d", mode);
    if ((rc = ensure_init())) return rc;
    if (n == 0) return HSX_OK;
    hipStream_t s = stream_of(stream);
    return dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        hipLaunchKernelGGL(k_inject<TS>, dim3(blocks_for(n)), dim3(256), 0, s, (TS *)u, g->n1 * g->n2 * g->n3, n, idx,
                           (const TS *)val,
                           (const TS *)w, scale, mode);
        HSX_LAUNCHED("k_inject");
        return HSX_OK;
    });
}

int hsx_gather(const hsx_grid *g, const void *u, int64_t n, const int64_t *idx, void *out, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if ((rc = check_ptr(u, "u")) || (rc = check_ptr(idx, "idx")) || (rc = check_ptr(out, "out"))) return rc;
    if (n < 0) return fail(HSX_ERR_INVALID, "n = %lld", (long long)n);
    if ((rc = ensure_init())) return rc;
    if (n == 0) return HSX_OK;
    hipStream_t s = stream_of(stream);
    return dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        hipLaunchKernelGGL(k_gather<TS>, dim3(blocks_for(n)), dim3(256), 0, s, (const TS *)u, g->n1 * g->n2 * g->n3, n,
                           idx, (TS *)out);
        HSX_LAUNCHED("k_gather");
        return HSX_OK;
    });
}

int hsx_imaging(const hsx_grid *g, double *grad, const void *mu, const void *u_next, const void *u, const void *u_prev,
                const int64_t range[6], double scale, int64_t src_idx, double src_val, void *stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if ((rc = check_range(g, range))) return rc;
    if ((rc = check_ptr(grad, "grad")) || (rc = check_ptr(mu, "mu")) || (rc = check_ptr(u_next, "u_next")) ||
        (rc = check_ptr(u, "u")) || (rc = check_ptr(u_prev, "u_prev")))
        return rc;
    if ((rc = ensure_init())) return rc;
    const Box b = box_of(range);
    hipStream_t s = stream_of(stream);
    const int64_t rows = b.n2() * b.n3();
    const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((rows + 3) / 4, 1 << 20));
    return dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        hipLaunchKernelGGL(k_imaging<TS>, dim3(nb), dim3(64, 4), 0, s, grad, (const TS *)mu, (const TS *)u_next,
                           (const TS *)u, (const TS *)u_prev, b, g->n1, g->n1 * g->n2, scale, src_idx, src_val);
        HSX_LAUNCHED("k_imaging");
        return HSX_OK;
    });
}

int hsx_ckpt_encode(const float *u, const float *u_prev, int64_t n, void *packed, int format, void *stream)
{
    int rc;
    if ((rc = check_ptr(u, "u")) || (rc = check_ptr(packed, "packed"))) return rc;
    if (format != HSX_CKPT_FP16 && format != HSX_CKPT_BF16) return fail(HSX_ERR_INVALID, "format %d", format);
    if (n <= 0) return fail(HSX_ERR_INVALID, "n = %lld", (long long)n);
    if ((uintptr_t)packed % 16) return fail(HSX_ERR_INVALID, "packed snapshot not 16-byte aligned");
    if ((rc = ensure_init())) return rc;
    hipStream_t s = stream_of(stream);
    unsigned *hdr = reinterpret_cast<unsigned *>(packed);
    hipLaunchKernelGGL(k_ckpt_clear, dim3(1), dim3(64), 0, s, hdr);
    HSX_LAUNCHED("k_ckpt_clear");
    hipLaunchKernelGGL(k_ckpt_maxabs, dim3(blocks_for_max(n)), dim3(256), 0, s, u, u_prev, n, hdr);
    HSX_LAUNCHED("k_ckpt_maxabs");
    hipLaunchKernelGGL(k_ckpt_encode, dim3(blocks_for(n)), dim3(256), 0, s, u, u_prev, n, (char *)packed, format);
    HSX_LAUNCHED("k_ckpt_encode");
    return HSX_OK;
}

int hsx_ckpt_decode(const void *packed, int64_t n, int format, const float *base, float *out, void *stream)
{
    int rc;
    if ((rc = check_ptr(packed, "packed")) || (rc = check_ptr(out, "out"))) return rc;
    if (format != HSX_CKPT_FP16 && format != HSX_CKPT_BF16) return fail(HSX_ERR_INVALID, "format %d", format);
    if (n <= 0) return fail(HSX_ERR_INVALID, "n = %lld", (long long)n);
    if ((uintptr_t)packed % 16) return fail(HSX_ERR_INVALID, "packed snapshot not 16-byte aligned");
    if ((rc = ensure_init())) return rc;
    hipStream_t s = stream_of(stream);
    hipLaunchKernelGGL(k_ckpt_decode, dim3(blocks_for(n)), dim3(256), 0, s, (const char *)packed, n, format, base,
                       out);
    HSX_LAUNCHED("k_ckpt_decode");
    return HSX_OK;
}

} // extern "C"
