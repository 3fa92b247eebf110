// Forward modeling of a shot and the gradient of the receiver misfit with
// respect to squared slowness (adjoint of the Ac2 propagator with
// checkpointing).  The reference has no adjoint (its ADJOINT mode is a stub,
// src/data_Acquisition.cpp:307-310); the forward step is the reference's
// Propagator_Ac2 step (src/propagator_Ac2.cpp:366-393) plus a point source.
//
// Discrete problem, with q^n the wavefield after the free surface B:
//   q^{n+1} = B( 2 q^n - q^{n-1} + c (L q^n) + F^n ),   c = dt^2 / m,
//   d^n_r = q^{n+1}(x_r),   J = 1/2 sum_{n,r} (d^n_r - dobs^n_r)^2,
// L the FD Laplacian with the antisymmetric halos of B, which is symmetric on
// the points B does not zero.  The adjoint field is carried as mu = c lambda:
//   mu^k = 2 mu^{k+1} - mu^{k+2} + c (L mu^{k+1}) + c e^k,  e^k = residual at step k-1,
// i.e. the forward step itself (free surface + fused kernel) with the residual
// injected at the receivers, scaled by c there.  The imaging condition is
//   dJ/dm = -(1/dt^2) sum_k mu^k (q^k - 2 q^{k-1} + q^{k-2} - F^{k-1}),
// accumulated in double over the points B does not touch.
#include <algorithm>
#include <vector>

#include "hsx_internal.h"

namespace hsx {

// res = d - dobs; partial[block] = sum of res^2 over the block (double)
template <typename TS>
__global__ void __launch_bounds__(256) k_residual(const TS *d, const TS *dobs, TS *res, int64_t n, double *partial)
{
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const TS r = d[i] - dobs[i];
        res[i] = r;
        acc += (double)r * (double)r;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

__global__ void k_sum_partials(const double *partial, int n, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += partial[i];
        out[0] = s;
    }
}

// grad[src_i] += scale * mu[src_i] * f_i (the source term of the imaging condition)
template <typename TS>
__global__ void k_imaging_src(double *grad, const TS *mu, int64_t nsrc, const int64_t *idx, const TS *f, double scale)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsrc) grad[idx[i]] += scale * (double)mu[idx[i]] * (double)f[i];
}

struct Dev {
    void *p = nullptr;
    ~Dev()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes)
    {
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        return e == hipSuccess ? HSX_OK : hip_fail(e, "hipMalloc (adjoint buffers)");
    }
    char *c() const { return (char *)p; }
};

static int shot_checks(const hsx_grid *g, const void *coef, const hsx_shot *sh)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if ((rc = check_ptr(coef, "coef")) || (rc = check_ptr(sh, "shot"))) return rc;
    if (sh->nt < 1) return fail(HSX_ERR_INVALID, "shot: nt = %lld", (long long)sh->nt);
    if (sh->nsrc < 0 || sh->nrec < 0) return fail(HSX_ERR_INVALID, "shot: negative source / receiver count");
    if (sh->nsrc > 0 && ((rc = check_ptr(sh->src_idx, "src_idx")) || (rc = check_ptr(sh->wavelet, "wavelet"))))
        return rc;
    if (sh->nrec > 0 && (rc = check_ptr(sh->rec_idx, "rec_idx"))) return rc;
    if (!(sh->dt > 0)) return fail(HSX_ERR_INVALID, "shot: dt must be > 0");
    return ensure_init();
}

static int zero(const hsx_grid *g, void *u, hipStream_t s)
{
    return hsx_array_op(g, HSX_ARRAY_FILL, u, nullptr, nullptr, 0.0, s);
}

// one forward step: prev <- B(cur) step, source, record
static int fwd_step(const hsx_grid *g, const void *coef, const hsx_shot *sh, void *out, const void *prev, void *cur,
                    int64_t n, void *traces, size_t es, hipStream_t s)
{
    int rc = hsx_free_surface(g, cur, 0x3F, s);
    if (!rc) rc = hsx_ac2_step_to(g, out, prev, cur, coef, s);
    if (!rc && sh->nsrc > 0)
        rc = hsx_inject(g, out, sh->nsrc, sh->src_idx, (const char *)sh->wavelet + (size_t)n * sh->nsrc * es, nullptr,
                        1.0, HSX_INJECT_ADD, s);
    if (!rc && sh->nrec > 0 && traces)
        rc = hsx_gather(g, out, sh->nrec, sh->rec_idx, (char *)traces + (size_t)n * sh->nrec * es, s);
    return rc;
}

} // namespace hsx

using namespace hsx;

extern "C" {

int hsx_forward(const hsx_grid *g, const void *coef, const hsx_shot *sh, void *traces, void *stream)
{
    int rc = shot_checks(g, coef, sh);
    if (rc) return rc;
    if ((rc = check_ptr(traces, "traces"))) return rc;
    const size_t es = g->storage == HSX_FP64 ? 8 : 4;
    const size_t bytes = (size_t)(g->n1 * g->n2 * g->n3) * es;
    hipStream_t s = stream_of(stream);
    Dev a, b;
    if ((rc = a.alloc(bytes)) || (rc = b.alloc(bytes))) return rc;
    if ((rc = zero(g, a.p, s)) || (rc = zero(g, b.p, s))) return rc;
    void *prev = a.p, *cur = b.p;
    for (int64_t n = 0; n < sh->nt; n++) {
        if ((rc = fwd_step(g, coef, sh, prev, prev, cur, n, traces, es, s))) return rc;
        std::swap(prev, cur);
    }
    HSX_HIP(hipStreamSynchronize(s));   // the buffers are freed on return
    return HSX_OK;
}

int hsx_gradient(const hsx_grid *g, const void *coef, const hsx_shot *sh, const void *d_obs,
                 const hsx_adjoint_opts *opt, double *grad, double *misfit, void *stream)
{
    int rc = shot_checks(g, coef, sh);
    if (rc) return rc;
    if ((rc = check_ptr(d_obs, "d_obs")) || (rc = check_ptr(grad, "grad")) || (rc = check_ptr(opt, "opts")))
        return rc;
    const int K = opt->ckpt_every;
    const int fmt = opt->ckpt_format;
    if (K < 0) return fail(HSX_ERR_INVALID, "ckpt_every %d < 0", K);
    if (fmt != HSX_CKPT_FP32 && fmt != HSX_CKPT_FP16 && fmt != HSX_CKPT_BF16)
        return fail(HSX_ERR_INVALID, "checkpoint format %d", fmt);
    if (fmt != HSX_CKPT_FP32 && g->storage != HSX_FP32)
        return fail(HSX_ERR_UNSUPPORTED, "16-bit checkpoints of FP64 wavefields");
    if (K == 0 && fmt != HSX_CKPT_FP32) return fail(HSX_ERR_INVALID, "full storage keeps every step in the storage type");

    const size_t es = g->storage == HSX_FP64 ? 8 : 4;
    const int64_t npoint = g->n1 * g->n2 * g->n3;
    const size_t bytes = (size_t)npoint * es;
    const int64_t nt = sh->nt, nrec = sh->nrec, ntr = nt * nrec;
    hipStream_t s = stream_of(stream);

    // imaging range: the inner points minus the boundary planes zeroed by B
    int64_t rim[6];
    for (int a = 0; a < 3; a++) {
        rim[2 * a] = g->inner[2 * a] + (a < g->dim ? 1 : 0);
        rim[2 * a + 1] = g->inner[2 * a + 1] - (a < g->dim ? 1 : 0);
        if (rim[2 * a + 1] < rim[2 * a]) return fail(HSX_ERR_INVALID, "grid too small for the imaging range");
    }
    const double img_scale = -1.0 / (sh->dt * sh->dt);

    // ---- buffers
    Dev traces, res, crec, partial, msum, m1, m2;
    if ((rc = traces.alloc(ntr * es)) || (rc = res.alloc(ntr * es)) || (rc = crec.alloc(nrec * es)) ||
        (rc = partial.alloc(4096 * sizeof(double))) || (rc = msum.alloc(sizeof(double))) || (rc = m1.alloc(bytes)) ||
        (rc = m2.alloc(bytes)))
        return rc;
    const int64_t nseg = K == 0 ? 0 : (nt + K - 1) / K;
    const size_t snap = fmt == HSX_CKPT_FP32 ? bytes : hsx_ckpt_bytes(npoint);
    std::vector<Dev> q(K == 0 ? nt + 2 : K + 2);   // wavefields q^{m}, m from the segment start - 1
    for (auto &x : q)
        if ((rc = x.alloc(bytes))) return rc;
    Dev ck;
    if (K > 0 && (rc = ck.alloc((size_t)nseg * 2 * snap))) return rc;
    Dev fa, fb;   // forward pair when checkpointing
    if (K > 0 && ((rc = fa.alloc(bytes)) || (rc = fb.alloc(bytes)))) return rc;

    auto save = [&](int64_t seg, const void *cur, const void *prev) -> int {
        char *dst = ck.c() + (size_t)seg * 2 * snap;
        if (fmt == HSX_CKPT_FP32) {
            HSX_HIP(hipMemcpyAsync(dst, cur, bytes, hipMemcpyDeviceToDevice, s));
            HSX_HIP(hipMemcpyAsync(dst + snap, prev, bytes, hipMemcpyDeviceToDevice, s));
            return HSX_OK;
        }
        int r = hsx_ckpt_encode((const float *)cur, nullptr, npoint, dst, fmt, s);
        if (!r) r = hsx_ckpt_encode((const float *)cur, (const float *)prev, npoint, dst + snap, fmt, s);
        return r;
    };
    auto restore = [&](int64_t seg, void *cur, void *prev) -> int {
        const char *src = ck.c() + (size_t)seg * 2 * snap;
        if (fmt == HSX_CKPT_FP32) {
            HSX_HIP(hipMemcpyAsync(cur, src, bytes, hipMemcpyDeviceToDevice, s));
            HSX_HIP(hipMemcpyAsync(prev, src + snap, bytes, hipMemcpyDeviceToDevice, s));
            return HSX_OK;
        }
        int r = hsx_ckpt_decode(src, npoint, fmt, nullptr, (float *)cur, s);
        if (!r) r = hsx_ckpt_decode(src + snap, npoint, fmt, (const float *)cur, (float *)prev, s);
        return r;
    };

    // ---- forward pass
    if (K == 0) {
        if ((rc = zero(g, q[0].p, s)) || (rc = zero(g, q[1].p, s))) return rc;
        for (int64_t n = 0; n < nt; n++)
            if ((rc = fwd_step(g, coef, sh, q[n + 2].p, q[n].p, q[n + 1].p, n, traces.p, es, s))) return rc;
    } else {
        if ((rc = zero(g, fa.p, s)) || (rc = zero(g, fb.p, s))) return rc;
        void *prev = fa.p, *cur = fb.p;
        for (int64_t n = 0; n < nt; n++) {
            if (n % K == 0 && (rc = save(n / K, cur, prev))) return rc;
            if ((rc = fwd_step(g, coef, sh, prev, prev, cur, n, traces.p, es, s))) return rc;
            std::swap(prev, cur);
        }
    }

    // ---- residual and misfit
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((ntr + 255) / 256, 4096));
    rc = dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        hipLaunchKernelGGL(k_residual<TS>, dim3(nb), dim3(256), 0, s, (const TS *)traces.p, (const TS *)d_obs,
                           (TS *)res.p, ntr, (double *)partial.p);
        HSX_LAUNCHED("k_residual");
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, (const double *)partial.p, nb, (double *)msum.p);
        HSX_LAUNCHED("k_sum_partials");
        return HSX_OK;
    });
    if (rc) return rc;
    if (nrec > 0 && (rc = hsx_gather(g, coef, nrec, sh->rec_idx, crec.p, s))) return rc;

    // ---- backward pass: mu^k for k = nt .. 1
    if ((rc = zero(g, m1.p, s)) || (rc = zero(g, m2.p, s))) return rc;
    void *mu1 = m1.p, *mu2 = m2.p;   // mu^{k+1}, mu^{k+2}
    auto back_step = [&](int64_t k, const void *qk, const void *qk1, const void *qk2) -> int {
        int r = hsx_free_surface(g, mu1, 0x3F, s);
        if (!r) r = hsx_ac2_step(g, mu2, mu1, coef, s);                    // mu2 <- mu^k
        if (!r && nrec > 0)
            r = hsx_inject(g, mu2, nrec, sh->rec_idx, res.c() + (size_t)(k - 1) * nrec * es, crec.p, 1.0,
                           HSX_INJECT_ADD, s);
        if (!r) r = hsx_imaging(g, grad, mu2, qk, qk1, qk2, rim, img_scale, -1, 0.0, s);
        if (!r && sh->nsrc > 0) {
            r = dispatch_types(g, [&](auto ts, auto) -> int {
                using TS = decltype(ts);
                hipLaunchKernelGGL(k_imaging_src<TS>, dim3((unsigned)((sh->nsrc + 63) / 64)), dim3(64), 0, s, grad,
                                   (const TS *)mu2, sh->nsrc, sh->src_idx,
                                   (const TS *)((const char *)sh->wavelet + (size_t)(k - 1) * sh->nsrc * es),
                                   -img_scale);
                HSX_LAUNCHED("k_imaging_src");
                return HSX_OK;
            });
        }
        std::swap(mu1, mu2);
        return r;
    };
    if (K == 0) {
        for (int64_t k = nt; k >= 1; k--)
            if ((rc = back_step(k, q[k + 1].p, q[k].p, q[k - 1].p))) return rc;
    } else {
        for (int64_t seg = nseg - 1; seg >= 0; seg--) {
            const int64_t n0 = seg * K, n1 = std::min<int64_t>(n0 + K, nt);
            // q[j] holds q^{n0 - 1 + j}
            if ((rc = restore(seg, q[1].p, q[0].p))) return rc;
            for (int64_t n = n0; n < n1; n++)
                if ((rc = fwd_step(g, coef, sh, q[n - n0 + 2].p, q[n - n0].p, q[n - n0 + 1].p, n, nullptr, es, s)))
                    return rc;
            for (int64_t k = n1; k > n0; k--)
                if ((rc = back_step(k, q[k - n0 + 1].p, q[k - n0].p, q[k - n0 - 1].p))) return rc;
        }
    }
    double m = 0.0;
    HSX_HIP(hipMemcpyAsync(&m, msum.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HSX_HIP(hipStreamSynchronize(s));   // buffers are freed on return
    if (misfit) *misfit = 0.5 * m;
    return HSX_OK;
}

} // extern "C"
