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
//   mu^k = 2 mu^{k+1} - (mu^{k+2} - c e^k) + c (L mu^{k+1}),  e^k = residual at step k-1,
// i.e. the forward step itself (free surface + fused kernel) with the residual,
// scaled by c, taken off mu^{k+2} at the receivers first.  The imaging condition is
//   dJ/dm = -(1/dt^2) sum_k mu^k (q^k - 2 q^{k-1} + q^{k-2} - F^{k-1}),
// accumulated in double over the points B does not touch, by the adjoint
// step's own kernel where the tiled kernels take the grid.
#include <algorithm>
#include <vector>

#include "hsx_internal.h"

namespace hsx {

// res = d - dobs (d NULL: res = dobs, the residual given by the caller);
// partial[block] = sum of res^2 over the block (double)
template <typename TS>
__global__ void __launch_bounds__(256) k_residual(const TS *d, const TS *dobs, TS *res, int64_t n, double *partial)
{
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const TS r = d ? d[i] - dobs[i] : dobs[i];
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

// Sources then receivers of one forward step in one launch (k_inject + k_gather
// in order): block 0 adds the sources (out += 1.0 * w, in double), waits, then
// records the receivers that sit on a source point; the other blocks record
// every other receiver, whose points no source writes.  nsrc <= SRCREC_MAX
// (the membership test is a loop over the sources).
constexpr int64_t SRCREC_MAX = 64;

__device__ __forceinline__ bool on_source(int64_t j, int64_t nsrc, const int64_t *src)
{
    for (int64_t k = 0; k < nsrc; k++)
        if (src[k] == j) return true;
    return false;
}

template <typename TS>
__global__ void __launch_bounds__(256) k_srcrec(TS *u, int64_t npoint, int64_t nsrc, const int64_t *src, const TS *w,
                                                int64_t nrec, const int64_t *rec, TS *tr)
{
    if (blockIdx.x == 0) {
        for (int64_t i = threadIdx.x; i < nsrc; i += blockDim.x) {
            if ((uint64_t)src[i] >= (uint64_t)npoint) continue;
            TS *p = u + src[i];
            *p = (TS)((double)*p + 1.0 * (double)w[i]);
        }
        __syncthreads();
        for (int64_t i = threadIdx.x; i < nrec; i += blockDim.x) {
            const int64_t j = rec[i];
            if (on_source(j, nsrc, src)) tr[i] = (uint64_t)j < (uint64_t)npoint ? u[j] : (TS)0;
        }
        return;
    }
    for (int64_t i = (int64_t)(blockIdx.x - 1) * blockDim.x + threadIdx.x; i < nrec;
         i += (int64_t)(gridDim.x - 1) * blockDim.x) {
        const int64_t j = rec[i];
        if (!on_source(j, nsrc, src)) tr[i] = (uint64_t)j < (uint64_t)npoint ? u[j] : (TS)0;
    }
}

// The source term of reverse step k (the imaging condition's -F part:
// grad[src_i] += scale * mu^k[src_i] * f_i, for the sources inside the imaging
// box b only) and the residual of step k-1 into
// mu^{k+1} (v[rec_i] += -1 * res_i * crec_i, in double, as k_inject) in one
// launch: blocks [0, nb_src) take the sources, the rest the receivers; the two
// parts touch different buffers.
template <typename TS>
__global__ void __launch_bounds__(256) k_imgsrc_residual(double *grad, const TS *mu, int64_t nsrc, const int64_t *sidx,
                                                         const TS *f, double scale, Box b, int64_t n1, int64_t n12,
                                                         uint32_t nb_src, TS *v, int64_t npoint, int64_t nrec,
                                                         const int64_t *ridx, const TS *res, const TS *crec)
{
    if (blockIdx.x < nb_src) {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= nsrc) return;
        const int64_t j = sidx[i];
        const int64_t i3 = j / n12, i2 = (j - i3 * n12) / n1, i1 = j - i3 * n12 - i2 * n1;
        if (i1 < b.s1 || i1 > b.e1 || i2 < b.s2 || i2 > b.e2 || i3 < b.s3 || i3 > b.e3) return;
        grad[j] += scale * (double)mu[j] * (double)f[i];
        return;
    }
    for (int64_t i = (int64_t)(blockIdx.x - nb_src) * blockDim.x + threadIdx.x; i < nrec;
         i += (int64_t)(gridDim.x - nb_src) * blockDim.x) {
        if ((uint64_t)ridx[i] >= (uint64_t)npoint) continue;
        double x = -1.0 * (double)res[i];
        x *= (double)crec[i];
        TS *p = v + ridx[i];
        *p = (TS)((double)*p + x);
    }
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
    if (!rc && sh->nsrc > 0 && sh->nsrc <= SRCREC_MAX && sh->nrec > 0 && traces) {
        const int64_t nrec = sh->nrec;
        const unsigned nb = 1 + (unsigned)std::min<int64_t>((nrec + 255) / 256, 1024);
        return dispatch_types(g, [&](auto ts, auto) -> int {
            using TS = decltype(ts);
            hipLaunchKernelGGL(k_srcrec<TS>, dim3(nb), dim3(256), 0, s, (TS *)out, g->n1 * g->n2 * g->n3, sh->nsrc,
                               sh->src_idx, (const TS *)((const char *)sh->wavelet + (size_t)n * sh->nsrc * es), nrec,
                               sh->rec_idx, (TS *)((char *)traces + (size_t)n * nrec * es));
            HSX_LAUNCHED("k_srcrec");
            return HSX_OK;
        });
    }
    if (!rc && sh->nsrc > 0)
        rc = hsx_inject(g, out, sh->nsrc, sh->src_idx, (const char *)sh->wavelet + (size_t)n * sh->nsrc * es, nullptr,
                        1.0, HSX_INJECT_ADD, s);
    if (!rc && sh->nrec > 0 && traces)
        rc = hsx_gather(g, out, sh->nrec, sh->rec_idx, (char *)traces + (size_t)n * sh->nrec * es, s);
    return rc;
}

} // namespace hsx

// Work buffers of hsx_gradient for one grid layout, nt, receiver count and
// checkpoint options, allocated once (hsx_adjoint_plan_create) and reused by
// every gradient of that geometry.
struct hsx_adjoint_plan {
    hsx_grid g;
    int64_t nt = 0, nrec = 0;
    hsx_adjoint_opts opt;
    size_t es = 4, bytes = 0, snap = 0;
    int64_t nseg = 0;
    hsx::Dev traces, res, crec, partial, msum, m1, m2, ck, fa, fb;
    std::vector<hsx::Dev> q;   // wavefields q^{m}, m from the segment start - 1
};

namespace hsx {

static int plan_checks(const hsx_grid *g, int64_t nt, int64_t nrec, const hsx_adjoint_opts *opt)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if ((rc = check_ptr(opt, "opts"))) return rc;
    if (nt < 1) return fail(HSX_ERR_INVALID, "nt = %lld", (long long)nt);
    if (nrec < 0) return fail(HSX_ERR_INVALID, "negative receiver count");
    const int K = opt->ckpt_every;
    const int fmt = opt->ckpt_format;
    if (K < 0) return fail(HSX_ERR_INVALID, "ckpt_every %d < 0", K);
    if (fmt != HSX_CKPT_FP32 && fmt != HSX_CKPT_FP16 && fmt != HSX_CKPT_BF16)
        return fail(HSX_ERR_INVALID, "checkpoint format %d", fmt);
    if (fmt != HSX_CKPT_FP32 && g->storage != HSX_FP32)
        return fail(HSX_ERR_UNSUPPORTED, "16-bit checkpoints of FP64 wavefields");
    if (K == 0 && fmt != HSX_CKPT_FP32) return fail(HSX_ERR_INVALID, "full storage keeps every step in the storage type");
    if (opt->flags & ~(HSX_ADJ_RESIDUAL | HSX_ADJ_AUTOTUNE))
        return fail(HSX_ERR_INVALID, "adjoint flags 0x%x", opt->flags);
    for (int a = 0; a < g->dim; a++)
        if (g->inner[2 * a + 1] - g->inner[2 * a] < 2) return fail(HSX_ERR_INVALID, "grid too small for the imaging range");
    return ensure_init();
}

static bool same_layout(const hsx_grid &a, const hsx_grid &b)
{
    if (a.dim != b.dim || a.fd_order != b.fd_order || a.storage != b.storage || a.compute != b.compute ||
        a.n1 != b.n1 || a.n2 != b.n2 || a.n3 != b.n3)
        return false;
    for (int k = 0; k < 6; k++)
        if (a.inner[k] != b.inner[k]) return false;
    return true;
}

static int plan_run(hsx_adjoint_plan *P, const hsx_grid *g, const void *coef, const hsx_shot *sh, const void *d_obs,
                    double *grad, double *misfit, hipStream_t s)
{
    int rc;
    const int K = P->opt.ckpt_every;
    const int fmt = P->opt.ckpt_format;
    const bool given_residual = (P->opt.flags & HSX_ADJ_RESIDUAL) != 0;
    const size_t es = P->es, bytes = P->bytes, snap = P->snap;
    const int64_t npoint = g->n1 * g->n2 * g->n3;
    const int64_t nt = sh->nt, nrec = sh->nrec, ntr = nt * nrec, nseg = P->nseg;

    // imaging range: the inner points minus the boundary planes zeroed by B
    int64_t rim[6];
    for (int a = 0; a < 3; a++) {
        rim[2 * a] = g->inner[2 * a] + (a < g->dim ? 1 : 0);
        rim[2 * a + 1] = g->inner[2 * a + 1] - (a < g->dim ? 1 : 0);
    }
    const Box rimb = box_of(rim);
    const double img_scale = -1.0 / (sh->dt * sh->dt);
    auto &q = P->q;

    auto save = [&](int64_t seg, const void *cur, const void *prev) -> int {
        char *dst = P->ck.c() + (size_t)seg * 2 * snap;
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
        const char *src = P->ck.c() + (size_t)seg * 2 * snap;
        if (fmt == HSX_CKPT_FP32) {
            HSX_HIP(hipMemcpyAsync(cur, src, bytes, hipMemcpyDeviceToDevice, s));
            HSX_HIP(hipMemcpyAsync(prev, src + snap, bytes, hipMemcpyDeviceToDevice, s));
            return HSX_OK;
        }
        int r = hsx_ckpt_decode(src, npoint, fmt, nullptr, (float *)cur, s);
        if (!r) r = hsx_ckpt_decode(src + snap, npoint, fmt, (const float *)cur, (float *)prev, s);
        return r;
    };

    // ---- forward pass (traces only when the residual is formed here)
    void *tr = given_residual ? nullptr : P->traces.p;
    if (K == 0) {
        if ((rc = zero(g, q[0].p, s)) || (rc = zero(g, q[1].p, s))) return rc;
        for (int64_t n = 0; n < nt; n++)
            if ((rc = fwd_step(g, coef, sh, q[n + 2].p, q[n].p, q[n + 1].p, n, tr, es, s))) return rc;
    } else {
        if ((rc = zero(g, P->fa.p, s)) || (rc = zero(g, P->fb.p, s))) return rc;
        void *prev = P->fa.p, *cur = P->fb.p;
        for (int64_t n = 0; n < nt; n++) {
            if (n % K == 0 && (rc = save(n / K, cur, prev))) return rc;
            if ((rc = fwd_step(g, coef, sh, prev, prev, cur, n, tr, es, s))) return rc;
            std::swap(prev, cur);
        }
    }

    // ---- residual and misfit
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((ntr + 255) / 256, 4096));
    rc = dispatch_types(g, [&](auto ts, auto) -> int {
        using TS = decltype(ts);
        hipLaunchKernelGGL(k_residual<TS>, dim3(nb), dim3(256), 0, s, (const TS *)tr, (const TS *)d_obs,
                           (TS *)P->res.p, ntr, (double *)P->partial.p);
        HSX_LAUNCHED("k_residual");
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, (const double *)P->partial.p, nb,
                           (double *)P->msum.p);
        HSX_LAUNCHED("k_sum_partials");
        return HSX_OK;
    });
    if (rc) return rc;
    if (nrec > 0 && (rc = hsx_gather(g, coef, nrec, sh->rec_idx, P->crec.p, s))) return rc;

    // ---- backward pass: mu^k for k = nt .. 1
    if ((rc = zero(g, P->m1.p, s)) || (rc = zero(g, P->m2.p, s))) return rc;
    void *mu1 = P->m1.p, *mu2 = P->m2.p;   // mu^{k+1}, mu^{k+2}
    // the imaging condition fused into the adjoint step's kernel where the
    // tiled kernels take the grid (exact arithmetic only), else its own pass
    const bool fuse_img = kernel_family() != HSX_KERNEL_REF && !contract_mode();
    // the residual of step k - 1 goes into mu^{k+1}: for k = nt here, for the
    // others in the launch that ends reverse step k
    auto res_at = [&](int64_t k) { return P->res.c() + (size_t)(k - 1) * nrec * es; };
    if (nrec > 0 &&
        (rc = hsx_inject(g, mu2, nrec, sh->rec_idx, res_at(nt), P->crec.p, -1.0, HSX_INJECT_ADD, s)))
        return rc;
    auto back_step = [&](int64_t k, const void *qk, const void *qk1, const void *qk2) -> int {
        // mu^k = 2 mu^{k+1} - (mu^{k+2} - c e^k) + c L mu^{k+1}: the residual
        // went into mu^{k+2} (mu2) before the step, so that the step's output is
        // mu^k complete and the imaging condition can be taken from it
        int r = hsx_free_surface(g, mu1, 0x3F, s);
        if (!r) {
            r = fuse_img ? tiled_ac2_step_img(g, mu2, mu1, coef, grad, qk, qk1, qk2, rim, img_scale, s)
                         : HSX_ERR_UNSUPPORTED;
            if (r == HSX_ERR_UNSUPPORTED) {
                r = hsx_ac2_step(g, mu2, mu1, coef, s);                   // mu2 <- mu^k
                if (!r) r = hsx_imaging(g, grad, mu2, qk, qk1, qk2, rim, img_scale, -1, 0.0, s);
            }
        }
        // the source term of step k (reads mu^k in mu2) and the residual of
        // step k - 1 into mu^{k+1} (mu1): one launch
        const int64_t ns = sh->nsrc, nr = k >= 2 ? nrec : 0;
        if (!r && (ns > 0 || nr > 0)) {
            r = dispatch_types(g, [&](auto ts, auto) -> int {
                using TS = decltype(ts);
                const uint32_t nbs = (uint32_t)((ns + 255) / 256);
                const uint32_t nbr = (uint32_t)std::min<int64_t>((nr + 255) / 256, 1024);
                hipLaunchKernelGGL(k_imgsrc_residual<TS>, dim3(nbs + nbr), dim3(256), 0, s, grad, (const TS *)mu2,
                                   ns, sh->src_idx,
                                   ns > 0 ? (const TS *)((const char *)sh->wavelet + (size_t)(k - 1) * ns * es)
                                          : (const TS *)nullptr,
                                   -img_scale, rimb, g->n1, g->n1 * g->n2, nbs, (TS *)mu1, npoint, nr, sh->rec_idx,
                                   nr > 0 ? (const TS *)res_at(k - 1) : (const TS *)nullptr, (const TS *)P->crec.p);
                HSX_LAUNCHED("k_imgsrc_residual");
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
    HSX_HIP(hipMemcpyAsync(&m, P->msum.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HSX_HIP(hipStreamSynchronize(s));
    if (misfit) *misfit = 0.5 * m;
    return HSX_OK;
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

int hsx_adjoint_plan_create(const hsx_grid *g, int64_t nt, int64_t nrec, const hsx_adjoint_opts *opt,
                            hsx_adjoint_plan **out)
{
    if (!out) return fail(HSX_ERR_INVALID, "null plan pointer");
    *out = nullptr;
    int rc = plan_checks(g, nt, nrec, opt);
    if (rc) return rc;
    if ((opt->flags & HSX_ADJ_AUTOTUNE) && g->dim == 3 && (rc = hsx_autotune(g, 5, nullptr, nullptr))) return rc;
    hsx_adjoint_plan *P = new hsx_adjoint_plan();
    P->g = *g;
    P->nt = nt;
    P->nrec = nrec;
    P->opt = *opt;
    P->es = g->storage == HSX_FP64 ? 8 : 4;
    const int64_t npoint = g->n1 * g->n2 * g->n3;
    P->bytes = (size_t)npoint * P->es;
    const int K = opt->ckpt_every;
    const int64_t ntr = nt * nrec;
    P->nseg = K == 0 ? 0 : (nt + K - 1) / K;
    P->snap = opt->ckpt_format == HSX_CKPT_FP32 ? P->bytes : hsx_ckpt_bytes(npoint);
    auto bail = [&](int code) {
        delete P;
        return code;
    };
    if ((rc = P->traces.alloc(ntr * P->es)) || (rc = P->res.alloc(ntr * P->es)) || (rc = P->crec.alloc(nrec * P->es)) ||
        (rc = P->partial.alloc(4096 * sizeof(double))) || (rc = P->msum.alloc(sizeof(double))) ||
        (rc = P->m1.alloc(P->bytes)) || (rc = P->m2.alloc(P->bytes)))
        return bail(rc);
    P->q.resize(K == 0 ? nt + 2 : K + 2);
    for (auto &x : P->q)
        if ((rc = x.alloc(P->bytes))) return bail(rc);
    if (K > 0 && ((rc = P->ck.alloc((size_t)P->nseg * 2 * P->snap)) || (rc = P->fa.alloc(P->bytes)) ||
                  (rc = P->fb.alloc(P->bytes))))
        return bail(rc);
    *out = P;
    return HSX_OK;
}

int hsx_adjoint_plan_destroy(hsx_adjoint_plan *P)
{
    if (!P) return HSX_OK;
    (void)hipDeviceSynchronize();   // no kernel of a gradient may still use the buffers
    delete P;
    return HSX_OK;
}

int hsx_gradient_plan(hsx_adjoint_plan *P, const hsx_grid *g, const void *coef, const hsx_shot *sh, const void *d_obs,
                      double *grad, double *misfit, void *stream)
{
    if (!P) return fail(HSX_ERR_INVALID, "null plan");
    int rc = shot_checks(g, coef, sh);
    if (rc) return rc;
    if ((rc = check_ptr(d_obs, "d_obs")) || (rc = check_ptr(grad, "grad"))) return rc;
    if (!same_layout(*g, P->g)) return fail(HSX_ERR_INVALID, "gradient: grid layout differs from the plan's");
    if (sh->nt != P->nt || sh->nrec != P->nrec)
        return fail(HSX_ERR_INVALID, "gradient: nt / receiver count %lld / %lld, plan %lld / %lld", (long long)sh->nt,
                    (long long)sh->nrec, (long long)P->nt, (long long)P->nrec);
    return plan_run(P, g, coef, sh, d_obs, grad, misfit, stream_of(stream));
}

int hsx_gradient(const hsx_grid *g, const void *coef, const hsx_shot *sh, const void *d_obs,
                 const hsx_adjoint_opts *opt, double *grad, double *misfit, void *stream)
{
    int rc = shot_checks(g, coef, sh);
    if (rc) return rc;
    if ((rc = check_ptr(d_obs, "d_obs")) || (rc = check_ptr(grad, "grad")) || (rc = check_ptr(opt, "opts")))
        return rc;
    hsx_adjoint_plan *P = nullptr;
    if ((rc = hsx_adjoint_plan_create(g, sh->nt, sh->nrec, opt, &P))) return rc;
    rc = plan_run(P, g, coef, sh, d_obs, grad, misfit, stream_of(stream));
    hsx_adjoint_plan_destroy(P);   // plan_run has synchronised the stream (or failed)
    return rc;
}

} // extern "C"
