// Runtime part of the C ABI: errors, device selection, memory, streams,
// events, descriptor validation and the FD coefficient table.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "hsx_internal.h"

namespace hsx {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(e == hipErrorOutOfMemory ? HSX_ERR_NOMEM : HSX_ERR_HIP, "%s: %s (%d)", what,
                hipGetErrorString(e), (int)e);
}

struct State {
    std::mutex mu;
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    int family = HSX_KERNEL_AUTO;
    int contract = 0;             // hsx_set_contract
    int cus = 0;                  // compute units of the device
    void *scratch[SCRATCH_SLOTS] = {nullptr};
    size_t scratch_bytes[SCRATCH_SLOTS] = {0};
};
static State g;

static int init_locked(int device)
{
    int n = 0;
    HSX_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(HSX_ERR_INVALID, "device %d not in [0,%d)", device, n);
    HSX_HIP(hipSetDevice(device));
    if (g.ready && g.device == device) return HSX_OK;
    if (g.ready) {
        // switching device: drop the old device's resources
        for (int i = 0; i < SCRATCH_SLOTS; i++) {
            if (g.scratch[i]) (void)hipFree(g.scratch[i]);
            g.scratch[i] = nullptr;
            g.scratch_bytes[i] = 0;
        }
        if (g.stream) (void)hipStreamDestroy(g.stream);
        g.stream = nullptr;
    }
    // default priority on purpose: streams of mixed priorities are not used here
    HSX_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    HSX_HIP(hipDeviceGetAttribute(&g.cus, hipDeviceAttributeMultiprocessorCount, device));
    g.device = device;
    g.ready = true;
    return HSX_OK;
}

int device_cus(void) { return g.cus > 0 ? g.cus : 1; }

int ensure_init(void)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (g.ready) {
        HSX_HIP(hipSetDevice(g.device));
        return HSX_OK;
    }
    return init_locked(0);
}

hipStream_t stream_of(void *s) { return s ? (hipStream_t)s : g.stream; }

int kernel_family(void) { return g.family; }

int contract_mode(void) { return g.contract; }

int scratch_get(int slot, size_t bytes, void **ptr)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (slot < 0 || slot >= SCRATCH_SLOTS) return fail(HSX_ERR_INVALID, "scratch slot %d", slot);
    if (g.scratch_bytes[slot] < bytes) {
        if (g.scratch[slot]) {
            HSX_HIP(hipDeviceSynchronize());
            HSX_HIP(hipFree(g.scratch[slot]));
            g.scratch[slot] = nullptr;
            g.scratch_bytes[slot] = 0;
        }
        size_t sz = bytes < 4096 ? 4096 : bytes;
        HSX_HIP(hipMalloc(&g.scratch[slot], sz));
        g.scratch_bytes[slot] = sz;
    }
    *ptr = g.scratch[slot];
    return HSX_OK;
}

int check_ptr(const void *p, const char *name)
{
    return p ? HSX_OK : fail(HSX_ERR_INVALID, "null pointer: %s", name);
}

int check_grid(const hsx_grid *g)
{
    if (!g) return fail(HSX_ERR_INVALID, "null grid descriptor");
    if (g->dim < 1 || g->dim > 3) return fail(HSX_ERR_INVALID, "dim %d not in 1..3", g->dim);
    if (g->fd_order < 2 || g->fd_order > 16 || (g->fd_order % 2))
        return fail(HSX_ERR_INVALID, "fd_order %d not in {2,4,...,16}", g->fd_order);
    if ((g->storage != HSX_FP32 && g->storage != HSX_FP64) || (g->compute != HSX_FP32 && g->compute != HSX_FP64))
        return fail(HSX_ERR_INVALID, "precision storage=%d compute=%d", g->storage, g->compute);
    const int64_t n[3] = {g->n1, g->n2, g->n3};
    const int R = g->fd_order / 2;
    for (int a = 0; a < 3; a++) {
        if (n[a] < 1) return fail(HSX_ERR_INVALID, "n%d = %lld", a + 1, (long long)n[a]);
        const int64_t s = g->inner[2 * a], e = g->inner[2 * a + 1];
        if (s < 0 || e < s || e >= n[a])
            return fail(HSX_ERR_INVALID, "inner range of axis %d [%lld,%lld] outside [0,%lld)", a + 1,
                        (long long)s, (long long)e, (long long)n[a]);
        if (a < g->dim) {
            // every stencil read stays inside the allocation
            if (s - R < 0 || e + R >= n[a])
                return fail(HSX_ERR_INVALID, "axis %d: inner range [%lld,%lld] +- halo %d outside [0,%lld)",
                            a + 1, (long long)s, (long long)e, R, (long long)n[a]);
            for (int h = 0; h < 2; h++) {
                const int64_t *hr = h == 0 ? g->halo1 : g->halo2;
                if (hr[2 * a] < 0 || hr[2 * a + 1] >= n[a] || hr[2 * a + 1] < hr[2 * a] - 1)
                    return fail(HSX_ERR_INVALID, "halo%d range of axis %d invalid", h + 1, a + 1);
            }
        } else if (n[a] != 1 || s != 0 || e != 0) {
            return fail(HSX_ERR_INVALID, "inactive axis %d must have n = 1 and range [0,0]", a + 1);
        }
    }
    return HSX_OK;
}

int check_range(const hsx_grid *g, const int64_t r[6])
{
    if (!r) return fail(HSX_ERR_INVALID, "null range");
    const int64_t n[3] = {g->n1, g->n2, g->n3};
    for (int a = 0; a < 3; a++)
        if (r[2 * a] < 0 || r[2 * a + 1] < r[2 * a] || r[2 * a + 1] >= n[a])
            return fail(HSX_ERR_INVALID, "range of axis %d [%lld,%lld] outside [0,%lld)", a + 1,
                        (long long)r[2 * a], (long long)r[2 * a + 1], (long long)n[a]);
    return HSX_OK;
}

// exact rationals of the centred D2 weights (src/fdm.h:276-507)
static const double FD_NUM_DEN[8][9][2] = {
    {{-2, 1}, {1, 1}},
    {{-5, 2}, {4, 3}, {-1, 12}},
    {{-49, 18}, {3, 2}, {-3, 20}, {1, 90}},
    {{-205, 72}, {8, 5}, {-1, 5}, {8, 315}, {-1, 560}},
    {{-5269, 1800}, {5, 3}, {-5, 21}, {5, 126}, {-5, 1008}, {1, 3150}},
    {{-5369, 1800}, {12, 7}, {-15, 56}, {10, 189}, {-1, 112}, {2, 1925}, {-1, 16632}},
    {{-266681, 88200}, {7, 4}, {-7, 24}, {7, 108}, {-7, 528}, {7, 3300}, {-7, 30888}, {1, 84084}},
    {{-1077749, 352800}, {16, 9}, {-14, 45}, {112, 1485}, {-7, 396}, {112, 32175}, {-2, 3861},
     {16, 315315}, {-1, 411840}},
};

double fd_coef_double(int fd_order, int j)
{
    const double *q = FD_NUM_DEN[fd_order / 2 - 1][j];
    return q[0] / q[1];
}

// staggered first-derivative weights A1..AR (src/fdm.h:29-268), index 0 unused
static const double FD1_NUM_DEN[8][9][2] = {
    {{0, 1}, {1, 1}},
    {{0, 1}, {9, 8}, {-1, 24}},
    {{0, 1}, {75, 64}, {-25, 384}, {3, 640}},
    {{0, 1}, {1225, 1024}, {-245, 3072}, {49, 5120}, {-5, 7168}},
    {{0, 1}, {19845, 16384}, {-735, 8192}, {567, 40960}, {-405, 229376}, {35, 294912}},
    {{0, 1}, {160083, 131072}, {-12705, 131072}, {22869, 1310720}, {-5445, 1835008}, {847, 2359296},
     {-63, 2883584}},
    {{0, 1}, {1288287, 1048576}, {-429429, 4194304}, {429429, 20971520}, {-61347, 14680064}, {13013, 18874368},
     {-3549, 46137344}, {231, 54525952}},
    {{0, 1}, {41409225, 33554432}, {-3578575, 33554432}, {3864861, 167772160}, {-1254825, 234881024},
     {325325, 301989888}, {-61425, 369098752}, {7425, 436207616}, {-143, 167772160}},
};

double fd1_coef_double(int fd_order, int j)
{
    const double *q = FD1_NUM_DEN[fd_order / 2 - 1][j];
    return q[1] != 0 ? q[0] / q[1] : 0.0;
}

} // namespace hsx

using namespace hsx;

extern "C" {

int hsx_abi_version(void) { return HSX_ABI_VERSION; }

const char *hsx_last_error(void) { return g_err; }

int hsx_init(int device)
{
    std::lock_guard<std::mutex> lk(g.mu);
    return init_locked(device);
}

int hsx_device_count(int *count)
{
    if (!count) return fail(HSX_ERR_INVALID, "null count");
    HSX_HIP(hipGetDeviceCount(count));
    return HSX_OK;
}

int hsx_device_info(int device, hsx_device_props *out)
{
    if (!out) return fail(HSX_ERR_INVALID, "null props");
    hipDeviceProp_t p;
    HSX_HIP(hipGetDeviceProperties(&p, device));
    memset(out, 0, sizeof(*out));
    strncpy(out->name, p.name, sizeof(out->name) - 1);
    strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
    out->compute_units = p.multiProcessorCount;
    out->clock_mhz = p.clockRate / 1000;
    out->mem_clock_mhz = p.memoryClockRate / 1000;
    out->bus_width_bits = p.memoryBusWidth;
    out->hbm_bytes = (int64_t)p.totalGlobalMem;
    out->lds_bytes_per_block = (int32_t)p.sharedMemPerBlock;
    out->l2_bytes = p.l2CacheSize;
    return HSX_OK;
}

int hsx_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(HSX_ERR_INVALID, "null ptr");
    int rc = ensure_init();
    if (rc) return rc;
    *ptr = nullptr;
    HSX_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return HSX_OK;
}

int hsx_free(void *ptr)
{
    if (!ptr) return HSX_OK;
    HSX_HIP(hipFree(ptr));
    return HSX_OK;
}

int hsx_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    HSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(stream)));
    HSX_HIP(hipStreamSynchronize(stream_of(stream)));
    return HSX_OK;
}

int hsx_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    HSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(stream)));
    HSX_HIP(hipStreamSynchronize(stream_of(stream)));
    return HSX_OK;
}

int hsx_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    HSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_of(stream)));
    return HSX_OK;
}

int hsx_stream_create(void **stream)
{
    if (!stream) return fail(HSX_ERR_INVALID, "null stream");
    int rc = ensure_init();
    if (rc) return rc;
    hipStream_t s;
    HSX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return HSX_OK;
}

int hsx_stream_destroy(void *stream)
{
    if (!stream) return HSX_OK;
    HSX_HIP(hipStreamDestroy((hipStream_t)stream));
    return HSX_OK;
}

int hsx_stream_sync(void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    HSX_HIP(hipStreamSynchronize(stream_of(stream)));
    return HSX_OK;
}

int hsx_event_create(void **event)
{
    if (!event) return fail(HSX_ERR_INVALID, "null event");
    int rc = ensure_init();
    if (rc) return rc;
    hipEvent_t e;
    HSX_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return HSX_OK;
}

int hsx_event_destroy(void *event)
{
    if (!event) return HSX_OK;
    HSX_HIP(hipEventDestroy((hipEvent_t)event));
    return HSX_OK;
}

int hsx_event_record(void *event, void *stream)
{
    if (!event) return fail(HSX_ERR_INVALID, "null event");
    HSX_HIP(hipEventRecord((hipEvent_t)event, stream_of(stream)));
    return HSX_OK;
}

int hsx_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!start || !stop || !ms) return fail(HSX_ERR_INVALID, "null event/ms");
    HSX_HIP(hipEventSynchronize((hipEvent_t)stop));
    HSX_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return HSX_OK;
}

int hsx_set_kernel(int family)
{
    if (family < HSX_KERNEL_AUTO || family > HSX_KERNEL_TILED)
        return fail(HSX_ERR_INVALID, "kernel family %d", family);
    g.family = family;
    return HSX_OK;
}

int hsx_get_kernel(void) { return g.family; }

int hsx_set_contract(int on)
{
    if (on != 0 && on != 1) return fail(HSX_ERR_INVALID, "contract %d not 0 or 1", on);
    g.contract = on;
    return HSX_OK;
}

int hsx_get_contract(void) { return g.contract; }

int hsx_set_tile(int lanes_n1, int rows_n2, int chunk_planes, int flags)
{
    return set_tile(lanes_n1, rows_n2, chunk_planes, flags);
}

} // extern "C"
