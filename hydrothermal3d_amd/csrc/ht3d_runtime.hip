// ht3d_runtime.hip -- context, device memory, copies and reductions.
//
// Replaces what ParallelStencil/CUDA.jl provide around the kernels of
// tduretz/HydroThermal3D: device selection (HydroThermal3D_v7.jl:9), @zeros and
// Data.Array/Array transfers (v7:87-121,153-157) and the host reductions
// maximum_g/minimum_g/mean_g/sum_g (tools/Macros.jl:20-25).
#include <cstdio>
#include <cstring>
#include <cmath>
#include <new>

#include "ht3d_internal.hpp"

namespace ht3d {

static Context g_ctx;
// the last error's text, in a fixed buffer: reporting an error allocates nothing
static thread_local char g_err[512];

Context &ctx() { return g_ctx; }

void set_error(const char *msg) noexcept { std::snprintf(g_err, sizeof g_err, "%s", msg ? msg : ""); }

int hip_fail(hipError_t e, const char *what) noexcept {
    std::snprintf(g_err, sizeof g_err, "%s: %s", what ? what : "", hipGetErrorString(e));
    return HT3D_ERR_HIP;
}

int arg_fail(const char *what) noexcept {
    set_error(what);
    return HT3D_ERR_ARG;
}

static int g_inject_throw = 0;   // ht3d_debug_inject_host_failure

void maybe_inject_failure() {
    if (g_inject_throw > 0) {
        --g_inject_throw;
        throw std::bad_alloc();
    }
}

int guard_failure(const char *entry, bool nomem) noexcept {
    std::snprintf(g_err, sizeof g_err, "%s: %s", entry,
                  nomem ? "host memory allocation failed" : "internal error (C++ exception)");
    return nomem ? HT3D_ERR_NOMEM : HT3D_ERR_INTERNAL;
}

int workspace(int slot, size_t bytes, void **out) {
    Context &c = g_ctx;
    if (slot < 0 || slot >= WS_COUNT) return arg_fail("workspace slot");
    if (bytes == 0) bytes = 256;
    if (c.ws_bytes[slot] < bytes) {
        if (c.ws[slot]) {
            hipError_t e = hipStreamSynchronize(c.stream);
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize(workspace)");
            e = hipFree(c.ws[slot]);
            if (e != hipSuccess) return hip_fail(e, "hipFree(workspace)");
            c.ws[slot] = nullptr;
            c.ws_bytes[slot] = 0;
        }
        hipError_t e = hipMalloc(&c.ws[slot], bytes);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(workspace)");
        c.ws_bytes[slot] = bytes;
    }
    *out = c.ws[slot];
    return HT3D_OK;
}

int after_launch(const char *name, bool force_sync) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, name);
    if (g_ctx.sync || force_sync) {
        e = hipStreamSynchronize(g_ctx.stream);
        if (e != hipSuccess) return hip_fail(e, name);
    }
    return HT3D_OK;
}

// ------------------------------------------------------------------ timing
void SegmentTimer::before_launch() {
    Context &c = g_ctx;
    if (!c.timing || open) return;
    (void)hipEventRecord(c.ev0, c.stream);
    open = true;
    count = 0;
}

void SegmentTimer::close() {
    Context &c = g_ctx;
    if (!c.timing || !open) return;
    (void)hipEventRecord(c.ev1, c.stream);
}

void SegmentTimer::collect(double bytes_per_launch, bool fp32) {
    Context &c = g_ctx;
    if (!c.timing || !open) return;
    float ms = 0.f;
    if (hipEventSynchronize(c.ev1) == hipSuccess && hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess) {
        if (fp32) {
            c.stats.ms_fp32 += ms;
            c.stats.launches_fp32 += count;
            c.stats.bytes_fp32 += bytes_per_launch * (double)count;
        } else {
            c.stats.ms += ms;
            c.stats.launches += count;
            c.stats.bytes += bytes_per_launch * (double)count;
        }
    }
    open = false;
    count = 0;
}

// ------------------------------------------------------------------ reductions
__global__ __launch_bounds__(1024) void finish_kernel(const double *__restrict__ partials, unsigned nblocks,
                                                       unsigned ops, double *__restrict__ result) {
    double h[4], l[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        h[q] = (ops >> q & 1u) ? -INFINITY : 0.0;
        l[q] = 0.0;
    }
    // four blocks' partials per thread and pass are loaded before they are
    // combined, so their loads overlap instead of waiting one after another
    const double2 *p2 = reinterpret_cast<const double2 *>(partials);
    for (unsigned b0 = threadIdx.x; b0 < nblocks; b0 += 4 * blockDim.x) {
        double2 v[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned b = b0 + u * blockDim.x;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                v[u][q] = b < nblocks ? p2[(size_t)b * 4 + q] : make_double2((ops >> q & 1u) ? -INFINITY : 0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (ops >> q & 1u) h[q] = (v[u][q].x > h[q] || v[u][q].x != v[u][q].x) ? v[u][q].x : h[q];
                else dd_combine(h[q], l[q], v[u][q].x, v[u][q].y);
            }
    }
    __shared__ double part[8];
    block_reduce4(h, l, ops, part, 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 4; ++q) result[q] = (ops >> q & 1u) ? part[2 * q] : part[2 * q] + part[2 * q + 1];
    }
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

int fill_zero(void *p, size_t bytes) {
    const size_t n = bytes / sizeof(double);
    if (!n) return HT3D_OK;
    unsigned nb = div_up(n, 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(nb), dim3(256), 0, g_ctx.stream, (double *)p, n);
    HIP_TRY(hipGetLastError());
    return HT3D_OK;
}

int finish_partials(const double *partials, unsigned nblocks, unsigned ops, double *result_dev) {
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(1024), 0, g_ctx.stream, partials, nblocks, ops, result_dev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "finish_kernel");
    return HT3D_OK;
}

// The copy is followed by an event, and the wait is on that event only: work
// queued after post_result4 keeps the GPU busy while the host waits.
static hipEvent_t g_res_event = nullptr;

int post_result4(const double *result_dev) {
    if (!g_res_event) HIP_TRY(hipEventCreateWithFlags(&g_res_event, hipEventDisableTiming));
    HIP_TRY(hipMemcpyAsync(g_ctx.pinned, result_dev, 4 * sizeof(double), hipMemcpyDeviceToHost, g_ctx.stream));
    HIP_TRY(hipEventRecord(g_res_event, g_ctx.stream));
    return HT3D_OK;
}

int wait_result4(double out[4]) {
    HIP_TRY(hipEventSynchronize(g_res_event));
    for (int q = 0; q < 4; ++q) out[q] = g_ctx.pinned[q];
    return HT3D_OK;
}

int read_result4(const double *result_dev, double out[4]) {
    HIP_TRY(hipMemcpyAsync(g_ctx.pinned, result_dev, 4 * sizeof(double), hipMemcpyDeviceToHost, g_ctx.stream));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    for (int q = 0; q < 4; ++q) out[q] = g_ctx.pinned[q];
    return HT3D_OK;
}

// generic 1D reduction over n doubles into partials (slot 0)
__global__ __launch_bounds__(256) void reduce_kernel(const double *__restrict__ a, size_t n, int op,
                                                     double *__restrict__ partials) {
    const bool is_max = (op == HT3D_MAX || op == HT3D_MAX_ABS);
    const bool is_min = (op == HT3D_MIN);
    double hi = (is_max || is_min) ? -INFINITY : 0.0, lo = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double v = a[i];
        if (op == HT3D_SUM_ABS || op == HT3D_MAX_ABS) v = fabs(v);
        if (is_min) v = -v;  // min as the max of negatives
        if (is_max || is_min) hi = (v > hi || v != v) ? v : hi;
        else dd_add(hi, lo, v);
    }
    const double h4[4] = {hi, 0.0, 0.0, 0.0}, l4[4] = {lo, 0.0, 0.0, 0.0};
    block_reduce4(h4, l4, (is_max || is_min) ? 1u : 0u, partials, blockIdx.x);
}

}  // namespace ht3d

using namespace ht3d;

extern "C" {

int ht3d_init(int device) {
    HT3D_API_BEGIN
    Context &c = ctx();
    if (c.initialised) {
        if (c.device == device) return HT3D_OK;
        ht3d_finalize();
    }
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return arg_fail("ht3d_init: no such device"), HT3D_ERR_ARG;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc((void **)&c.pinned, 64 * sizeof(double), hipHostMallocDefault));
    HIP_TRY(hipEventCreate(&c.ev0));
    HIP_TRY(hipEventCreate(&c.ev1));
    if (int rc = noalloc_setup(device)) return rc;
    c.device = device;
    c.initialised = true;
    c.sync = true;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_timing_enable(int on) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    ctx().timing = on != 0;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_timing_get(ht3d_timing_stats *out) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!out) return arg_fail("ht3d_timing_get: null pointer");
    *out = ctx().stats;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_timing_reset(void) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    ctx().stats = ht3d_timing_stats{};
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_finalize(void) {
    HT3D_API_BEGIN
    Context &c = ctx();
    if (!c.initialised) return HT3D_OK;
    (void)hipStreamSynchronize(c.stream);
    for (int s = 0; s < WS_COUNT; ++s) {
        if (c.ws[s]) (void)hipFree(c.ws[s]);
        c.ws[s] = nullptr;
        c.ws_bytes[s] = 0;
    }
    if (c.ev0) (void)hipEventDestroy(c.ev0);
    if (c.ev1) (void)hipEventDestroy(c.ev1);
    c.ev0 = c.ev1 = nullptr;
    if (c.pinned) (void)hipHostFree(c.pinned);
    c.pinned = nullptr;
    (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    c.initialised = false;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_debug_inject_host_failure(int count) {
    g_inject_throw = count > 0 ? count : 0;
    return HT3D_OK;
}

const char *ht3d_strerror(int code) {
    switch (code) {
    case HT3D_OK: return "ok";
    case HT3D_ERR_ARG: return "invalid argument";
    case HT3D_ERR_HIP: return "HIP runtime error";
    case HT3D_ERR_NOMEM: return "out of memory";
    case HT3D_ERR_STATE: return "library not initialised";
    case HT3D_ERR_INTERNAL: return "internal error";
    case HT3D_NOT_CONVERGED: return "pseudo-transient solve reached nitmax";
    case HT3D_NAN: return "NaN in residual norm";
    case HT3D_DIVERGED: return "pseudo-transient solve diverged";
    default: return "unknown status";
    }
}

const char *ht3d_last_error(void) { return g_err; }

int ht3d_set_sync(int synchronous) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    ctx().sync = synchronous != 0;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_sync(void) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    HIP_TRY(hipStreamSynchronize(ctx().stream));
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_device_name(char *buf, size_t len) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, ctx().device));
    if (buf && len) {
        // some runtimes leave the marketing name empty: the architecture names the part then
        std::snprintf(buf, len, "%s (%s, %d CUs, %.1f GiB)", p.name[0] ? p.name : "AMD GPU", p.gcnArchName,
                      p.multiProcessorCount, (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    }
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_alloc(size_t bytes, void **dptr) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!dptr) return arg_fail("ht3d_alloc: null out pointer");
    *dptr = nullptr;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 8);
    if (e == hipErrorOutOfMemory) return hip_fail(e, "hipMalloc"), HT3D_ERR_NOMEM;
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_free(void *dptr) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!dptr) return HT3D_OK;
    HIP_TRY(hipStreamSynchronize(ctx().stream));
    HIP_TRY(hipFree(dptr));
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_h2d(void *dst, const void *src, size_t bytes) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!bytes) return HT3D_OK;
    if (!dst || !src) return arg_fail("ht3d_h2d: null pointer");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx().stream));
    HIP_TRY(hipStreamSynchronize(ctx().stream));
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_d2h(void *dst, const void *src, size_t bytes) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!bytes) return HT3D_OK;
    if (!dst || !src) return arg_fail("ht3d_d2h: null pointer");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().stream));
    HIP_TRY(hipStreamSynchronize(ctx().stream));
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_d2d(void *dst, const void *src, size_t bytes) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!bytes) return HT3D_OK;
    if (!dst || !src) return arg_fail("ht3d_d2d: null pointer");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx().stream));
    return after_launch("ht3d_d2d");
    HT3D_API_END
}

int ht3d_memset(void *dptr, int value, size_t bytes) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!bytes) return HT3D_OK;
    if (!dptr) return arg_fail("ht3d_memset: null pointer");
    HIP_TRY(hipMemsetAsync(dptr, value, bytes, ctx().stream));
    return after_launch("ht3d_memset");
    HT3D_API_END
}

int ht3d_reduce(double *out_host, const double *a, size_t n, int op) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!out_host || !a || op < HT3D_SUM || op > HT3D_MAX_ABS) return arg_fail("ht3d_reduce: bad argument");
    if (n == 0) {
        *out_host = (op == HT3D_MAX || op == HT3D_MAX_ABS) ? -INFINITY : (op == HT3D_MIN ? INFINITY : 0.0);
        return HT3D_OK;
    }
    unsigned nb = div_up(n, 256);
    if (nb > 2048) nb = 2048;
    void *part = nullptr, *res = nullptr;
    int rc = workspace(WS_MISC, (size_t)nb * 8 * sizeof(double), &part);
    if (rc) return rc;
    rc = workspace(WS_RESULT, 64 * sizeof(double), &res);
    if (rc) return rc;
    hipLaunchKernelGGL(reduce_kernel, dim3(nb), dim3(256), 0, ctx().stream, a, n, op, (double *)part);
    HIP_TRY(hipGetLastError());
    const bool mx = (op == HT3D_MAX || op == HT3D_MAX_ABS || op == HT3D_MIN);
    rc = finish_partials((const double *)part, nb, mx ? 1u : 0u, (double *)res);
    if (rc) return rc;
    double r4[4];
    rc = read_result4((const double *)res, r4);
    if (rc) return rc;
    *out_host = (op == HT3D_MIN) ? -r4[0] : r4[0];
    return HT3D_OK;
    HT3D_API_END
}

}  // extern "C"
