// ht3d_internal.hpp -- shared runtime state and helpers of libht3d.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <new>
#include <string>

#include "../../include/ht3d.h"

namespace ht3d {

// ------------------------------------------------------------------ context
enum WsSlot {
    WS_PARTIALS = 0, WS_RESULT = 1, WS_XB = 2, WS_F0 = 3, WS_FCIT = 4, WS_MISC = 5, WS_DIAG = 6,
    WS_F32_X = 7, WS_F32_XB = 8, WS_F32_F0 = 9, WS_F32_R = 10, WS_F32_K = 11, WS_F32_D = 12, WS_DC_RES = 13,
    WS_F0B = 14, WS_F32_F0B = 15, WS_F32_DIAG = 16, WS_V18_CLS = 17, WS_V18_G = 18, WS_V18_GY = 19,
    WS_TUNE_X = 20, WS_BAND = 21, WS_F32_KV = 22, WS_COUNT = 23
};

struct Context {
    bool initialised = false;
    int device = 0;
    hipStream_t stream = nullptr;
    bool sync = true;               // per-kernel entries synchronise (ParallelStencil semantics)
    double *pinned = nullptr;       // small pinned host buffer for scalar read-back
    // live timing of the iteration kernels
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ht3d_timing_stats stats{};
    // grow-only device workspaces
    void *ws[WS_COUNT] = {nullptr};
    size_t ws_bytes[WS_COUNT] = {0};
};

Context &ctx();
void set_error(const char *msg) noexcept;
int hip_fail(hipError_t e, const char *what) noexcept;
int arg_fail(const char *what) noexcept;

// Every extern "C" entry runs its body through api_guard: a C++ exception
// (std::bad_alloc from a host container, or anything else) becomes a status
// code and the error text, instead of unwinding into the caller (ctypes,
// Julia's ccall).  maybe_inject_failure throws on the next guarded calls
// after ht3d_debug_inject_host_failure(n), so tests can force that path.
void maybe_inject_failure();
int guard_failure(const char *entry, bool nomem) noexcept;
template <typename F>
inline int api_guard(const char *entry, F &&body) noexcept {
    try {
        maybe_inject_failure();
        return body();
    } catch (const std::bad_alloc &) {
        return guard_failure(entry, true);
    } catch (...) {
        return guard_failure(entry, false);
    }
}
#define HT3D_API_BEGIN return ::ht3d::api_guard(__func__, [&]() -> int {
#define HT3D_API_END });

// Brackets runs of back-to-back iteration launches with events (when timing is on).
struct SegmentTimer {
    bool open = false;
    long long count = 0;
    void before_launch();
    // call after the last launch of a run; `sync_pending` false means the caller
    // synchronises the stream before calling collect()
    void close();
    void collect(double bytes_per_launch, bool fp32);
};

// returns a device workspace of at least `bytes` for slot `slot` (grow-only)
int workspace(int slot, size_t bytes, void **out);


#define HT3D_CHECK_INIT()                                                     \
    do {                                                                      \
        if (!::ht3d::ctx().initialised) return ::ht3d::arg_fail("ht3d_init not called"), HT3D_ERR_STATE; \
    } while (0)

#define HIP_TRY(expr)                                                         \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) return ::ht3d::hip_fail(e_, #expr);             \
    } while (0)

// Two timing events owned by a tuner, destroyed on every return path.
struct TimerEvents {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int create() {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        return 0;
    }
    TimerEvents() = default;
    TimerEvents(const TimerEvents &) = delete;
    TimerEvents &operator=(const TimerEvents &) = delete;
    ~TimerEvents() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// after a launch: check the launch, then synchronise if requested
int after_launch(const char *name, bool force_sync = false);

// ------------------------------------------------------------------ grids
struct Grid {
    int nx, ny, nz;
    __host__ __device__ size_t c(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
    __host__ __device__ size_t e(int i, int j, int k) const {  // E array, (i,j,k) in E coordinates
        return ((size_t)k * (ny + 2) + j) * (nx + 2) + i;
    }
    __host__ __device__ size_t fx(int i, int j, int k) const { return ((size_t)k * ny + j) * (nx + 1) + i; }
    __host__ __device__ size_t fy(int i, int j, int k) const { return ((size_t)k * (ny + 1) + j) * nx + i; }
    __host__ __device__ size_t fz(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
    __host__ __device__ size_t v(int i, int j, int k) const { return ((size_t)k * (ny + 1) + j) * (nx + 1) + i; }
    __host__ __device__ size_t x6(int i, int j, int k) const { return ((size_t)k * (ny + 6) + j) * (nx + 6) + i; }
    size_t ncell() const { return (size_t)nx * ny * nz; }
    size_t next() const { return (size_t)(nx + 2) * (ny + 2) * (nz + 2); }
};

inline bool grid_ok(int nx, int ny, int nz) { return nx >= 3 && ny >= 3 && nz >= 3; }

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ------------------------------------------------------------------ reductions
// Sums are accumulated in double-double (hi + lo, TwoSum with no FMA
// contraction), so a sum of N terms is accurate to about N * 2^-106 of the sum
// of magnitudes and its final rounding to double does not depend on the
// order of accumulation except in vanishingly rare ties.  The CPU oracle
// accumulates the same way, which makes GPU and CPU solves take identical
// convergence decisions and lambda_min updates.  Maxima are exact.
__host__ __device__ inline void dd_add(double &hi, double &lo, double x) {
    const double s = hi + x;
    const double bp = s - hi;
    const double e = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += e;
}

__host__ __device__ inline void dd_combine(double &hi, double &lo, double h2, double l2) {
    const double s = hi + h2;
    const double bp = s - hi;
    double e = (hi - (s - bp)) + (h2 - bp);
    e += lo + l2;
    hi = s + e;
    lo = e - (hi - s);
}

template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, 64);
        v = (o > v || o != o) ? o : v;
    }
    return v;
}

__device__ inline void wave_dd(double &hi, double &lo) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double h2 = __shfl_xor(hi, off, 64), l2 = __shfl_xor(lo, off, 64);
        dd_combine(hi, lo, h2, l2);
    }
}

// Block reduction of 4 slots; bit q of `ops` set = max slot (value in hi),
// clear = double-double sum slot (hi, lo).  Writes 8 doubles per block:
// partials[8*block + 2*q + {0,1}].  Thread count: multiple of 64, <= 1024.
__device__ inline void block_reduce4(const double hi_in[4], const double lo_in[4], unsigned ops, double *partials,
                                     unsigned block_linear) {
    __shared__ double red[16][8];
    const unsigned tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const unsigned nthreads = blockDim.x * blockDim.y * blockDim.z;
    const unsigned lane = tid & 63u, wid = tid >> 6;
    double h[4], l[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        h[q] = hi_in[q];
        l[q] = lo_in[q];
        if (ops >> q & 1u) h[q] = wave_max(h[q]);
        else wave_dd(h[q], l[q]);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            red[wid][2 * q] = h[q];
            red[wid][2 * q + 1] = l[q];
        }
    }
    __syncthreads();
    if (tid == 0) {
        const unsigned nw = nthreads >> 6;
        for (int q = 0; q < 4; ++q) {
            double a = red[0][2 * q], b = red[0][2 * q + 1];
            for (unsigned w = 1; w < nw; ++w) {
                if (ops >> q & 1u) {
                    const double o = red[w][2 * q];
                    a = (o > a || o != o) ? o : a;
                } else {
                    dd_combine(a, b, red[w][2 * q], red[w][2 * q + 1]);
                }
            }
            partials[(size_t)block_linear * 8 + 2 * q] = a;
            partials[(size_t)block_linear * 8 + 2 * q + 1] = (ops >> q & 1u) ? 0.0 : b;
        }
    }
}

// Reduce `nblocks` x 8 partials into result[4] (single block launch).
int finish_partials(const double *partials, unsigned nblocks, unsigned ops, double *result_dev);
// Zero `bytes` (a multiple of 8) at p with a fill kernel on the library stream.
int fill_zero(void *p, size_t bytes);

// Device buffers of a DYREL solve handed in by the caller instead of taken
// from the workspaces (the allocation-free entry, ht3d_noalloc.hip).
struct SolveBuffers {
    double *xb, *f0, *f0b, *fst, *dst;   // E, then four cell arrays
    double *partials;                    // max_blocks x 8 doubles
    double *result;                      // 64 doubles
    size_t max_blocks;
};
int dyrel_solve_impl(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep, double *nu_out,
                     double nu_div, const SolveBuffers *given = nullptr);
// per-device address of the allocation-free entry's scratch array (ht3d_init)
int noalloc_setup(int device);
// Read back 4 doubles from the device result buffer (synchronises the stream).
int read_result4(const double *result_dev, double out[4]);
// read_result4 in two halves: queue the copy, and later wait for it (work
// queued in between runs before the wait returns)
int post_result4(const double *result_dev);
int wait_result4(double out[4]);

}  // namespace ht3d
