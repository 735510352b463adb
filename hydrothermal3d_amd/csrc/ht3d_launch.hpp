// ht3d_launch.hpp -- launch helpers of the per-kernel entries (one thread per
// output point, x fastest), shared by ht3d_kernels.hip and ht3d_dim.hip.
#pragma once

#include "ht3d_internal.hpp"

namespace ht3d {

// 3D launch: block (64,4,1); grid covers (n1,n2,n3).
template <typename Fn>
__global__ __launch_bounds__(256) void for3d(int n1, int n2, int n3, Fn fn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i < n1 && j < n2 && k < n3) fn(i, j, k);
}

template <typename Fn>
static int launch3d(const char *name, int n1, int n2, int n3, Fn fn) {
    if (n1 <= 0 || n2 <= 0 || n3 <= 0) return HT3D_OK;
    dim3 blk(64, 4, 1), grd(div_up(n1, 64), div_up(n2, 4), n3);
    hipLaunchKernelGGL(for3d<Fn>, grd, blk, 0, ctx().stream, n1, n2, n3, fn);
    return after_launch(name);
}

#define CHECK_GRID(nx, ny, nz)                                                           \
    do {                                                                                 \
        HT3D_CHECK_INIT();                                                               \
        if (!grid_ok(nx, ny, nz)) return arg_fail("grid extents must be >= 3"), HT3D_ERR_ARG; \
    } while (0)

#define CHECK_PTR(p)                                                                     \
    do {                                                                                 \
        if (!(p)) return arg_fail("null device pointer: " #p), HT3D_ERR_ARG;             \
    } while (0)

}  // namespace ht3d
