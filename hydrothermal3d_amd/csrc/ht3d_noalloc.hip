// ht3d_noalloc.hip -- the DYREL solve without any allocation.
//
// ht3d_dyrel_solve_noalloc runs the fused DYREL solve of ht3d_dyrel_solve with
// every device buffer it needs (the ping-pong copy of the unknown, the DYREL
// momentum F0 and its pair twin, the F of the last lambda iteration, the stored
// Gershgorin diagonal, the block partials and the result slots) carved out of
// one file-scope device array, `scratch`.  It makes no hipMalloc or hipFree
// call and runs no tuning pass, for callers that cannot allocate (serving
// loops, allocation-free harnesses).  Every call on a device shares that
// device's array and leaves its contents behind; calls that overlap on one
// device race.
#include "ht3d_internal.hpp"

namespace ht3d {

// limits of the scratch array: cells of the grid, entries of its extended
// (ghosted) array and block partials of one iteration launch
constexpr size_t kNaMaxCells = (size_t)1 << 21;      // 128^3
constexpr size_t kNaMaxExt = (size_t)1 << 22;        // E entries
constexpr size_t kNaMaxBlocks = (size_t)1 << 16;     // at least the 4096 of the Gershgorin pass
constexpr size_t kNaDoubles = kNaMaxExt + 4 * kNaMaxCells + 8 * kNaMaxBlocks + 64;   // 105 MB

static __device__ __align__(256) double scratch[kNaDoubles];

// The array's device address per device, taken once at ht3d_init with the
// device current (host code never passes the symbol itself).
constexpr int kNaMaxDevices = 64;
static double *g_scratch[kNaMaxDevices] = {nullptr};

int noalloc_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return arg_fail("noalloc_setup: device index out of range");
    if (g_scratch[device]) return HT3D_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = (double *)ptr;
    return HT3D_OK;
}

}  // namespace ht3d

using namespace ht3d;

extern "C" int ht3d_dyrel_solve_noalloc(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p,
                                        ht3d_dyrel_report *rep) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!f || !p) return arg_fail("ht3d_dyrel_solve_noalloc: null argument");
    const int dev = ctx().device;
    double *base = dev >= 0 && dev < kNaMaxDevices ? g_scratch[dev] : nullptr;
    if (!base) return arg_fail("ht3d_dyrel_solve_noalloc: no scratch address for this device"), HT3D_ERR_STATE;
    if (f->ncx < 3 || f->ncy < 3 || f->ncz < 3) return arg_fail("ht3d_dyrel_solve_noalloc: extents must be >= 3");
    const size_t N = (size_t)f->ncx * f->ncy * f->ncz;
    const size_t NE = (size_t)(f->ncx + 2) * (f->ncy + 2) * (f->ncz + 2);
    if (N > kNaMaxCells || NE > kNaMaxExt)
        return arg_fail("ht3d_dyrel_solve_noalloc: the grid exceeds the scratch limit of 2^21 cells and 2^22 "
                        "extended-array entries");
    SolveBuffers b;
    b.xb = base;
    b.f0 = b.xb + kNaMaxExt;
    b.f0b = b.f0 + kNaMaxCells;
    b.fst = b.f0b + kNaMaxCells;
    b.dst = b.fst + kNaMaxCells;
    b.partials = b.dst + kNaMaxCells;
    b.result = b.partials + 8 * kNaMaxBlocks;
    b.max_blocks = kNaMaxBlocks;
    return dyrel_solve_impl(f, p, rep, nullptr, 1.0, &b);
    HT3D_API_END
}
