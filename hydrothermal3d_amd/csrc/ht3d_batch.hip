// ht3d_batch.hip -- many independent closed-box problems stepped together.
//
// A batch groups ht3d_hrl models of any grid size, Rayleigh number and
// permeability and steps them in lockstep: thermal solves of all problems,
// then Darcy solves, then velocity, time step and advection per problem.
//
// Solves.  Every problem's domain is cut into the items of the fused
// pseudo-iteration (a patch of 64 x TY lanes over the row-major (x, y) plane,
// CPL x-cells per lane, KZ planes deep); the (problem, item) pairs of the
// problems still iterating form a packed work list.  One persistent launch per
// pseudo-iteration walks that list (block b takes items b, b + G, ...), each
// item reading its problem's operator, pointers and DYREL scalars from a
// device table.  The DYREL control of HydroThermal3D_v18.jl:448-478 -- norms,
// nF_ini, the Rayleigh-quotient lambda_min update, NaN/divergence/convergence
// -- runs on the device after each check iteration, one block per problem, and
// writes the new h1/h2 back into the table; a problem that stops is marked
// inactive and its items are skipped.  Because nothing on the host has to
// change between checks, each run of pseudo-iterations between two checks is
// captured once as a HIP graph and replayed; the host only reads the
// per-problem states one block behind the device and repacks the work list
// when problems have finished.  The FP32 defect-correction form replays
// [FP64 anchor residual + decision, FP32 sweeps, x += e] blocks the same way.
// All problems of a batch share the DYREL parameters (the schedule is global),
// which keeps them in lockstep; results equal those of the single-problem
// solves (same kernel body, same double-double reductions).
//
// No reference counterpart: the reference steps one problem per process.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "ht3d_hrl.hpp"
#include "ht3d_iter.hpp"

namespace ht3d {

unsigned gershgorin_blocks(size_t N);
int launch_gershgorin(const double *kx, const double *ky, const double *kz, int nx, int ny, int nz, double dx,
                      double dy, double dz, double rc_tr_dt, double *partials, unsigned nb, double *Dout);
unsigned nu_blocks(int nx, int nz);
int launch_nu_partials(const double *T, const double *ky, int nx, int ny, int nz, double _dy, double *partials,
                       unsigned nb);
int copy_ghosts(double *dst, const double *src, int nx, int ny, int nz);
int bc_fill_table(double *A, int nx, int ny, int nz, const int type[6], const double val[6]);
int to_float(float *out, const double *in, size_t n);
int faces_f32(float *k32, const ht3d_dyrel_field *f);

// per-problem solver state, device resident
struct BState {
    ht3d_dyrel_report r;
    double c;
    int active, outer, inner, pad;
};

// tiles of the batched launches (rows of lanes per block, planes per item,
// x-cells per lane); the 2- and 4-cell tiles (vector loads, bit-identical
// cells) are offered when every problem's ncx is a multiple of that count, and
// 4 only to the FP32 sweeps.  The list drives every batch kernel's dispatch.
#define HT3D_BATCH_TILES(X) \
    X(0, 2, 8, 1) X(1, 4, 4, 1) X(2, 4, 8, 1) X(3, 8, 4, 1) X(4, 4, 2, 1) X(5, 1, 8, 1) X(6, 4, 4, 2) X(7, 4, 8, 2) \
    X(8, 2, 8, 2) X(9, 4, 4, 4) X(10, 4, 8, 4) X(11, 2, 8, 4) X(12, 4, 16, 2) X(13, 8, 8, 2) X(14, 2, 16, 2) \
    X(15, 4, 16, 4) X(16, 4, 2, 2) X(17, 8, 2, 2) X(18, 2, 4, 2) X(19, 8, 4, 2) X(20, 4, 2, 4) X(21, 8, 2, 4) \
    X(22, 2, 4, 4)
#define HT3D_TILE_CFG(N, TY, KZ, CPL) {TY, KZ, CPL},
static constexpr TileCfg kBTiles[] = {HT3D_BATCH_TILES(HT3D_TILE_CFG)};
#undef HT3D_TILE_CFG
static constexpr int kNumBTiles = sizeof(kBTiles) / sizeof(kBTiles[0]);
static constexpr int kBDefaultTile = 1;

// A batch item is a patch of 64*TY consecutive lane-groups (CPL x-cells each)
// of a problem's row-major (x, y) plane, marched over KZ planes (patch_cell):
// the small grids of a batch (32 to 128 cells wide) keep every lane busy,
// where a tile of whole 64*CPL-cell rows would leave up to 3/4 of them idle.
static inline unsigned bpatch_count(int nx, int ny, int tile) {
    return div_up((size_t)(nx / kBTiles[tile].cpl) * ny, 64 * kBTiles[tile].ty);
}
// items of one problem under batch tile `tile`
static inline unsigned btile_count(int nx, int ny, int nz, int tile) {
    return bpatch_count(nx, ny, tile) * div_up(nz, kBTiles[tile].kz);
}

// item l of a problem -> its patch and its chunk of planes
__device__ __forceinline__ void item_of(unsigned l, int npatch, int &p, int &bz) {
    p = (int)(l % (unsigned)npatch);
    bz = (int)(l / (unsigned)npatch);
}

// ---------------------------------------------------------------- kernels
// KV: a problem whose table has the vertex field (a.kv) runs iter_tile_kv
// (face coefficients formed from it: 6 words per cell instead of 8, the same
// results bit for bit); the others run iter_tile_v
template <typename T, bool TR, int TY, int KZ, int CPL, bool KV>
__global__ __launch_bounds__(64 * TY) void batch_iter_kernel(const IterArgs<T> *__restrict__ tab,
                                                             const uint2 *__restrict__ items,
                                                             const unsigned *__restrict__ n_items,
                                                             const BState *__restrict__ st,
                                                             double *__restrict__ partials, int phase,
                                                             unsigned flags) {
    const unsigned n = *n_items;
    for (unsigned w = blockIdx.x; w < n; w += gridDim.x) {
        const unsigned t = xcd_tile(w, n, flags & FL_REVERSE);
        const uint2 it = items[t];
        if (!st[it.x].active) continue;
        const IterArgs<T> &a = tab[it.x];
        int p, bz, i, j;
        item_of(it.y, a.nbx, p, bz);
        patch_cell<TY, CPL>(p, a.nx / CPL, i, j);
        T *x0 = const_cast<T *>(a.Xin), *x1 = a.Xout;
        // FP64 sweeps: non-temporal store of the unknown, cached face coefficients
        // (the single solve's variant 6; 9 % faster per batched FP64 step); the
        // FP32 sweeps measured no difference and keep the plain policy
        constexpr bool V6 = sizeof(T) == 8;
        if constexpr (CPL == 1)
            iter_tile<T, TR, false, V6, V6, TY, KZ>(a, phase ? x1 : x0, phase ? x0 : x1, flags, i, j, bz, partials, t);
        else if (KV && a.kv)
            iter_tile_kv<T, TR, false, V6, false, TY, KZ, CPL>(a, phase ? x1 : x0, phase ? x0 : x1, flags, i, j, bz,
                                                              partials, t);
        else
            iter_tile_v<T, TR, false, V6, V6, TY, KZ, CPL>(a, phase ? x1 : x0, phase ? x0 : x1, flags, i, j, bz,
                                                         partials, t);
        if (flags & FL_CHECK) __syncthreads();  // the block reduction's shared array is reused by the next item
    }
}

// FP64 anchor residual of the defect-correction form, on the same tiles as the
// sweeps (dc_residual_tile: the single solve's residual body; with KV, from the
// vertex field for the problems that have it)
template <bool TR, int TY, int KZ, int CPL, bool KV>
__global__ __launch_bounds__(64 * TY) void batch_dc_residual_kernel(const DcArgs *__restrict__ dtab,
                                                                    const IterArgs<float> *__restrict__ tab32,
                                                                    const uint2 *__restrict__ items,
                                                                    const unsigned *__restrict__ n_items,
                                                                    const BState *__restrict__ st,
                                                                    double *__restrict__ partials) {
    const unsigned n = *n_items;
    for (unsigned w = blockIdx.x; w < n; w += gridDim.x) {
        const unsigned t = xcd_tile(w, n);
        const uint2 it = items[t];
        if (!st[it.x].active) continue;
        int p, bz, i, j;
        item_of(it.y, tab32[it.x].nbx, p, bz);
        patch_cell<TY, CPL>(p, dtab[it.x].nx / CPL, i, j);
        double s_hi = 0.0, s_lo = 0.0, mx = 0.0;
        if (KV && dtab[it.x].kv) dc_residual_tile<TR, TY, KZ, CPL, KV>(dtab[it.x], i, j, bz, s_hi, s_lo, mx);
        else dc_residual_tile<TR, TY, KZ, CPL>(dtab[it.x], i, j, bz, s_hi, s_lo, mx);
        const double h4[4] = {s_hi, mx, 0.0, 0.0}, l4[4] = {s_lo, 0.0, 0.0, 0.0};
        block_reduce4(h4, l4, 0x2u, partials, t);
        __syncthreads();
    }
}

// x += e (FP64 += FP32) and e = 0 on the interior; `phase` selects the e buffer
template <int TY, int KZ, int CPL>
__global__ __launch_bounds__(64 * TY) void batch_dc_update_kernel(const DcArgs *__restrict__ dtab,
                                                                  const IterArgs<float> *__restrict__ tab32,
                                                                  const uint2 *__restrict__ items,
                                                                  const unsigned *__restrict__ n_items,
                                                                  const BState *__restrict__ st, int phase) {
    const unsigned n = *n_items;
    for (unsigned w = blockIdx.x; w < n; w += gridDim.x) {
        const unsigned t = xcd_tile(w, n);
        const uint2 it = items[t];
        if (!st[it.x].active) continue;
        const IterArgs<float> &a = tab32[it.x];
        double *X = const_cast<double *>(dtab[it.x].X);
        float *e = phase ? a.Xout : const_cast<float *>(a.Xin);
        int p, bz, i, j;
        item_of(it.y, a.nbx, p, bz);
        patch_cell<TY, CPL>(p, a.nx / CPL, i, j);
        const int k0 = bz * KZ, k1 = min(k0 + KZ, a.nz);
        if (j < a.ny)
            for (int k = k0; k < k1; ++k) {
                const size_t idx =
                    (size_t)(i + 1) + (size_t)(j + 1) * (a.nx + 2) + (size_t)(k + 1) * (a.nx + 2) * (a.ny + 2);
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    X[idx + q] += (double)e[idx + q];
                    e[idx + q] = 0.0f;
                }
            }
    }
}

// Reduce 4 partial slots over [b, e) with the whole block (1024 threads, the
// same shape as the single-problem finish), result on thread 0.
__device__ inline void range_reduce4(const double *__restrict__ partials, unsigned b, unsigned e, unsigned ops,
                                     double out[4]) {
    double h[4], l[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        h[q] = (ops >> q & 1u) ? -INFINITY : 0.0;
        l[q] = 0.0;
    }
    for (unsigned s = b + threadIdx.x; s < e; s += blockDim.x) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double v = partials[(size_t)s * 8 + 2 * q];
            if (ops >> q & 1u) h[q] = (v > h[q] || v != v) ? v : h[q];
            else dd_combine(h[q], l[q], v, partials[(size_t)s * 8 + 2 * q + 1]);
        }
    }
    __shared__ double part[8];
    block_reduce4(h, l, ops, part, 0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = (ops >> q & 1u) ? part[2 * q] : part[2 * q] + part[2 * q + 1];
}

__device__ inline void dyrel_coeffs(double lmin, double dtau, double cfact, double &c, double &h1, double &h2) {
    c = 2.0 * sqrt(lmin) * cfact;
    h1 = (2 - c * dtau) / (2 + c * dtau);
    h2 = 2 * dtau / (2 + c * dtau);
}

// lambda_max (Gershgorin) -> DYREL scalars, HydroThermal3D_v18.jl:433-439
template <typename T>
__global__ __launch_bounds__(1024) void batch_setup_kernel(BState *__restrict__ st, IterArgs<T> *__restrict__ tab,
                                                           const double *__restrict__ gpart,
                                                           const uint2 *__restrict__ grange, ht3d_dyrel_params p) {
    const int q = blockIdx.x;
    const uint2 r = grange[q];
    double r4[4];
    range_reduce4(gpart, r.x, r.x + r.y, 1u, r4);
    if (threadIdx.x) return;
    BState s;
    memset(&s, 0, sizeof s);
    s.r.status = HT3D_NOT_CONVERGED;
    s.r.lmax = r4[0];
    s.r.lmin = s.r.lmax / p.lmin_div;
    s.r.dtau = 2.0 / sqrt(s.r.lmax) * p.CFL;
    dyrel_coeffs(s.r.lmin, s.r.dtau, p.cfact, s.c, s.r.h1, s.r.h2);
    s.active = 1;
    st[q] = s;
    tab[q].dtau = (T)s.r.dtau;
    tab[q].h1 = (T)s.r.h1;
    tab[q].h2 = (T)s.r.h2;
}

__global__ void batch_activate_kernel(BState *st, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) st[q].active = 1;
}

// Check iteration of the FP64 solve: norms and the control flow of
// HydroThermal3D_v18.jl:457-476 for every active problem.
__global__ __launch_bounds__(1024) void batch_check_kernel(BState *__restrict__ st, IterArgs<double> *__restrict__ tab,
                                                           const uint2 *__restrict__ range,
                                                           const double *__restrict__ partials,
                                                           const unsigned *__restrict__ base, ht3d_dyrel_params p,
                                                           int offset, int lam) {
    const int q = blockIdx.x;
    if (!st[q].active) return;
    const uint2 rg = range[q];
    double r4[4];
    range_reduce4(partials, rg.x, rg.x + rg.y, 0x2u, r4);
    if (threadIdx.x) return;
    BState &s = st[q];
    const int iter = (int)*base + offset;
    const double N = (double)((size_t)tab[q].nx * tab[q].ny * tab[q].nz);
    s.r.n_checks++;
    double nF = p.norm == 0 ? r4[0] / N / sqrt(N) : r4[1];
    if (r4[0] != r4[0]) nF = r4[0];
    if (iter == 1) s.r.nF_ini = nF;
    s.r.nF_abs = nF;
    s.r.nF_rel = nF / s.r.nF_ini;
    s.r.iters = iter;
    if (nF != nF) { s.r.status = HT3D_NAN; s.active = 0; return; }
    if (s.r.nF_rel > 100) { s.r.status = HT3D_DIVERGED; s.active = 0; return; }
    if (lam && iter > 1) {
        double lmin = fabs(r4[2]);
        lmin /= fabs(r4[3]);
        s.r.lmin = lmin;
        dyrel_coeffs(lmin, s.r.dtau, p.cfact, s.c, s.r.h1, s.r.h2);
        tab[q].h1 = s.r.h1;
        tab[q].h2 = s.r.h2;
        s.r.n_lambda++;
    }
    if (nF < p.tol || s.r.nF_rel < p.tol) { s.r.status = HT3D_OK; s.active = 0; }
}

// nitmax reached: problems still iterating stop unconverged
__global__ void batch_final_kernel(BState *st, int n, int nitmax) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || !st[q].active) return;
    st[q].r.iters = nitmax;
    st[q].r.status = HT3D_NOT_CONVERGED;
    st[q].active = 0;
}

// Anchor of the defect-correction form: FP64 residual norm and the control of
// the FP64 solve; `run` FP32 sweeps follow for the problems that continue.  A
// problem that stops here gets one more update after the solve
// (dc_final_kernel), as the FP64 check iteration applies its own: it counts
// as an iteration.
__global__ __launch_bounds__(1024) void batch_anchor_kernel(BState *__restrict__ st,
                                                            const IterArgs<float> *__restrict__ tab32,
                                                            const uint2 *__restrict__ range,
                                                            const double *__restrict__ partials, ht3d_dyrel_params p,
                                                            int run) {
    const int q = blockIdx.x;
    if (!st[q].active) return;
    const uint2 rg = range[q];
    double r4[4];
    range_reduce4(partials, rg.x, rg.x + rg.y, 0x2u, r4);
    if (threadIdx.x) return;
    BState &s = st[q];
    const double N = (double)((size_t)tab32[q].nx * tab32[q].ny * tab32[q].nz);
    s.r.n_checks++;
    double nF = p.norm == 0 ? r4[0] / N / sqrt(N) : r4[1];
    if (r4[0] != r4[0]) nF = r4[0];
    if (s.outer == 0) s.r.nF_ini = nF;
    s.r.nF_abs = nF;
    s.r.nF_rel = nF / s.r.nF_ini;
    bool stop = true;
    if (nF != nF) s.r.status = HT3D_NAN;
    else if (s.r.nF_rel > 100) s.r.status = HT3D_DIVERGED;
    else if (nF < p.tol || s.r.nF_rel < p.tol) s.r.status = HT3D_OK;
    else stop = s.inner >= p.nitmax - 1;
    if (stop) {
        s.r.iters = s.inner + 1;
        s.active = 0;
        return;
    }
    s.outer++;
    s.inner += run;
    s.r.iters = s.inner;
}

// lambda iteration of the FP32 sweeps: Rayleigh quotient (HydroThermal3D_v18.jl:467-473)
__global__ __launch_bounds__(1024) void batch_lambda32_kernel(BState *__restrict__ st,
                                                              IterArgs<float> *__restrict__ tab32,
                                                              const uint2 *__restrict__ range,
                                                              const double *__restrict__ partials, ht3d_dyrel_params p,
                                                              int rayleigh) {
    const int q = blockIdx.x;
    if (!st[q].active) return;
    const uint2 rg = range[q];
    double r4[4];
    range_reduce4(partials, rg.x, rg.x + rg.y, 0x2u, r4);
    if (threadIdx.x) return;
    BState &s = st[q];
    if (rayleigh && r4[3] > 0.0) {
        s.r.lmin = fabs(r4[2]) / fabs(r4[3]);
        dyrel_coeffs(s.r.lmin, s.r.dtau, p.cfact, s.c, s.r.h1, s.r.h2);
        tab32[q].h1 = (float)s.r.h1;
        tab32[q].h2 = (float)s.r.h2;
        s.r.n_lambda++;
    }
}

__global__ void batch_advance_kernel(unsigned *base, unsigned L) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *base += L;
}

__global__ void batch_set_kernel(unsigned *p, unsigned v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(256) void batch_zero_kernel(uint4 *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Pair tiles of the batch (rows of lanes per block, planes per item, x-cells
// per lane); the 2- and 4-cell tiles need every problem's ncx a multiple of
// the count, and 4 cells per lane are for the FP32 sweeps only (in FP64 they
// exhaust the registers).
#define HT3D_BATCH_PAIR_TILES(X)                                                                                     \
    X(0, 8, 16, 1) X(1, 8, 32, 1) X(2, 16, 8, 1) X(3, 16, 16, 1) X(4, 16, 32, 1) X(5, 8, 16, 2) X(6, 8, 32, 2)       \
    X(7, 16, 8, 2) X(8, 16, 16, 2) X(9, 16, 32, 2) X(10, 8, 16, 4) X(11, 16, 8, 4) X(12, 16, 16, 4)
#define HT3D_PAIR_CFG(N, TY, KZ, CPL) {TY, KZ, CPL},
static constexpr TileCfg kBPairTiles[] = {HT3D_BATCH_PAIR_TILES(HT3D_PAIR_CFG)};
#undef HT3D_PAIR_CFG
static constexpr int kNumBPairTiles = sizeof(kBPairTiles) / sizeof(kBPairTiles[0]);

// rows of a band (iteration n+1; the inner ones run iteration n+2) of pair
// tile `pt` on a grid nx cells wide: 64*TY lanes of CPL cells.  The tile fits
// when a lane's cells stay in one row and every band has an inner row.
static inline int pair_band_rows(int pt, int nx) { return 64 * kBPairTiles[pt].ty * kBPairTiles[pt].cpl / nx; }
static inline bool pair_tile_fits(int pt, int nx) {
    return nx % kBPairTiles[pt].cpl == 0 && pair_band_rows(pt, nx) >= 3;
}
static inline unsigned pair_items(int nx, int ny, int nz, int pt) {
    return div_up(ny, pair_band_rows(pt, nx) - 2) * div_up(nz, kBPairTiles[pt].kz);
}

// The band-pair kernels (batch_pair_kernel, batch_pair_kernel_v) live in
// ht3d_iter.hpp: the single-problem defect-correction solve runs them too.

// ---------------------------------------------------------------- launchers
struct BLaunch {  // device pointers shared by every launch of a solve phase
    const void *tab;       // IterArgs<T> table of the phase
    const DcArgs *dtab;
    const uint2 *items;
    const unsigned *n_items;
    BState *st;
    double *partials;
    unsigned grid;
    unsigned threads;
    const uint2 *pitems;       // work list of the pair launches
    const unsigned *n_pitems;
    unsigned pgrid;
};

// kv: the vertex-coefficient form (bkv_ok: tiles with 2 cells per lane, and
// 4 in FP32)
template <typename T, bool TR>
static void launch_iter_b(int tile, const BLaunch &L, int phase, unsigned flags, bool kv) {
    const IterArgs<T> *tab = (const IterArgs<T> *)L.tab;
    hipStream_t s = ctx().stream;
    // 4 cells per lane: FP32 only (btile_ok)
#define BI_(N, TY, KZ, CPL)                                                                                          \
    case N:                                                                                                          \
        if constexpr (CPL > 1 && (CPL < 4 || sizeof(T) == 4)) {                                                      \
            if (kv)                                                                                                  \
                hipLaunchKernelGGL((batch_iter_kernel<T, TR, TY, KZ, CPL, true>), dim3(L.grid), dim3(64 * TY), 0, s,  \
                                   tab, L.items, L.n_items, L.st, L.partials, phase, flags);                        \
            else                                                                                                     \
                hipLaunchKernelGGL((batch_iter_kernel<T, TR, TY, KZ, CPL, false>), dim3(L.grid), dim3(64 * TY), 0, s, \
                                   tab, L.items, L.n_items, L.st, L.partials, phase, flags);                        \
        } else if constexpr (CPL < 4 || sizeof(T) == 4) {                                                            \
            hipLaunchKernelGGL((batch_iter_kernel<T, TR, TY, KZ, CPL, false>), dim3(L.grid), dim3(64 * TY), 0, s,     \
                               tab, L.items, L.n_items, L.st, L.partials, phase, flags);                            \
        }                                                                                                            \
        break;
    switch (tile) { HT3D_BATCH_TILES(BI_) default: break; }
#undef BI_
}

template <typename T, bool TR>
static void launch_pair_b(int pt, const BLaunch &L, int phase, int fpar, long long dF0, unsigned flags) {
    const IterArgs<T> *tab = (const IterArgs<T> *)L.tab;
    hipStream_t s = ctx().stream;
#define BP_(N, TY, KZ, CPL)                                                                                          \
    case N:                                                                                                          \
        if constexpr (CPL == 1)                                                                                      \
            hipLaunchKernelGGL((batch_pair_kernel<T, TR, TY, KZ, BState>), dim3(L.pgrid), dim3(64 * TY), 0, s, tab,  \
                               L.pitems, L.n_pitems, L.st, phase, fpar, dF0, flags);                                \
        else if constexpr (CPL < 4 || sizeof(T) == 4)                                                                \
            hipLaunchKernelGGL((batch_pair_kernel_v<T, TR, TY, KZ, CPL, BState>), dim3(L.pgrid), dim3(64 * TY), 0, s, \
                               tab, L.pitems, L.n_pitems, L.st, phase, fpar, dF0, flags);                           \
        break;
    switch (pt) { HT3D_BATCH_PAIR_TILES(BP_) default: break; }
#undef BP_
}

template <bool TR>
static void launch_res_b(int tile, const BLaunch &L, bool kv) {
    const IterArgs<float> *tab = (const IterArgs<float> *)L.tab;
    hipStream_t s = ctx().stream;
#define BR_(N, TY, KZ, CPL)                                                                                          \
    case N:                                                                                                          \
        if (CPL > 1 && kv)                                                                                           \
            hipLaunchKernelGGL((batch_dc_residual_kernel<TR, TY, KZ, CPL, CPL != 1>), dim3(L.grid), dim3(64 * TY), 0, \
                               s, L.dtab, tab, L.items, L.n_items, L.st, L.partials);                               \
        else                                                                                                         \
            hipLaunchKernelGGL((batch_dc_residual_kernel<TR, TY, KZ, CPL, false>), dim3(L.grid), dim3(64 * TY), 0, s, \
                               L.dtab, tab, L.items, L.n_items, L.st, L.partials);                                  \
        break;
    switch (tile) { HT3D_BATCH_TILES(BR_) default: break; }
#undef BR_
}

static void launch_upd_b(int tile, const BLaunch &L, int phase) {
    const IterArgs<float> *tab = (const IterArgs<float> *)L.tab;
    hipStream_t s = ctx().stream;
#define BU_(N, TY, KZ, CPL)                                                                                          \
    case N:                                                                                                          \
        hipLaunchKernelGGL((batch_dc_update_kernel<TY, KZ, CPL>), dim3(L.grid), dim3(64 * TY), 0, s, L.dtab, tab,   \
                           L.items, L.n_items, L.st, phase);                                                         \
        break;
    switch (tile) { HT3D_BATCH_TILES(BU_) default: break; }
#undef BU_
}

// one operation of a block of pseudo-iterations
enum BOpKind : int { OP_ITER64, OP_CHECK64, OP_FINAL64, OP_ITER32, OP_LAM32, OP_ANCHOR_RES, OP_ANCHOR, OP_UPDATE,
                     OP_ADVANCE, OP_PAIR64, OP_PAIR32 };
struct BOp {
    int kind;
    unsigned flags;
    int a, b;  // phase / offset / run / L, lambda / rayleigh / F0 buffer of a pair
};

}  // namespace ht3d

using namespace ht3d;

struct ht3d_batch {
    int nprob = 0;
    std::vector<ht3d_hrl *> m;
    int ncu = 256;
    // DYREL controls shared by the batch
    ht3d_dyrel_params pT{}, pP{};
    int precision = HT3D_FP64;
    int graphs = 1;
    // per-problem solver buffers, offsets into contiguous regions
    std::vector<size_t> oc, oe, of, ov;    // cell, E, face and vertex offsets (elements)
    size_t sc = 0, se = 0, sf = 0, sv = 0; // totals
    double *xb = nullptr, *f0 = nullptr, *fst = nullptr, *f0b = nullptr;
    float *e0 = nullptr, *e1 = nullptr, *f032 = nullptr, *fst32 = nullptr, *r32 = nullptr, *k32 = nullptr,
          *f032b = nullptr;  // f0b, f032b: the second F0 buffers of the pair launches
    float *kv32 = nullptr;   // FP32 vertex fields of the problems that solve with one
    // vertex-coefficient form per [kind][fp32] (timed like the tiles), and the
    // one of the launches being enqueued
    int kvm[2][2] = {{0, 0}, {0, 0}};
    int kv_cur = 0;
    // tables, per solve kind (0 thermal, 1 Darcy)
    IterArgs<double> *d_it64[2] = {nullptr, nullptr};
    IterArgs<float> *d_it32[2] = {nullptr, nullptr};
    DcArgs *d_dc[2] = {nullptr, nullptr};
    std::vector<IterArgs<double>> h_it64[2];
    std::vector<IterArgs<float>> h_it32[2];
    std::vector<DcArgs> h_dc[2];
    int tile64[2] = {-1, -1}, tile32[2] = {-1, -1};
    // launch grid per [kind][fp32]: 0 persistent (a few blocks per CU, each
    // walking items b, b + G, ...), 1 one block per item; gmode_cur is the one
    // the launches being enqueued use
    int gmode[2][2] = {{0, 0}, {0, 0}};
    int gmode_cur = 0;
    // pair tile (-1: single iterations only) and its launch grid per [kind][fp32];
    // ptile_cur / pgmode_cur are those of the solve being enqueued
    int ptile[2][2] = {{-1, -1}, {-1, -1}}, pgmode[2][2] = {{0, 0}, {0, 0}};
    int ptile_cur = -1, pgmode_cur = 0;
    // check-interval tuner inputs per [kind][fp32]: sweep and check costs (us), last iterations
    double t_it[2][2] = {{-1, -1}, {-1, -1}}, t_chk[2][2] = {{-1, -1}, {-1, -1}};
    double last_iters[2][2] = {{-1, -1}, {-1, -1}};
    // work list, partials, states
    uint2 *d_items = nullptr, *d_range = nullptr, *d_grange = nullptr, *d_pitems = nullptr;
    unsigned *d_nitems = nullptr;  // [0] item count, [1] iteration base, [2] pair item count
    size_t max_items = 0, max_pitems = 0;
    double *d_part = nullptr, *d_gpart = nullptr, *d_nupart = nullptr, *d_nures = nullptr;
    std::vector<unsigned> g_off, g_nb, nu_off, nu_nb;
    BState *d_state = nullptr;
    // pinned host staging
    BState *h_state[2] = {nullptr, nullptr};
    uint2 *h_items[2] = {nullptr, nullptr}, *h_range[2] = {nullptr, nullptr}, *h_pitems[2] = {nullptr, nullptr};
    unsigned *h_nitems[2] = {nullptr, nullptr};
    void *h_tabs = nullptr;
    size_t tabs_bytes = 0;
    double *h_nures = nullptr;
    WenoItem *d_weno = nullptr, *h_weno = nullptr;   // the batched advection's table
    int stage = 0;
    hipEvent_t ev[2] = {nullptr, nullptr}, ev0 = nullptr, ev1 = nullptr;
    std::map<std::string, hipGraphExec_t> cache;
    ht3d_batch_stats stats{};
    std::vector<void *> allocs;
};

namespace ht3d {

static int b_alloc(ht3d_batch *b, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(batch)");
    b->allocs.push_back(*p);
    return HT3D_OK;
}

static int zero_dev(void *p, size_t bytes) {
    const size_t n = bytes / 16;
    if (!n) return HT3D_OK;
    unsigned nb = div_up(n, 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(batch_zero_kernel, dim3(nb), dim3(256), 0, ctx().stream, (uint4 *)p, n);
    HIP_TRY(hipGetLastError());
    return HT3D_OK;
}

static size_t round16(size_t elems, size_t esize) { return ((elems * esize + 255) / 256) * 256 / esize; }

static void clear_graphs(ht3d_batch *b) {
    for (auto &kv : b->cache) (void)hipGraphExecDestroy(kv.second);
    b->cache.clear();
}

// items of the problems with `act[q]` set, for a tile configuration
static size_t build_items(ht3d_batch *b, int tile, const std::vector<char> &act, uint2 *items, uint2 *range) {
    size_t n = 0;
    for (int q = 0; q < b->nprob; ++q) {
        const ht3d_hrl *m = b->m[q];
        const unsigned cnt = btile_count(m->nx, m->ny, m->nz, tile);
        range[q] = make_uint2((unsigned)n, act[q] ? cnt : 0u);
        if (!act[q]) continue;
        for (unsigned l = 0; l < cnt; ++l) items[n++] = make_uint2((unsigned)q, l);
    }
    return n;
}

// pair items (bands x plane chunks) of the problems with `act[q]` set
static size_t build_pitems(ht3d_batch *b, int pt, const std::vector<char> &act, uint2 *items) {
    size_t n = 0;
    for (int q = 0; q < b->nprob; ++q) {
        if (!act[q]) continue;
        const ht3d_hrl *m = b->m[q];
        const unsigned cnt = pair_items(m->nx, m->ny, m->nz, pt);
        for (unsigned l = 0; l < cnt; ++l) items[n++] = make_uint2((unsigned)q, l);
    }
    return n;
}

static unsigned persistent_grid_threads(const ht3d_batch *b, size_t items, size_t threads) {
    size_t g = (size_t)b->ncu * (2048 / threads);
    if (g > items) g = items;
    if (g >= 8) g &= ~(size_t)7;  // a multiple of 8 keeps item w on XCD w mod 8
    return (unsigned)(g ? g : 1);
}

static unsigned persistent_grid(const ht3d_batch *b, size_t items, int tile) {
    return persistent_grid_threads(b, items, 64 * (size_t)kBTiles[tile].ty);
}

// upload the work list of the active problems (pinned staging, stream ordered)
static int upload_items(ht3d_batch *b, int tile, const std::vector<char> &act, size_t *n_out) {
    const int s = b->stage;
    b->stage ^= 1;
    const size_t n = build_items(b, tile, act, b->h_items[s], b->h_range[s]);
    const size_t np = b->ptile_cur >= 0 ? build_pitems(b, b->ptile_cur, act, b->h_pitems[s]) : 0;
    b->h_nitems[s][0] = (unsigned)n;
    b->h_nitems[s][2] = (unsigned)np;
    hipStream_t st = ctx().stream;  // the staging half used here was last read before the previous poll's event
    if (n) HIP_TRY(hipMemcpyAsync(b->d_items, b->h_items[s], n * sizeof(uint2), hipMemcpyHostToDevice, st));
    if (np) HIP_TRY(hipMemcpyAsync(b->d_pitems, b->h_pitems[s], np * sizeof(uint2), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_range, b->h_range[s], b->nprob * sizeof(uint2), hipMemcpyHostToDevice, st));
    // the item count ([0]) and the pair item count ([2]), not the iteration base ([1]) between them
    HIP_TRY(hipMemcpyAsync(b->d_nitems, b->h_nitems[s], sizeof(unsigned), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_nitems + 2, b->h_nitems[s] + 2, sizeof(unsigned), hipMemcpyHostToDevice, st));
    if (n_out) *n_out = n;
    return HT3D_OK;
}

// fill the host tables of a solve kind from the problems' field descriptions
static void fill_tables(ht3d_batch *b, int kind, const std::vector<ht3d_dyrel_field> &f, int tile64, int tile32,
                        int pt64 = -1, int pt32 = -1) {
    for (int q = 0; q < b->nprob; ++q) {
        const ht3d_dyrel_field &d = f[q];
        const int nx = d.ncx, ny = d.ncy, nz = d.ncz;
        const double rc_tr_dt = d.rhoC * (double)d.transient / d.dt;
        IterArgs<double> &a = b->h_it64[kind][q];
        std::memset(&a, 0, sizeof a);
        a.Xin = d.X; a.Xout = b->xb + b->oe[q];
        a.R = d.R; a.kx = d.kx; a.ky = d.ky; a.kz = d.kz; a.kv = d.kv; a.Dst = nullptr;
        a.F0 = b->f0 + b->oc[q]; a.Fst = b->fst + b->oc[q]; a.partials = nullptr;
        a.nx = nx; a.ny = ny; a.nz = nz;
        a.nbx = (int)bpatch_count(nx, ny, tile64); a.nby = 1;   // patches of the (x, y) plane
        a.nbz = (int)div_up(nz, kBTiles[tile64].kz);
        a.nband = pt64 >= 0 ? (int)div_up(ny, pair_band_rows(pt64, nx) - 2) : 0;
        a._dx = 1.0 / d.dx; a._dy = 1.0 / d.dy; a._dz = 1.0 / d.dz; a._dt = 1.0 / d.dt;
        a.dx = d.dx; a.dy = d.dy; a.dz = d.dz; a.rc_tr_dt = rc_tr_dt;
        set_inv_h2(a, d.dx, d.dy, d.dz);
        for (int k = 0; k < 6; ++k) { a.bct[k] = d.bc_type[k]; a.bcv[k] = d.bc_val[k]; }

        const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz;
        IterArgs<float> &g = b->h_it32[kind][q];
        std::memset(&g, 0, sizeof g);
        g.Xin = b->e0 + b->oe[q]; g.Xout = b->e1 + b->oe[q];
        g.R = b->r32 + b->oc[q];
        g.kx = b->k32 + b->of[q]; g.ky = g.kx + nfx; g.kz = g.ky + nfy; g.Dst = nullptr;
        g.kv = d.kv ? b->kv32 + b->ov[q] : nullptr;
        g.F0 = b->f032 + b->oc[q]; g.Fst = b->fst32 + b->oc[q]; g.partials = nullptr;
        g.nx = nx; g.ny = ny; g.nz = nz;
        g.nbx = (int)bpatch_count(nx, ny, tile32); g.nby = 1;
        g.nbz = (int)div_up(nz, kBTiles[tile32].kz);
        g.nband = pt32 >= 0 ? (int)div_up(ny, pair_band_rows(pt32, nx) - 2) : 0;
        g._dx = (float)a._dx; g._dy = (float)a._dy; g._dz = (float)a._dz; g._dt = (float)a._dt;
        g.dx = (float)d.dx; g.dy = (float)d.dy; g.dz = (float)d.dz; g.rc_tr_dt = (float)rc_tr_dt;
        set_inv_h2(g, d.dx, d.dy, d.dz);
        for (int k = 0; k < 6; ++k) { g.bct[k] = d.bc_type[k]; g.bcv[k] = 0.0f; }  // homogeneous for the correction

        DcArgs &c = b->h_dc[kind][q];
        std::memset(&c, 0, sizeof c);
        c.X = d.X; c.R = d.R; c.kx = d.kx; c.ky = d.ky; c.kz = d.kz; c.kv = d.kv; c.r32 = (float *)g.R;
        c.partials = nullptr;
        c.nx = nx; c.ny = ny; c.nz = nz;
        c._dx = a._dx; c._dy = a._dy; c._dz = a._dz; c._dt = a._dt;
        for (int k = 0; k < 6; ++k) { c.bct[k] = d.bc_type[k]; c.bcv[k] = d.bc_val[k]; }
    }
}

static int upload_tables(ht3d_batch *b, int kind) {
    const size_t n64 = b->nprob * sizeof(IterArgs<double>), n32 = b->nprob * sizeof(IterArgs<float>),
                 ndc = b->nprob * sizeof(DcArgs);
    char *h = (char *)b->h_tabs + (size_t)kind * (n64 + n32 + ndc);
    // the staging of this kind was last read by a copy before the previous
    // solve of this kind, which has been synchronised since
    std::memcpy(h, b->h_it64[kind].data(), n64);
    std::memcpy(h + n64, b->h_it32[kind].data(), n32);
    std::memcpy(h + n64 + n32, b->h_dc[kind].data(), ndc);
    hipStream_t st = ctx().stream;
    HIP_TRY(hipMemcpyAsync(b->d_it64[kind], h, n64, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_it32[kind], h + n64, n32, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_dc[kind], h + n64 + n32, ndc, hipMemcpyHostToDevice, st));
    return HT3D_OK;
}

static BLaunch make_launch(ht3d_batch *b, int kind, bool fp32, int tile) {
    BLaunch L;
    L.tab = fp32 ? (const void *)b->d_it32[kind] : (const void *)b->d_it64[kind];
    L.dtab = b->d_dc[kind];
    L.items = b->d_items;
    L.n_items = b->d_nitems;
    L.st = b->d_state;
    L.partials = b->d_part;
    size_t n = 0;
    for (int q = 0; q < b->nprob; ++q) {
        const ht3d_hrl *m = b->m[q];
        n += btile_count(m->nx, m->ny, m->nz, tile);
    }
    L.grid = b->gmode_cur ? (unsigned)(n ? n : 1) : persistent_grid(b, n, tile);
    L.threads = 64 * kBTiles[tile].ty;
    L.pitems = b->d_pitems;
    L.n_pitems = b->d_nitems + 2;
    L.pgrid = 1;
    if (b->ptile_cur >= 0) {
        size_t np = 0;
        for (int q = 0; q < b->nprob; ++q) np += pair_items(b->m[q]->nx, b->m[q]->ny, b->m[q]->nz, b->ptile_cur);
        L.pgrid = b->pgmode_cur ? (unsigned)(np ? np : 1)
                                : persistent_grid_threads(b, np, 64 * (size_t)kBPairTiles[b->ptile_cur].ty);
    }
    return L;
}

// launch one block of operations on the library stream
static int enqueue_ops(ht3d_batch *b, const std::vector<BOp> &ops, int kind, bool transient, int tile,
                       const ht3d_dyrel_params &p) {
    const bool fp32 = b->precision == HT3D_FP32_DC;
    const BLaunch L = make_launch(b, kind, fp32, tile);
    hipStream_t s = ctx().stream;
    const unsigned nq = (unsigned)b->nprob;
    for (const BOp &o : ops) {
        switch (o.kind) {
        case OP_ITER64:
            if (transient) launch_iter_b<double, true>(tile, L, o.a, o.flags, b->kv_cur);
            else launch_iter_b<double, false>(tile, L, o.a, o.flags, b->kv_cur);
            break;
        case OP_ITER32:
            if (transient) launch_iter_b<float, true>(tile, L, o.a, o.flags, b->kv_cur);
            else launch_iter_b<float, false>(tile, L, o.a, o.flags, b->kv_cur);
            break;
        case OP_PAIR64:
            if (b->ptile_cur < 0) return arg_fail("batch: pair launch without a pair tile");
            if (transient) launch_pair_b<double, true>(b->ptile_cur, L, o.a, o.b, (long long)(b->f0b - b->f0), o.flags);
            else launch_pair_b<double, false>(b->ptile_cur, L, o.a, o.b, (long long)(b->f0b - b->f0), o.flags);
            break;
        case OP_PAIR32:
            if (b->ptile_cur < 0) return arg_fail("batch: pair launch without a pair tile");
            if (transient) launch_pair_b<float, true>(b->ptile_cur, L, o.a, o.b, (long long)(b->f032b - b->f032), o.flags);
            else launch_pair_b<float, false>(b->ptile_cur, L, o.a, o.b, (long long)(b->f032b - b->f032), o.flags);
            break;
        case OP_CHECK64:
            hipLaunchKernelGGL(batch_check_kernel, dim3(nq), dim3(1024), 0, s, b->d_state, b->d_it64[kind], b->d_range,
                               b->d_part, b->d_nitems + 1, p, o.a, o.b);
            break;
        case OP_FINAL64:
            hipLaunchKernelGGL(batch_final_kernel, dim3(div_up(nq, 256)), dim3(256), 0, s, b->d_state, (int)nq,
                               p.nitmax);
            break;
        case OP_ANCHOR_RES:
            if (transient) launch_res_b<true>(tile, L, b->kv_cur);
            else launch_res_b<false>(tile, L, b->kv_cur);
            break;
        case OP_ANCHOR:
            hipLaunchKernelGGL(batch_anchor_kernel, dim3(nq), dim3(1024), 0, s, b->d_state, b->d_it32[kind], b->d_range,
                               b->d_part, p, o.a);
            break;
        case OP_LAM32:
            hipLaunchKernelGGL(batch_lambda32_kernel, dim3(nq), dim3(1024), 0, s, b->d_state, b->d_it32[kind],
                               b->d_range, b->d_part, p, o.b);
            break;
        case OP_UPDATE: launch_upd_b(tile, L, o.a); break;
        case OP_ADVANCE:
            hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, s, b->d_nitems + 1, (unsigned)o.a);
            break;
        default: return arg_fail("batch: unknown operation");
        }
        HIP_TRY(hipGetLastError());
    }
    return HT3D_OK;
}

// launch a block, from a cached graph when graphs are on
static int run_block(ht3d_batch *b, const std::vector<BOp> &ops, int kind, bool transient, int tile,
                     const ht3d_dyrel_params &p) {
    ++b->stats.blocks;
    if (!b->graphs) return enqueue_ops(b, ops, kind, transient, tile, p);
    std::string key;
    key.reserve(ops.size() * 16 + 16);
    const int hdr[8] = {kind, transient ? 1 : 0, tile, b->precision, b->gmode_cur, b->ptile_cur, b->pgmode_cur,
                        b->kv_cur};
    key.append((const char *)hdr, sizeof hdr);
    for (const BOp &o : ops) key.append((const char *)&o, sizeof o);
    auto it = b->cache.find(key);
    hipStream_t s = ctx().stream;
    if (it == b->cache.end()) {
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_ops(b, ops, kind, transient, tile, p);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture(batch block)");
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate(batch block)");
        it = b->cache.emplace(key, ex).first;
        ++b->stats.graphs;
    }
    HIP_TRY(hipGraphLaunch(it->second, s));
    ++b->stats.graph_launches;
    return HT3D_OK;
}

static inline bool lam_it(int it, int nitout) { return it <= 2 || (nitout > 0 && it % nitout == 0); }

// a tile with CPL x-cells per lane needs every problem's ncx to be a multiple
// of CPL; 4 cells per lane only in FP32 (in FP64 they exhaust the registers)
static bool btile_ok(const ht3d_batch *b, int tile, bool fp32) {
    const int cpl = kBTiles[tile].cpl;
    if (cpl > (fp32 ? 4 : 2)) return false;
    for (const ht3d_hrl *m : b->m)
        if (m->nx % cpl) return false;
    return true;
}

static int g_force_btile = -1;   // ht3d_batch_tune_force
static int g_force_bgrid = -1;   // ht3d_batch_tune_force_grid
static int g_bpairs = -1;        // ht3d_batch_tune_pairs
static int g_bkv = -1;           // ht3d_batch_tune_vertex

// the vertex-coefficient form can run on batch tile t: 2 cells per lane (and
// 4 in FP32), and some problem solves with a vertex field
static bool bkv_ok(int t, bool fp32, const std::vector<ht3d_dyrel_field> &f) {
    const int cpl = kBTiles[t].cpl;
    if (!(cpl == 2 || (cpl == 4 && fp32))) return false;
    for (const ht3d_dyrel_field &d : f)
        if (d.kv) return true;
    return false;
}
static int g_force_bptile = -1;  // ht3d_batch_tune_force_pair

// The pair tile of a solve kind (-1: none).  Every pair tile that fits all
// problems is timed with both launch grids on the whole work list, like the
// single tiles; the fastest is kept when its time per iteration is at least
// 7 % below `single_us` (the single-iteration launch), or whenever pairs are
// forced.  The margin: in the solves, where runs of pairs stop at every check
// and the work list shrinks as problems converge, FP64 pairs timed 2-5 % faster
// than single launches on the whole list made the step 2-4 % slower, FP64
// pairs with two x-cells per lane timed 10-13 % faster made it 2 % faster, and
// the FP32 pairs, timed 20-30 % faster, make it 12-24 % faster.
// Without timing (a pinned single tile, or the autotuner off) pairs are used
// only when forced: the pinned pair tile, else the first that fits.
static int autotune_pairs(ht3d_batch *b, int kind, bool transient, bool fp32, const std::vector<ht3d_dyrel_field> &f,
                          int tile, double single_us, bool timed, int *pt_out, int *pg_out) {
    *pt_out = -1;
    *pg_out = g_force_bgrid > 0 ? 1 : 0;
    if (g_bpairs == 0) return HT3D_OK;
    std::vector<int> cand;
    for (int t = 0; t < kNumBPairTiles; ++t) {
        if (g_force_bptile >= 0 && t != g_force_bptile) continue;
        if (kBPairTiles[t].cpl == 4 && !fp32) continue;   // 4 cells per lane: FP32 only
        bool fits = true;
        for (const ht3d_hrl *m : b->m) fits = fits && pair_tile_fits(t, m->nx);
        if (fits) cand.push_back(t);
    }
    if (cand.empty()) return HT3D_OK;
    if (!timed) {
        if (g_bpairs == 1) *pt_out = cand[0];
        return HT3D_OK;
    }
    hipStream_t s = ctx().stream;
    std::vector<char> all(b->nprob, 1);
    std::vector<double> tbest(2 * cand.size(), 1e300);
    for (int round = 0; round < 3; ++round)
        for (size_t cg = 0; cg < 2 * cand.size(); ++cg) {
            const int gm = (int)(cg & 1);
            if (g_force_bgrid >= 0 && gm != g_force_bgrid) continue;
            b->ptile_cur = cand[cg >> 1];
            b->pgmode_cur = gm;
            fill_tables(b, kind, f, tile, tile, b->ptile_cur, b->ptile_cur);
            int rc;
            if ((rc = upload_tables(b, kind)) || (rc = upload_items(b, tile, all, nullptr))) return rc;
            // every problem active: the anchor timed before may have stopped some,
            // and a stopped problem's items are skipped
            hipLaunchKernelGGL(batch_activate_kernel, dim3(div_up(b->nprob, 256)), dim3(256), 0, s, b->d_state,
                               b->nprob);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(s));  // the staging is reused by the next candidate
            const std::vector<BOp> ops(1, BOp{fp32 ? OP_PAIR32 : OP_PAIR64, 0u, 0, 0});
            if ((rc = enqueue_ops(b, ops, kind, transient, tile, ht3d_dyrel_params{}))) return rc;
            HIP_TRY(hipEventRecord(b->ev0, s));
            for (int r = 0; r < 4; ++r)
                if ((rc = enqueue_ops(b, ops, kind, transient, tile, ht3d_dyrel_params{}))) return rc;
            HIP_TRY(hipEventRecord(b->ev1, s));
            HIP_TRY(hipEventSynchronize(b->ev1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            if (ms < tbest[cg]) tbest[cg] = ms;
        }
    size_t bc = 0;
    for (size_t cg = 1; cg < tbest.size(); ++cg)
        if (tbest[cg] < tbest[bc]) bc = cg;
    b->ptile_cur = -1;
    const double pair_us = 1e3 * tbest[bc] / 8;   // per iteration
    if (g_bpairs == 1 || pair_us < 0.93 * single_us) {
        *pt_out = cand[bc >> 1];
        *pg_out = (int)(bc & 1);
        b->t_it[kind][fp32 ? 1 : 0] = pair_us;
    }
    return HT3D_OK;
}

// time `reps` plain batched sweeps of every tile and keep the fastest
static int autotune_batch(ht3d_batch *b, int kind, bool transient, bool fp32, const std::vector<ht3d_dyrel_field> &f,
                          int *tile_out, int *gmode_out) {
    *gmode_out = g_force_bgrid > 0 ? 1 : 0;
    const int pk = fp32 ? 1 : 0;
    if (g_force_btile >= 0 || !g_autotune) {   // a pinned tile the batch cannot run falls back to the default
        *tile_out = (g_force_btile >= 0 && btile_ok(b, g_force_btile, fp32)) ? g_force_btile : kBDefaultTile;
        b->kvm[kind][pk] = g_bkv > 0 && bkv_ok(*tile_out, fp32, f);
        return autotune_pairs(b, kind, transient, fp32, f, *tile_out, 0.0, false, &b->ptile[kind][pk],
                              &b->pgmode[kind][pk]);
    }
    hipStream_t s = ctx().stream;
    b->ptile_cur = -1;   // the single tiles are timed alone
    if (fp32) {   // the timed anchor below adds e (the phase-0 buffer, which the
                  // phase-0 sweeps only read) to the problems' unknowns: it must
                  // hold zeros, and the sweeps should run on defined data
        int rc;
        if ((rc = zero_dev(b->e0, b->se * sizeof(float))) || (rc = zero_dev(b->e1, b->se * sizeof(float))) ||
            (rc = zero_dev(b->r32, b->sc * sizeof(float))) || (rc = zero_dev(b->f032, b->sc * sizeof(float))) ||
            (rc = zero_dev(b->fst32, b->sc * sizeof(float))))
            return rc;
    }
    hipLaunchKernelGGL(batch_activate_kernel, dim3(div_up(b->nprob, 256)), dim3(256), 0, s, b->d_state, b->nprob);
    HIP_TRY(hipGetLastError());
    std::vector<char> all(b->nprob, 1);
    double best = 1e300;
    int best_tile = kBDefaultTile, best_gm = 0, best_kv = 0;
    ht3d_dyrel_params p{};
    // every (tile, grid mode, vertex form) is timed in three interleaved
    // rounds, keeping its best: one pass ranked tiles a few percent apart by
    // noise.  tg = 4 tile + 2 kv + grid mode
    std::vector<double> tbest(4 * kNumBTiles, 1e300);
    for (int round = 0; round < 3; ++round)
        for (int tg = 0; tg < 4 * kNumBTiles; ++tg) {
            const int t = tg >> 2, kv = (tg >> 1) & 1;
            if (!btile_ok(b, t, fp32) || (g_force_bgrid >= 0 && (tg & 1) != g_force_bgrid)) continue;
            if (kv ? (g_bkv == 0 || !bkv_ok(t, fp32, f)) : g_bkv == 1 && bkv_ok(t, fp32, f)) continue;
            b->gmode_cur = tg & 1;
            b->kv_cur = kv;
            fill_tables(b, kind, f, t, t);
            int rc;
            if ((rc = upload_tables(b, kind))) return rc;
            if ((rc = upload_items(b, t, all, nullptr))) return rc;
            HIP_TRY(hipStreamSynchronize(s));  // the staging is reused by the next candidate
            const std::vector<BOp> ops(1, BOp{fp32 ? OP_ITER32 : OP_ITER64, 0u, 0, 0});
            if ((rc = enqueue_ops(b, ops, kind, transient, t, p))) return rc;
            HIP_TRY(hipEventRecord(b->ev0, s));
            const int reps = 4;
            for (int r = 0; r < reps; ++r)
                if ((rc = enqueue_ops(b, ops, kind, transient, t, p))) return rc;
            HIP_TRY(hipEventRecord(b->ev1, s));
            HIP_TRY(hipEventSynchronize(b->ev1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            if (ms < tbest[tg]) tbest[tg] = ms;
        }
    for (int tg = 0; tg < 4 * kNumBTiles; ++tg)
        if (tbest[tg] < best) {
            best = tbest[tg];
            best_tile = tg >> 2;
            best_kv = (tg >> 1) & 1;
            best_gm = tg & 1;
        }
    *tile_out = best_tile;
    *gmode_out = best_gm;
    b->gmode_cur = best_gm;
    b->kvm[kind][pk] = best_kv;
    b->kv_cur = best_kv;
    // costs for the check-interval tuner: one sweep, and what a check (FP64) or
    // an anchor with its update (defect correction) adds, on the chosen tile
    {
        fill_tables(b, kind, f, best_tile, best_tile);
        int rc;
        if ((rc = upload_tables(b, kind)) || (rc = upload_items(b, best_tile, all, nullptr))) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        auto time_ops = [&](const std::vector<BOp> &ops, double *us) -> int {
            const int reps = 4;
            int r2 = HT3D_OK;
            for (int r = -1; r < reps && !r2; ++r) {
                if (r == 0) HIP_TRY(hipEventRecord(b->ev0, s));
                hipLaunchKernelGGL(batch_activate_kernel, dim3(div_up(b->nprob, 256)), dim3(256), 0, s, b->d_state,
                                   b->nprob);
                HIP_TRY(hipGetLastError());
                r2 = enqueue_ops(b, ops, kind, transient, best_tile, p);
            }
            if (r2) return r2;
            HIP_TRY(hipEventRecord(b->ev1, s));
            HIP_TRY(hipEventSynchronize(b->ev1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            *us = 1e3 * ms / reps;
            return HT3D_OK;
        };
        double plain = 0, chk = 0;
        if ((rc = time_ops({BOp{fp32 ? OP_ITER32 : OP_ITER64, 0u, 0, 0}}, &plain))) return rc;
        if (fp32) rc = time_ops({BOp{OP_ANCHOR_RES, 0u, 0, 0}, BOp{OP_ANCHOR, 0u, 0, 0}, BOp{OP_UPDATE, 0u, 0, 0}}, &chk);
        else rc = time_ops({BOp{OP_ITER64, FL_CHECK, 0, 0}, BOp{OP_CHECK64, 0u, 1, 0}}, &chk);
        if (rc) return rc;
        const double t_sweep = 1e3 * best / 4;
        b->t_it[kind][fp32 ? 1 : 0] = t_sweep;
        b->t_chk[kind][fp32 ? 1 : 0] = fp32 ? fmax(chk - (plain - t_sweep), 1.0) : fmax(chk - plain, 1.0);
    }
    return autotune_pairs(b, kind, transient, fp32, f, best_tile, 1e3 * best / 4, true, &b->ptile[kind][pk],
                          &b->pgmode[kind][pk]);
}

// Poll loop shared by both precisions: `next` appends the next block's
// operations and returns false when there is none; the device states are read
// one block behind the launches and the work list is repacked when problems
// have stopped.
template <typename NextFn>
static int poll_loop(ht3d_batch *b, int kind, bool transient, int tile, const ht3d_dyrel_params &p, NextFn next) {
    hipStream_t s = ctx().stream;
    std::vector<char> act(b->nprob, 1);
    int n_active = b->nprob;
    long long nblk = 0;
    std::vector<BOp> ops;
    HIP_TRY(hipEventRecord(b->ev0, s));
    for (;;) {
        ops.clear();
        if (!next(ops)) break;
        int rc = run_block(b, ops, kind, transient, tile, p);
        if (rc) return rc;
        const int slot = (int)(nblk & 1);
        HIP_TRY(hipMemcpyAsync(b->h_state[slot], b->d_state, b->nprob * sizeof(BState), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipEventRecord(b->ev[slot], s));
        if (nblk > 0) {
            const int prev = slot ^ 1;
            HIP_TRY(hipEventSynchronize(b->ev[prev]));
            int n = 0;
            for (int q = 0; q < b->nprob; ++q) {
                act[q] = (char)(b->h_state[prev][q].active != 0);
                n += act[q];
            }
            if (n == 0) { ++nblk; break; }
            if (n < n_active) {
                n_active = n;
                if ((rc = upload_items(b, tile, act, nullptr))) return rc;
                ++b->stats.repacks;
            }
        }
        ++nblk;
    }
    HIP_TRY(hipEventRecord(b->ev1, s));
    HIP_TRY(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->stats.ms_solve += ms;
    return HT3D_OK;
}

// The batched solve of one kind; reports and (thermal) Nusselt sums per problem.
static int batch_solve(ht3d_batch *b, int kind, const ht3d_dyrel_params &p_in, const std::vector<ht3d_dyrel_field> &f,
                       std::vector<ht3d_dyrel_report> &rep, std::vector<int> &outer, std::vector<double> *nusum) {
    hipStream_t s = ctx().stream;
    const bool transient = f[0].transient != 0;
    const bool fp32 = b->precision == HT3D_FP32_DC;
    int rc;
    if (fp32)  // FP32 face coefficients (and vertex fields) of the correction equation, as the single-problem solve forms them
        for (int q = 0; q < b->nprob; ++q) {
            const ht3d_dyrel_field &d = f[q];
            if ((rc = faces_f32(b->k32 + b->of[q], &d))) return rc;
            if (d.kv && (rc = to_float(b->kv32 + b->ov[q], d.kv, (size_t)(d.ncx + 1) * (d.ncy + 1) * (d.ncz + 1))))
                return rc;
        }
    int &tile = fp32 ? b->tile32[kind] : b->tile64[kind];
    if (tile < 0 && (rc = autotune_batch(b, kind, transient, fp32, f, &tile, &b->gmode[kind][fp32 ? 1 : 0]))) return rc;
    b->gmode_cur = b->gmode[kind][fp32 ? 1 : 0];
    b->kv_cur = b->kvm[kind][fp32 ? 1 : 0];
    b->ptile_cur = b->ptile[kind][fp32 ? 1 : 0];
    b->pgmode_cur = b->pgmode[kind][fp32 ? 1 : 0];
    ht3d_dyrel_params p = p_in;
    const int pk = fp32 ? 1 : 0;
    if (p.ncheck < 0)   // the check-interval tuner (the anchor interval in the defect-correction form)
        p.ncheck = tuned_check_interval(b->last_iters[kind][pk], b->t_it[kind][pk], b->t_chk[kind][pk], p.nitout);
    fill_tables(b, kind, f, tile, tile, b->ptile_cur, b->ptile_cur);
    if ((rc = upload_tables(b, kind))) return rc;
    std::vector<char> all(b->nprob, 1);
    if ((rc = upload_items(b, tile, all, nullptr))) return rc;

    // set-up: Gershgorin lambda_max and the DYREL scalars per problem
    for (int q = 0; q < b->nprob; ++q) {
        const ht3d_dyrel_field &d = f[q];
        const double rc_tr_dt = d.rhoC * (double)d.transient / d.dt;
        if ((rc = launch_gershgorin(d.kx, d.ky, d.kz, d.ncx, d.ncy, d.ncz, d.dx, d.dy, d.dz, rc_tr_dt,
                                    b->d_gpart + (size_t)8 * b->g_off[q], b->g_nb[q], nullptr)))
            return rc;
    }
    if (fp32) {
        hipLaunchKernelGGL(batch_setup_kernel<float>, dim3(b->nprob), dim3(1024), 0, s, b->d_state, b->d_it32[kind],
                           b->d_gpart, b->d_grange, p);
        HIP_TRY(hipGetLastError());
        if ((rc = zero_dev(b->e0, b->se * sizeof(float))) || (rc = zero_dev(b->e1, b->se * sizeof(float))) ||
            (rc = zero_dev(b->f032, b->sc * sizeof(float))) || (rc = zero_dev(b->fst32, b->sc * sizeof(float))))
            return rc;
    } else {
        hipLaunchKernelGGL(batch_setup_kernel<double>, dim3(b->nprob), dim3(1024), 0, s, b->d_state, b->d_it64[kind],
                           b->d_gpart, b->d_grange, p);
        HIP_TRY(hipGetLastError());
        if ((rc = zero_dev(b->f0, b->sc * sizeof(double))) || (rc = zero_dev(b->fst, b->sc * sizeof(double))))
            return rc;
    }
    hipLaunchKernelGGL(batch_set_kernel, dim3(1), dim3(64), 0, s, b->d_nitems + 1, 0u);  // iteration base
    HIP_TRY(hipGetLastError());

    // A run of `m` plain iterations from `first` (no check, lambda update or
    // stored F): pairs when the solve has a pair tile, always an even number
    // of them so that X and F0 are back in the buffers a single launch would
    // have left them in, then single iterations.  `xph` is the buffer holding
    // the current iterate.
    auto rev = [&](int it) { return (g_alternate && (it & 1) == 0) ? FL_REVERSE : 0u; };
    auto emit_run = [&](std::vector<BOp> &ops, int first, int m, int &xph) {
        int npair = b->ptile_cur >= 0 ? m / 2 : 0;
        if (npair & 1) --npair;
        int it = first, fpar = 0;
        for (int k = 0; k < npair; ++k, it += 2) {
            ops.push_back(BOp{fp32 ? OP_PAIR32 : OP_PAIR64, rev(it), xph, fpar});
            xph ^= 1;
            fpar ^= 1;
        }
        for (; it < first + m; ++it) {
            ops.push_back(BOp{fp32 ? OP_ITER32 : OP_ITER64, rev(it), xph, 0});
            xph ^= 1;
        }
    };
    if (!fp32) {
        int it0 = 0;
        rc = poll_loop(b, kind, transient, tile, p, [&](std::vector<BOp> &ops) {
            if (it0 >= p.nitmax) return false;
            auto lam = [&](int it) { return lam_it(it, p.nitout); };
            auto chk = [&](int it) { return lam(it) || (p.ncheck > 0 && it % p.ncheck == 0); };
            auto store = [&](int it) { return it < p.nitmax && lam_it(it + 1, p.nitout); };
            auto plain = [&](int it) { return !chk(it) && !store(it) && it < p.nitmax; };
            int end = it0;   // the block runs to the next check, or to nitmax
            do ++end; while (!chk(end) && end < p.nitmax);
            int xph = it0 & 1, L = 0;
            for (int it = it0 + 1; it <= end;) {
                if (plain(it)) {
                    int m = 1;
                    while (it + m <= end && plain(it + m)) ++m;
                    emit_run(ops, it, m, xph);
                    it += m;
                    L += m;
                    continue;
                }
                ++L;
                const unsigned fl = (chk(it) ? FL_CHECK : 0u) | (lam(it) ? FL_LAMBDA : 0u) |
                                    (store(it) ? FL_STORE_F : 0u) | rev(it);
                ops.push_back(BOp{OP_ITER64, fl, xph, 0});
                xph ^= 1;
                if (chk(it)) ops.push_back(BOp{OP_CHECK64, 0u, L, lam(it) ? 1 : 0});
                ++it;
            }
            if (end == p.nitmax) ops.push_back(BOp{OP_FINAL64, 0u, 0, 0});
            ops.push_back(BOp{OP_ADVANCE, 0u, L, 0});
            it0 = end;
            b->stats.iterations += L;
            return true;
        });
    } else {
        const int anchor = p.ncheck > 0 ? p.ncheck : (p.nitout > 0 ? p.nitout : 50);
        int i0 = 0;
        bool last = false;
        rc = poll_loop(b, kind, transient, tile, p, [&](std::vector<BOp> &ops) {
            if (last) return false;
            // the first run is one sweep short: the anchors test the iterates
            // x_{k*anchor - 1}, as the FP64 check iterations do, the last one
            // at nitmax - 1 at the latest (the single solve's schedule)
            const int want = (i0 == 0 && anchor > 1) ? anchor - 1 : anchor;
            const int run = std::max(0, std::min(want, p.nitmax - 1 - i0));
            ops.push_back(BOp{OP_ANCHOR_RES, 0u, 0, 0});
            ops.push_back(BOp{OP_ANCHOR, 0u, run, 0});
            auto lam = [&](int it) { return lam_it(it, p.nitout); };
            auto store = [&](int it) { return it < p.nitmax && lam_it(it + 1, p.nitout); };
            auto plain = [&](int it) { return !lam(it) && !store(it) && it < p.nitmax; };
            int xph = i0 & 1;
            for (int it = i0 + 1; it <= i0 + run;) {
                if (plain(it)) {
                    int m = 1;
                    while (it + m <= i0 + run && plain(it + m)) ++m;
                    emit_run(ops, it, m, xph);
                    it += m;
                    continue;
                }
                const unsigned fl = (lam(it) ? (FL_CHECK | FL_LAMBDA) : 0u) | (store(it) ? FL_STORE_F : 0u) | rev(it);
                ops.push_back(BOp{OP_ITER32, fl, xph, 0});
                xph ^= 1;
                if (lam(it)) ops.push_back(BOp{OP_LAM32, 0u, 0, it > 1 ? 1 : 0});
                ++it;
            }
            if (run > 0) ops.push_back(BOp{OP_UPDATE, 0u, (i0 + run) & 1, 0});
            else last = true;
            i0 += run;
            b->stats.iterations += run;
            return true;
        });
    }
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(b->h_state[0], b->d_state, b->nprob * sizeof(BState), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));

    // finalise per problem; Nu sums from the state the reference evaluates it on
    double maxit = 0;
    for (int q = 0; q < b->nprob; ++q) maxit = fmax(maxit, (double)b->h_state[0][q].r.iters);
    b->last_iters[kind][pk] = maxit;
    for (int q = 0; q < b->nprob; ++q) {
        const BState &st = b->h_state[0][q];
        const ht3d_dyrel_field &d = f[q];
        rep[q] = st.r;
        rep[q].ncheck = p.ncheck > 0 ? p.ncheck : p.nitout;
        outer[q] = st.outer;
        const double *nu_src = d.X;
        if (!fp32) {
            double *buf[2] = {d.X, b->xb + b->oe[q]};
            const int it = st.r.iters > 0 ? st.r.iters : 1;
            double *res = buf[it & 1], *prev = buf[(it - 1) & 1];
            if ((rc = bc_fill_table(prev, d.ncx, d.ncy, d.ncz, d.bc_type, d.bc_val))) return rc;
            if ((rc = copy_ghosts(res, prev, d.ncx, d.ncy, d.ncz))) return rc;
            nu_src = prev;
            if (nusum) {
                if ((rc = launch_nu_partials(prev, d.ky, d.ncx, d.ncy, d.ncz, 1.0 / d.dy, b->d_nupart + (size_t)8 * b->nu_off[q],
                                             b->nu_nb[q])) ||
                    (rc = finish_partials(b->d_nupart + (size_t)8 * b->nu_off[q], b->nu_nb[q], 0u, b->d_nures + 4 * q)))
                    return rc;
            }
            if (res != d.X)
                HIP_TRY(hipMemcpyAsync(d.X, res, (size_t)(d.ncx + 2) * (d.ncy + 2) * (d.ncz + 2) * sizeof(double),
                                       hipMemcpyDeviceToDevice, s));
        } else {
            // ghosts and Nu from the tested iterate, then the interior one
            // update further (the single defect-correction solve's result)
            if ((rc = bc_fill_table(d.X, d.ncx, d.ncy, d.ncz, d.bc_type, d.bc_val))) return rc;
            if (nusum) {
                if ((rc = launch_nu_partials(nu_src, d.ky, d.ncx, d.ncy, d.ncz, 1.0 / d.dy,
                                             b->d_nupart + (size_t)8 * b->nu_off[q], b->nu_nb[q])) ||
                    (rc = finish_partials(b->d_nupart + (size_t)8 * b->nu_off[q], b->nu_nb[q], 0u, b->d_nures + 4 * q)))
                    return rc;
            }
            // {h1, h2, dtau} of the problem, as the solve left them in its device table
            const float *hsc = &b->d_it32[kind][q].h1;
            const size_t N = (size_t)d.ncx * d.ncy * d.ncz;
            const unsigned nb = (unsigned)std::min<size_t>(div_up(N, 256), 2048);
            hipLaunchKernelGGL(dc_final_kernel<float>, dim3(nb), dim3(256), 0, s, b->h_it32[kind][q], hsc, d.X);
            HIP_TRY(hipGetLastError());
        }
        // algorithmic bytes: 8 N words per pseudo-iteration, 6 N FP64 words per anchor
        const double N = (double)d.ncx * d.ncy * d.ncz;
        b->stats.bytes_solve += fp32 ? 8.0 * N * 4.0 * st.r.iters + 6.0 * N * 8.0 * st.r.n_checks
                                     : 8.0 * N * 8.0 * st.r.iters;
    }
    if (nusum) {
        HIP_TRY(hipMemcpyAsync(b->h_nures, b->d_nures, (size_t)4 * b->nprob * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int q = 0; q < b->nprob; ++q) (*nusum)[q] = b->h_nures[4 * q];
    }
    return HT3D_OK;
}

}  // namespace ht3d

extern "C" {

// A kernel trace under rocprofv3 ended in a host-side segmentation fault inside
// the runtime's frames below hipGraphLaunch of a newly captured block graph,
// while the same graphs run cleanly without the profiler (and the eager form
// runs cleanly under it).  So when a rocprofiler tool library is configured
// (rocprofv3 sets ROCP_TOOL_LIBRARIES), the batch launches its blocks kernel
// by kernel: the same kernels on the same work lists, bit-identical results.
// HT3D_BATCH_GRAPHS_UNDER_PROFILER=1 keeps the graphs anyway.
static bool profiler_tool_configured() {
    const char *t = std::getenv("ROCP_TOOL_LIBRARIES");
    const char *force = std::getenv("HT3D_BATCH_GRAPHS_UNDER_PROFILER");
    return t && t[0] && !(force && force[0] == '1');
}

int ht3d_batch_create(ht3d_hrl *const *models, int nprob, ht3d_batch **out) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!models || nprob < 1 || !out) return arg_fail("ht3d_batch_create: bad argument");
    for (int q = 0; q < nprob; ++q)
        if (!models[q]) return arg_fail("ht3d_batch_create: null model");
    ht3d_batch *b = new (std::nothrow) ht3d_batch();
    if (!b) return HT3D_ERR_NOMEM;
    b->nprob = nprob;
    b->m.assign(models, models + nprob);
    b->pT = models[0]->pT;
    b->pP = models[0]->pP;
    b->precision = models[0]->precision;
    b->graphs = !profiler_tool_configured();
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx().device) == hipSuccess && ncu > 0)
        b->ncu = ncu;
    // contiguous per-problem regions (each problem's part 256-byte aligned)
    size_t maxit = 0, gsum = 0, nusum = 0;
    for (int q = 0; q < nprob; ++q) {
        const ht3d_hrl *m = models[q];
        const size_t c = (size_t)m->nx * m->ny * m->nz, E = (size_t)(m->nx + 2) * (m->ny + 2) * (m->nz + 2);
        const size_t F = (size_t)(m->nx + 1) * m->ny * m->nz + (size_t)m->nx * (m->ny + 1) * m->nz +
                         (size_t)m->nx * m->ny * (m->nz + 1);
        const size_t V = (size_t)(m->nx + 1) * (m->ny + 1) * (m->nz + 1);
        b->oc.push_back(b->sc); b->oe.push_back(b->se); b->of.push_back(b->sf); b->ov.push_back(b->sv);
        b->sc += round16(c, 8); b->se += round16(E, 8); b->sf += round16(F, 8); b->sv += round16(V, 8);
        b->g_off.push_back((unsigned)gsum); b->g_nb.push_back(gershgorin_blocks(c)); gsum += b->g_nb.back();
        b->nu_off.push_back((unsigned)nusum); b->nu_nb.push_back(nu_blocks(m->nx, m->nz)); nusum += b->nu_nb.back();
    }
    for (int t = 0; t < kNumBTiles; ++t) {
        size_t n = 0;
        for (int q = 0; q < nprob; ++q) {
            const ht3d_hrl *m = models[q];
            n += btile_count(m->nx, m->ny, m->nz, t);
        }
        if (n > maxit) maxit = n;
    }
    b->max_items = maxit;
    size_t maxp = 0;   // the largest pair work list among the pair tiles that fit every problem
    for (int t = 0; t < kNumBPairTiles; ++t) {
        size_t n = 0;
        bool fits = true;
        for (int q = 0; q < nprob && fits; ++q) {
            const ht3d_hrl *m = models[q];
            fits = pair_tile_fits(t, m->nx);
            if (fits) n += pair_items(m->nx, m->ny, m->nz, t);
        }
        if (fits && n > maxp) maxp = n;
    }
    b->max_pitems = maxp;
    int rc = HT3D_OK;
    auto A = [&](void **p, size_t bytes) { if (!rc) rc = b_alloc(b, p, bytes); };
    A((void **)&b->xb, b->se * sizeof(double));
    A((void **)&b->f0, b->sc * sizeof(double));
    A((void **)&b->f0b, b->sc * sizeof(double));
    A((void **)&b->f032b, b->sc * sizeof(float));
    A((void **)&b->fst, b->sc * sizeof(double));
    A((void **)&b->e0, b->se * sizeof(float));
    A((void **)&b->e1, b->se * sizeof(float));
    A((void **)&b->f032, b->sc * sizeof(float));
    A((void **)&b->fst32, b->sc * sizeof(float));
    A((void **)&b->r32, b->sc * sizeof(float));
    A((void **)&b->k32, b->sf * sizeof(float));
    A((void **)&b->kv32, b->sv * sizeof(float));
    for (int k = 0; k < 2; ++k) {
        A((void **)&b->d_it64[k], nprob * sizeof(IterArgs<double>));
        A((void **)&b->d_it32[k], nprob * sizeof(IterArgs<float>));
        A((void **)&b->d_dc[k], nprob * sizeof(DcArgs));
        b->h_it64[k].resize(nprob);
        b->h_it32[k].resize(nprob);
        b->h_dc[k].resize(nprob);
    }
    A((void **)&b->d_items, maxit * sizeof(uint2));
    A((void **)&b->d_pitems, (maxp ? maxp : 1) * sizeof(uint2));
    A((void **)&b->d_range, nprob * sizeof(uint2));
    A((void **)&b->d_grange, nprob * sizeof(uint2));
    A((void **)&b->d_nitems, 64);
    A((void **)&b->d_part, maxit * 8 * sizeof(double));
    A((void **)&b->d_gpart, gsum * 8 * sizeof(double));
    A((void **)&b->d_nupart, nusum * 8 * sizeof(double));
    A((void **)&b->d_nures, (size_t)4 * nprob * sizeof(double));
    A((void **)&b->d_state, nprob * sizeof(BState));
    A((void **)&b->d_weno, nprob * sizeof(WenoItem));
    if (rc) { ht3d_batch_destroy(b); return rc; }
    b->tabs_bytes = 2 * (size_t)nprob * (sizeof(IterArgs<double>) + sizeof(IterArgs<float>) + sizeof(DcArgs));
    hipError_t e = hipSuccess;
    for (int s = 0; s < 2 && e == hipSuccess; ++s) {
        e = hipHostMalloc((void **)&b->h_state[s], nprob * sizeof(BState), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_items[s], (maxit ? maxit : 1) * sizeof(uint2), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_pitems[s], (maxp ? maxp : 1) * sizeof(uint2), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_range[s], nprob * sizeof(uint2), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_nitems[s], 64, hipHostMallocDefault);
        if (e == hipSuccess) std::memset(b->h_nitems[s], 0, 64);
    }
    if (e == hipSuccess) e = hipHostMalloc(&b->h_tabs, b->tabs_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_nures, (size_t)4 * nprob * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_weno, nprob * sizeof(WenoItem), hipHostMallocDefault);
    for (int s = 0; s < 2 && e == hipSuccess; ++s) e = hipEventCreate(&b->ev[s]);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) { ht3d_batch_destroy(b); return hip_fail(e, "ht3d_batch_create"); }
    // Gershgorin partial ranges are fixed
    std::vector<uint2> gr(nprob);
    for (int q = 0; q < nprob; ++q) gr[q] = make_uint2(b->g_off[q], b->g_nb[q]);
    std::memcpy(b->h_range[0], gr.data(), nprob * sizeof(uint2));
    e = hipMemcpyAsync(b->d_grange, b->h_range[0], nprob * sizeof(uint2), hipMemcpyHostToDevice, ctx().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx().stream);
    if (e != hipSuccess) { ht3d_batch_destroy(b); return hip_fail(e, "ht3d_batch_create"); }
    *out = b;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_destroy(ht3d_batch *b) {
    HT3D_API_BEGIN
    if (!b) return HT3D_OK;
    if (ctx().initialised) (void)hipStreamSynchronize(ctx().stream);
    clear_graphs(b);
    for (void *p : b->allocs) (void)hipFree(p);
    for (int s = 0; s < 2; ++s) {
        if (b->h_state[s]) (void)hipHostFree(b->h_state[s]);
        if (b->h_items[s]) (void)hipHostFree(b->h_items[s]);
        if (b->h_pitems[s]) (void)hipHostFree(b->h_pitems[s]);
        if (b->h_range[s]) (void)hipHostFree(b->h_range[s]);
        if (b->h_nitems[s]) (void)hipHostFree(b->h_nitems[s]);
        if (b->ev[s]) (void)hipEventDestroy(b->ev[s]);
    }
    if (b->h_tabs) (void)hipHostFree(b->h_tabs);
    if (b->h_nures) (void)hipHostFree(b->h_nures);
    if (b->h_weno) (void)hipHostFree(b->h_weno);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_set_params(ht3d_batch *b, const ht3d_dyrel_params *pT, const ht3d_dyrel_params *pP, int precision,
                          int use_graphs) {
    HT3D_API_BEGIN
    if (!b) return arg_fail("ht3d_batch_set_params: null batch");
    if (precision != HT3D_FP64 && precision != HT3D_FP32_DC) return arg_fail("ht3d_batch_set_params: bad precision");
    for (const ht3d_dyrel_params *p : {pT, pP})
        if (p && (p->nitmax < 1 || p->nitout < 1 || !(p->tol >= 0.0)))
            return arg_fail("ht3d_batch_set_params: bad DYREL parameters");
    if (pT) b->pT = *pT;
    if (pP) b->pP = *pP;
    b->precision = precision;
    b->graphs = use_graphs != 0 && !profiler_tool_configured();
    clear_graphs(b);  // the graphs carry the parameters as launch arguments
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_step(ht3d_batch *b, ht3d_step_report *reps) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!b) return arg_fail("ht3d_batch_step: null batch");
    const auto t0 = std::chrono::steady_clock::now();
    SyncGuard guard;
    const int n = b->nprob;
    std::vector<ht3d_step_report> r(n);
    std::vector<ht3d_dyrel_field> f(n);
    std::vector<ht3d_dyrel_report> rep(n);
    std::vector<int> outer(n);
    std::vector<double> nus(n, 0.0);
    int rc;
    for (int q = 0; q < n; ++q) {
        std::memset(&r[q], 0, sizeof r[q]);
        r[q].dt_used = b->m[q]->dt;
    }
    int worst = HT3D_OK;
    auto solve_status = [&](int which) {
        for (int q = 0; q < n; ++q) {
            const int st = rep[q].status;
            (which ? r[q].darcy : r[q].thermal) = rep[q];
            if (b->precision == HT3D_FP32_DC) {
                (which ? r[q].dc_outer_darcy : r[q].dc_outer_thermal) = outer[q];
                (which ? r[q].inner_darcy : r[q].inner_thermal) = rep[q].iters;
            }
            if ((st == HT3D_NAN || st == HT3D_DIVERGED) && worst != HT3D_NAN) worst = st;
            if (st == HT3D_NOT_CONVERGED && worst == HT3D_OK) worst = st;
        }
        return worst != HT3D_NAN && worst != HT3D_DIVERGED;
    };
    do {
        // thermal solves, Nu
        for (int q = 0; q < n; ++q)
            if ((rc = hrl_pre_thermal(b->m[q], &f[q]))) return rc;
        if ((rc = batch_solve(b, 0, b->pT, f, rep, outer, &nus))) return rc;
        if (!solve_status(0)) break;
        for (int q = 0; q < n; ++q) hrl_set_nu(b->m[q], nus[q]);
        // Darcy solves
        for (int q = 0; q < n; ++q)
            if ((rc = hrl_pre_darcy(b->m[q], &f[q]))) return rc;
        if ((rc = batch_solve(b, 1, b->pP, f, rep, outer, nullptr))) return rc;
        if (!solve_status(1)) break;
        // velocity of every problem, then one read-back of all the max|V|
        // (a read-back per problem would drain the stream once per problem),
        // then the time steps and the first advection sub-step of every
        // problem in six launches; further sub-steps (the constant-dt rule)
        // per problem.  The read-back's synchronisation also guarantees that
        // the previous step's table upload has finished before h_weno is
        // refilled.
        for (int q = 0; q < n; ++q)
            if ((rc = hrl_velocity_dev(b->m[q], b->d_nures + 4 * q))) return rc;
        hipStream_t s = ctx().stream;
        HIP_TRY(hipMemcpyAsync(b->h_nures, b->d_nures, (size_t)4 * n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        std::vector<int> n_adv(n);
        std::vector<double> dt_a(n);
        unsigned nblocks = 0;
        for (int q = 0; q < n; ++q) {
            const ht3d_hrl *m = b->m[q];
            if ((rc = hrl_advect_plan(b->m[q], b->h_nures[4 * q], &r[q], &n_adv[q], &dt_a[q]))) return rc;
            const unsigned nb = weno_blocks(m->nx, m->ny, m->nz);
            if (nblocks > 0x7fffffffu - nb) return arg_fail("ht3d_batch_step: advection grid too large");
            b->h_weno[q] = hrl_weno_item(m, dt_a[q], nblocks);
            nblocks += nb;
        }
        HIP_TRY(hipMemcpyAsync(b->d_weno, b->h_weno, (size_t)n * sizeof(WenoItem), hipMemcpyHostToDevice, s));
        if ((rc = weno_advect_batch(b->d_weno, n, nblocks))) return rc;
        for (int q = 0; q < n; ++q)
            for (int sub = 1; sub < n_adv[q]; ++sub)
                if ((rc = hrl_weno(b->m[q], dt_a[q]))) return rc;
        for (int q = 0; q < n; ++q) {
            r[q].Nu = b->m[q]->Nu;
            r[q].status = (r[q].thermal.status == HT3D_NOT_CONVERGED || r[q].darcy.status == HT3D_NOT_CONVERGED)
                              ? HT3D_NOT_CONVERGED : HT3D_OK;
        }
    } while (false);
    for (int q = 0; q < n; ++q) {
        if (worst == HT3D_NAN || worst == HT3D_DIVERGED) {
            const int a = r[q].thermal.status, d = r[q].darcy.status;
            r[q].status = (a == HT3D_NAN || a == HT3D_DIVERGED) ? a : (d == HT3D_NAN || d == HT3D_DIVERGED) ? d : HT3D_OK;
        }
        if (reps) reps[q] = r[q];
    }
    if ((rc = after_launch("ht3d_batch_step", true))) return rc;
    b->stats.ms_step += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ++b->stats.steps;
    return worst;
    HT3D_API_END
}

int ht3d_batch_stats_get(const ht3d_batch *b, ht3d_batch_stats *s) {
    HT3D_API_BEGIN
    if (!b || !s) return arg_fail("ht3d_batch_stats_get: bad argument");
    *s = b->stats;
    const int tiles[4] = {b->tile64[0], b->tile64[1], b->tile32[0], b->tile32[1]};
    for (int q = 0; q < 4; ++q) s->tile_cpl[q] = tiles[q] >= 0 ? kBTiles[tiles[q]].cpl : 0;
    s->tile_rows[0] = b->tile64[0] >= 0 ? kBTiles[b->tile64[0]].ty : 0;
    s->tile_rows[1] = b->tile64[1] >= 0 ? kBTiles[b->tile64[1]].ty : 0;
    s->tile_rows[2] = b->tile32[0] >= 0 ? kBTiles[b->tile32[0]].ty : 0;
    s->tile_rows[3] = b->tile32[1] >= 0 ? kBTiles[b->tile32[1]].ty : 0;
    s->tile_z[0] = b->tile64[0] >= 0 ? kBTiles[b->tile64[0]].kz : 0;
    s->tile_z[1] = b->tile64[1] >= 0 ? kBTiles[b->tile64[1]].kz : 0;
    s->tile_z[2] = b->tile32[0] >= 0 ? kBTiles[b->tile32[0]].kz : 0;
    s->tile_z[3] = b->tile32[1] >= 0 ? kBTiles[b->tile32[1]].kz : 0;
    for (int q = 0; q < 4; ++q) {
        s->grid_mode[q] = b->gmode[q & 1][q >> 1];
        s->vertex[q] = b->kvm[q & 1][q >> 1];
        const int pt = b->ptile[q & 1][q >> 1];
        s->pair_rows[q] = pt >= 0 ? kBPairTiles[pt].ty : 0;
        s->pair_z[q] = pt >= 0 ? kBPairTiles[pt].kz : 0;
        s->pair_cpl[q] = pt >= 0 ? kBPairTiles[pt].cpl : 0;
    }
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_tune_force(int tile) {
    HT3D_API_BEGIN
    if (tile >= kNumBTiles) return arg_fail("ht3d_batch_tune_force: tile index past the end of the table");
    g_force_btile = tile < 0 ? -1 : tile;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_tile_count(void) { return kNumBTiles; }

int ht3d_batch_tune_vertex(int mode) {
    HT3D_API_BEGIN
    if (mode < -1 || mode > 1) return arg_fail("ht3d_batch_tune_vertex: mode must be -1, 0 or 1");
    g_bkv = mode;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_tune_force_grid(int mode) {
    HT3D_API_BEGIN
    if (mode > 1) return arg_fail("ht3d_batch_tune_force_grid: mode must be -1, 0 or 1");
    g_force_bgrid = mode < 0 ? -1 : mode;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_tune_pairs(int mode) {
    HT3D_API_BEGIN
    if (mode < -1 || mode > 1) return arg_fail("ht3d_batch_tune_pairs: mode must be -1, 0 or 1");
    g_bpairs = mode;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_tune_force_pair(int tile) {
    HT3D_API_BEGIN
    if (tile >= kNumBPairTiles) return arg_fail("ht3d_batch_tune_force_pair: tile index past the end of the table");
    g_force_bptile = tile < 0 ? -1 : tile;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_pair_tile_count(void) { return kNumBPairTiles; }

int ht3d_batch_pair_tile_info(int tile, int *rows, int *planes, int *cells_per_lane) {
    HT3D_API_BEGIN
    if (tile < 0 || tile >= kNumBPairTiles) return arg_fail("ht3d_batch_pair_tile_info: tile index out of range");
    if (rows) *rows = kBPairTiles[tile].ty;
    if (planes) *planes = kBPairTiles[tile].kz;
    if (cells_per_lane) *cells_per_lane = kBPairTiles[tile].cpl;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_batch_stats_reset(ht3d_batch *b) {
    HT3D_API_BEGIN
    if (!b) return arg_fail("ht3d_batch_stats_reset: null batch");
    std::memset(&b->stats, 0, sizeof b->stats);
    return HT3D_OK;
    HT3D_API_END
}

}  // extern "C"
