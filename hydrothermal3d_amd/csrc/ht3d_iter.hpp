// ht3d_iter.hpp -- the fused DYREL pseudo-iteration kernel, its tile table and
// the autotuner, shared by the FP64 solve (ht3d_dyrel.hip) and the FP32 inner
// sweeps of the defect-correction solve (ht3d_dc.hip).
#pragma once

#include <vector>

#include "ht3d_internal.hpp"

namespace ht3d {

enum : unsigned { FL_CHECK = 1u, FL_LAMBDA = 2u, FL_STORE_F = 4u };

template <typename T>
struct IterArgs {
    const T *__restrict__ Xin;
    T *__restrict__ Xout;
    const T *__restrict__ R;
    const T *__restrict__ kx;
    const T *__restrict__ ky;
    const T *__restrict__ kz;
    const T *__restrict__ Dst;   // stored Gershgorin diagonal (DSTORED variants)
    T *__restrict__ F0;
    T *__restrict__ Fst;         // F of the previous lambda iteration (read) / this one (write)
    double *__restrict__ partials;
    int nx, ny, nz;
    int nbx, nby, nbz;           // tile grid
    T _dx, _dy, _dz, _dt;
    T dx, dy, dz, rc_tr_dt;
    T h1, h2, dtau;
    int bct[6];
    T bcv[6];
    unsigned flags;
};

// block id -> logical tile id, making the blocks of one XCD (id mod 8) contiguous
__device__ __forceinline__ unsigned xcd_tile(unsigned bid, unsigned nblocks) {
    const unsigned q = nblocks / 8u, r = nblocks % 8u;
    const unsigned x = bid % 8u, l = bid / 8u;
    return x * q + (x < r ? x : r) + l;
}

template <typename T>
__device__ __forceinline__ T ghost(int type, T val, T inner) {
    return type ? (T)2 * val - inner : inner;
}

// One fused pseudo-iteration.  Every neighbour address lies inside the E array
// (the ghost layer is part of its allocation), so all loads are unconditional
// and the boundary condition is applied by selecting the ghost value computed
// from the centre cell; offsets are 32-bit (E arrays stay below 2^32 words).
//
// iter_tile is one (x,y) tile of 64 x TY cells marched over KZ planes from
// plane bz*KZ; `slot` is where a check iteration writes its block partials.
// Xin/Xout and flags come separately so the batched kernel can take them from
// its launch while the rest of `a` lives in a per-problem table.
template <typename T, bool TRANSIENT, bool DSTORED, bool NTX, int TY, int KZ>
__device__ __forceinline__ void iter_tile(const IterArgs<T> &a, const T *__restrict__ Xin, T *__restrict__ Xout,
                                          unsigned flags, int bx, int by, int bz, double *__restrict__ partials,
                                          unsigned slot) {
    const int i = bx * 64 + (int)(threadIdx.x & 63u);
    const int j = by * TY + (int)(threadIdx.x >> 6);
    const int k0 = bz * KZ;
    const int k1 = min(k0 + KZ, a.nz);
    const uint32_t nx = a.nx, ny = a.ny;
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t sC2 = nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = nx * (ny + 1);

    double s_hi = 0.0, s_lo = 0.0, acc_max = 0.0, n_hi = 0.0, n_lo = 0.0, d_hi = 0.0, d_lo = 0.0;
    const bool check = flags & FL_CHECK, lam = flags & FL_LAMBDA, store = flags & FL_STORE_F;

    if ((uint32_t)i < nx && (uint32_t)j < ny) {
        const bool wlo = i == 0, whi = (uint32_t)i == nx - 1, slo = j == 0, shi = (uint32_t)j == ny - 1;
        uint32_t e = (uint32_t)(i + 1) + (uint32_t)(j + 1) * sE1 + (uint32_t)(k0 + 1) * sE2;
        uint32_t c = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sC2;
        uint32_t fx = (uint32_t)i + (uint32_t)j * sX1 + (uint32_t)k0 * sX2;
        uint32_t fy = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sY2;
        T xc = Xin[e];
        T xm = Xin[e - sE2];
        if (k0 == 0) xm = ghost(a.bct[4], a.bcv[4], xc);
        T kzB = a.kz[c];
#pragma unroll 2
        for (int k = k0; k < k1; ++k) {
            T xp = Xin[e + sE2];
            T xW = Xin[e - 1], xE = Xin[e + 1];
            T xS = Xin[e - sE1], xN = Xin[e + sE1];
            // streams read once per iteration: non-temporal, so the L2 keeps the
            // X planes whose neighbours the other waves of the tile still need
            const T kW = __builtin_nontemporal_load(a.kx + fx), kE = __builtin_nontemporal_load(a.kx + fx + 1);
            const T kS = __builtin_nontemporal_load(a.ky + fy), kN = __builtin_nontemporal_load(a.ky + fy + nx);
            const T kF = __builtin_nontemporal_load(a.kz + c + sC2);
            const T kB = kzB;
            const T r = __builtin_nontemporal_load(a.R + c);
            const T f0o = __builtin_nontemporal_load(a.F0 + c);
            if (k + 1 == a.nz) xp = ghost(a.bct[5], a.bcv[5], xc);
            if (wlo) xW = ghost(a.bct[0], a.bcv[0], xc);
            if (whi) xE = ghost(a.bct[1], a.bcv[1], xc);
            if (slo) xS = ghost(a.bct[2], a.bcv[2], xc);
            if (shi) xN = ghost(a.bct[3], a.bcv[3], xc);
            // ComputeFlux! (kernels_HT3D.jl:112-114)
            const T qW = -kW * (a._dx * (xc - xW)), qE = -kE * (a._dx * (xE - xc));
            const T qS = -kS * (a._dy * (xc - xS)), qN = -kN * (a._dy * (xN - xc));
            const T qB = -kB * (a._dz * (xc - xm)), qF = -kF * (a._dz * (xp - xc));
            // ResidualDiffusion(Transient)! (kernels_HT3D.jl:131, Alban_v6:87)
            T f;
            if (TRANSIENT)
                f = r + a._dt * xc + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
            else
                f = r + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
            // GershgorinPoisson! diagonal D (stored once per solve, or recomputed
            // from the face coefficients already in registers) and "F /= PC"
            T D;
            if (DSTORED)
                D = a.Dst[c];
            else
                D = (kB / a.dz + kF / a.dz) / a.dz + (kN / a.dy + kS / a.dy) / a.dy + (kE / a.dx + kW / a.dx) / a.dx +
                    a.rc_tr_dt;
            f = f / D;
            // DYRELUpdate! (kernels_HT3D_dimensional.jl:323-324)
            const T f0 = a.h1 * f0o + a.h2 * f;
            __builtin_nontemporal_store(f0, a.F0 + c);
            if (NTX) __builtin_nontemporal_store(xc - a.dtau * f0, Xout + e);
            else Xout[e] = xc - a.dtau * f0;
            if (check) {
                const double fd = fabs((double)(f * D));
                dd_add(s_hi, s_lo, fd);
                acc_max = (fd > acc_max || fd != fd) ? fd : acc_max;
                if (lam) {  // RayleighQuotientNumerator! / Denominator! terms
                    const double d = (double)(a.dtau * f0);
                    dd_add(n_hi, n_lo, (-d) * ((double)f - (double)a.Fst[c]));
                    dd_add(d_hi, d_lo, d * d);
                }
            }
            if (store) a.Fst[c] = f;
            xm = xc;
            xc = xp;
            kzB = kF;
            e += sE2;
            c += sC2;
            fx += sX2;
            fy += sY2;
        }
    }
    if (check) {
        const double h4[4] = {s_hi, acc_max, n_hi, d_hi}, l4[4] = {s_lo, 0.0, n_lo, d_lo};
        block_reduce4(h4, l4, 0x2u, partials, slot);
    }
}

template <typename T, bool TRANSIENT, bool DSTORED, bool NTX, int TY, int KZ>
__global__ __launch_bounds__(64 * TY) void dyrel_iter_kernel(IterArgs<T> a) {
    const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
    const unsigned t = xcd_tile(blockIdx.x, nblocks);
    const int bx = t % a.nbx, by = (t / a.nbx) % a.nby, bz = t / (a.nbx * a.nby);
    iter_tile<T, TRANSIENT, DSTORED, NTX, TY, KZ>(a, a.Xin, a.Xout, a.flags, bx, by, bz, a.partials, blockIdx.x);
}

// ------------------------------------------------------------------ host driver
// Tile configurations (rows of 64 cells per block, z cells per thread); the
// autotuner picks one per (grid, operator, precision, D mode).
struct TileCfg {
    int ty, kz;
};
static constexpr TileCfg kTiles[] = {{2, 8}, {2, 16}, {4, 4}, {4, 8}, {4, 16}, {4, 32}, {8, 4}, {8, 8},
                                      {8, 16}, {4, 2}, {8, 2}, {16, 2}, {16, 4}, {1, 8}};
static constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
static constexpr int kDefaultTile = 2;

template <typename T, bool TR, bool DS, bool NX, int TY, int KZ>
inline void launch_t(const IterArgs<T> &a, unsigned nb, hipStream_t st) {
    hipLaunchKernelGGL((dyrel_iter_kernel<T, TR, DS, NX, TY, KZ>), dim3(nb), dim3(64 * TY), 0, st, a);
}

template <typename T, bool TR, bool DS, bool NX>
inline void launch_tile(int tile, const IterArgs<T> &a, unsigned nb, hipStream_t st) {
    switch (tile) {
    case 0: launch_t<T, TR, DS, NX, 2, 8>(a, nb, st); break;
    case 1: launch_t<T, TR, DS, NX, 2, 16>(a, nb, st); break;
    case 2: launch_t<T, TR, DS, NX, 4, 4>(a, nb, st); break;
    case 3: launch_t<T, TR, DS, NX, 4, 8>(a, nb, st); break;
    case 4: launch_t<T, TR, DS, NX, 4, 16>(a, nb, st); break;
    case 5: launch_t<T, TR, DS, NX, 4, 32>(a, nb, st); break;
    case 6: launch_t<T, TR, DS, NX, 8, 4>(a, nb, st); break;
    case 7: launch_t<T, TR, DS, NX, 8, 8>(a, nb, st); break;
    case 8: launch_t<T, TR, DS, NX, 8, 16>(a, nb, st); break;
    case 9: launch_t<T, TR, DS, NX, 4, 2>(a, nb, st); break;
    case 10: launch_t<T, TR, DS, NX, 8, 2>(a, nb, st); break;
    case 11: launch_t<T, TR, DS, NX, 16, 2>(a, nb, st); break;
    case 12: launch_t<T, TR, DS, NX, 16, 4>(a, nb, st); break;
    default: launch_t<T, TR, DS, NX, 1, 8>(a, nb, st); break;
    }
}

// variant bits: 1 = read the stored diagonal instead of recomputing it,
//               2 = non-temporal store of the updated unknown
struct IterCfg {
    int tile = kDefaultTile;
    int variant = 0;
};

template <typename T>
inline void set_tiling(IterArgs<T> &a, int tile) {
    a.nbx = (int)div_up(a.nx, 64);
    a.nby = (int)div_up(a.ny, kTiles[tile].ty);
    a.nbz = (int)div_up(a.nz, kTiles[tile].kz);
}

template <typename T, bool TR>
inline void launch_variant(const IterArgs<T> &a, const IterCfg &cfg, unsigned nb, hipStream_t st) {
    switch (cfg.variant & 3) {
    case 0: launch_tile<T, TR, false, false>(cfg.tile, a, nb, st); break;
    case 1: launch_tile<T, TR, true, false>(cfg.tile, a, nb, st); break;
    case 2: launch_tile<T, TR, false, true>(cfg.tile, a, nb, st); break;
    default: launch_tile<T, TR, true, true>(cfg.tile, a, nb, st); break;
    }
}

template <typename T>
inline int launch_iter(const IterArgs<T> &a, bool transient, const IterCfg &cfg) {
    const unsigned nb = (unsigned)a.nbx * a.nby * a.nbz;
    hipStream_t st = ctx().stream;
    if (transient) launch_variant<T, true>(a, cfg, nb, st);
    else launch_variant<T, false>(a, cfg, nb, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "dyrel_iter_kernel");
    return HT3D_OK;
}

// ---------------------------------------------------------------- autotuner
struct TuneKey {
    int nx, ny, nz, transient, prec;
    bool operator==(const TuneKey &o) const {
        return nx == o.nx && ny == o.ny && nz == o.nz && transient == o.transient && prec == o.prec;
    }
};
struct TuneEntry {
    TuneKey key;
    IterCfg cfg;
    double us;
};
extern std::vector<TuneEntry> g_tune;
extern bool g_autotune;
extern int g_force_tile, g_force_variant;

inline const TuneEntry *tune_find(const TuneKey &k) {
    for (const auto &t : g_tune)
        if (t.key == k) return &t;
    return nullptr;
}

// Times every tile configuration (both D modes) on the solve's own read-only
// inputs, writing only scratch outputs (Xout, F0), and keeps the fastest.
template <typename T>
inline int autotune(IterArgs<T> a, bool transient, bool can_store_d, int prec, IterCfg *out) {
    const TuneKey key{a.nx, a.ny, a.nz, transient ? 1 : 0, prec};
    if (const TuneEntry *t = tune_find(key)) {
        *out = t->cfg;
        if (!can_store_d) out->variant &= ~1;
        return HT3D_OK;
    }
    Context &cx = ctx();
    a.flags = 0;
    IterCfg best;
    double best_us = 1e300;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int reps = 8;
    for (int var = 0; var < 4; ++var) {
        if ((var & 1) && !can_store_d) continue;
        for (int tile = 0; tile < kNumTiles; ++tile) {
            IterCfg cfg{tile, var};
            set_tiling(a, tile);
            int rc = launch_iter(a, transient, cfg);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(e0, cx.stream));
            for (int r = 0; r < reps; ++r)
                if ((rc = launch_iter(a, transient, cfg))) return rc;
            HIP_TRY(hipEventRecord(e1, cx.stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / reps;
            if (us < best_us) {
                best_us = us;
                best = cfg;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g_tune.push_back(TuneEntry{key, best, best_us});
    *out = best;
    return HT3D_OK;
}

IterCfg &last_cfg();

template <typename T>
inline int choose_cfg(IterArgs<T> &a, bool transient, bool can_store_d, int prec, IterCfg *cfg) {
    IterCfg c;
    if (g_force_tile >= 0) {
        c.tile = g_force_tile;
        c.variant = g_force_variant & (can_store_d ? 3 : 2);
    } else if (g_autotune) {
        int rc = autotune(a, transient, can_store_d, prec, &c);
        if (rc) return rc;
    }
    set_tiling(a, c.tile);
    *cfg = c;
    return HT3D_OK;
}


// ------------------------------------------------------------ defect correction
// FP64 residual of the dimensionless operators at one cell, with on-the-fly
// ghost values: r = R [+ x/dt] + div q(x) (ResidualDiffusion(Transient)!,
// kernels_HT3D.jl:131, HydroThermal3D_Alban_v6.jl:87); r is also stored in FP32
// as the right-hand side of the correction equation.
struct DcArgs {
    const double *__restrict__ X;   // E
    const double *__restrict__ R;   // c
    const double *__restrict__ kx;
    const double *__restrict__ ky;
    const double *__restrict__ kz;
    float *__restrict__ r32;        // c: FP32 right-hand side of the correction equation
    double *__restrict__ partials;
    int nx, ny, nz;
    double _dx, _dy, _dz, _dt;
    int bct[6];
    double bcv[6];
};

template <bool TR>
__device__ __forceinline__ double dc_residual_cell(const DcArgs &a, int i, int j, int k) {
    const uint32_t nx = a.nx, ny = a.ny;
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t e = (uint32_t)(i + 1) + (uint32_t)(j + 1) * sE1 + (uint32_t)(k + 1) * sE2;
    const uint32_t c = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k * nx * ny;
    const uint32_t fx = (uint32_t)i + (uint32_t)j * (nx + 1) + (uint32_t)k * (nx + 1) * ny;
    const uint32_t fy = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k * nx * (ny + 1);
    const double xc = a.X[e];
    double xW = a.X[e - 1], xE = a.X[e + 1], xS = a.X[e - sE1], xN = a.X[e + sE1], xB = a.X[e - sE2],
           xF = a.X[e + sE2];
    if (i == 0) xW = ghost(a.bct[0], a.bcv[0], xc);
    if ((uint32_t)i == nx - 1) xE = ghost(a.bct[1], a.bcv[1], xc);
    if (j == 0) xS = ghost(a.bct[2], a.bcv[2], xc);
    if ((uint32_t)j == ny - 1) xN = ghost(a.bct[3], a.bcv[3], xc);
    if (k == 0) xB = ghost(a.bct[4], a.bcv[4], xc);
    if (k == a.nz - 1) xF = ghost(a.bct[5], a.bcv[5], xc);
    const double kW = a.kx[fx], kE = a.kx[fx + 1], kS = a.ky[fy], kN = a.ky[fy + nx], kB = a.kz[c],
                 kF = a.kz[c + nx * ny];
    const double qW = -kW * (a._dx * (xc - xW)), qE = -kE * (a._dx * (xE - xc));
    const double qS = -kS * (a._dy * (xc - xS)), qN = -kN * (a._dy * (xN - xc));
    const double qB = -kB * (a._dz * (xc - xB)), qF = -kF * (a._dz * (xF - xc));
    double f;
    if (TR) f = a.R[c] + a._dt * xc + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
    else f = a.R[c] + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
    a.r32[c] = (float)f;
    return f;
}

}  // namespace ht3d
