// ht3d_iter.hpp -- the fused DYREL pseudo-iteration kernels (one iteration per
// launch, and two per launch for runs of plain iterations), their tile tables
// and the autotuner, shared by the FP64 solve (ht3d_dyrel.hip), the FP32 inner
// sweeps of the defect-correction solve (ht3d_dc.hip) and the batched solves
// (ht3d_batch.hip).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ht3d_internal.hpp"

namespace ht3d {

// FL_REVERSE: sweep each XCD's slab of tiles backwards.  The solves alternate
// it from launch to launch, so a launch starts on the planes the previous one
// touched last, which are still in the MALL (256 MB, about half of an
// iteration's working set at 256x128x256); results do not depend on it.
enum : unsigned { FL_CHECK = 1u, FL_LAMBDA = 2u, FL_STORE_F = 4u, FL_REVERSE = 8u };

extern int g_alternate;   // alternate FL_REVERSE between launches (default 1)

template <typename T>
struct IterArgs {
    const T *__restrict__ Xin;
    T *__restrict__ Xout;
    const T *__restrict__ R;
    const T *__restrict__ kx;
    const T *__restrict__ ky;
    const T *__restrict__ kz;
    const T *__restrict__ kv;    // vertex field whose face averages kx, ky, kz are (variant bit 8), or null
    const T *__restrict__ Dst;   // stored Gershgorin diagonal (DSTORED variants)
    T *__restrict__ F0;
    T *__restrict__ Fst;         // F of the previous lambda iteration (read) / this one (write)
    double *__restrict__ partials;
    int nx, ny, nz;
    int nbx, nby, nbz;           // tile grid
    int nband;                   // batch: bands of rows per plane chunk of the pair launches
    T _dx, _dy, _dz, _dt;
    T dx, dy, dz, rc_tr_dt;
    T idx2, idy2, idz2;          // 1/dx^2, 1/dy^2, 1/dz^2 (the FP32 diagonal)
    T h1, h2, dtau;
    const T *hv;                 // when not null: h1, h2 read from the device (hv[0], hv[1])
    int bct[6];
    T bcv[6];
    unsigned flags;
};

// The first cell (i, j) of this lane's CPL x-cells: in tile (bx, by) of
// 64*CPL x TY cells (one row of the tile per wave) ...
template <int TY, int CPL>
__device__ __forceinline__ void tile_cell(int bx, int by, int &i, int &j) {
    i = bx * (64 * CPL) + CPL * (int)(threadIdx.x & 63u);
    j = by * TY + (int)(threadIdx.x >> 6);
}
// ... or in patch p of the (x, y) plane seen as a row-major sequence of
// groups of CPL x-cells, gx = nx / CPL groups per row: a block covers the
// 64*TY consecutive groups from p*64*TY, so a wave spans several rows of a
// narrow grid and keeps all its lanes busy (j >= ny past the last row)
template <int TY, int CPL>
__device__ __forceinline__ void patch_cell(int p, int gx, int &i, int &j) {
    const unsigned g = (unsigned)p * (64u * TY) + threadIdx.x;
    j = (int)(g / (unsigned)gx);
    i = (int)(g - (unsigned)j * (unsigned)gx) * CPL;
}

// block id -> logical tile id, making the blocks of one XCD (id mod 8) contiguous;
// with `rev` each XCD walks its slab from the far end
__device__ __forceinline__ unsigned xcd_tile(unsigned bid, unsigned nblocks, bool rev = false) {
    const unsigned q = nblocks / 8u, r = nblocks % 8u;
    const unsigned x = bid % 8u, l = bid / 8u;
    const unsigned len = q + (x < r ? 1u : 0u);
    return x * q + (x < r ? x : r) + (rev ? len - 1u - l : l);
}

template <typename T>
__device__ __forceinline__ T ghost(int type, T val, T inner) {
    return type ? (T)2 * val - inner : inner;
}

// Residual of one cell from its six neighbours and face coefficients:
// ComputeFlux! (kernels_HT3D.jl:112-114) then ResidualDiffusion(Transient)!
// (kernels_HT3D.jl:131, HydroThermal3D_Alban_v6.jl:87), in the reference's order.
template <typename T, bool TRANSIENT>
__device__ __forceinline__ T cell_residual(const IterArgs<T> &a, T xc, T xW, T xE, T xS, T xN, T xB, T xF, T kW,
                                           T kE, T kS, T kN, T kB, T kF, T r) {
    if constexpr (sizeof(T) == 4) {
        // FP32 sweeps: _dx (qE - qW) = _dx^2 (kW (xc - xW) - kE (xE - xc)), fused
        const T tx = __builtin_fmaf(kW, xc - xW, -(kE * (xE - xc)));
        const T ty = __builtin_fmaf(kS, xc - xS, -(kN * (xN - xc)));
        const T tz = __builtin_fmaf(kB, xc - xB, -(kF * (xF - xc)));
        const T s = TRANSIENT ? __builtin_fmaf(a._dt, xc, r) : r;
        return __builtin_fmaf(a.idx2, tx, __builtin_fmaf(a.idy2, ty, __builtin_fmaf(a.idz2, tz, s)));
    }
    const T qW = -kW * (a._dx * (xc - xW)), qE = -kE * (a._dx * (xE - xc));
    const T qS = -kS * (a._dy * (xc - xS)), qN = -kN * (a._dy * (xN - xc));
    const T qB = -kB * (a._dz * (xc - xB)), qF = -kF * (a._dz * (xF - xc));
    if (TRANSIENT) return r + a._dt * xc + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
    return r + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
}

// The preconditioned residual F/D, the DYREL momentum h1 F0 + h2 F
// (DYRELUpdate!, kernels_HT3D_dimensional.jl:323-324) and the update
// x - dtau F0.  FP64 evaluates them as written, one rounding per operation and
// an IEEE division: bit parity with the oracle.  The FP32 sweeps of the defect
// correction have no reference counterpart and, at large grids, are limited by
// their instruction count rather than by HBM: they multiply by the hardware
// reciprocal (1 ulp) and fuse the multiply-adds.  Every FP32 kernel (single,
// pair, vector and batched) goes through these same helpers, so they stay bit
// for bit identical to each other.
template <typename T>
__device__ __forceinline__ T precond(T f, T D) {
    if constexpr (sizeof(T) == 4) return f * __builtin_amdgcn_rcpf(D);
    else return f / D;
}
template <typename T>
__device__ __forceinline__ T momentum(T h1, T f0o, T h2, T f) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(h1, f0o, h2 * f);
    else return h1 * f0o + h2 * f;
}
template <typename T>
__device__ __forceinline__ T step_x(T xc, T dtau, T f0) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(-dtau, f0, xc);
    else return xc - dtau * f0;
}

// GershgorinPoisson! diagonal D of one cell (kernels_HT3D_dimensional.jl:207-221).
// FP64 keeps the reference's divisions (bit parity with the oracle); the FP32
// sweeps of the defect correction, which have no reference counterpart, use
// products with 1/h^2, which leaves them bound by memory instead of by nine
// FP32 divisions per cell.
template <typename T>
__device__ __forceinline__ T cell_diag(const IterArgs<T> &a, T kW, T kE, T kS, T kN, T kB, T kF) {
    if constexpr (sizeof(T) == 4)
        return __builtin_fmaf(kB + kF, a.idz2, __builtin_fmaf(kN + kS, a.idy2, __builtin_fmaf(kE + kW, a.idx2, a.rc_tr_dt)));
    return (kB / a.dz + kF / a.dz) / a.dz + (kN / a.dy + kS / a.dy) / a.dy + (kE / a.dx + kW / a.dx) / a.dx +
           a.rc_tr_dt;
}

// the 1/h^2 factors of the FP32 diagonal, rounded once from FP64
template <typename T>
inline void set_inv_h2(IterArgs<T> &a, double dx, double dy, double dz) {
    a.idx2 = (T)(1.0 / (dx * dx));
    a.idy2 = (T)(1.0 / (dy * dy));
    a.idz2 = (T)(1.0 / (dz * dz));
}

// face coefficients: cached (KT) or streamed past the caches
template <bool KT, typename T>
__device__ __forceinline__ T ld_k(const T *p) {
    if (KT) return *p;
    return __builtin_nontemporal_load(p);
}

// One fused pseudo-iteration.  Every neighbour address lies inside the E array
// (the ghost layer is part of its allocation), so all loads are unconditional
// and the boundary condition is applied by selecting the ghost value computed
// from the centre cell; offsets are 32-bit (E arrays stay below 2^32 words).
//
// iter_tile marches the lane's cell (i, j) over the KZ planes from plane
// bz*KZ (tile_cell: a tile of 64 x TY cells; patch_cell: a patch of 64*TY
// consecutive cells of the row-major (x, y) plane); `slot` is where a check
// iteration writes its block partials.
// Xin/Xout and flags come separately so the batched kernel can take them from
// its launch while the rest of `a` lives in a per-problem table.
template <typename T, bool TRANSIENT, bool DSTORED, bool NTX, bool KT, int TY, int KZ>
__device__ __forceinline__ void iter_tile(const IterArgs<T> &a, const T *__restrict__ Xin, T *__restrict__ Xout,
                                          unsigned flags, int i, int j, int bz, double *__restrict__ partials,
                                          unsigned slot) {
    const int k0 = bz * KZ;
    const int k1 = min(k0 + KZ, a.nz);
    const uint32_t nx = a.nx, ny = a.ny;
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t sC2 = nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = nx * (ny + 1);

    double s_hi = 0.0, s_lo = 0.0, acc_max = 0.0, n_hi = 0.0, n_lo = 0.0, d_hi = 0.0, d_lo = 0.0;
    const bool check = flags & FL_CHECK, lam = flags & FL_LAMBDA, store = flags & FL_STORE_F;
    const T h1 = a.hv ? a.hv[0] : a.h1, h2 = a.hv ? a.hv[1] : a.h2;

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
        auto plane = [&](int k) {
            T xp = Xin[e + sE2];
            T xW = Xin[e - 1], xE = Xin[e + 1];
            T xS = Xin[e - sE1], xN = Xin[e + sE1];
            // streams read once per iteration: non-temporal, so the L2 keeps the
            // X planes whose neighbours the other waves of the tile still need
            // (the KT variants let the face coefficients be cached)
            const T kW = ld_k<KT>(a.kx + fx), kE = ld_k<KT>(a.kx + fx + 1);
            const T kS = ld_k<KT>(a.ky + fy), kN = ld_k<KT>(a.ky + fy + nx);
            const T kF = ld_k<KT>(a.kz + c + sC2);
            const T kB = kzB;
            const T r = __builtin_nontemporal_load(a.R + c);
            const T f0o = __builtin_nontemporal_load(a.F0 + c);
            if (k + 1 == a.nz) xp = ghost(a.bct[5], a.bcv[5], xc);
            if (wlo) xW = ghost(a.bct[0], a.bcv[0], xc);
            if (whi) xE = ghost(a.bct[1], a.bcv[1], xc);
            if (slo) xS = ghost(a.bct[2], a.bcv[2], xc);
            if (shi) xN = ghost(a.bct[3], a.bcv[3], xc);
            T f = cell_residual<T, TRANSIENT>(a, xc, xW, xE, xS, xN, xm, xp, kW, kE, kS, kN, kB, kF, r);
            // diagonal D (stored once per solve, or recomputed from the face
            // coefficients already in registers) and "F /= PC"
            const T D = DSTORED ? a.Dst[c] : cell_diag(a, kW, kE, kS, kN, kB, kF);
            f = precond<T>(f, D);
            // DYRELUpdate! (kernels_HT3D_dimensional.jl:323-324)
            const T f0 = momentum<T>(h1, f0o, h2, f);
            __builtin_nontemporal_store(f0, a.F0 + c);
            if (NTX) __builtin_nontemporal_store(step_x<T>(xc, a.dtau, f0), Xout + e);
            else Xout[e] = step_x<T>(xc, a.dtau, f0);
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
        };
        // a full chunk of KZ planes is unrolled completely, so the loads of later
        // planes can be issued ahead of earlier planes' arithmetic (about 1 %
        // faster); the last, partial chunk of the z range keeps a rolled loop
        if (k1 - k0 == KZ) {
#pragma unroll
            for (int q = 0; q < KZ; ++q) plane(k0 + q);
        } else {
            for (int k = k0; k < k1; ++k) plane(k);
        }
    }
    if (check) {
        const double h4[4] = {s_hi, acc_max, n_hi, d_hi}, l4[4] = {s_lo, 0.0, n_lo, d_lo};
        block_reduce4(h4, l4, 0x2u, partials, slot);
    }
}

// CPL consecutive x-cells as one vector (element alignment only: the E arrays'
// rows start at odd offsets, and gfx950 loads and stores vectors at element
// alignment).
template <typename T, int CPL>
struct CellVec {
    typedef T type __attribute__((ext_vector_type(CPL), aligned(sizeof(T))));
};

template <typename T, int CPL>
__device__ __forceinline__ typename CellVec<T, CPL>::type ldv(const T *p) {
    return *reinterpret_cast<const typename CellVec<T, CPL>::type *>(p);
}
template <typename T, int CPL>
__device__ __forceinline__ typename CellVec<T, CPL>::type ldv_nt(const T *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const typename CellVec<T, CPL>::type *>(p));
}
template <bool KT, typename T, int CPL>
__device__ __forceinline__ typename CellVec<T, CPL>::type ldv_k(const T *p) {
    if (KT) return ldv<T, CPL>(p);
    return ldv_nt<T, CPL>(p);
}
template <typename T, int CPL>
__device__ __forceinline__ void stv(typename CellVec<T, CPL>::type v, T *p) {
    *reinterpret_cast<typename CellVec<T, CPL>::type *>(p) = v;
}
template <typename T, int CPL>
__device__ __forceinline__ void stv_nt(typename CellVec<T, CPL>::type v, T *p) {
    __builtin_nontemporal_store(v, reinterpret_cast<typename CellVec<T, CPL>::type *>(p));
}

// iter_tile with CPL (2 or 4) x-neighbouring cells per lane, from (i, j) (a
// tile_cell block covers 64*CPL x TY cells).  The centre, lateral and z rows of X, the face
// coefficients, R, F0 (and D, F) arrive as CPL-element vector loads, which
// divides the load instructions per cell by about CPL; the west neighbour of
// the first cell and the east neighbour of the last are single loads, and a
// face between two cells of a lane is loaded once.  Each cell is computed with
// iter_tile's operations in the same order, so X and F0 are bit-identical to
// iter_tile; the check sums are the same double-double sums in another order.
// Needs nx % CPL == 0 (a lane's cells never cross the end of a row).
template <typename T, bool TRANSIENT, bool DSTORED, bool NTX, bool KT, int TY, int KZ, int CPL>
__device__ __forceinline__ void iter_tile_v(const IterArgs<T> &a, const T *__restrict__ Xin, T *__restrict__ Xout,
                                            unsigned flags, int i, int j, int bz, double *__restrict__ partials,
                                            unsigned slot) {
    typedef typename CellVec<T, CPL>::type V;
    const int k0 = bz * KZ;
    const int k1 = min(k0 + KZ, a.nz);
    const uint32_t nx = a.nx, ny = a.ny;
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t sC2 = nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = nx * (ny + 1);

    double s_hi = 0.0, s_lo = 0.0, acc_max = 0.0, n_hi = 0.0, n_lo = 0.0, d_hi = 0.0, d_lo = 0.0;
    const bool check = flags & FL_CHECK, lam = flags & FL_LAMBDA, store = flags & FL_STORE_F;
    const T h1 = a.hv ? a.hv[0] : a.h1, h2 = a.hv ? a.hv[1] : a.h2;

    if ((uint32_t)i < nx && (uint32_t)j < ny) {
        const bool wlo = i == 0, whi = (uint32_t)i + CPL == nx, slo = j == 0, shi = (uint32_t)j == ny - 1;
        uint32_t e = (uint32_t)(i + 1) + (uint32_t)(j + 1) * sE1 + (uint32_t)(k0 + 1) * sE2;
        uint32_t c = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sC2;
        uint32_t fx = (uint32_t)i + (uint32_t)j * sX1 + (uint32_t)k0 * sX2;
        uint32_t fy = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sY2;
        V xc = ldv<T, CPL>(Xin + e);
        V xm = ldv<T, CPL>(Xin + e - sE2);
        if (k0 == 0) {
#pragma unroll
            for (int q = 0; q < CPL; ++q) xm[q] = ghost(a.bct[4], a.bcv[4], (T)xc[q]);
        }
        V kzB = ldv_k<KT, T, CPL>(a.kz + c);
        auto plane = [&](int k) {
            V xp = ldv<T, CPL>(Xin + e + sE2);
            T xW = Xin[e - 1], xE = Xin[e + CPL];
            V xS = ldv<T, CPL>(Xin + e - sE1), xN = ldv<T, CPL>(Xin + e + sE1);
            // x-faces: fx (west of the first cell), then fx+1 .. fx+CPL
            const T kW = ld_k<KT>(a.kx + fx);
            const V kE = ldv_k<KT, T, CPL>(a.kx + fx + 1);
            const V kS = ldv_k<KT, T, CPL>(a.ky + fy), kN = ldv_k<KT, T, CPL>(a.ky + fy + nx);
            const V kF = ldv_k<KT, T, CPL>(a.kz + c + sC2);
            const V kB = kzB;
            const V r = ldv_nt<T, CPL>(a.R + c);
            const V f0o = ldv_nt<T, CPL>(a.F0 + c);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                if (k + 1 == a.nz) xp[q] = ghost(a.bct[5], a.bcv[5], (T)xc[q]);
                if (slo) xS[q] = ghost(a.bct[2], a.bcv[2], (T)xc[q]);
                if (shi) xN[q] = ghost(a.bct[3], a.bcv[3], (T)xc[q]);
            }
            if (wlo) xW = ghost(a.bct[0], a.bcv[0], (T)xc[0]);
            if (whi) xE = ghost(a.bct[1], a.bcv[1], (T)xc[CPL - 1]);
            V f, D, g, xo;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T w = q == 0 ? xW : (T)xc[q - 1], ea = q == CPL - 1 ? xE : (T)xc[q + 1];
                const T kw = q == 0 ? kW : (T)kE[q - 1];
                f[q] = cell_residual<T, TRANSIENT>(a, xc[q], w, ea, xS[q], xN[q], xm[q], xp[q], kw, kE[q], kS[q], kN[q],
                                                   kB[q], kF[q], r[q]);
                if (!DSTORED) D[q] = cell_diag(a, kw, (T)kE[q], (T)kS[q], (T)kN[q], (T)kB[q], (T)kF[q]);
            }
            if (DSTORED) D = ldv<T, CPL>(a.Dst + c);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                f[q] = precond<T>(f[q], D[q]);
                g[q] = momentum<T>(h1, f0o[q], h2, f[q]);   // DYRELUpdate! (kernels_HT3D_dimensional.jl:323-324)
                xo[q] = step_x<T>(xc[q], a.dtau, g[q]);
            }
            stv_nt<T, CPL>(g, a.F0 + c);
            if (NTX) stv_nt<T, CPL>(xo, Xout + e);
            else stv<T, CPL>(xo, Xout + e);
            if (check) {
                V fst = xo;   // only read on lambda iterations
                if (lam) fst = ldv<T, CPL>(a.Fst + c);
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const double fd = fabs((double)(f[q] * D[q]));
                    dd_add(s_hi, s_lo, fd);
                    acc_max = (fd > acc_max || fd != fd) ? fd : acc_max;
                    if (lam) {   // RayleighQuotientNumerator! / Denominator! terms
                        const double d = (double)(a.dtau * g[q]);
                        dd_add(n_hi, n_lo, (-d) * ((double)f[q] - (double)fst[q]));
                        dd_add(d_hi, d_lo, d * d);
                    }
                }
            }
            if (store) stv<T, CPL>(f, a.Fst + c);
            xm = xc;
            xc = xp;
            kzB = kF;
            e += sE2;
            c += sC2;
            fx += sX2;
            fy += sY2;
        };
        if (k1 - k0 == KZ) {
#pragma unroll
            for (int q = 0; q < KZ; ++q) plane(k0 + q);
        } else {
            for (int k = k0; k < k1; ++k) plane(k);
        }
    }
    if (check) {
        const double h4[4] = {s_hi, acc_max, n_hi, d_hi}, l4[4] = {s_lo, 0.0, n_lo, d_lo};
        block_reduce4(h4, l4, 0x2u, partials, slot);
    }
}

// Face coefficients formed from a vertex field.  In the closed-box model the
// three face arrays are the face averages of one vertex field kv
// ((nx+1)(ny+1)(nz+1) values, x fastest): kx = @av_yza(kv), ky = @av_xza(kv),
// kz = @av_xya(kv) (InitThermal! / InitDarcy!, kernels_HT3D.jl:82-107).  The
// iteration then needs one read-only word per cell for its coefficients
// instead of three: 6 words per cell in all instead of 8.  av4 adds the four
// vertex values in the order the averaging kernels add them (left to right,
// then times 0.25), so every face value, and every result, is the same bit for
// bit as with the face arrays.
template <typename T>
__device__ __forceinline__ T av4(T p, T q, T r, T s) {
    return (p + q + r + s) * (T)0.25;
}

// iter_tile_v with the face coefficients formed from the vertex field a.kv.
// The lane keeps the vertex values of its plane k, rows j and j+1, x
// positions i .. i+CPL (2(CPL+1) values), and loads those of plane k+1 as one
// CPL-vector and one scalar per row; with a0/a1 (plane k) and b0/b1 (plane
// k+1) on rows j/j+1, the faces of cell i+q are
//   west  kx(i+q)   = av4(a0[q], a1[q], b0[q], b1[q])      (@av_yza)
//   south ky(j)     = av4(a0[q], a0[q+1], b0[q], b0[q+1])  (@av_xza)
//   north ky(j+1)   = av4(a1[q], a1[q+1], b1[q], b1[q+1])
//   back  kz(k)     = av4(a0[q], a0[q+1], a1[q], a1[q+1])  (@av_xya; the front
//   front kz(k+1)   = av4(b0[q], b0[q+1], b1[q], b1[q+1])   face of plane k-1)
// Every vertex index stays inside the vertex array (i + CPL <= nx, j + 1 <= ny,
// k + 1 <= nz).  DSTORED reads the Gershgorin diagonal stored once per solve
// (one word more, but none of the nine FP64 divisions that recomputing it
// costs).  ROLL unrolls a chunk's z march by two planes instead of all KZ:
// fully unrolled, the compiler hoists later planes' loads and the FP32
// four-cell kernel needs 137 VGPRs (3 waves per SIMD), by two 127 (4 waves).
// Needs nx % CPL == 0, CPL >= 2.
template <typename T, bool TRANSIENT, bool DSTORED, bool NTX, bool ROLL, int TY, int KZ, int CPL>
__device__ __forceinline__ void iter_tile_kv(const IterArgs<T> &a, const T *__restrict__ Xin, T *__restrict__ Xout,
                                             unsigned flags, int i, int j, int bz, double *__restrict__ partials,
                                             unsigned slot) {
    typedef typename CellVec<T, CPL>::type V;
    const int k0 = bz * KZ;
    const int k1 = min(k0 + KZ, a.nz);
    const uint32_t nx = a.nx, ny = a.ny;
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t sC2 = nx * ny, sV1 = nx + 1, sV2 = sV1 * (ny + 1);

    double s_hi = 0.0, s_lo = 0.0, acc_max = 0.0, n_hi = 0.0, n_lo = 0.0, d_hi = 0.0, d_lo = 0.0;
    const bool check = flags & FL_CHECK, lam = flags & FL_LAMBDA, store = flags & FL_STORE_F;
    const T h1 = a.hv ? a.hv[0] : a.h1, h2 = a.hv ? a.hv[1] : a.h2;

    if ((uint32_t)i < nx && (uint32_t)j < ny) {
        const bool wlo = i == 0, whi = (uint32_t)i + CPL == nx, slo = j == 0, shi = (uint32_t)j == ny - 1;
        uint32_t e = (uint32_t)(i + 1) + (uint32_t)(j + 1) * sE1 + (uint32_t)(k0 + 1) * sE2;
        uint32_t c = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sC2;
        uint32_t v = (uint32_t)i + (uint32_t)j * sV1 + (uint32_t)k0 * sV2;
        const T *__restrict__ kv = a.kv;
        T a0[CPL + 1], a1[CPL + 1];
        auto vrow = [&](uint32_t o, T *out) {   // CPL + 1 vertex values from kv[o]
            const V w = ldv<T, CPL>(kv + o);
#pragma unroll
            for (int q = 0; q < CPL; ++q) out[q] = w[q];
            out[CPL] = kv[o + CPL];
        };
        vrow(v, a0);
        vrow(v + sV1, a1);
        V xc = ldv<T, CPL>(Xin + e);
        V xm = ldv<T, CPL>(Xin + e - sE2);
        if (k0 == 0) {
#pragma unroll
            for (int q = 0; q < CPL; ++q) xm[q] = ghost(a.bct[4], a.bcv[4], (T)xc[q]);
        }
        T kzB[CPL];
#pragma unroll
        for (int q = 0; q < CPL; ++q) kzB[q] = av4(a0[q], a0[q + 1], a1[q], a1[q + 1]);
        auto plane = [&](int k) {
            V xp = ldv<T, CPL>(Xin + e + sE2);
            T xW = Xin[e - 1], xE = Xin[e + CPL];
            V xS = ldv<T, CPL>(Xin + e - sE1), xN = ldv<T, CPL>(Xin + e + sE1);
            T b0[CPL + 1], b1[CPL + 1];
            vrow(v + sV2, b0);
            vrow(v + sV2 + sV1, b1);
            const V r = ldv_nt<T, CPL>(a.R + c);
            const V f0o = ldv_nt<T, CPL>(a.F0 + c);
            T kX[CPL + 1], kS[CPL], kN[CPL], kF[CPL];
#pragma unroll
            for (int q = 0; q <= CPL; ++q) kX[q] = av4(a0[q], a1[q], b0[q], b1[q]);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                kS[q] = av4(a0[q], a0[q + 1], b0[q], b0[q + 1]);
                kN[q] = av4(a1[q], a1[q + 1], b1[q], b1[q + 1]);
                kF[q] = av4(b0[q], b0[q + 1], b1[q], b1[q + 1]);
            }
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                if (k + 1 == a.nz) xp[q] = ghost(a.bct[5], a.bcv[5], (T)xc[q]);
                if (slo) xS[q] = ghost(a.bct[2], a.bcv[2], (T)xc[q]);
                if (shi) xN[q] = ghost(a.bct[3], a.bcv[3], (T)xc[q]);
            }
            if (wlo) xW = ghost(a.bct[0], a.bcv[0], (T)xc[0]);
            if (whi) xE = ghost(a.bct[1], a.bcv[1], (T)xc[CPL - 1]);
            V f, D, g, xo;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T w = q == 0 ? xW : (T)xc[q - 1], ea = q == CPL - 1 ? xE : (T)xc[q + 1];
                f[q] = cell_residual<T, TRANSIENT>(a, xc[q], w, ea, xS[q], xN[q], xm[q], xp[q], kX[q], kX[q + 1], kS[q],
                                                   kN[q], kzB[q], kF[q], r[q]);
                if (!DSTORED) D[q] = cell_diag(a, kX[q], kX[q + 1], kS[q], kN[q], kzB[q], kF[q]);
            }
            if (DSTORED) D = ldv_nt<T, CPL>(a.Dst + c);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                f[q] = precond<T>(f[q], D[q]);
                g[q] = momentum<T>(h1, f0o[q], h2, f[q]);   // DYRELUpdate! (kernels_HT3D_dimensional.jl:323-324)
                xo[q] = step_x<T>(xc[q], a.dtau, g[q]);
            }
            stv_nt<T, CPL>(g, a.F0 + c);
            if (NTX) stv_nt<T, CPL>(xo, Xout + e);
            else stv<T, CPL>(xo, Xout + e);
            if (check) {
                V fst = xo;   // only read on lambda iterations
                if (lam) fst = ldv<T, CPL>(a.Fst + c);
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const double fd = fabs((double)(f[q] * D[q]));
                    dd_add(s_hi, s_lo, fd);
                    acc_max = (fd > acc_max || fd != fd) ? fd : acc_max;
                    if (lam) {   // RayleighQuotientNumerator! / Denominator! terms
                        const double d = (double)(a.dtau * g[q]);
                        dd_add(n_hi, n_lo, (-d) * ((double)f[q] - (double)fst[q]));
                        dd_add(d_hi, d_lo, d * d);
                    }
                }
            }
            if (store) stv<T, CPL>(f, a.Fst + c);
            xm = xc;
            xc = xp;
#pragma unroll
            for (int q = 0; q <= CPL; ++q) {
                a0[q] = b0[q];
                a1[q] = b1[q];
            }
#pragma unroll
            for (int q = 0; q < CPL; ++q) kzB[q] = kF[q];
            e += sE2;
            c += sC2;
            v += sV2;
        };
        if (!ROLL && k1 - k0 == KZ) {
#pragma unroll
            for (int q = 0; q < KZ; ++q) plane(k0 + q);
        } else if (k1 - k0 == KZ) {
#pragma unroll 2
            for (int q = 0; q < KZ; ++q) plane(k0 + q);
        } else {
            for (int k = k0; k < k1; ++k) plane(k);
        }
    }
    if (check) {
        const double h4[4] = {s_hi, acc_max, n_hi, d_hi}, l4[4] = {s_lo, 0.0, n_lo, d_lo};
        block_reduce4(h4, l4, 0x2u, partials, slot);
    }
}

template <typename T, bool TRANSIENT, bool DSTORED, bool NTX, bool ROLL, int TY, int KZ, int CPL>
__global__ __launch_bounds__(64 * TY) void dyrel_iter_kv_kernel(IterArgs<T> a) {
    const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
    const unsigned t = xcd_tile(blockIdx.x, nblocks, a.flags & FL_REVERSE);
    const int bx = t % a.nbx, by = (t / a.nbx) % a.nby, bz = t / (a.nbx * a.nby);
    int i, j;
    tile_cell<TY, CPL>(bx, by, i, j);
    iter_tile_kv<T, TRANSIENT, DSTORED, NTX, ROLL, TY, KZ, CPL>(a, a.Xin, a.Xout, a.flags, i, j, bz, a.partials,
                                                                blockIdx.x);
}

// CPL: cells per lane along x (1: iter_tile; 2 or 4: iter_tile_v, nx % CPL == 0)
template <typename T, bool TRANSIENT, bool DSTORED, bool NTX, bool KT, int TY, int KZ, int CPL>
__global__ __launch_bounds__(64 * TY) void dyrel_iter_kernel(IterArgs<T> a) {
    const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
    const unsigned t = xcd_tile(blockIdx.x, nblocks, a.flags & FL_REVERSE);
    const int bx = t % a.nbx, by = (t / a.nbx) % a.nby, bz = t / (a.nbx * a.nby);
    int i, j;
    tile_cell<TY, CPL>(bx, by, i, j);
    if constexpr (CPL > 1)
        iter_tile_v<T, TRANSIENT, DSTORED, NTX, KT, TY, KZ, CPL>(a, a.Xin, a.Xout, a.flags, i, j, bz, a.partials,
                                                                 blockIdx.x);
    else
        iter_tile<T, TRANSIENT, DSTORED, NTX, KT, TY, KZ>(a, a.Xin, a.Xout, a.flags, i, j, bz, a.partials, blockIdx.x);
}

// Two pseudo-iterations per launch, for runs of plain iterations (no check, no
// stored F, no lambda update between them).  Tiles overlap: a block of 64 x TY
// threads computes iteration n+1 on its cells [bx*62-1, +64) x [by*(TY-2)-1, +TY)
// and iteration n+2 on the inner 62 x (TY-2) of them.  The intermediate iterate
// reaches its x/y neighbours through a triple-buffered LDS plane and its z
// neighbours through a register queue, iteration n+1 running one plane ahead of
// n+2 along the z march over the planes [bz*KZ, +KZ) (iteration n+1 also on the
// plane before and after).  Iteration n+2 reuses the face coefficients, R and D
// its cell loaded for n+1.  F0 is read from F0in and written to a.F0 (an
// overlapping tile must still see the old F0).  Every operation is the single
// iteration's in the same order: two single launches and one pair launch give
// bit-identical X and F0.
template <typename T, bool TRANSIENT, bool DSTORED, bool KT, int TY, int KZ>
__global__ __launch_bounds__(64 * TY) void dyrel_pair_kernel(IterArgs<T> a, const T *__restrict__ F0in) {
    __shared__ T L[3][TY][64];
    const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
    const unsigned t = xcd_tile(blockIdx.x, nblocks, a.flags & FL_REVERSE);
    const int bx = t % a.nbx, by = (t / a.nbx) % a.nby, bz = t / (a.nbx * a.nby);
    const int lx = (int)(threadIdx.x & 63u), ly = (int)(threadIdx.x >> 6);
    const int i = bx * 62 - 1 + lx, j = by * (TY - 2) - 1 + ly;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = bz * KZ, k1 = min(k0 + KZ, nz);
    const bool in_dom = i >= 0 && i < nx && j >= 0 && j < ny;
    const bool inner = in_dom && lx > 0 && lx < 63 && ly > 0 && ly < TY - 1;   // computes iteration n+2
    const int ic = min(max(i, 0), nx - 1), jc = min(max(j, 0), ny - 1);      // every load in bounds
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t sC2 = (uint32_t)nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = (uint32_t)nx * (ny + 1);
    const bool wlo = ic == 0, whi = ic == nx - 1, slo = jc == 0, shi = jc == ny - 1;
    const int ps = max(k0 - 1, 0), pe = min(k1, nz - 1);   // planes of iteration n+1
    const T h1 = a.hv ? a.hv[0] : a.h1, h2 = a.hv ? a.hv[1] : a.h2;
    uint32_t e = (uint32_t)(ic + 1) + (uint32_t)(jc + 1) * sE1 + (uint32_t)(ps + 1) * sE2;
    uint32_t c = (uint32_t)ic + (uint32_t)jc * nx + (uint32_t)ps * sC2;
    uint32_t fx = (uint32_t)ic + (uint32_t)jc * sX1 + (uint32_t)ps * sX2;
    uint32_t fy = (uint32_t)ic + (uint32_t)jc * nx + (uint32_t)ps * sY2;
    T x0c = a.Xin[e];
    T x0m = a.Xin[e - sE2];
    if (ps == 0) x0m = ghost(a.bct[4], a.bcv[4], x0c);
    T kzB = a.kz[c];
    // plane p-1 as iteration n+1 left it: the iterate, its neighbours in z, and
    // the coefficients, R, D and F0 the n+2 update needs
    T x1m = 0, x1c = 0, pkW = 0, pkE = 0, pkS = 0, pkN = 0, pkB = 0, pkF = 0, pr = 0, pD = 1, pf0 = 0;
    uint32_t pe_idx = 0, pc = 0;
    for (int p = ps; p <= pe + 1; ++p) {
        T x1p;
        if (p <= pe) {   // iteration n+1 on plane p
            T xp = a.Xin[e + sE2];
            T xW = a.Xin[e - 1], xE = a.Xin[e + 1];
            T xS = a.Xin[e - sE1], xN = a.Xin[e + sE1];
            const T kW = ld_k<KT>(a.kx + fx), kE = ld_k<KT>(a.kx + fx + 1);
            const T kS = ld_k<KT>(a.ky + fy), kN = ld_k<KT>(a.ky + fy + nx);
            const T kF = ld_k<KT>(a.kz + c + sC2);
            const T kB = kzB;
            const T r = __builtin_nontemporal_load(a.R + c);
            const T f0o = __builtin_nontemporal_load(F0in + c);
            if (p + 1 == nz) xp = ghost(a.bct[5], a.bcv[5], x0c);
            if (wlo) xW = ghost(a.bct[0], a.bcv[0], x0c);
            if (whi) xE = ghost(a.bct[1], a.bcv[1], x0c);
            if (slo) xS = ghost(a.bct[2], a.bcv[2], x0c);
            if (shi) xN = ghost(a.bct[3], a.bcv[3], x0c);
            T f = cell_residual<T, TRANSIENT>(a, x0c, xW, xE, xS, xN, x0m, xp, kW, kE, kS, kN, kB, kF, r);
            const T D = DSTORED ? __builtin_nontemporal_load(a.Dst + c) : cell_diag(a, kW, kE, kS, kN, kB, kF);
            f = precond<T>(f, D);
            const T f0 = momentum<T>(h1, f0o, h2, f);
            x1p = step_x<T>(x0c, a.dtau, f0);
            L[p % 3][ly][lx] = x1p;
            x0m = x0c;
            x0c = xp;
            kzB = kF;
            __syncthreads();
            if (p - 1 >= k0 && inner) {   // iteration n+2 on plane p-1
                const int q = (p - 1) % 3;
                T yW = L[q][ly][lx - 1], yE = L[q][ly][lx + 1], yS = L[q][ly - 1][lx], yN = L[q][ly + 1][lx];
                T yB = x1m;
                if (p - 1 == 0) yB = ghost(a.bct[4], a.bcv[4], x1c);
                if (wlo) yW = ghost(a.bct[0], a.bcv[0], x1c);
                if (whi) yE = ghost(a.bct[1], a.bcv[1], x1c);
                if (slo) yS = ghost(a.bct[2], a.bcv[2], x1c);
                if (shi) yN = ghost(a.bct[3], a.bcv[3], x1c);
                T g = cell_residual<T, TRANSIENT>(a, x1c, yW, yE, yS, yN, yB, x1p, pkW, pkE, pkS, pkN, pkB, pkF, pr);
                g = precond<T>(g, pD);
                const T g0 = momentum<T>(h1, pf0, h2, g);
                __builtin_nontemporal_store(g0, a.F0 + pc);
                a.Xout[pe_idx] = step_x<T>(x1c, a.dtau, g0);
            }
            x1m = x1c;
            x1c = x1p;
            pkW = kW; pkE = kE; pkS = kS; pkN = kN; pkB = kB; pkF = kF; pr = r; pD = D; pf0 = f0;
            pe_idx = e;
            pc = c;
            e += sE2;
            c += sC2;
            fx += sX2;
            fy += sY2;
        } else {         // past the top plane: iteration n+2 on plane nz-1 of the top chunk
            __syncthreads();
            if (p - 1 >= k0 && p - 1 < k1 && inner) {
                const int q = (p - 1) % 3;
                T yW = L[q][ly][lx - 1], yE = L[q][ly][lx + 1], yS = L[q][ly - 1][lx], yN = L[q][ly + 1][lx];
                T yB = x1m;
                const T yF = ghost(a.bct[5], a.bcv[5], x1c);
                if (p - 1 == 0) yB = ghost(a.bct[4], a.bcv[4], x1c);
                if (wlo) yW = ghost(a.bct[0], a.bcv[0], x1c);
                if (whi) yE = ghost(a.bct[1], a.bcv[1], x1c);
                if (slo) yS = ghost(a.bct[2], a.bcv[2], x1c);
                if (shi) yN = ghost(a.bct[3], a.bcv[3], x1c);
                T g = cell_residual<T, TRANSIENT>(a, x1c, yW, yE, yS, yN, yB, yF, pkW, pkE, pkS, pkN, pkB, pkF, pr);
                g = precond<T>(g, pD);
                const T g0 = momentum<T>(h1, pf0, h2, g);
                __builtin_nontemporal_store(g0, a.F0 + pc);
                a.Xout[pe_idx] = step_x<T>(x1c, a.dtau, g0);
            }
        }
    }
}

// dyrel_pair_kernel with CPL (2 or 4) x-cells per lane.  The overlap is one
// whole lane on each side: a block's lanes hold the cells
// [CPL*(bx*62-1), +64*CPL) of TY rows, lanes 1..62 of rows 1..TY-2 compute
// iteration n+2, so every lane's first cell is a multiple of CPL and its cells
// lie all inside or all outside the grid (nx % CPL == 0).  A lane's cells go
// through the LDS plane as one vector; x neighbours inside a lane stay in
// registers.  Same operations in the same order as the single iteration: X
// and F0 are bit-identical to two single launches.
template <typename T, bool TRANSIENT, bool DSTORED, bool KT, int TY, int KZ, int CPL>
__global__ __launch_bounds__(64 * TY) void dyrel_pair_kernel_v(IterArgs<T> a, const T *__restrict__ F0in) {
    typedef typename CellVec<T, CPL>::type V;
    __shared__ T L[3][TY][64 * CPL];
    const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
    const unsigned t = xcd_tile(blockIdx.x, nblocks, a.flags & FL_REVERSE);
    const int bx = t % a.nbx, by = (t / a.nbx) % a.nby, bz = t / (a.nbx * a.nby);
    const int lx = (int)(threadIdx.x & 63u), ly = (int)(threadIdx.x >> 6);
    const int i = CPL * (bx * 62 - 1 + lx), j = by * (TY - 2) - 1 + ly;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = bz * KZ, k1 = min(k0 + KZ, nz);
    const bool in_dom = i >= 0 && i < nx && j >= 0 && j < ny;
    const bool inner = in_dom && lx > 0 && lx < 63 && ly > 0 && ly < TY - 1;   // computes iteration n+2
    const int ic = min(max(i, 0), nx - CPL), jc = min(max(j, 0), ny - 1);    // every load in bounds
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t sC2 = (uint32_t)nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = (uint32_t)nx * (ny + 1);
    const bool wlo = ic == 0, whi = ic + CPL == nx, slo = jc == 0, shi = jc == ny - 1;
    const int ps = max(k0 - 1, 0), pe = min(k1, nz - 1);   // planes of iteration n+1
    const T h1 = a.hv ? a.hv[0] : a.h1, h2 = a.hv ? a.hv[1] : a.h2;
    uint32_t e = (uint32_t)(ic + 1) + (uint32_t)(jc + 1) * sE1 + (uint32_t)(ps + 1) * sE2;
    uint32_t c = (uint32_t)ic + (uint32_t)jc * nx + (uint32_t)ps * sC2;
    uint32_t fx = (uint32_t)ic + (uint32_t)jc * sX1 + (uint32_t)ps * sX2;
    uint32_t fy = (uint32_t)ic + (uint32_t)jc * nx + (uint32_t)ps * sY2;
    V x0c = ldv<T, CPL>(a.Xin + e);
    V x0m = ldv<T, CPL>(a.Xin + e - sE2);
    if (ps == 0) {
#pragma unroll
        for (int q = 0; q < CPL; ++q) x0m[q] = ghost(a.bct[4], a.bcv[4], (T)x0c[q]);
    }
    V kzB = ldv_k<KT, T, CPL>(a.kz + c);
    // plane p-1 as iteration n+1 left it: the iterate, its neighbours in z, and
    // the coefficients, R, D and F0 the n+2 update needs
    V x1m = {}, x1c = {}, pkE = {}, pkS = {}, pkN = {}, pkB = {}, pkF = {}, pr = {}, pD = {}, pf0 = {};
    T pkW = 0;
    uint32_t pe_idx = 0, pc = 0;
    // iteration n+2 on plane p-1 of this lane (yF: the n+1 iterate of plane p)
    auto second = [&](int p, const V &yF) {
        const int qb = (p - 1) % 3;
        T yW = L[qb][ly][CPL * lx - 1], yE = L[qb][ly][CPL * lx + CPL];
        V yS = *reinterpret_cast<const V *>(&L[qb][ly - 1][CPL * lx]);
        V yN = *reinterpret_cast<const V *>(&L[qb][ly + 1][CPL * lx]);
        V yB = x1m;
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            if (p - 1 == 0) yB[q] = ghost(a.bct[4], a.bcv[4], (T)x1c[q]);
            if (slo) yS[q] = ghost(a.bct[2], a.bcv[2], (T)x1c[q]);
            if (shi) yN[q] = ghost(a.bct[3], a.bcv[3], (T)x1c[q]);
        }
        if (wlo) yW = ghost(a.bct[0], a.bcv[0], (T)x1c[0]);
        if (whi) yE = ghost(a.bct[1], a.bcv[1], (T)x1c[CPL - 1]);
        V g0, xo;
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            const T w = q == 0 ? yW : (T)x1c[q - 1], ea = q == CPL - 1 ? yE : (T)x1c[q + 1];
            const T kw = q == 0 ? pkW : (T)pkE[q - 1];
            T g = cell_residual<T, TRANSIENT>(a, x1c[q], w, ea, yS[q], yN[q], yB[q], yF[q], kw, pkE[q], pkS[q], pkN[q],
                                              pkB[q], pkF[q], pr[q]);
            g = precond<T>(g, pD[q]);
            g0[q] = momentum<T>(h1, pf0[q], h2, g);
            xo[q] = step_x<T>(x1c[q], a.dtau, g0[q]);
        }
        stv_nt<T, CPL>(g0, a.F0 + pc);
        stv<T, CPL>(xo, a.Xout + pe_idx);
    };
    for (int p = ps; p <= pe + 1; ++p) {
        if (p <= pe) {   // iteration n+1 on plane p
            V xp = ldv<T, CPL>(a.Xin + e + sE2);
            T xW = a.Xin[e - 1], xE = a.Xin[e + CPL];
            V xS = ldv<T, CPL>(a.Xin + e - sE1), xN = ldv<T, CPL>(a.Xin + e + sE1);
            const T kW = ld_k<KT>(a.kx + fx);
            const V kE = ldv_k<KT, T, CPL>(a.kx + fx + 1);
            const V kS = ldv_k<KT, T, CPL>(a.ky + fy), kN = ldv_k<KT, T, CPL>(a.ky + fy + nx);
            const V kF = ldv_k<KT, T, CPL>(a.kz + c + sC2);
            const V kB = kzB;
            const V r = ldv_nt<T, CPL>(a.R + c);
            const V f0o = ldv_nt<T, CPL>(F0in + c);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                if (p + 1 == nz) xp[q] = ghost(a.bct[5], a.bcv[5], (T)x0c[q]);
                if (slo) xS[q] = ghost(a.bct[2], a.bcv[2], (T)x0c[q]);
                if (shi) xN[q] = ghost(a.bct[3], a.bcv[3], (T)x0c[q]);
            }
            if (wlo) xW = ghost(a.bct[0], a.bcv[0], (T)x0c[0]);
            if (whi) xE = ghost(a.bct[1], a.bcv[1], (T)x0c[CPL - 1]);
            V D, f0, x1p;
            if (DSTORED) D = ldv_nt<T, CPL>(a.Dst + c);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T w = q == 0 ? xW : (T)x0c[q - 1], ea = q == CPL - 1 ? xE : (T)x0c[q + 1];
                const T kw = q == 0 ? kW : (T)kE[q - 1];
                T f = cell_residual<T, TRANSIENT>(a, x0c[q], w, ea, xS[q], xN[q], x0m[q], xp[q], kw, kE[q], kS[q], kN[q],
                                                  kB[q], kF[q], r[q]);
                if (!DSTORED) D[q] = cell_diag(a, kw, (T)kE[q], (T)kS[q], (T)kN[q], (T)kB[q], (T)kF[q]);
                f = precond<T>(f, D[q]);
                f0[q] = momentum<T>(h1, f0o[q], h2, f);
                x1p[q] = step_x<T>(x0c[q], a.dtau, f0[q]);
            }
            *reinterpret_cast<V *>(&L[p % 3][ly][CPL * lx]) = x1p;
            x0m = x0c;
            x0c = xp;
            kzB = kF;
            __syncthreads();
            if (p - 1 >= k0 && inner) second(p, x1p);
            x1m = x1c;
            x1c = x1p;
            pkW = kW; pkE = kE; pkS = kS; pkN = kN; pkB = kB; pkF = kF; pr = r; pD = D; pf0 = f0;
            pe_idx = e;
            pc = c;
            e += sE2;
            c += sC2;
            fx += sX2;
            fy += sY2;
        } else {         // past the top plane: iteration n+2 on plane nz-1 of the top chunk
            __syncthreads();
            if (p - 1 >= k0 && p - 1 < k1 && inner) {
                V yF;
#pragma unroll
                for (int q = 0; q < CPL; ++q) yF[q] = ghost(a.bct[5], a.bcv[5], (T)x1c[q]);
                second(p, yF);
            }
        }
    }
}

// ------------------------------------------------------------------ host driver
// Tile configurations (rows of 64 cells per block, z cells per thread); the
// autotuner picks one per (grid, operator, precision, D mode).
struct TileCfg {
    int ty, kz;
    int cpl = 1;    // cells per lane along x
    int band = 0;   // pair tiles: 1 = bands of whole rows (batch_pair_kernel), 0 = overlapped 2-D tiles
};
static constexpr TileCfg kTiles[] = {{2, 8}, {2, 16}, {4, 4}, {4, 8}, {4, 16}, {4, 32}, {8, 4}, {8, 8},
                                      {8, 16}, {4, 2}, {8, 2}, {16, 2}, {16, 4}, {1, 8},
                                      {4, 4, 2}, {4, 8, 2}, {2, 8, 2}, {8, 4, 2}, {4, 16, 2}, {2, 16, 2},
                                      {4, 4, 4}, {4, 8, 4}, {2, 8, 4}, {4, 16, 4}, {2, 16, 4}};
static constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
static constexpr int kDefaultTile = 2;

template <typename T, bool TR, bool DS, bool NX, bool KT, int TY, int KZ, int CPL = 1>
inline void launch_t(const IterArgs<T> &a, unsigned nb, hipStream_t st) {
    hipLaunchKernelGGL((dyrel_iter_kernel<T, TR, DS, NX, KT, TY, KZ, CPL>), dim3(nb), dim3(64 * TY), 0, st, a);
}

template <typename T, bool TR, bool DS, bool NX, bool KT>
inline void launch_tile(int tile, const IterArgs<T> &a, unsigned nb, hipStream_t st) {
    switch (tile) {
    case 0: launch_t<T, TR, DS, NX, KT, 2, 8>(a, nb, st); break;
    case 1: launch_t<T, TR, DS, NX, KT, 2, 16>(a, nb, st); break;
    case 2: launch_t<T, TR, DS, NX, KT, 4, 4>(a, nb, st); break;
    case 3: launch_t<T, TR, DS, NX, KT, 4, 8>(a, nb, st); break;
    case 4: launch_t<T, TR, DS, NX, KT, 4, 16>(a, nb, st); break;
    case 5: launch_t<T, TR, DS, NX, KT, 4, 32>(a, nb, st); break;
    case 6: launch_t<T, TR, DS, NX, KT, 8, 4>(a, nb, st); break;
    case 7: launch_t<T, TR, DS, NX, KT, 8, 8>(a, nb, st); break;
    case 8: launch_t<T, TR, DS, NX, KT, 8, 16>(a, nb, st); break;
    case 9: launch_t<T, TR, DS, NX, KT, 4, 2>(a, nb, st); break;
    case 10: launch_t<T, TR, DS, NX, KT, 8, 2>(a, nb, st); break;
    case 11: launch_t<T, TR, DS, NX, KT, 16, 2>(a, nb, st); break;
    case 12: launch_t<T, TR, DS, NX, KT, 16, 4>(a, nb, st); break;
    case 13: launch_t<T, TR, DS, NX, KT, 1, 8>(a, nb, st); break;
    case 14: launch_t<T, TR, DS, NX, KT, 4, 4, 2>(a, nb, st); break;
    case 15: launch_t<T, TR, DS, NX, KT, 4, 8, 2>(a, nb, st); break;
    case 16: launch_t<T, TR, DS, NX, KT, 2, 8, 2>(a, nb, st); break;
    case 17: launch_t<T, TR, DS, NX, KT, 8, 4, 2>(a, nb, st); break;
    case 18: launch_t<T, TR, DS, NX, KT, 4, 16, 2>(a, nb, st); break;
    case 19: launch_t<T, TR, DS, NX, KT, 2, 16, 2>(a, nb, st); break;
    case 20: launch_t<T, TR, DS, NX, KT, 4, 4, 4>(a, nb, st); break;
    case 21: launch_t<T, TR, DS, NX, KT, 4, 8, 4>(a, nb, st); break;
    case 22: launch_t<T, TR, DS, NX, KT, 2, 8, 4>(a, nb, st); break;
    case 23: launch_t<T, TR, DS, NX, KT, 4, 16, 4>(a, nb, st); break;
    default: launch_t<T, TR, DS, NX, KT, 2, 16, 4>(a, nb, st); break;
    }
}

// variant bits: 1 = read the stored diagonal instead of recomputing it,
//               2 = non-temporal store of the updated unknown,
//               4 = cached (not streamed) face-coefficient loads,
//               8 = face coefficients from the vertex field (dyrel_iter_kv_kernel;
//                   with it bit 1 applies in FP64 only, and bit 4 means a z
//                   march unrolled by two planes, in FP32 only)
struct IterCfg {
    int tile = kDefaultTile;
    int variant = 0;
};

// A tile with CPL cells per lane needs nx % CPL == 0; otherwise its twin
// (same rows and planes) with the most cells per lane that nx allows is used.
inline int tile_for_grid(int tile, int nx) {
    if (nx % kTiles[tile].cpl == 0) return tile;
    for (int cpl = kTiles[tile].cpl / 2; cpl >= 1; cpl /= 2) {
        if (nx % cpl) continue;
        for (int t = 0; t < kNumTiles; ++t)
            if (kTiles[t].cpl == cpl && kTiles[t].ty == kTiles[tile].ty && kTiles[t].kz == kTiles[tile].kz) return t;
    }
    return kDefaultTile;
}

template <typename T>
inline void set_tiling(IterArgs<T> &a, int tile) {
    a.nbx = (int)div_up(a.nx, 64 * kTiles[tile].cpl);
    a.nby = (int)div_up(a.ny, kTiles[tile].ty);
    a.nbz = (int)div_up(a.nz, kTiles[tile].kz);
}

// Vertex-coefficient kernels (variant bit 8, dyrel_iter_kv_kernel), when the
// solve has the vertex field: the tiles with two cells per lane, and in FP32
// also those with four.
template <typename T>
inline bool kv_tile_ok(int tile, int nx) {
    const int cpl = kTiles[tile].cpl;
    return (cpl == 2 || (cpl == 4 && sizeof(T) == 4)) && nx % cpl == 0;
}
template <typename T>
inline bool kv_cfg_ok(const IterArgs<T> &a, const IterCfg &cfg) {
    return !(cfg.variant & 8) || (a.kv && kv_tile_ok<T>(cfg.tile, a.nx));
}
// drops the vertex bit where it cannot run (no vertex field, or another
// tile), and bit 4 with it; with the vertex bit, bit 1 (stored diagonal)
// applies in FP64 only and bit 4 (z march unrolled by two) in FP32 only
template <typename T>
inline void kv_normalize(const IterArgs<T> &a, IterCfg &cfg) {
    if ((cfg.variant & 8) && !(a.kv && kv_tile_ok<T>(cfg.tile, a.nx))) cfg.variant &= ~(8 | 4);
    if (cfg.variant & 8) cfg.variant &= sizeof(T) == 8 ? ~4 : ~1;
}

template <typename T, bool TR, bool DS, bool NX, bool RL, int TY, int KZ, int CPL>
inline void launch_kv(const IterArgs<T> &a, unsigned nb, hipStream_t st) {
    hipLaunchKernelGGL((dyrel_iter_kv_kernel<T, TR, DS, NX, RL, TY, KZ, CPL>), dim3(nb), dim3(64 * TY), 0, st, a);
}

template <typename T, bool TR, bool DS, bool NX, bool RL>
inline void launch_kv_tile(int tile, const IterArgs<T> &a, unsigned nb, hipStream_t st) {
    switch (tile) {
    case 14: launch_kv<T, TR, DS, NX, RL, 4, 4, 2>(a, nb, st); break;
    case 15: launch_kv<T, TR, DS, NX, RL, 4, 8, 2>(a, nb, st); break;
    case 16: launch_kv<T, TR, DS, NX, RL, 2, 8, 2>(a, nb, st); break;
    case 17: launch_kv<T, TR, DS, NX, RL, 8, 4, 2>(a, nb, st); break;
    case 18: launch_kv<T, TR, DS, NX, RL, 4, 16, 2>(a, nb, st); break;
    case 19: launch_kv<T, TR, DS, NX, RL, 2, 16, 2>(a, nb, st); break;
    default:
        if constexpr (sizeof(T) == 4) {   // 20-24 (kv_tile_ok admits 14-24 in FP32, 14-19 in FP64)
            switch (tile) {
            case 20: launch_kv<T, TR, DS, NX, RL, 4, 4, 4>(a, nb, st); break;
            case 21: launch_kv<T, TR, DS, NX, RL, 4, 8, 4>(a, nb, st); break;
            case 22: launch_kv<T, TR, DS, NX, RL, 2, 8, 4>(a, nb, st); break;
            case 23: launch_kv<T, TR, DS, NX, RL, 4, 16, 4>(a, nb, st); break;
            default: launch_kv<T, TR, DS, NX, RL, 2, 16, 4>(a, nb, st); break;
            }
        }
        break;
    }
}

template <typename T, bool TR>
inline void launch_variant(const IterArgs<T> &a, const IterCfg &cfg, unsigned nb, hipStream_t st) {
    if (cfg.variant & 8) {   // FP64: bits 1 and 2; FP32: bits 2 and 4 (kv_normalize)
        if constexpr (sizeof(T) == 8) {
            switch (cfg.variant & 3) {
            case 0: launch_kv_tile<T, TR, false, false, false>(cfg.tile, a, nb, st); break;
            case 1: launch_kv_tile<T, TR, true, false, false>(cfg.tile, a, nb, st); break;
            case 2: launch_kv_tile<T, TR, false, true, false>(cfg.tile, a, nb, st); break;
            default: launch_kv_tile<T, TR, true, true, false>(cfg.tile, a, nb, st); break;
            }
        } else {
            switch (cfg.variant & 6) {
            case 0: launch_kv_tile<T, TR, false, false, false>(cfg.tile, a, nb, st); break;
            case 2: launch_kv_tile<T, TR, false, true, false>(cfg.tile, a, nb, st); break;
            case 4: launch_kv_tile<T, TR, false, false, true>(cfg.tile, a, nb, st); break;
            default: launch_kv_tile<T, TR, false, true, true>(cfg.tile, a, nb, st); break;
            }
        }
        return;
    }
    switch (cfg.variant & 7) {
    case 0: launch_tile<T, TR, false, false, false>(cfg.tile, a, nb, st); break;
    case 1: launch_tile<T, TR, true, false, false>(cfg.tile, a, nb, st); break;
    case 2: launch_tile<T, TR, false, true, false>(cfg.tile, a, nb, st); break;
    case 3: launch_tile<T, TR, true, true, false>(cfg.tile, a, nb, st); break;
    case 4: launch_tile<T, TR, false, false, true>(cfg.tile, a, nb, st); break;
    case 5: launch_tile<T, TR, true, false, true>(cfg.tile, a, nb, st); break;
    case 6: launch_tile<T, TR, false, true, true>(cfg.tile, a, nb, st); break;
    default: launch_tile<T, TR, true, true, true>(cfg.tile, a, nb, st); break;
    }
}

template <typename T>
inline int launch_iter(const IterArgs<T> &a, bool transient, const IterCfg &cfg) {
    if (!kv_cfg_ok(a, cfg)) return arg_fail("dyrel_iter_kernel: vertex-coefficient variant on a tile or solve without it");
    const unsigned nb = (unsigned)a.nbx * a.nby * a.nbz;
    hipStream_t st = ctx().stream;
    if (transient) launch_variant<T, true>(a, cfg, nb, st);
    else launch_variant<T, false>(a, cfg, nb, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "dyrel_iter_kernel");
    return HT3D_OK;
}

// --------------------------------------------------------- two-iteration launches
// Pair tiles (rows of the overlapped block, planes per block, x-cells per
// lane); variant bits: 1 = stored diagonal, 2 = cached face-coefficient loads.
// The last five are bands of whole rows (64*rows lanes of CPL cells, the
// band-pair kernels of the batch with a one-entry table), FP32 sweeps only,
// without variants.  tile < 0: no pairs.
static constexpr TileCfg kPairTiles[] = {{8, 16}, {8, 32}, {16, 16}, {16, 32}, {16, 64},
                                          {8, 16, 2}, {8, 32, 2}, {16, 32, 2},
                                          {8, 16, 4}, {8, 32, 4}, {16, 32, 4},
                                          {8, 16, 4, 1}, {16, 16, 4, 1}, {16, 8, 4, 1}, {16, 16, 2, 1}, {16, 16, 1, 1}};
static constexpr int kNumPairTiles = sizeof(kPairTiles) / sizeof(kPairTiles[0]);

struct PairCfg {
    int tile = -1;
    int variant = 0;
};

// Whether a pair tile can run on a grid of x extent nx: CPL cells per lane
// need nx % CPL == 0; in FP64 the vector tiles need more registers than a
// 1024-thread block has (4 cells per lane, or 2 with 16 rows) and are not used.
template <typename T>
inline bool pair_tile_ok(int tile, int nx) {
    const int cpl = kPairTiles[tile].cpl;
    if (kPairTiles[tile].band)   // whole rows in a band of at least 3 (one inner row), FP32 only
        return sizeof(T) == 4 && nx % cpl == 0 && 64 * kPairTiles[tile].ty * cpl / nx >= 3;
    if (sizeof(T) == 8 && (cpl == 4 || (cpl == 2 && kPairTiles[tile].ty > 8))) return false;
    return nx % cpl == 0;
}

template <typename T>
inline void set_pair_tiling(IterArgs<T> &a, int tile) {
    if (kPairTiles[tile].band) {   // bands of rb rows overlapping by one row, chunks of kz planes
        const int rb = 64 * kPairTiles[tile].ty * kPairTiles[tile].cpl / a.nx;
        a.nband = (int)div_up(a.ny, rb - 2);
        a.nbx = a.nby = 1;
        a.nbz = (int)div_up(a.nz, kPairTiles[tile].kz);
        return;
    }
    a.nbx = (int)div_up(a.nx, 62 * kPairTiles[tile].cpl);
    a.nby = (int)div_up(a.ny, kPairTiles[tile].ty - 2);
    a.nbz = (int)div_up(a.nz, kPairTiles[tile].kz);
}

template <typename T, bool TR, bool DS, bool KT>
inline void launch_pair_tile(int tile, const IterArgs<T> &a, const T *F0in, unsigned nb, hipStream_t st) {
#define PL_(TY, KZ) hipLaunchKernelGGL((dyrel_pair_kernel<T, TR, DS, KT, TY, KZ>), dim3(nb), dim3(64 * TY), 0, st, a, F0in)
#define PV_(TY, KZ, CPL) \
    hipLaunchKernelGGL((dyrel_pair_kernel_v<T, TR, DS, KT, TY, KZ, CPL>), dim3(nb), dim3(64 * TY), 0, st, a, F0in)
    switch (tile) {
    case 0: PL_(8, 16); break;
    case 1: PL_(8, 32); break;
    case 2: PL_(16, 16); break;
    case 3: PL_(16, 32); break;
    case 4: PL_(16, 64); break;
    case 5: PV_(8, 16, 2); break;
    case 6: PV_(8, 32, 2); break;
    case 7:
        if constexpr (sizeof(T) == 4) PV_(16, 32, 2);   // pair_tile_ok keeps FP64 off it
        break;
    default:
        if constexpr (sizeof(T) == 4) {   // pair_tile_ok keeps FP64 off these
            if (tile == 8) PV_(8, 16, 4);
            else if (tile == 9) PV_(8, 32, 4);
            else if (tile == 10) PV_(16, 32, 4);
        }
        break;   // band tiles go through launch_band_pair
    }
#undef PL_
#undef PV_
}

// a: pair tiling set, a.F0 = the F0 buffer written, F0in = the one read
template <typename T>
inline int launch_pair(const IterArgs<T> &a, const T *F0in, bool transient, const PairCfg &pc) {
    if (pc.tile < 0 || pc.tile >= kNumPairTiles || kPairTiles[pc.tile].band)
        return arg_fail("launch_pair: not an overlapped 2-D pair tile (band tiles go through launch_band_pair)");
    const unsigned nb = (unsigned)a.nbx * a.nby * a.nbz;
    hipStream_t st = ctx().stream;
    const bool ds = pc.variant & 1, kt = pc.variant & 2;
    if (transient) {
        if (ds) { if (kt) launch_pair_tile<T, true, true, true>(pc.tile, a, F0in, nb, st); else launch_pair_tile<T, true, true, false>(pc.tile, a, F0in, nb, st); }
        else { if (kt) launch_pair_tile<T, true, false, true>(pc.tile, a, F0in, nb, st); else launch_pair_tile<T, true, false, false>(pc.tile, a, F0in, nb, st); }
    } else {
        if (ds) { if (kt) launch_pair_tile<T, false, true, true>(pc.tile, a, F0in, nb, st); else launch_pair_tile<T, false, true, false>(pc.tile, a, F0in, nb, st); }
        else { if (kt) launch_pair_tile<T, false, false, true>(pc.tile, a, F0in, nb, st); else launch_pair_tile<T, false, false, false>(pc.tile, a, F0in, nb, st); }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "dyrel_pair_kernel");
    return HT3D_OK;
}

// Gershgorin diagonal of every cell in T, with the formula and order the
// iteration kernels use (so the stored-D variants stay bit-identical)
template <typename T>
__global__ __launch_bounds__(256) void diag_kernel(IterArgs<T> a, T *__restrict__ D) {
    const uint32_t nx = a.nx, ny = a.ny;
    const size_t N = (size_t)nx * ny * a.nz;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(c % nx), j = (uint32_t)((c / nx) % ny), k = (uint32_t)(c / ((size_t)nx * ny));
        const size_t fx = i + (size_t)j * (nx + 1) + (size_t)k * (nx + 1) * ny;
        const size_t fy = i + (size_t)j * nx + (size_t)k * nx * (ny + 1);
        D[c] = cell_diag(a, a.kx[fx], a.kx[fx + 1], a.ky[fy], a.ky[fy + nx], a.kz[c], a.kz[c + (size_t)nx * ny]);
    }
}

// The last update of a defect-correction solve, after the anchor that stopped
// it.  The FP64 solve stops at a check iteration: its stopping test evaluates
// the residual of the iterate that iteration starts from, and the iteration
// still applies its DYREL update, so the solve returns the tested iterate plus
// one update, with ghosts from the tested iterate (in the reference the BC
// kernel runs before the update, HydroThermal3D_v18.jl:448-456).  The anchor
// has just set x = x + e and e = 0, so the FP32 sweep that would follow it
// reduces to a pointwise step from the anchor's right-hand side R:
// F = R / D, F0 = h1 F0 + h2 F, x += -dtau F0 -- the sweep kernels' operations
// on e = 0, bit for bit.  hsc = {h1, h2, dtau} on the device (the solve's
// current DYREL scalars).  Only interior cells of X are written: its ghosts
// stay as the caller filled them from the tested iterate.
template <typename T>
__global__ __launch_bounds__(256) void dc_final_kernel(IterArgs<T> a, const T *__restrict__ hsc,
                                                       double *__restrict__ X) {
    const uint32_t nx = a.nx, ny = a.ny;
    const size_t N = (size_t)nx * ny * a.nz;
    const T h1 = hsc[0], h2 = hsc[1], dtau = hsc[2];
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(c % nx), j = (uint32_t)((c / nx) % ny), k = (uint32_t)(c / ((size_t)nx * ny));
        const size_t fx = i + (size_t)j * (nx + 1) + (size_t)k * (nx + 1) * ny;
        const size_t fy = i + (size_t)j * nx + (size_t)k * nx * (ny + 1);
        const T D = cell_diag(a, a.kx[fx], a.kx[fx + 1], a.ky[fy], a.ky[fy + nx], a.kz[c], a.kz[c + (size_t)nx * ny]);
        const T f0 = momentum<T>(h1, a.F0[c], h2, precond<T>(a.R[c], D));
        a.F0[c] = f0;
        X[(i + 1) + (size_t)(j + 1) * (nx + 2) + (size_t)(k + 1) * (nx + 2) * (ny + 2)] += (double)step_x<T>((T)0, dtau, f0);
    }
}

// ------------------------------------------------ band pairs (batch and DC)
// Per-problem state seen by the band-pair kernels: they skip the items of a
// problem whose `active` is 0 (the batch's BState, or OneActive for a single
// problem).
struct OneActive {
    int active;
};

// Two plain pseudo-iterations per launch, on bands of whole rows.  An item
// (problem p, band-chunk index) of the work list is a band of RB = 64*TY/nx
// rows of problem p (the batch's grids are 32 to 128 cells wide; the
// single-problem solve uses a one-entry table), marched over KZ planes.  Iteration n+1 runs on the band's rows
// [bb*RI - 1, bb*RI + RI + 1), iteration n+2 on the RI = RB - 2 inner rows, so
// neighbouring bands overlap by one row on each side.  A lane holds cell
// (tid mod nx, row tid / nx) of the band: its x neighbours are lanes tid -+ 1
// and its y neighbours lanes tid -+ nx, reached through a triple-buffered LDS
// plane; the z neighbours go through a register queue, n+1 running one plane
// ahead of n+2 (the scheme of dyrel_pair_kernel).  Lanes past the last whole
// row, and rows outside the grid, compute on clamped (in-bounds) cells and
// write nothing.  F0 is read from one of the two F0 buffers (fpar) and
// written to the other, since an overlapping band still needs the old F0; X
// goes from the `phase` buffer to the other.  Every cell is computed with the
// single iteration's operations in the same order: two single launches and one
// pair launch give bit-identical X and F0.
template <typename T, bool TR, int TY, int KZ, typename S>
__global__ __launch_bounds__(64 * TY) void batch_pair_kernel(const IterArgs<T> *__restrict__ tab,
                                                             const uint2 *__restrict__ items,
                                                             const unsigned *__restrict__ n_items,
                                                             const S *__restrict__ st, int phase, int fpar,
                                                             long long dF0, unsigned flags) {
    constexpr int NT = 64 * TY;
    __shared__ T L[3][NT];
    const unsigned n = *n_items;
    const int tid = (int)threadIdx.x;
    for (unsigned w = blockIdx.x; w < n; w += gridDim.x) {
        const unsigned t = xcd_tile(w, n, flags & FL_REVERSE);
        const uint2 it = items[t];
        if (!st[it.x].active) continue;   // the whole block skips the item
        const IterArgs<T> &a = tab[it.x];
        const int nx = a.nx, ny = a.ny, nz = a.nz;
        const int RB = NT / nx, RI = RB - 2;
        const int bb = (int)(it.y % (unsigned)a.nband), bz = (int)(it.y / (unsigned)a.nband);
        const int r = tid / nx, i = tid - r * nx;
        const int j = bb * RI - 1 + r;
        const bool in_dom = r < RB && j >= 0 && j < ny;
        const bool inner = in_dom && r >= 1 && r <= RB - 2;   // computes iteration n+2
        const int jc = min(max(j, 0), ny - 1);                // every load in bounds
        // Pointers read from the table are generic to the compiler, whose flat
        // accesses count against the LDS wait counter as well: typed as global,
        // the loads and stores of the march do not hold up the LDS waits.
        typedef __attribute__((address_space(1))) T GT;
        typedef __attribute__((address_space(1))) const T GCT;
        GCT *__restrict__ Xin = (GCT *)(phase ? a.Xout : a.Xin);
        GT *__restrict__ Xout = (GT *)(phase ? const_cast<T *>(a.Xin) : a.Xout);
        GCT *__restrict__ F0in = (GCT *)(fpar ? a.F0 + dF0 : a.F0);
        GT *__restrict__ F0out = (GT *)(fpar ? a.F0 : a.F0 + dF0);
        GCT *__restrict__ kxg = (GCT *)a.kx, *__restrict__ kyg = (GCT *)a.ky, *__restrict__ kzg = (GCT *)a.kz;
        GCT *__restrict__ Rg = (GCT *)a.R;
        const int k0 = bz * KZ, k1 = min(k0 + KZ, nz);
        const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
        const uint32_t sC2 = (uint32_t)nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = (uint32_t)nx * (ny + 1);
        const bool wlo = i == 0, whi = i == nx - 1, slo = jc == 0, shi = jc == ny - 1;
        const int ps = max(k0 - 1, 0), pe = min(k1, nz - 1);   // planes of iteration n+1
        const T h1 = a.hv ? a.hv[0] : a.h1, h2 = a.hv ? a.hv[1] : a.h2;
        uint32_t e = (uint32_t)(i + 1) + (uint32_t)(jc + 1) * sE1 + (uint32_t)(ps + 1) * sE2;
        uint32_t c = (uint32_t)i + (uint32_t)jc * nx + (uint32_t)ps * sC2;
        uint32_t fx = (uint32_t)i + (uint32_t)jc * sX1 + (uint32_t)ps * sX2;
        uint32_t fy = (uint32_t)i + (uint32_t)jc * nx + (uint32_t)ps * sY2;
        T x0c = Xin[e];
        T x0m = Xin[e - sE2];
        if (ps == 0) x0m = ghost(a.bct[4], a.bcv[4], x0c);
        T kzB = kzg[c];
        T x1m = 0, x1c = 0, pkW = 0, pkE = 0, pkS = 0, pkN = 0, pkB = 0, pkF = 0, pr = 0, pD = 1, pf0 = 0;
        uint32_t pe_idx = 0, pc = 0;
        // iteration n+2 on plane p-1, from the n+1 plane in L[(p-1) % 3]
        auto second = [&](int p, T yF) {
            const int q = (p - 1) % 3;
            T yW = L[q][tid - 1], yE = L[q][tid + 1], yS = L[q][tid - nx], yN = L[q][tid + nx];
            T yB = x1m;
            if (p - 1 == 0) yB = ghost(a.bct[4], a.bcv[4], x1c);
            if (wlo) yW = ghost(a.bct[0], a.bcv[0], x1c);
            if (whi) yE = ghost(a.bct[1], a.bcv[1], x1c);
            if (slo) yS = ghost(a.bct[2], a.bcv[2], x1c);
            if (shi) yN = ghost(a.bct[3], a.bcv[3], x1c);
            T g = cell_residual<T, TR>(a, x1c, yW, yE, yS, yN, yB, yF, pkW, pkE, pkS, pkN, pkB, pkF, pr);
            g = precond<T>(g, pD);
            const T g0 = momentum<T>(h1, pf0, h2, g);
            __builtin_nontemporal_store(g0, F0out + pc);
            Xout[pe_idx] = step_x<T>(x1c, a.dtau, g0);
        };
        for (int p = ps; p <= pe + 1; ++p) {
            if (p <= pe) {   // iteration n+1 on plane p
                T xp = Xin[e + sE2];
                T xW = Xin[e - 1], xE = Xin[e + 1];
                T xS = Xin[e - sE1], xN = Xin[e + sE1];
                const T kW = kxg[fx], kE = kxg[fx + 1];
                const T kS = kyg[fy], kN = kyg[fy + nx];
                const T kF = kzg[c + sC2];
                const T kB = kzB;
                const T rr = __builtin_nontemporal_load(Rg + c);
                const T f0o = __builtin_nontemporal_load(F0in + c);
                if (p + 1 == nz) xp = ghost(a.bct[5], a.bcv[5], x0c);
                if (wlo) xW = ghost(a.bct[0], a.bcv[0], x0c);
                if (whi) xE = ghost(a.bct[1], a.bcv[1], x0c);
                if (slo) xS = ghost(a.bct[2], a.bcv[2], x0c);
                if (shi) xN = ghost(a.bct[3], a.bcv[3], x0c);
                T f = cell_residual<T, TR>(a, x0c, xW, xE, xS, xN, x0m, xp, kW, kE, kS, kN, kB, kF, rr);
                const T D = cell_diag(a, kW, kE, kS, kN, kB, kF);
                f = precond<T>(f, D);
                const T f0 = momentum<T>(h1, f0o, h2, f);
                const T x1p = step_x<T>(x0c, a.dtau, f0);
                L[p % 3][tid] = x1p;
                x0m = x0c;
                x0c = xp;
                kzB = kF;
                __syncthreads();
                if (p - 1 >= k0 && inner) second(p, x1p);
                x1m = x1c;
                x1c = x1p;
                pkW = kW; pkE = kE; pkS = kS; pkN = kN; pkB = kB; pkF = kF; pr = rr; pD = D; pf0 = f0;
                pe_idx = e;
                pc = c;
                e += sE2;
                c += sC2;
                fx += sX2;
                fy += sY2;
            } else {         // past the top plane: iteration n+2 on plane nz-1 of the top chunk
                __syncthreads();
                if (p - 1 >= k0 && p - 1 < k1 && inner) second(p, ghost(a.bct[5], a.bcv[5], x1c));
            }
        }
        __syncthreads();   // L is rewritten by the block's next item
    }
}

// batch_pair_kernel with CPL (2 or 4) x-cells per lane.  A band row holds
// gx = nx / CPL lanes, so a band of 64*TY lanes is RB = 64*TY / gx rows: CPL
// times as many as with one cell per lane, which divides the share of overlap
// rows that iteration n+1 recomputes (on a 128-wide grid with 1024 lanes, 16
// rows per band and 14 inner against 8 and 6).  X, the face coefficients, R and
// F0 arrive as CPL-element vector loads (the pattern of iter_tile_v), a lane's
// cells go through the LDS plane as one vector, and x neighbours inside a lane
// stay in registers; a lane's first cell is at x = CPL * (lane mod gx), so a
// row's first and last lanes meet the walls.  Same operations in the same
// order as the single iteration: bit-identical X and F0.  Needs nx % CPL == 0.
// KV: iteration n+1 forms the face coefficients from the vertex field a.kv,
// as iter_tile_kv does (the same values bit for bit); every item of the
// launch must have one.
template <typename T, bool TR, int TY, int KZ, int CPL, typename S, bool KV = false>
__global__ __launch_bounds__(64 * TY) void batch_pair_kernel_v(const IterArgs<T> *__restrict__ tab,
                                                               const uint2 *__restrict__ items,
                                                               const unsigned *__restrict__ n_items,
                                                               const S *__restrict__ st, int phase, int fpar,
                                                               long long dF0, unsigned flags) {
    typedef typename CellVec<T, CPL>::type V;
    constexpr int NT = 64 * TY;
    __shared__ T L[3][NT * CPL];
    const unsigned n = *n_items;
    const int tid = (int)threadIdx.x;
    const int s0 = tid * CPL;   // this lane's slot in an LDS plane
    for (unsigned w = blockIdx.x; w < n; w += gridDim.x) {
        const unsigned t = xcd_tile(w, n, flags & FL_REVERSE);
        const uint2 it = items[t];
        if (!st[it.x].active) continue;   // the whole block skips the item
        const IterArgs<T> &a = tab[it.x];
        const int nx = a.nx, ny = a.ny, nz = a.nz;
        const int gx = nx / CPL, RB = NT / gx, RI = RB - 2;
        const int bb = (int)(it.y % (unsigned)a.nband), bz = (int)(it.y / (unsigned)a.nband);
        const int r = tid / gx, i = (tid - r * gx) * CPL;
        const int j = bb * RI - 1 + r;
        const bool in_dom = r < RB && j >= 0 && j < ny;
        const bool inner = in_dom && r >= 1 && r <= RB - 2;   // computes iteration n+2
        const int jc = min(max(j, 0), ny - 1);                // every load in bounds
        // global-typed pointers and vectors (see batch_pair_kernel)
        typedef __attribute__((address_space(1))) T GT;
        typedef __attribute__((address_space(1))) const T GCT;
        typedef __attribute__((address_space(1))) V GV;
        typedef __attribute__((address_space(1))) const V GCV;
        GCT *__restrict__ Xin = (GCT *)(phase ? a.Xout : a.Xin);
        GT *__restrict__ Xout = (GT *)(phase ? const_cast<T *>(a.Xin) : a.Xout);
        GCT *__restrict__ F0in = (GCT *)(fpar ? a.F0 + dF0 : a.F0);
        GT *__restrict__ F0out = (GT *)(fpar ? a.F0 : a.F0 + dF0);
        GCT *__restrict__ kxg = (GCT *)a.kx, *__restrict__ kyg = (GCT *)a.ky, *__restrict__ kzg = (GCT *)a.kz;
        GCT *__restrict__ Rg = (GCT *)a.R;
        auto ld = [](GCT *p) { return *reinterpret_cast<GCV *>(p); };
        auto ld_nt = [](GCT *p) { return __builtin_nontemporal_load(reinterpret_cast<GCV *>(p)); };
        const int k0 = bz * KZ, k1 = min(k0 + KZ, nz);
        const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
        const uint32_t sC2 = (uint32_t)nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = (uint32_t)nx * (ny + 1);
        const bool wlo = i == 0, whi = i + CPL == nx, slo = jc == 0, shi = jc == ny - 1;
        const int ps = max(k0 - 1, 0), pe = min(k1, nz - 1);   // planes of iteration n+1
        const T h1 = a.hv ? a.hv[0] : a.h1, h2 = a.hv ? a.hv[1] : a.h2;
        uint32_t e = (uint32_t)(i + 1) + (uint32_t)(jc + 1) * sE1 + (uint32_t)(ps + 1) * sE2;
        uint32_t c = (uint32_t)i + (uint32_t)jc * nx + (uint32_t)ps * sC2;
        uint32_t fx = (uint32_t)i + (uint32_t)jc * sX1 + (uint32_t)ps * sX2;
        uint32_t fy = (uint32_t)i + (uint32_t)jc * nx + (uint32_t)ps * sY2;
        V x0c = ld(Xin + e);
        V x0m = ld(Xin + e - sE2);
        if (ps == 0) {
#pragma unroll
            for (int q = 0; q < CPL; ++q) x0m[q] = ghost(a.bct[4], a.bcv[4], (T)x0c[q]);
        }
        // KV: vertex values of plane p on rows jc and jc + 1 (iter_tile_kv)
        const uint32_t sV1 = nx + 1, sV2 = sV1 * (ny + 1);
        uint32_t v = (uint32_t)i + (uint32_t)jc * sV1 + (uint32_t)ps * sV2;
        GCT *__restrict__ kvg = (GCT *)a.kv;
        T a0[CPL + 1], a1[CPL + 1];
        auto vrow = [&](uint32_t o, T *out) {
            const V w = ld(kvg + o);
#pragma unroll
            for (int q = 0; q < CPL; ++q) out[q] = w[q];
            out[CPL] = kvg[o + CPL];
        };
        V kzB;
        if constexpr (KV) {
            vrow(v, a0);
            vrow(v + sV1, a1);
#pragma unroll
            for (int q = 0; q < CPL; ++q) kzB[q] = av4(a0[q], a0[q + 1], a1[q], a1[q + 1]);
        } else {
            kzB = ld(kzg + c);
        }
        // plane p-1 as iteration n+1 left it: the iterate, its neighbours in z,
        // and the coefficients, R, D and F0 the n+2 update needs
        V x1m = {}, x1c = {}, pkE = {}, pkS = {}, pkN = {}, pkB = {}, pkF = {}, pr = {}, pD = {}, pf0 = {};
        T pkW = 0;
        uint32_t pe_idx = 0, pc = 0;
        // iteration n+2 on plane p-1 of this lane (yF: the n+1 iterate of plane p)
        auto second = [&](int p, const V &yF) {
            const int qb = (p - 1) % 3;
            T yW = L[qb][s0 - 1], yE = L[qb][s0 + CPL];
            V yS = *reinterpret_cast<const V *>(&L[qb][s0 - gx * CPL]);
            V yN = *reinterpret_cast<const V *>(&L[qb][s0 + gx * CPL]);
            V yB = x1m;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                if (p - 1 == 0) yB[q] = ghost(a.bct[4], a.bcv[4], (T)x1c[q]);
                if (slo) yS[q] = ghost(a.bct[2], a.bcv[2], (T)x1c[q]);
                if (shi) yN[q] = ghost(a.bct[3], a.bcv[3], (T)x1c[q]);
            }
            if (wlo) yW = ghost(a.bct[0], a.bcv[0], (T)x1c[0]);
            if (whi) yE = ghost(a.bct[1], a.bcv[1], (T)x1c[CPL - 1]);
            V g0, xo;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const T wv = q == 0 ? yW : (T)x1c[q - 1], ev = q == CPL - 1 ? yE : (T)x1c[q + 1];
                const T kw = q == 0 ? pkW : (T)pkE[q - 1];
                T g = cell_residual<T, TR>(a, x1c[q], wv, ev, yS[q], yN[q], yB[q], yF[q], kw, pkE[q], pkS[q], pkN[q],
                                           pkB[q], pkF[q], pr[q]);
                g = precond<T>(g, pD[q]);
                g0[q] = momentum<T>(h1, pf0[q], h2, g);
                xo[q] = step_x<T>(x1c[q], a.dtau, g0[q]);
            }
            __builtin_nontemporal_store(g0, reinterpret_cast<GV *>(F0out + pc));
            *reinterpret_cast<GV *>(Xout + pe_idx) = xo;
        };
        for (int p = ps; p <= pe + 1; ++p) {
            if (p <= pe) {   // iteration n+1 on plane p
                V xp = ld(Xin + e + sE2);
                T xW = Xin[e - 1], xE = Xin[e + CPL];
                V xS = ld(Xin + e - sE1), xN = ld(Xin + e + sE1);
                T kW;
                V kE, kS, kN, kF;
                if constexpr (KV) {
                    T b0[CPL + 1], b1[CPL + 1];
                    vrow(v + sV2, b0);
                    vrow(v + sV2 + sV1, b1);
                    kW = av4(a0[0], a1[0], b0[0], b1[0]);
#pragma unroll
                    for (int q = 0; q < CPL; ++q) {
                        kE[q] = av4(a0[q + 1], a1[q + 1], b0[q + 1], b1[q + 1]);
                        kS[q] = av4(a0[q], a0[q + 1], b0[q], b0[q + 1]);
                        kN[q] = av4(a1[q], a1[q + 1], b1[q], b1[q + 1]);
                        kF[q] = av4(b0[q], b0[q + 1], b1[q], b1[q + 1]);
                    }
#pragma unroll
                    for (int q = 0; q <= CPL; ++q) {
                        a0[q] = b0[q];
                        a1[q] = b1[q];
                    }
                } else {
                    kW = kxg[fx];
                    kE = ld(kxg + fx + 1);
                    kS = ld(kyg + fy);
                    kN = ld(kyg + fy + nx);
                    kF = ld(kzg + c + sC2);
                }
                const V kB = kzB;
                const V rr = ld_nt(Rg + c);
                const V f0o = ld_nt(F0in + c);
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    if (p + 1 == nz) xp[q] = ghost(a.bct[5], a.bcv[5], (T)x0c[q]);
                    if (slo) xS[q] = ghost(a.bct[2], a.bcv[2], (T)x0c[q]);
                    if (shi) xN[q] = ghost(a.bct[3], a.bcv[3], (T)x0c[q]);
                }
                if (wlo) xW = ghost(a.bct[0], a.bcv[0], (T)x0c[0]);
                if (whi) xE = ghost(a.bct[1], a.bcv[1], (T)x0c[CPL - 1]);
                V D, f0, x1p;
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const T wv = q == 0 ? xW : (T)x0c[q - 1], ev = q == CPL - 1 ? xE : (T)x0c[q + 1];
                    const T kw = q == 0 ? kW : (T)kE[q - 1];
                    T f = cell_residual<T, TR>(a, x0c[q], wv, ev, xS[q], xN[q], x0m[q], xp[q], kw, kE[q], kS[q], kN[q],
                                               kB[q], kF[q], rr[q]);
                    D[q] = cell_diag(a, kw, (T)kE[q], (T)kS[q], (T)kN[q], (T)kB[q], (T)kF[q]);
                    f = precond<T>(f, D[q]);
                    f0[q] = momentum<T>(h1, f0o[q], h2, f);
                    x1p[q] = step_x<T>(x0c[q], a.dtau, f0[q]);
                }
                *reinterpret_cast<V *>(&L[p % 3][s0]) = x1p;
                x0m = x0c;
                x0c = xp;
                kzB = kF;
                __syncthreads();
                if (p - 1 >= k0 && inner) second(p, x1p);
                x1m = x1c;
                x1c = x1p;
                pkW = kW; pkE = kE; pkS = kS; pkN = kN; pkB = kB; pkF = kF; pr = rr; pD = D; pf0 = f0;
                pe_idx = e;
                pc = c;
                e += sE2;
                c += sC2;
                fx += sX2;
                fy += sY2;
                v += sV2;
            } else {         // past the top plane: iteration n+2 on plane nz-1 of the top chunk
                __syncthreads();
                if (p - 1 >= k0 && p - 1 < k1 && inner) {
                    V yF;
#pragma unroll
                    for (int q = 0; q < CPL; ++q) yF[q] = ghost(a.bct[5], a.bcv[5], (T)x1c[q]);
                    second(p, yF);
                }
            }
        }
        __syncthreads();   // L is rewritten by the block's next item
    }
}

// One problem's device-side table, work list and state for the band-pair
// kernels, in a workspace slot of the solve.  The table's Xin/Xout are the two
// iterate buffers and its F0 the first of two F0 buffers dF0 elements apart
// in one allocation; a launch picks the direction with `phase` and `fpar`.
template <typename T>
struct BandCtx {
    const IterArgs<T> *tab = nullptr;
    const uint2 *items = nullptr;
    const unsigned *n_items = nullptr;
    const OneActive *st = nullptr;
    unsigned n = 0;
    const T *x0 = nullptr, *f0 = nullptr;
    long long dF0 = 0;
};

// Upload the one-entry table `a` (band tiling set by set_pair_tiling), the
// work list (every band x plane-chunk of the problem) and an active state.
// Stream-ordered, then waited for: launches still reading the slot's previous
// contents finish first, and the host staging is free on return.
template <typename T>
inline int band_setup(const IterArgs<T> &a, long long dF0, int slot, BandCtx<T> *out) {
    const unsigned n = (unsigned)a.nband * (unsigned)a.nbz;
    const size_t o_st = (sizeof(IterArgs<T>) + 255) / 256 * 256, o_n = o_st + 64, o_items = o_st + 256;
    const size_t bytes = o_items + (size_t)n * sizeof(uint2);
    void *p = nullptr;
    if (int rc = workspace(slot, bytes, &p)) return rc;
    std::vector<char> h(bytes, 0);
    std::memcpy(h.data(), &a, sizeof(IterArgs<T>));
    const OneActive act{1};
    std::memcpy(h.data() + o_st, &act, sizeof act);
    std::memcpy(h.data() + o_n, &n, sizeof n);
    uint2 *it = reinterpret_cast<uint2 *>(h.data() + o_items);
    for (unsigned q = 0; q < n; ++q) it[q] = uint2{0u, q};
    hipStream_t st = ctx().stream;
    HIP_TRY(hipMemcpyAsync(p, h.data(), bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    char *d = (char *)p;
    out->tab = (const IterArgs<T> *)d;
    out->st = (const OneActive *)(d + o_st);
    out->n_items = (const unsigned *)(d + o_n);
    out->items = (const uint2 *)(d + o_items);
    out->n = n;
    out->x0 = a.Xin;
    out->f0 = a.F0;
    out->dF0 = dF0;
    return HT3D_OK;
}

// Two iterations on bands of whole rows (pair tiles 11-15): from Xin (one of
// the table's two iterate buffers) and F0in (one of its two F0 buffers) into
// the other two.
// kv: the kernels that form the coefficients from the table's vertex field
// (tiles with 2 or 4 cells per lane)
template <typename T>
inline int launch_band_pair(const BandCtx<T> &b, int tile, bool transient, const T *Xin, const T *F0in,
                            unsigned flags, bool kv = false) {
    const int phase = Xin == b.x0 ? 0 : 1, fpar = F0in == b.f0 ? 0 : 1;
    const unsigned fl = flags & FL_REVERSE;
    hipStream_t st = ctx().stream;
    if constexpr (sizeof(T) == 4) {   // pair_tile_ok keeps FP64 off the band tiles
#define BK_(TR, TY, KZ, CPL)                                                                                         \
    do {                                                                                                             \
        if constexpr (CPL == 1)                                                                                      \
            hipLaunchKernelGGL((batch_pair_kernel<T, TR, TY, KZ, OneActive>), dim3(b.n), dim3(64 * TY), 0, st,       \
                               b.tab, b.items, b.n_items, b.st, phase, fpar, b.dF0, fl);                            \
        else if (kv)                                                                                                 \
            hipLaunchKernelGGL((batch_pair_kernel_v<T, TR, TY, KZ, CPL, OneActive, true>), dim3(b.n), dim3(64 * TY), \
                               0, st, b.tab, b.items, b.n_items, b.st, phase, fpar, b.dF0, fl);                     \
        else                                                                                                         \
            hipLaunchKernelGGL((batch_pair_kernel_v<T, TR, TY, KZ, CPL, OneActive>), dim3(b.n), dim3(64 * TY), 0,    \
                               st, b.tab, b.items, b.n_items, b.st, phase, fpar, b.dF0, fl);                        \
    } while (0)
#define BT_(TY, KZ, CPL) do { if (transient) BK_(true, TY, KZ, CPL); else BK_(false, TY, KZ, CPL); } while (0)
        switch (tile) {
        case 11: BT_(8, 16, 4); break;
        case 12: BT_(16, 16, 4); break;
        case 13: BT_(16, 8, 4); break;
        case 14: BT_(16, 16, 2); break;
        case 15: BT_(16, 16, 1); break;
        default: return arg_fail("launch_band_pair: not a band tile");
        }
#undef BT_
#undef BK_
    } else {
        return arg_fail("launch_band_pair: band tiles are for the FP32 sweeps");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch_pair_kernel (band pairs)");
    return HT3D_OK;
}

// ---------------------------------------------------------------- autotuner
// HT3D_TUNE_VERBOSE=1: the tuner prints each candidate's time to stderr
inline bool tune_verbose() {
    static const bool v = [] {
        const char *e = std::getenv("HT3D_TUNE_VERBOSE");
        return e && e[0] == '1';
    }();
    return v;
}

struct TuneKey {
    int nx, ny, nz, transient, prec;
    int kv = 0;   // the solve has the vertex field (variant bit 8 can run)
    bool operator==(const TuneKey &o) const {
        return nx == o.nx && ny == o.ny && nz == o.nz && transient == o.transient && prec == o.prec && kv == o.kv;
    }
};
struct TuneEntry {
    TuneKey key;
    IterCfg cfg;
    double us;
    PairCfg pair;       // fastest pair configuration (timed per iteration)
    double pair_us;
    double check_us = -1.0;   // extra cost of one convergence check (check launch, finish, read-back)
    double last_iters = -1.0; // iterations of the last solve with this key
    double anchor_us = -1.0;  // defect-correction anchor: FP64 residual, finish, read-back
    int res_tile = -1;        // defect-correction residual tile (ht3d_dc.hip)
};
extern std::vector<TuneEntry> g_tune;
extern bool g_autotune;
extern int g_force_tile, g_force_variant;
extern int g_force_pair_tile;      // >= 0: pin this pair tile (no pairs where it cannot run)
extern int g_force_pair_variant;
extern int g_pairs;     // 0: never, 1: when the tuner finds them faster, 2: always (best pair tile)

inline TuneEntry *tune_find(const TuneKey &k) {
    for (auto &t : g_tune)
        if (t.key == k) return &t;
    return nullptr;
}

// Check-interval tuner (a requested interval < 0).  A solve of I iterations
// checked every L costs I t_it + (I/L) t_chk + (L/2) t_it on average (it stops at
// the first check after convergence), which is least at L* = sqrt(2 I t_chk / t_it).
// L is taken among the divisors of the lambda interval (so checks stay aligned
// with the lambda refreshes), from the previous solve's I; 50 before any history.
inline int tuned_check_interval(double iters_prev, double t_it, double t_chk, int nitout) {
    static const int cand[] = {10, 20, 25, 50, 100, 200, 250, 500, 1000};
    int fallback = nitout;
    for (int c : cand)
        if (c <= nitout && nitout % c == 0 && c <= 50) fallback = c;
    if (!(iters_prev > 0) || !(t_it > 0) || !(t_chk > 0)) return fallback;
    const double L = sqrt(2.0 * iters_prev * t_chk / t_it);
    int best = -1;
    double bd = 1e300;
    for (int c : cand) {
        if (c > nitout || nitout % c) continue;
        const double d = fabs(log((double)c / L));
        if (d < bd) { bd = d; best = c; }
    }
    return best > 0 ? best : nitout;
}

// Times every single-iteration tile and variant, and every pair tile and
// variant, on the solve's own read-only inputs, writing only scratch outputs
// (Xout, F0, F0b), and keeps the fastest of each.  Dst (when can_store_d) must
// hold the diagonal; F0b (when not null) is a second F0 buffer for the pairs.
template <typename T>
inline int autotune(IterArgs<T> a, bool transient, bool can_store_d, int prec, T *F0b, TuneEntry *out) {
    const TuneKey key{a.nx, a.ny, a.nz, transient ? 1 : 0, prec, a.kv ? 1 : 0};
    if (const TuneEntry *t = tune_find(key)) {
        *out = *t;
        return HT3D_OK;
    }
    Context &cx = ctx();
    a.flags = 0;
    TuneEntry te;
    te.key = key;
    te.us = 1e300;
    te.pair_us = 1e300;
    TimerEvents ev;
    if (int rc_ = ev.create()) return rc_;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    // The timed launches ping-pong the iterate between two buffers, as the
    // solve does (what the MALL holds at the next launch depends on it): a
    // scratch copy of the input and the solve's second buffer, so the solve's
    // own X is never written.
    const size_t ne = (size_t)(a.nx + 2) * (a.ny + 2) * (a.nz + 2);
    T *xs[2] = {nullptr, a.Xout};
    {
        void *p = nullptr;
        if (int rc = workspace(WS_TUNE_X, ne * sizeof(T), &p)) return rc;
        xs[0] = (T *)p;
        HIP_TRY(hipMemcpyAsync(xs[0], a.Xin, ne * sizeof(T), hipMemcpyDeviceToDevice, cx.stream));
    }
    int pp = 0;
    auto flip = [&](IterArgs<T> &b) {
        b.Xin = xs[pp];
        b.Xout = xs[pp ^ 1];
        pp ^= 1;
    };
    int reps = 8;
    auto timed = [&](auto &&launch, double *us) -> int {
        int rc = launch();
        if (rc) return rc;
        HIP_TRY(hipEventRecord(e0, cx.stream));
        for (int r = 0; r < reps; ++r)
            if ((rc = launch())) return rc;
        HIP_TRY(hipEventRecord(e1, cx.stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        *us = 1e3 * ms / reps;
        return HT3D_OK;
    };
    // stage 1: every (tile, variant) once; stage 2: the fastest few again, in
    // interleaved rounds with more launches each, keeping each one's best time
    // (a single short timing is noisy at the 1 % level that separates the
    // leading tiles)
    struct Cand {
        IterCfg cfg;
        double us;
    };
    std::vector<Cand> cands;
    auto time_cfg = [&](const IterCfg &cfg, double *us) -> int {
        set_tiling(a, cfg.tile);
        int rc = timed([&] {
            if (g_alternate) a.flags ^= FL_REVERSE;
            flip(a);
            return launch_iter(a, transient, cfg);
        }, us);
        a.flags = 0;
        return rc;
    };
    // Grids under 2^18 cells (64^3-class and smaller) are launch-bound and
    // solved in a few milliseconds, about what a full tuning pass takes: there
    // the tuner tries the shapes such grids pick when the whole table is timed
    // (4x2 and 8x2 planes, 4x4, 1x8) with the cached-coefficient variants
    // (4-7), and two pair tiles (for the forced-pairs mode; timing never
    // picks them there).
    const bool small = (size_t)a.nx * a.ny * a.nz < ((size_t)1 << 18);
    auto short_tile = [](int tile) { return tile == 9 || tile == 10 || tile == 2 || tile == 13; };
    for (int var = 0; var < 8; ++var) {
        if ((var & 1) && !can_store_d) continue;
        if (small && var < 4) continue;
        for (int tile = 0; tile < kNumTiles; ++tile) {
            if (tile_for_grid(tile, a.nx) != tile) continue;
            if (small && !short_tile(tile)) continue;
            Cand cd{IterCfg{tile, var}, 0.0};
            if (int rc = time_cfg(cd.cfg, &cd.us)) return rc;
            cands.push_back(cd);
        }
    }
    if (a.kv && !small)   // the vertex-coefficient kernels (FP64: variants 8-11, FP32: 8, 10, 12, 14)
        for (int var = 8; var <= 15; ++var) {
            if (sizeof(T) == 8 ? (var & 4) || ((var & 1) && !can_store_d) : (var & 1)) continue;
            for (int tile = 0; tile < kNumTiles; ++tile) {
                if (!kv_tile_ok<T>(tile, a.nx)) continue;
                Cand cd{IterCfg{tile, var}, 0.0};
                if (int rc = time_cfg(cd.cfg, &cd.us)) return rc;
                cands.push_back(cd);
            }
        }
    std::sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.us < y.us; });
    const size_t nfin = cands.size() < 4 ? cands.size() : 4;
    std::vector<double> fin(nfin, 1e300);
    reps = 16;
    for (int round = 0; round < 3; ++round)
        for (size_t q = 0; q < nfin; ++q) {
            double us = 0;
            if (int rc = time_cfg(cands[q].cfg, &us)) return rc;
            fin[q] = us < fin[q] ? us : fin[q];
        }
    reps = 8;
    for (size_t q = 0; q < nfin; ++q)
        if (fin[q] < te.us) { te.us = fin[q]; te.cfg = cands[q].cfg; }
    if (tune_verbose())
        fprintf(stderr, "[ht3d tune] %dx%dx%d %s prec %d: single tile %d variant %d %.2f us\n", a.nx, a.ny, a.nz,
                transient ? "heat" : "darcy", prec, te.cfg.tile, te.cfg.variant, te.us);
    if (a.partials) {   // cost of a check: check launch + finish kernel + read-back, less a plain launch
        void *res = nullptr;
        int rc = workspace(WS_RESULT, 64 * sizeof(double), &res);
        if (rc) return rc;
        IterArgs<T> c = a;
        c.flags = FL_CHECK;
        set_tiling(c, te.cfg.tile);
        const unsigned nb = (unsigned)c.nbx * c.nby * c.nbz;
        double r4[4];
        double best = 1e300;
        for (int r = 0; r < 4; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = launch_iter(c, transient, te.cfg)) || (rc = finish_partials(c.partials, nb, 0x2u, (double *)res)) ||
                (rc = read_result4((const double *)res, r4)))
                return rc;
            const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
            best = us < best ? us : best;
        }
        te.check_us = best - te.us > 1.0 ? best - te.us : 1.0;
    }
    if (F0b && a.ny >= 3) {
        for (int var = 0; var < 4; ++var) {
            if ((var & 1) && !can_store_d) continue;
            if (small && var < 2) continue;
            for (int tile = 0; tile < kNumPairTiles; ++tile) {
                if (kPairTiles[tile].band || !pair_tile_ok<T>(tile, a.nx)) continue;   // bands: timed below
                if (small && tile != 0 && tile != 5) continue;
                PairCfg pc{tile, var};
                IterArgs<T> b = a;
                b.F0 = F0b;
                set_pair_tiling(b, tile);
                double us = 0;
                int rc = timed([&] {
                    if (g_alternate) b.flags ^= FL_REVERSE;
                    flip(b);
                    return launch_pair(b, (const T *)a.F0, transient, pc);
                }, &us);
                if (rc) return rc;
                us *= 0.5;   // per iteration
                if (tune_verbose())
                    fprintf(stderr, "[ht3d tune] %dx%dx%d %s prec %d: pair tile %d variant %d %.2f us per iteration\n",
                            a.nx, a.ny, a.nz, transient ? "heat" : "darcy", prec, tile, var, us);
                if (us < te.pair_us) { te.pair_us = us; te.pair = pc; }
            }
        }
        // band tiles (FP32): F0b must follow a.F0 in the same allocation; with
        // the vertex field also their vertex-coefficient form (variant 8)
        if (sizeof(T) == 4 && F0b > a.F0 && !small) {
            for (int tk = 0; tk < 2 * kNumPairTiles; ++tk) {
                const int tile = tk >> 1, kvv = tk & 1;
                if (!kPairTiles[tile].band || !pair_tile_ok<T>(tile, a.nx)) continue;
                if (kvv && (!a.kv || kPairTiles[tile].cpl < 2)) continue;
                IterArgs<T> b = a;
                b.Xin = xs[0];
                b.Xout = xs[1];
                set_pair_tiling(b, tile);
                BandCtx<T> bc;
                if (int rc = band_setup(b, (long long)(F0b - a.F0), WS_BAND, &bc)) return rc;
                int fp = 0;
                unsigned fl = 0;
                double us = 0;
                int rc = timed([&] {
                    if (g_alternate) fl ^= FL_REVERSE;
                    const T *xin = xs[pp], *f0in = fp ? F0b : a.F0;
                    pp ^= 1;
                    fp ^= 1;
                    return launch_band_pair(bc, tile, transient, xin, f0in, fl, kvv != 0);
                }, &us);
                if (rc) return rc;
                us *= 0.5;
                if (tune_verbose())
                    fprintf(stderr, "[ht3d tune] %dx%dx%d %s prec %d: band pair tile %d%s %.2f us per iteration\n",
                            a.nx, a.ny, a.nz, transient ? "heat" : "darcy", prec, tile, kvv ? " (vertex)" : "", us);
                if (us < te.pair_us) { te.pair_us = us; te.pair = PairCfg{tile, kvv ? 8 : 0}; }
            }
        }
    }
    g_tune.push_back(te);
    *out = te;
    return HT3D_OK;
}

IterCfg &last_cfg();
PairCfg &last_pair();

// The single-iteration configuration (its tiling set on `a`) and the pair
// configuration to use (tile < 0: none).
template <typename T>
inline int choose_cfg(IterArgs<T> &a, bool transient, bool can_store_d, int prec, T *F0b, IterCfg *cfg,
                      PairCfg *pair) {
    IterCfg c;
    PairCfg pc;
    if (g_force_tile >= 0) {
        c.tile = tile_for_grid(g_force_tile, a.nx);
        c.variant = g_force_variant & (can_store_d ? 15 : 14);
        kv_normalize(a, c);
    } else if (g_autotune) {
        TuneEntry te;
        int rc = autotune(a, transient, can_store_d, prec, F0b, &te);
        if (rc) return rc;
        c = te.cfg;
        if (!can_store_d) c.variant &= ~1;
        kv_normalize(a, c);
        if (te.pair.tile >= 0 && (g_pairs == 2 || (g_pairs == 1 && te.pair_us < te.us))) pc = te.pair;
        if (!can_store_d) pc.variant &= ~1;
        if (!a.kv) pc.variant &= ~8;
    }
    if (g_force_pair_tile >= 0) {   // pinned: that tile, or no pairs where it cannot run
        const bool ok = pair_tile_ok<T>(g_force_pair_tile, a.nx) && a.ny >= 3;
        pc.tile = ok ? g_force_pair_tile : -1;
        pc.variant = kPairTiles[g_force_pair_tile].band
                         ? ((g_force_pair_variant & 8) && a.kv && kPairTiles[g_force_pair_tile].cpl >= 2 ? 8 : 0)
                         : g_force_pair_variant & (can_store_d ? 3 : 2);
    }
    if (!F0b || g_pairs == 0) pc.tile = -1;
    set_tiling(a, c.tile);
    *cfg = c;
    *pair = pc;
    return HT3D_OK;
}

// ------------------------------------------------------------ defect correction
// Inputs of the FP64 residual of the dimensionless operators, with on-the-fly
// ghost values: r = R [+ x/dt] + div q(x); r is also stored in FP32 as the
// right-hand side of the correction equation.
struct DcArgs {
    const double *__restrict__ X;   // E
    const double *__restrict__ R;   // c
    const double *__restrict__ kx;
    const double *__restrict__ ky;
    const double *__restrict__ kz;
    const double *__restrict__ kv;  // V or null: the vertex field whose face averages kx, ky, kz are
    float *__restrict__ r32;        // c: FP32 right-hand side of the correction equation
    double *__restrict__ partials;
    int nx, ny, nz;
    double _dx, _dy, _dz, _dt;
    int bct[6];
    double bcv[6];
};

// The FP64 defect-correction residual of the lane's CPL cells from (i, j)
// (tile_cell or patch_cell) marched over the planes [bz*KZ, +KZ): the load
// pattern of iter_tile_v
// (X(k-1), X(k), X(k+1) and kz(k), kz(k+1) in registers, CPL-element vector
// loads, a face between two cells of a lane loaded once); per cell r = R
// [+ x/dt] + div q(x) in the order of ComputeFlux! and
// ResidualDiffusion(Transient)! (kernels_HT3D.jl:112-114,131,
// HydroThermal3D_Alban_v6.jl:87).  Writes r32, adds |r| to the double-double
// sum (s_hi, s_lo) and to the max mx.  It reads five words and writes half a word per
// cell, the fused iteration's stencil without the update; a full chunk of KZ
// planes is unrolled so later planes' loads issue ahead.  With KV the face
// coefficients are formed from the vertex field a.kv as in iter_tile_kv (the
// same values bit for bit): three words read per cell instead of five.
// Needs nx % CPL == 0 (and CPL >= 2 with KV).
template <bool TR, int TY, int KZ, int CPL, bool KV = false>
__device__ __forceinline__ void dc_residual_tile(const DcArgs &a, int i, int j, int bz, double &s_hi, double &s_lo,
                                                 double &mx) {
    typedef typename CellVec<double, CPL>::type V;
    typedef typename CellVec<float, CPL>::type FV;
    const int k0 = bz * KZ, k1 = min(k0 + KZ, a.nz);
    const uint32_t nx = a.nx, ny = a.ny;
    if ((uint32_t)i >= nx || (uint32_t)j >= ny) return;
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2), sC2 = nx * ny, sX2 = (nx + 1) * ny, sY2 = nx * (ny + 1);
    const bool wlo = i == 0, whi = (uint32_t)i + CPL == nx, slo = j == 0, shi = (uint32_t)j == ny - 1;
    uint32_t e = (uint32_t)(i + 1) + (uint32_t)(j + 1) * sE1 + (uint32_t)(k0 + 1) * sE2;
    uint32_t c = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sC2;
    uint32_t fx = (uint32_t)i + (uint32_t)j * (nx + 1) + (uint32_t)k0 * sX2;
    uint32_t fy = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sY2;
    const uint32_t sV1 = nx + 1, sV2 = sV1 * (ny + 1);
    uint32_t v = (uint32_t)i + (uint32_t)j * sV1 + (uint32_t)k0 * sV2;
    double a0[CPL + 1], a1[CPL + 1];   // KV: vertex values of plane k, rows j and j + 1
    auto vrow = [&](uint32_t o, double *out) {
        const V w = ldv<double, CPL>(a.kv + o);
#pragma unroll
        for (int q = 0; q < CPL; ++q) out[q] = w[q];
        out[CPL] = a.kv[o + CPL];
    };
    V xc = ldv<double, CPL>(a.X + e), xB = ldv<double, CPL>(a.X + e - sE2), kB;
    if constexpr (KV) {
        vrow(v, a0);
        vrow(v + sV1, a1);
#pragma unroll
        for (int q = 0; q < CPL; ++q) kB[q] = av4(a0[q], a0[q + 1], a1[q], a1[q + 1]);
    } else {
        kB = ldv<double, CPL>(a.kz + c);
    }
    if (k0 == 0) {
#pragma unroll
        for (int q = 0; q < CPL; ++q) xB[q] = ghost(a.bct[4], a.bcv[4], (double)xc[q]);
    }
    auto plane = [&](int k) {
        V xF = ldv<double, CPL>(a.X + e + sE2);
        double xW = a.X[e - 1], xE = a.X[e + CPL];
        V xS = ldv<double, CPL>(a.X + e - sE1), xN = ldv<double, CPL>(a.X + e + sE1);
        double kW;
        V kE, kS, kN, kF;
        if constexpr (KV) {
            double b0[CPL + 1], b1[CPL + 1];
            vrow(v + sV2, b0);
            vrow(v + sV2 + sV1, b1);
            kW = av4(a0[0], a1[0], b0[0], b1[0]);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                kE[q] = av4(a0[q + 1], a1[q + 1], b0[q + 1], b1[q + 1]);
                kS[q] = av4(a0[q], a0[q + 1], b0[q], b0[q + 1]);
                kN[q] = av4(a1[q], a1[q + 1], b1[q], b1[q + 1]);
                kF[q] = av4(b0[q], b0[q + 1], b1[q], b1[q + 1]);
            }
#pragma unroll
            for (int q = 0; q <= CPL; ++q) {
                a0[q] = b0[q];
                a1[q] = b1[q];
            }
        } else {
            kW = a.kx[fx];
            kE = ldv<double, CPL>(a.kx + fx + 1);
            kS = ldv<double, CPL>(a.ky + fy);
            kN = ldv<double, CPL>(a.ky + fy + nx);
            kF = ldv<double, CPL>(a.kz + c + sC2);
        }
        const V R = ldv_nt<double, CPL>(a.R + c);
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            if (k + 1 == a.nz) xF[q] = ghost(a.bct[5], a.bcv[5], (double)xc[q]);
            if (slo) xS[q] = ghost(a.bct[2], a.bcv[2], (double)xc[q]);
            if (shi) xN[q] = ghost(a.bct[3], a.bcv[3], (double)xc[q]);
        }
        if (wlo) xW = ghost(a.bct[0], a.bcv[0], (double)xc[0]);
        if (whi) xE = ghost(a.bct[1], a.bcv[1], (double)xc[CPL - 1]);
        FV r;
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            const double x0 = xc[q], w = q == 0 ? xW : (double)xc[q - 1], ea = q == CPL - 1 ? xE : (double)xc[q + 1];
            const double kw = q == 0 ? kW : (double)kE[q - 1];
            const double qW = -kw * (a._dx * (x0 - w)), qE = -kE[q] * (a._dx * (ea - x0));
            const double qS = -kS[q] * (a._dy * (x0 - xS[q])), qN = -kN[q] * (a._dy * (xN[q] - x0));
            const double qB = -kB[q] * (a._dz * (x0 - xB[q])), qF = -kF[q] * (a._dz * (xF[q] - x0));
            double f;
            if (TR) f = R[q] + a._dt * x0 + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
            else f = R[q] + a._dx * (qE - qW) + a._dy * (qN - qS) + a._dz * (qF - qB);
            r[q] = (float)f;
            const double af = fabs(f);
            dd_add(s_hi, s_lo, af);
            mx = (af > mx || af != af) ? af : mx;
        }
        stv_nt<float, CPL>(r, a.r32 + c);
        xB = xc;
        xc = xF;
        kB = kF;
        e += sE2;
        c += sC2;
        fx += sX2;
        fy += sY2;
        v += sV2;
    };
    if (k1 - k0 == KZ) {
#pragma unroll
        for (int q = 0; q < KZ; ++q) plane(k0 + q);
    } else {
        for (int k = k0; k < k1; ++k) plane(k);
    }
}

}  // namespace ht3d
