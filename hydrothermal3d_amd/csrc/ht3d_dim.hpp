// ht3d_dim.hpp -- device closures of the dimensional (v18) model and the
// ghost-fill rules of its immersed-surface boundary conditions, shared by the
// per-kernel entries (ht3d_dim.hip) and the fused v18 DYREL solves
// (ht3d_v18.hip).  Evaluation order: see ht3d_dim.hip.
#pragma once

#include <cmath>

#include "ht3d_internal.hpp"

namespace ht3d {
namespace dim {

// ---------------------------------------------------------------- closures
__device__ __forceinline__ double pow4(double x) {
    const double x2 = x * x;
    const double e2 = fma(x, x, -x2);
    const double x4 = x2 * x2;
    const double e4 = fma(x2, x2, -x4);
    const double lo = e4 + (x2 * 2.0) * e2;
    return x4 + lo;
}

__device__ __forceinline__ double kTs(double T) { return 3.1138 - 0.0023 * T; }
__device__ __forceinline__ double kTf(double T) { return -0.069 + 0.0012 * T; }
__device__ __forceinline__ double Cf(double T) { return 1000.0 + 7.5 * T; }
__device__ __forceinline__ double Cs(double T) { return 0.5915 * T + 636.14; }
__device__ __forceinline__ double rhos(double T) { return 2800.0 * (1.0 - (0.000024 * (T - 293.0))); }

// fluid density (kg/m^3) of temperature in C and pressure in Pa: fifteen terms
__device__ __forceinline__ double rhof_C(double Tc, double p) {
    const double p2 = p * p, p3 = p2 * p, T2 = Tc * Tc, T3 = T2 * Tc;
    double r = 1006.0;
    r = r + (7.424e-7 * p);
    r = r + (-0.3922 * Tc);
    r = r + (-4.441e-15 * p2);
    r = r + ((4.547e-9 * p) * Tc);
    r = r + (-0.003774 * T2);
    r = r + (1.451e-23 * p3);
    r = r + ((-1.793e-17 * p2) * Tc);
    r = r + ((7.485e-12 * p) * T2);
    r = r + (2.955e-6 * T3);
    r = r + (-1.463e-32 * pow4(p));
    r = r + ((1.361e-26 * p3) * Tc);
    r = r + ((4.018e-21 * p2) * T2);
    r = r + ((-7.372e-15 * p) * T3);
    r = r + (5.698e-11 * pow4(Tc));
    return r;
}

// its pressure derivative (the closure as printed by the reference's notebook)
__device__ __forceinline__ double drhodP_C(double Tc, double p) {
    const double p2 = p * p, p3 = p2 * p, T2 = Tc * Tc, T3 = T2 * Tc;
    double d = -7.372e-15 * T3;
    d = d + ((8.036e-21 * T2) * p);
    d = d + (7.485e-12 * T2);
    d = d + ((4.083e-26 * Tc) * p2);
    d = d - ((3.586e-17 * Tc) * p);
    d = d + (4.547e-9 * Tc);
    d = d - (5.852e-32 * p3);
    d = d + (4.353e-23 * p2);
    d = d - (8.882e-15 * p);
    d = d + 7.424e-7;
    return d;
}

__device__ __forceinline__ double muf(double T) { return 2.414e-5 * pow(10.0, 247.8 / (T - 140.0)); }
__device__ __forceinline__ double kF(double y, double delta) { return 5e-16 * exp(y / delta); }
__device__ __forceinline__ double phi_of(double phase, double phi0) { return phase == 2.0 ? 3.0 * phi0 : phi0; }

// Mean of the eight extended centroids around vertex (i,j,k) (0-based vertex
// index; the centroids are E indices i..i+1, j..j+1, k..k+1), the upper z plane
// summed first as in kernels_HT3D_dimensional.jl:42-43.
__device__ __forceinline__ double vertex_mean(const double *A, const Grid &g, int i, int j, int k) {
    const double up = ((A[g.e(i + 1, j + 1, k + 1)] + A[g.e(i + 1, j, k + 1)]) + A[g.e(i, j + 1, k + 1)]) + A[g.e(i, j, k + 1)];
    const double lo = ((A[g.e(i + 1, j + 1, k)] + A[g.e(i + 1, j, k)]) + A[g.e(i, j + 1, k)]) + A[g.e(i, j, k)];
    return 0.125 * up + 0.125 * lo;
}

// ParallelStencil @av: x, then y, then z neighbours, times 0.125.  idx(i,j,k)
// gives the linear index in the source array.
template <typename Idx>
__device__ __forceinline__ double cube_mean(const double *A, Idx idx, int i, int j, int k) {
    return (A[idx(i, j, k)] + A[idx(i + 1, j, k)] + A[idx(i, j + 1, k)] + A[idx(i + 1, j + 1, k)] +
            A[idx(i, j, k + 1)] + A[idx(i + 1, j, k + 1)] + A[idx(i, j + 1, k + 1)] + A[idx(i + 1, j + 1, k + 1)]) *
           0.125;
}

// ---------------------------------------------------------------- ghost fills
// Rules of SetTemperatureBCs! / SetPressureBCs!, in the priority that the
// order of the kernel body gives them (later statements override earlier).
enum BcRule { R_KEEP, R_AIR, R_XLO, R_XHI, R_BOT, R_IMM, R_ZLO, R_ZHI };

struct BcArgs {
    const double *A;      // values before the fill (read only)
    const double *ky, *wN, *ph;
    double top, qyS_dy;
    Grid g;
    int temperature;
};

// (i,j,k) are 0-based E indices.
__device__ __forceinline__ int bc_rule(const BcArgs &a, int i, int j, int k) {
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
        if (k == 0) return R_ZLO;
        if (k == nz + 1) return R_ZHI;
    }
    const double ph = a.ph[a.g.e(i, j, k)];
    if (ph == -1.0 && j > 0) return R_IMM;
    const int kx_hi = a.temperature ? nz + 1 : nz;   // the temperature x fill runs to the top z ghost
    if (j >= 1 && j <= ny && k >= 1 && k <= kx_hi) {
        if (i == 0) return R_XLO;
        if (i == nx + 1) return R_XHI;
    }
    if (j == 0 && i >= 1 && i <= nx && k >= 1 && k <= nz) return R_BOT;
    if (ph == 1.0) return R_AIR;
    return R_KEEP;
}

// Moves (i,j,k) to the source cell of rule r (a copy, the immersed rule's cell
// below, or the basal rule's cell above).
__device__ __forceinline__ void bc_source(const BcArgs &a, int r, int &i, int &j, int &k) {
    switch (r) {
        case R_ZLO: k = 1; break;
        case R_ZHI: k = a.g.nz; break;
        case R_IMM: j = j - 1; break;
        case R_XLO: i = 1; break;
        case R_XHI: i = a.g.nx; break;
        default: j = 1; break;   // R_BOT
    }
}

// Final value of cell (i,j,k): each rule reads its source cell's final value
// (the order of dependencies), so no cell is read while another thread of the
// launch rewrites it.  The chain of rules is followed to a cell with no rule
// (its stored value) or to air (the surface value); chains longer than five
// rules (stacked immersed cells) end at the stored value, as in the CPU oracle.
// The transforms along the chain (immersed Dirichlet, basal flux) are then
// applied from the innermost out, each re-found by walking the chain again:
// the arithmetic of a recursive evaluation in a few registers.
__device__ __forceinline__ double bc_value(const BcArgs &a, int i, int j, int k) {
    int ci = i, cj = j, ck = k, d = 0;
    double v;
#pragma unroll 1
    for (;;) {
        if (d > 4) { v = a.A[a.g.e(ci, cj, ck)]; break; }
        const int r = bc_rule(a, ci, cj, ck);
        if (r == R_AIR) { v = a.top; break; }
        if (r == R_KEEP) { v = a.A[a.g.e(ci, cj, ck)]; break; }
        bc_source(a, r, ci, cj, ck);
        ++d;
    }
#pragma unroll 1
    for (int L = d - 1; L >= 0; --L) {
        ci = i; cj = j; ck = k;
#pragma unroll 1
        for (int q = 0; q < L; ++q) bc_source(a, bc_rule(a, ci, cj, ck), ci, cj, ck);
        const int r = bc_rule(a, ci, cj, ck);
        if (r == R_IMM) {
            const double w = a.wN[(size_t)ck * (a.g.nx + 2) + ci];
            v = (a.top - (1.0 - w) * v) / w;
        } else if (r == R_BOT && a.temperature) {
            const double kS = a.ky[a.g.fy(ci - 1, 0, ck - 1)];
            v = a.qyS_dy / kS + v;
        }
    }
    return v;
}

}  // namespace dim
}  // namespace ht3d
