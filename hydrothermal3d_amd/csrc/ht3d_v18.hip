// ht3d_v18.hip -- fused DYREL solves of the dimensional (v18) model: the heat
// loop of Dimensional/HydroThermal3D_v18.jl:432-478 and the Darcy loop of
// :489-537, one kernel launch per pseudo-iteration.
//
// One pseudo-iteration of the reference is four to five launches:
//   heat:  SetTemperatureBCs!, ComputeFlux!, ResidualTemperatureLinearised!, DYRELUpdate!
//   Darcy: ϕρf, SetPressureBCs!, ComputeDarcyFlux!, ResidualFluidPressure!, DYRELUpdate!
// (kernels_HT3D_dimensional.jl:162-326).  Here one kernel does all of it per
// cell, in the reference's order of operations, so the result equals the
// unfused sequence bit for bit:
//   * the ghost/air/immersed values the BC kernel would write are computed from
//     the neighbours when needed (the same rule evaluator as the ghost-fill
//     entry), so the unknown's ghost layer is only written at the end;
//   * the ghost rules are applied once to the unknown before the first
//     iteration (as the reference's first BC launch does), so air cells hold
//     the surface value; the update leaves them there (F = 0 gives F0 = 0), so
//     an air neighbour is a plain load;
//   * a per-cell class byte, built once per solve from the phases, selects the
//     path: F = 0 (air and immersed cells: only the update, no coefficient or
//     neighbour loads), the fast path (plain neighbour loads; walls by select:
//     Neumann copies, basal heat flux), or the exact rule evaluation (cells
//     next to an immersed-surface cell, whose value follows the cell below it
//     every iteration);
//   * heat: the Gershgorin diagonal is recomputed from the face coefficients
//     and rho*C already loaded (it is what GershgorinPoisson! stores in PC);
//     Darcy: phi*rho_f of the cell is evaluated from the EOS each iteration
//     (the reference's ϕρf launch), g*<rho_f> on the y faces is formed once
//     per solve, and the stored diagonal is read;
//   * the once-read streams (class byte, old T or T, rho*C or phi*rho_f_old, D)
//     are loaded and the unknown stored non-temporally, which leaves the L2 to
//     the planes of the unknown that neighbouring waves re-read;
//   * check iterations accumulate |F*PC|, max|F*PC| and the Rayleigh-quotient
//     sums in double-double block partials, like the dimensionless solve.
// Memory left as the reference leaves it: the unknown with ghosts, air and
// immersed cells filled from the state before the last update; PC and G;
// for the Darcy solve phi*rho_f of the last iteration in the rhoC buffer
// (at immersed cells it is evaluated from the final immersed value).
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "ht3d_dim.hpp"
#include "ht3d_internal.hpp"
#include "ht3d_iter.hpp"

namespace ht3d {

int copy_ghosts(double *dst, const double *src, int nx, int ny, int nz);   // ht3d_dyrel.hip

namespace v18 {

enum : uint8_t { C_ZERO = 1, C_SRC = 2, C_SLOW = 4, C_FAULT = 8 };

struct Args {
    const double *__restrict__ Xin;
    double *__restrict__ Xout;
    const double *__restrict__ kx;
    const double *__restrict__ ky;
    const double *__restrict__ kz;
    const double *__restrict__ kv;     // vertex field whose face averages kx, ky, kz are (KV kernels)
    const uint8_t *__restrict__ cls;
    double *__restrict__ F0;
    double *__restrict__ Fst;
    double *__restrict__ partials;
    const double *__restrict__ aux;    // heat: T of the previous step (c); Darcy: temperature (E)
    const double *__restrict__ aux2;   // heat: rho*C (c); Darcy: phi*rho_f of the previous step (c)
    const double *__restrict__ Dst;    // Darcy: Gershgorin diagonal
    const double *__restrict__ gy;     // Darcy: g * @av_xza(rho_f) on the y faces
    dim::BcArgs bc;                    // ghost rules; bc.A is the launch's Xin
    int nx, ny, nz, nbx, nby, nbz;
    double _dx, _dy, _dz, _dt, dx, dy, dz, dt;
    double transient, tr_dt, Qt, qyS_dy;
    double phi_i, s_rho, s_T, s_sigma;
    double h1, h2, dtau;
    unsigned flags;
};

// Class byte of each cell (phases do not change during a solve), and the
// number of F = 0 cells in *nzero.  Air neighbours and ghosts need no rule
// evaluation once the rules have been applied to the unknown; immersed-surface
// cells do (their value follows the cell below every iteration).
__global__ __launch_bounds__(256) void class_kernel(const double *__restrict__ ph, int nx, int ny, int nz,
                                                    uint8_t *__restrict__ cls, unsigned long long *__restrict__ nzero) {
    __shared__ unsigned block_zeros;
    if (threadIdx.x == 0) block_zeros = 0;
    __syncthreads();
    const size_t N = (size_t)nx * ny * nz;
    const Grid g{nx, ny, nz};
    unsigned zeros = 0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((size_t)nx * ny));
        const int ie = i + 1, je = j + 1, ke = k + 1;
        const double p = ph[g.e(ie, je, ke)];
        uint8_t c = 0;
        if (fabs(p) == 1.0) { c |= C_ZERO; ++zeros; }
        if (p == 3.0) c |= C_SRC;
        if (p == 2.0) c |= C_FAULT;
        auto imm = [&](int a, int b, int d) { return ph[g.e(a, b, d)] == -1.0; };
        // every neighbour but the bottom ghost (always the basal rule) and the
        // z ghosts (always copies of the cell); an immersed x ghost overrides
        // the copy, an immersed top ghost its stored value
        const bool slow = imm(ie - 1, je, ke) || imm(ie + 1, je, ke) || (j > 0 && imm(ie, je - 1, ke)) ||
                          imm(ie, je + 1, ke) || (k > 0 && imm(ie, je, ke - 1)) || (k < nz - 1 && imm(ie, je, ke + 1));
        if (slow) c |= C_SLOW;
        cls[t] = c;
    }
    if (zeros) atomicAdd(&block_zeros, zeros);
    __syncthreads();
    if (threadIdx.x == 0 && block_zeros) atomicAdd(nzero, (unsigned long long)block_zeros);
}

// g * @av_xza(rho_f) on the y faces: the gravity term of ComputeDarcyFlux!
// (kernels_HT3D_dimensional.jl:269), formed once per solve.
__global__ __launch_bounds__(256) void gravity_kernel(const double *__restrict__ rhofv, double g_, int nx, int ny,
                                                      int nz, double *__restrict__ gy) {
    const size_t n = (size_t)nx * (ny + 1) * nz;
    const Grid g{nx, ny, nz};
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % (ny + 1)), k = (int)(t / ((size_t)nx * (ny + 1)));
        gy[t] = g_ * ((rhofv[g.v(i, j, k)] + rhofv[g.v(i + 1, j, k)] + rhofv[g.v(i, j, k + 1)] +
                       rhofv[g.v(i + 1, j, k + 1)]) * 0.25);
    }
}

// Copies the cells that a ghost-fill rule sets (ghosts, air, immersed) from src to dst.
__global__ __launch_bounds__(256) void copy_special_kernel(dim::BcArgs a, double *__restrict__ dst) {
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const size_t n = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % (nx + 2)), j = (int)((t / (nx + 2)) % (ny + 2)),
                  k = (int)(t / ((size_t)(nx + 2) * (ny + 2)));
        if (dim::bc_rule(a, i, j, k) != dim::R_KEEP) dst[t] = a.A[t];
    }
}

// One equation cell of the fused v18 iteration: the rules for its neighbours
// (immersed-surface neighbours re-evaluated, walls and the basal heat flux
// applied), its residual F and its diagonal D (returned in D); returns F / D.
// t1: heat T of the previous step / Darcy temperature of the cell; t2: heat
// rho*C / Darcy phi*rho_f of the previous step; Dv: Darcy stored diagonal;
// gyS, gyN: Darcy gravity terms on the cell's y faces.  The one- and two-cell
// kernels both call it, so they evaluate the same operations in the same order.
template <bool DARCY>
__device__ __forceinline__ double v18_cell(const Args &a, uint8_t cl, int i, int j, int k, double xc, double w,
                                           double ea, double s, double n, double b, double fr, double kW, double kE,
                                           double kS, double kN, double kB, double kF, double t1, double t2,
                                           double Dv, double gyS, double gyN, double &D) {
    if (cl & C_SLOW) {   // next to an immersed-surface cell: every neighbour from the rules
        const int ie = i + 1, je = j + 1, ke = k + 1;
        w = dim::bc_value(a.bc, ie - 1, je, ke);
        ea = dim::bc_value(a.bc, ie + 1, je, ke);
        s = dim::bc_value(a.bc, ie, je - 1, ke);
        n = dim::bc_value(a.bc, ie, je + 1, ke);
        b = dim::bc_value(a.bc, ie, je, ke - 1);
        fr = dim::bc_value(a.bc, ie, je, ke + 1);
    } else {             // walls: copies of the cell, basal heat flux
        if (i == 0) w = xc;
        if (i == a.nx - 1) ea = xc;
        if (k == 0) b = xc;
        if (k == a.nz - 1) fr = xc;
        if (j == 0) s = DARCY ? xc : a.qyS_dy / kS + xc;
    }
    const double qW = -kW * (a._dx * (xc - w)), qE = -kE * (a._dx * (ea - xc));
    const double qB = -kB * (a._dz * (xc - b)), qF = -kF * (a._dz * (fr - xc));
    double f;
    if (DARCY) {
        // phi*rho_f of the cell (kernels_HT3D_dimensional.jl:225-235), then
        // ComputeDarcyFlux! and ResidualFluidPressure! (:266-289)
        const double phi = (cl & C_FAULT) ? 3.0 * a.phi_i : a.phi_i;
        const double prf = phi * dim::rhof_C(t1 * a.s_T - 273.15, xc * a.s_sigma) / a.s_rho;
        const double qS = -kS * (a._dy * (xc - s) - gyS);
        const double qN = -kN * (a._dy * (n - xc) - gyN);
        f = a.tr_dt * (prf - t2);
        f += (qE - qW) * a._dx;
        f += (qN - qS) * a._dy;
        f += (qF - qB) * a._dz;
        D = Dv;
    } else {
        // ComputeFlux! and ResidualTemperatureLinearised! (:140-181)
        const double rc = t2;
        const double qS = -kS * (a._dy * (xc - s)), qN = -kN * (a._dy * (n - xc));
        f = a.transient * rc * a._dt * (xc - t1);
        f += (qE - qW) * a._dx;
        f += (qN - qS) * a._dy;
        f += (qF - qB) * a._dz;
        f -= a.transient * ((cl & C_SRC) ? a.Qt : 0.0);
        D = (kB / a.dz + kF / a.dz) / a.dz + (kN / a.dy + kS / a.dy) / a.dy + (kE / a.dx + kW / a.dx) / a.dx +
            rc * a.transient / a.dt;
    }
    return f / D;
}

// The fused v18 pseudo-iteration with CPL (1 or 2) x-cells per lane: a block
// covers 64*CPL x TY cells marched over KZ planes.  With CPL = 2 every stream
// comes in as 2-element vector loads (nx even); the loads of the neighbours,
// coefficients and operator streams are skipped when all the lane's cells are
// air or immersed (F = 0).  KV (CPL = 2): the face coefficients are formed
// from the vertex field a.kv (kx, ky, kz = InitConductivity! of it,
// kernels_HT3D_dimensional.jl:133-138) with the averaging kernel's additions
// (av4, as iter_tile_kv does), one vertex word per cell instead of three face
// words; its rows are loaded on every plane, air or not, to carry the march.
template <bool DARCY, int TY, int KZ, int CPL, bool KV = false>
__global__ __launch_bounds__(64 * TY) void iter_kernel(Args a) {
    typedef typename CellVec<double, CPL>::type V;
    const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
    const unsigned t = xcd_tile(blockIdx.x, nblocks, a.flags & FL_REVERSE);
    const int bx = t % a.nbx, by = (t / a.nbx) % a.nby, bz = t / (a.nbx * a.nby);
    const int i = bx * (64 * CPL) + CPL * (int)(threadIdx.x & 63u);
    const int j = by * TY + (int)(threadIdx.x >> 6);
    const int k0 = bz * KZ;
    const int k1 = min(k0 + KZ, a.nz);
    const uint32_t nx = a.nx, ny = a.ny;
    const uint32_t sE1 = nx + 2, sE2 = sE1 * (ny + 2);
    const uint32_t sC2 = nx * ny, sX1 = nx + 1, sX2 = sX1 * ny, sY2 = nx * (ny + 1);
    const bool check = a.flags & FL_CHECK, lam = a.flags & FL_LAMBDA, store = a.flags & FL_STORE_F;
    double s_hi = 0.0, s_lo = 0.0, acc_max = 0.0, n_hi = 0.0, n_lo = 0.0, d_hi = 0.0, d_lo = 0.0;

    if ((uint32_t)i < nx && (uint32_t)j < ny) {
        uint32_t e = (uint32_t)(i + 1) + (uint32_t)(j + 1) * sE1 + (uint32_t)(k0 + 1) * sE2;
        uint32_t c = (uint32_t)i + (uint32_t)j * nx + (uint32_t)k0 * sC2;
        uint32_t fx = (uint32_t)i + (uint32_t)j * sX1 + (uint32_t)k0 * sX2;
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
        V xc = ldv<double, CPL>(a.Xin + e);
        V xm = ldv<double, CPL>(a.Xin + e - sE2);
        V kzB;
        if constexpr (KV) {
            vrow(v, a0);
            vrow(v + sV1, a1);
#pragma unroll
            for (int q = 0; q < CPL; ++q) kzB[q] = av4(a0[q], a0[q + 1], a1[q], a1[q + 1]);
        } else {
            kzB = ldv<double, CPL>(a.kz + c);
        }
        // the Darcy loop unrolled by 4 (the next planes' loads issue ahead of this
        // plane's arithmetic: 1 % faster); the heat loop measured no gain
        constexpr int kUnrollZ = DARCY ? 4 : 1;
#pragma unroll kUnrollZ
        for (int k = k0; k < k1; ++k) {
            const V xp = ldv<double, CPL>(a.Xin + e + sE2);
            V kF;
            double b0[CPL + 1], b1[CPL + 1];
            if constexpr (KV) {
                vrow(v + sV2, b0);
                vrow(v + sV2 + sV1, b1);
#pragma unroll
                for (int q = 0; q < CPL; ++q) kF[q] = av4(b0[q], b0[q + 1], b1[q], b1[q + 1]);
            } else {
                kF = ldv<double, CPL>(a.kz + c + sC2);
            }
            uint8_t cl[CPL];
            if constexpr (CPL == 1) {
                cl[0] = __builtin_nontemporal_load(a.cls + c);
            } else {   // two class bytes in one load (c is even: nx is even)
                const uint16_t c2 = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(a.cls + c));
                cl[0] = (uint8_t)(c2 & 0xffu);
                cl[CPL - 1] = (uint8_t)(c2 >> 8);
            }
            const V f0o = ldv_nt<double, CPL>(a.F0 + c);
            V f, D;
            uint8_t all = C_ZERO;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                f[q] = 0.0;
                D[q] = 1.0;
                all &= cl[q];
            }
            if (!(all & C_ZERO)) {   // air and immersed cells: F = 0, nothing else to load
                double kW;
                V kE, kS, kN;
                if constexpr (KV) {
                    kW = av4(a0[0], a1[0], b0[0], b1[0]);
#pragma unroll
                    for (int q = 0; q < CPL; ++q) {
                        kE[q] = av4(a0[q + 1], a1[q + 1], b0[q + 1], b1[q + 1]);
                        kS[q] = av4(a0[q], a0[q + 1], b0[q], b0[q + 1]);
                        kN[q] = av4(a1[q], a1[q + 1], b1[q], b1[q + 1]);
                    }
                } else {
                    kW = a.kx[fx];
                    kE = ldv<double, CPL>(a.kx + fx + 1);
                    kS = ldv<double, CPL>(a.ky + fy);
                    kN = ldv<double, CPL>(a.ky + fy + nx);
                }
                const double xW = a.Xin[e - 1], xE = a.Xin[e + CPL];
                const V xS = ldv<double, CPL>(a.Xin + e - sE1), xN = ldv<double, CPL>(a.Xin + e + sE1);
                V t1, t2, Dv = {}, gS = {}, gN = {};
                if (DARCY) {
                    t1 = ldv_nt<double, CPL>(a.aux + e);
                    t2 = ldv_nt<double, CPL>(a.aux2 + c);
                    Dv = ldv_nt<double, CPL>(a.Dst + c);
                    gS = ldv<double, CPL>(a.gy + fy);
                    gN = ldv<double, CPL>(a.gy + fy + nx);
                } else {
                    t1 = ldv_nt<double, CPL>(a.aux + c);
                    t2 = ldv_nt<double, CPL>(a.aux2 + c);
                }
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    if (cl[q] & C_ZERO) continue;
                    const double w = q == 0 ? xW : (double)xc[q - 1], ea = q == CPL - 1 ? xE : (double)xc[q + 1];
                    const double kw = q == 0 ? kW : (double)kE[q - 1];
                    double d;
                    f[q] = v18_cell<DARCY>(a, cl[q], i + q, j, k, xc[q], w, ea, xS[q], xN[q], xm[q], xp[q], kw, kE[q],
                                           kS[q], kN[q], kzB[q], kF[q], t1[q], t2[q], Dv[q], gS[q], gN[q], d);
                    D[q] = d;
                }
            }
            // DYRELUpdate! (kernels_HT3D_dimensional.jl:323-324)
            V g, xo;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                g[q] = a.h1 * f0o[q] + a.h2 * f[q];
                xo[q] = xc[q] - a.dtau * g[q];
            }
            stv_nt<double, CPL>(g, a.F0 + c);
            stv_nt<double, CPL>(xo, a.Xout + e);
            if (check) {
                V fst = xo;   // only read on lambda iterations
                if (lam) fst = ldv<double, CPL>(a.Fst + c);
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const double fd = fabs(f[q] * D[q]);
                    dd_add(s_hi, s_lo, fd);
                    acc_max = (fd > acc_max || fd != fd) ? fd : acc_max;
                    if (lam) {
                        const double d = a.dtau * g[q];
                        dd_add(n_hi, n_lo, (-d) * (f[q] - fst[q]));
                        dd_add(d_hi, d_lo, d * d);
                    }
                }
            }
            if (store) stv<double, CPL>(f, a.Fst + c);
            xm = xc;
            xc = xp;
            kzB = kF;
            if constexpr (KV) {
#pragma unroll
                for (int q = 0; q <= CPL; ++q) {
                    a0[q] = b0[q];
                    a1[q] = b1[q];
                }
            }
            e += sE2;
            c += sC2;
            fx += sX2;
            fy += sY2;
            v += sV2;
        }
    }
    if (check) {
        const double h4[4] = {s_hi, acc_max, n_hi, d_hi}, l4[4] = {s_lo, 0.0, n_lo, d_lo};
        block_reduce4(h4, l4, 0x2u, a.partials, blockIdx.x);
    }
}

struct Tile {
    int ty, kz;
    int cpl = 1;       // x-cells per lane (2: nx even only)
    bool kv = false;   // face coefficients from the vertex field (2 cells per lane; the solve must have one)
};
static constexpr Tile kV18Tiles[] = {{4, 4},          {4, 8},          {8, 4},          {2, 8},
                                     {8, 8},          {4, 16},         {4, 4, 2},       {4, 8, 2},
                                     {2, 8, 2},       {4, 16, 2},      {4, 4, 2, true}, {4, 8, 2, true},
                                     {2, 8, 2, true}, {4, 16, 2, true}};
static constexpr int kNumV18Tiles = sizeof(kV18Tiles) / sizeof(kV18Tiles[0]);

template <bool DARCY>
static void launch(int tile, Args &a, hipStream_t st) {
    const Tile tl = kV18Tiles[tile];
    a.nbx = (int)div_up(a.nx, 64 * tl.cpl);
    a.nby = (int)div_up(a.ny, tl.ty);
    a.nbz = (int)div_up(a.nz, tl.kz);
    const unsigned nb = (unsigned)a.nbx * a.nby * a.nbz;
    a.bc.A = a.Xin;
    switch (tile) {
    case 0: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 4, 1>), dim3(nb), dim3(256), 0, st, a); break;
    case 1: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 8, 1>), dim3(nb), dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL((iter_kernel<DARCY, 8, 4, 1>), dim3(nb), dim3(512), 0, st, a); break;
    case 3: hipLaunchKernelGGL((iter_kernel<DARCY, 2, 8, 1>), dim3(nb), dim3(128), 0, st, a); break;
    case 4: hipLaunchKernelGGL((iter_kernel<DARCY, 8, 8, 1>), dim3(nb), dim3(512), 0, st, a); break;
    case 5: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 16, 1>), dim3(nb), dim3(256), 0, st, a); break;
    case 6: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 4, 2>), dim3(nb), dim3(256), 0, st, a); break;
    case 7: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 8, 2>), dim3(nb), dim3(256), 0, st, a); break;
    case 8: hipLaunchKernelGGL((iter_kernel<DARCY, 2, 8, 2>), dim3(nb), dim3(128), 0, st, a); break;
    case 9: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 16, 2>), dim3(nb), dim3(256), 0, st, a); break;
    case 10: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 4, 2, true>), dim3(nb), dim3(256), 0, st, a); break;
    case 11: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 8, 2, true>), dim3(nb), dim3(256), 0, st, a); break;
    case 12: hipLaunchKernelGGL((iter_kernel<DARCY, 2, 8, 2, true>), dim3(nb), dim3(128), 0, st, a); break;
    default: hipLaunchKernelGGL((iter_kernel<DARCY, 4, 16, 2, true>), dim3(nb), dim3(256), 0, st, a); break;
    }
}

// A 2-cell tile needs an even nx, and a vertex-field tile the vertex field;
// otherwise its twin with the same rows and planes is used (the face-array
// 2-cell tile, or the 1-cell tile).
static int tile_for_grid(int tile, int nx, bool has_kv) {
    const Tile &tl = kV18Tiles[tile];
    if (tl.cpl == 1 || (nx % 2 == 0 && (!tl.kv || has_kv))) return tile;
    const int cpl = nx % 2 == 0 ? 2 : 1;
    for (int t = 0; t < kNumV18Tiles; ++t)
        if (!kV18Tiles[t].kv && kV18Tiles[t].cpl == cpl && kV18Tiles[t].ty == tl.ty && kV18Tiles[t].kz == tl.kz)
            return t;
    return 0;
}

static unsigned max_blocks(int nx, int ny, int nz) {
    unsigned m = 0;
    for (const Tile &t : kV18Tiles) {
        const unsigned nb = div_up(nx, 64 * t.cpl) * div_up(ny, t.ty) * div_up(nz, t.kz);
        m = nb > m ? nb : m;
    }
    return m;
}

// Tile choice per (grid, operator): each candidate is timed on plain
// iterations of the solve's own inputs, ping-ponged between two scratch
// buffers (the unknown and F0 of the solve are untouched).
static std::map<std::tuple<int, int, int, int>, int> &tile_cache() {
    static std::map<std::tuple<int, int, int, int>, int> m;
    return m;
}
int g_v18_force_tile = -1;
void clear_tiles() { tile_cache().clear(); }

template <bool DARCY>
static int choose_tile(const Args &base, double *scratch_x, double *scratch_f0, int *tile_out) {
    const bool has_kv = base.kv != nullptr;
    if (g_v18_force_tile >= 0 && g_v18_force_tile < kNumV18Tiles) {
        *tile_out = tile_for_grid(g_v18_force_tile, base.nx, has_kv);
        return HT3D_OK;
    }
    const auto key = std::make_tuple(base.nx, base.ny, base.nz, (DARCY ? 1 : 0) + (has_kv ? 2 : 0));
    auto it = tile_cache().find(key);
    if (it != tile_cache().end()) { *tile_out = it->second; return HT3D_OK; }
    hipStream_t st = ctx().stream;
    const size_t N = (size_t)base.nx * base.ny * base.nz;
    HIP_TRY(hipMemsetAsync(scratch_f0, 0, N * sizeof(double), st));
    TimerEvents ev;
    if (int rc_ = ev.create()) return rc_;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    // the timed launches ping-pong the iterate between a scratch copy of the
    // input and scratch_x, as the solve does (the MALL contents at each launch
    // depend on it); each tile keeps its best of three interleaved rounds
    const size_t NE = (size_t)(base.nx + 2) * (base.ny + 2) * (base.nz + 2);
    void *tx = nullptr;
    if (int rc = workspace(WS_TUNE_X, NE * sizeof(double), &tx)) return rc;
    HIP_TRY(hipMemcpyAsync(tx, base.Xin, NE * sizeof(double), hipMemcpyDeviceToDevice, st));
    double *xs[2] = {(double *)tx, scratch_x};
    int pp = 0;
    float tbest[kNumV18Tiles];
    for (int tile = 0; tile < kNumV18Tiles; ++tile) tbest[tile] = 1e30f;
    for (int round = 0; round < 3; ++round)
        for (int tile = 0; tile < kNumV18Tiles; ++tile) {
            if (tile_for_grid(tile, base.nx, has_kv) != tile) continue;
            Args a = base;
            a.F0 = scratch_f0;
            a.flags = 0;
            auto go = [&] {
                a.Xin = xs[pp];
                a.Xout = xs[pp ^ 1];
                pp ^= 1;
                if (g_alternate) a.flags ^= FL_REVERSE;
                launch<DARCY>(tile, a, st);
            };
            go();   // warm-up
            HIP_TRY(hipEventRecord(e0, st));
            for (int r = 0; r < 6; ++r) go();
            HIP_TRY(hipEventRecord(e1, st));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            tbest[tile] = ms < tbest[tile] ? ms : tbest[tile];
        }
    int best = 0;
    for (int tile = 1; tile < kNumV18Tiles; ++tile)
        if (tbest[tile] < tbest[best]) best = tile;
    HIP_TRY(hipGetLastError());
    tile_cache()[key] = best;
    *tile_out = best;
    return HT3D_OK;
}

int g_v18_last_tile = -1;

static inline bool lambda_iter(int iter, int nitout) { return iter <= 2 || (nitout > 0 && iter % nitout == 0); }

template <bool DARCY>
static int solve(const ht3d_v18_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep) {
    HT3D_CHECK_INIT();
    const char *name = DARCY ? "ht3d_v18_darcy_solve" : "ht3d_v18_thermal_solve";
    if (!f || !p) return arg_fail("v18 solve: null argument");
    const int nx = f->ncx, ny = f->ncy, nz = f->ncz;
    if (!grid_ok(nx, ny, nz)) return arg_fail("v18 solve: extents must be >= 3");
    if (!f->X || !f->kx || !f->ky || !f->kz || !f->phc_ex || !f->wN || !f->rhoC)
        return arg_fail("v18 solve: null field pointer");
    if (DARCY ? (!f->Tc_ex || !f->phirhof0 || !f->rhofv) : !f->Tc_old)
        return arg_fail("v18 solve: null field pointer");
    if (p->nitmax < 1 || p->nitout < 1 || !(p->tol >= 0.0)) return arg_fail("v18 solve: bad parameters");
    if ((size_t)(nx + 2) * (ny + 2) * (nz + 2) >= ((size_t)1 << 32)) return arg_fail("v18 solve: grid too large");
    Context &cx = ctx();
    hipStream_t st = cx.stream;
    const size_t N = (size_t)nx * ny * nz, NE = (size_t)(nx + 2) * (ny + 2) * (nz + 2);

    void *xb, *f0, *f0b, *fst, *part, *res, *cls, *gy = nullptr, *pcw, *gw;
    int rc;
    if ((rc = workspace(WS_XB, NE * sizeof(double), &xb))) return rc;
    if ((rc = workspace(WS_F0, N * sizeof(double), &f0))) return rc;
    if ((rc = workspace(WS_F0B, N * sizeof(double), &f0b))) return rc;
    if ((rc = workspace(WS_FCIT, N * sizeof(double), &fst))) return rc;
    if ((rc = workspace(WS_PARTIALS, (size_t)max_blocks(nx, ny, nz) * 8 * sizeof(double), &part))) return rc;
    if ((rc = workspace(WS_RESULT, 64 * sizeof(double), &res))) return rc;
    if ((rc = workspace(WS_V18_CLS, N, &cls))) return rc;
    if ((rc = workspace(WS_DIAG, N * sizeof(double), &pcw))) return rc;
    if ((rc = workspace(WS_V18_G, N * sizeof(double), &gw))) return rc;
    if (DARCY && (rc = workspace(WS_V18_GY, (size_t)nx * (ny + 1) * nz * sizeof(double), &gy))) return rc;
    double *PC = f->PC ? f->PC : (double *)pcw;
    double *G = f->G ? f->G : (double *)gw;

    // set-up: GershgorinPoisson! and lambda_max (v18:432-439 / :495-502)
    if ((rc = ht3d_GershgorinPoisson(G, PC, f->rhoC, f->kx, f->ky, f->kz, f->transient, f->dt, f->dx, f->dy, f->dz,
                                     nx, ny, nz)))
        return rc;
    double lmax = 0.0;
    if ((rc = ht3d_reduce(&lmax, G, N, HT3D_MAX))) return rc;
    double lmin = lmax / p->lmin_div;
    const double dtau = 2.0 / sqrt(lmax) * p->CFL;
    double c = 2.0 * sqrt(lmin) * p->cfact;
    double h1 = (2 - c * dtau) / (2 + c * dtau), h2 = 2 * dtau / (2 + c * dtau);

    const unsigned nb_lin = div_up(N, 256) > 4096 ? 4096 : div_up(N, 256);
    unsigned long long *nzero_dev = (unsigned long long *)((double *)res + 32);
    HIP_TRY(hipMemsetAsync(nzero_dev, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(class_kernel, dim3(nb_lin), dim3(256), 0, st, f->phc_ex, nx, ny, nz, (uint8_t *)cls, nzero_dev);
    if (DARCY)
        hipLaunchKernelGGL(gravity_kernel, dim3(nb_lin), dim3(256), 0, st, f->rhofv, f->g, nx, ny, nz, (double *)gy);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(f0, 0, N * sizeof(double), st));    // ResetA!(Fc, Fc0)
    HIP_TRY(hipMemsetAsync(fst, 0, N * sizeof(double), st));
    unsigned long long nzero = 0;
    HIP_TRY(hipMemcpyAsync(&nzero, nzero_dev, sizeof nzero, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // the ghost rules on the initial unknown (what the first BC launch of the
    // reference does): air cells and air ghosts then hold the surface value
    if (DARCY)
        rc = ht3d_dim_SetPressureBCs(f->X, f->wN, f->phc_ex, 0.0, f->top, 0, 0, 0, 0, 0, 0, 0, 0, nx, ny, nz);
    else
        rc = ht3d_dim_SetTemperatureBCs(f->X, f->ky, f->wN, f->phc_ex, f->qyS, 1.0 / f->dy, 0, 0, f->top, 0, 0, 0, 0,
                                        0, 0, 0, 0, 0, 0, nx, ny, nz);
    if (rc) return rc;
    // the second buffer needs the ghosts (no iteration writes them)
    if ((rc = copy_ghosts((double *)xb, f->X, nx, ny, nz))) return rc;

    Args a;
    std::memset(&a, 0, sizeof a);
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.kx = f->kx; a.ky = f->ky; a.kz = f->kz; a.kv = f->kv;
    a.cls = (const uint8_t *)cls;
    a.Fst = (double *)fst;
    a.partials = (double *)part;
    a.aux = DARCY ? f->Tc_ex : f->Tc_old;
    a.aux2 = DARCY ? f->phirhof0 : f->rhoC;
    a.Dst = PC;
    a.gy = (const double *)gy;
    a._dx = 1.0 / f->dx; a._dy = 1.0 / f->dy; a._dz = 1.0 / f->dz;
    a.dx = f->dx; a.dy = f->dy; a.dz = f->dz; a.dt = f->dt;
    a._dt = 1.0 / f->dt;
    a.transient = f->transient;
    a.tr_dt = f->transient * a._dt;
    a.Qt = f->Qt;
    a.qyS_dy = f->qyS / a._dy;
    a.phi_i = f->phi_i; a.s_rho = f->scale_rho; a.s_T = f->scale_T; a.s_sigma = f->scale_sigma;
    a.bc = dim::BcArgs{f->X, f->ky, f->wN, f->phc_ex, f->top, a.qyS_dy, Grid{nx, ny, nz}, DARCY ? 0 : 1};
    a.dtau = dtau; a.h1 = h1; a.h2 = h2;
    a.Xin = f->X;
    a.F0 = (double *)f0;
    int tile = 0;
    if ((rc = choose_tile<DARCY>(a, (double *)xb, (double *)f0b, &tile))) return rc;   // writes xb's interior only
    g_v18_last_tile = tile;

    const int ncheck = p->ncheck < 0 ? tuned_check_interval(-1.0, -1.0, -1.0, p->nitout) : p->ncheck;
    double *buf[2] = {f->X, (double *)xb};
    int cur = 0;
    bool rev = false;   // sweep direction of the next launch (alternated: MALL reuse)
    SegmentTimer timer;
    // algorithmic bytes per iteration: for a cell with an equation the unknown
    // and F0 read and written, five (heat) or seven (Darcy) read-only words and
    // the class byte; for an air or immersed cell the unknown, F0, kz and the
    // class byte (its update is X -= dtau*F0 with F0 = h1*F0)
    const double bytes_per_iter = ((DARCY ? 11.0 : 9.0) * 8.0 + 1.0) * (double)(N - nzero) + 41.0 * (double)nzero;
    double nF_abs = 0.0, nF_rel = 0.0, nF_ini = 0.0;
    int status = HT3D_NOT_CONVERGED, iter, n_checks = 0, n_lambda = 0, done_iters = 0;
    for (iter = 1; iter <= p->nitmax; ++iter) {
        const bool lam = lambda_iter(iter, p->nitout);
        const bool chk = lam || (ncheck > 0 && iter % ncheck == 0);
        const bool store = iter < p->nitmax && lambda_iter(iter + 1, p->nitout);
        a.Xin = buf[cur];
        a.Xout = buf[cur ^ 1];
        a.h1 = h1; a.h2 = h2;
        a.flags = (chk ? FL_CHECK : 0u) | (lam ? FL_LAMBDA : 0u) | (store ? FL_STORE_F : 0u) |
                  (rev ? FL_REVERSE : 0u);
        rev = g_alternate && !rev;
        timer.before_launch();
        launch<DARCY>(tile, a, st);
        HIP_TRY(hipGetLastError());
        ++timer.count;
        cur ^= 1;
        done_iters = iter;
        if (chk || iter == p->nitmax) timer.close();
        if (chk) {
            ++n_checks;
            const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
            if ((rc = finish_partials((double *)part, nblocks, 0x2u, (double *)res))) return rc;
            double r4[4];
            if ((rc = read_result4((double *)res, r4))) return rc;
            timer.collect(bytes_per_iter, false);
            nF_abs = (p->norm == 0 ? r4[0] / (double)N / sqrt((double)N) : r4[1]) * f->norm_scale;
            if (r4[0] != r4[0]) nF_abs = r4[0];
            if (iter == 1) nF_ini = nF_abs;
            nF_rel = nF_abs / nF_ini;
            if (std::isnan(nF_abs)) { status = HT3D_NAN; break; }
            if (nF_rel > 100) { status = HT3D_DIVERGED; break; }
            if (lam && iter > 1) {
                lmin = fabs(r4[2]);
                lmin /= fabs(r4[3]);
                c = 2.0 * sqrt(lmin) * p->cfact;
                h1 = (2 - c * dtau) / (2 + c * dtau);
                h2 = 2 * dtau / (2 + c * dtau);
                ++n_lambda;
            }
            if (nF_abs < p->tol || nF_rel < p->tol) { status = HT3D_OK; break; }
        }
    }
    timer.close();
    timer.collect(bytes_per_iter, false);

    // Finalise as the reference leaves memory: the BC of the last iteration
    // applied to its input (ghosts, air and immersed cells), the update on the rest.
    double *res_buf = buf[cur], *prev = buf[cur ^ 1];
    if (DARCY) {   // ϕρf of the last iteration (before its BC and update)
        if ((rc = ht3d_PhiRhof(f->rhoC, f->Tc_ex, prev, f->phc_ex, f->phi_i, f->scale_rho, f->scale_T,
                               f->scale_sigma, nx, ny, nz)))
            return rc;
        if ((rc = ht3d_dim_SetPressureBCs(prev, f->wN, f->phc_ex, 0.0, f->top, 0, 0, 0, 0, 0, 0, 0, 0, nx, ny, nz)))
            return rc;
    } else {
        if ((rc = ht3d_dim_SetTemperatureBCs(prev, f->ky, f->wN, f->phc_ex, f->qyS, a._dy, 0, 0, f->top, 0, 0, 0, 0,
                                             0, 0, 0, 0, 0, 0, nx, ny, nz)))
            return rc;
    }
    dim::BcArgs bc = a.bc;
    bc.A = prev;
    const unsigned nb_e = div_up(NE, 256) > 4096 ? 4096 : div_up(NE, 256);
    hipLaunchKernelGGL(copy_special_kernel, dim3(nb_e), dim3(256), 0, st, bc, res_buf);
    HIP_TRY(hipGetLastError());
    if (res_buf != f->X) HIP_TRY(hipMemcpyAsync(f->X, res_buf, NE * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (rep) {
        rep->ncheck = ncheck;
        rep->iters = done_iters;
        rep->status = status;
        rep->n_checks = n_checks;
        rep->n_lambda = n_lambda;
        rep->lmax = lmax; rep->lmin = lmin; rep->dtau = dtau; rep->h1 = h1; rep->h2 = h2;
        rep->nF_abs = nF_abs; rep->nF_rel = nF_rel; rep->nF_ini = nF_ini;
    }
    rc = after_launch(name, true);
    if (rc) return rc;
    return status;
}

}  // namespace v18
}  // namespace ht3d

extern "C" {

int ht3d_v18_thermal_solve(const ht3d_v18_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep) {
    HT3D_API_BEGIN
    return ht3d::v18::solve<false>(f, p, rep);
    HT3D_API_END
}

int ht3d_v18_darcy_solve(const ht3d_v18_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep) {
    HT3D_API_BEGIN
    return ht3d::v18::solve<true>(f, p, rep);
    HT3D_API_END
}

int ht3d_v18_tile_count(void) { return ht3d::v18::kNumV18Tiles; }

int ht3d_v18_tune(int force_tile, int *last_tile) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (force_tile >= ht3d::v18::kNumV18Tiles) return ht3d::arg_fail("ht3d_v18_tune: tile index"), HT3D_ERR_ARG;
    ht3d::v18::g_v18_force_tile = force_tile;
    if (last_tile) *last_tile = ht3d::v18::g_v18_last_tile;
    return HT3D_OK;
    HT3D_API_END
}

}  // extern "C"
