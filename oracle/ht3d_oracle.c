/*
 * ht3d_oracle.c -- CPU oracle for the HydroThermal3D solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library
 * (hydrothermal3d_amd/libht3d.so) is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it; the
 * product path never does.
 *
 * It restates, unfused and one loop nest per reference kernel, the Julia /
 * ParallelStencil kernels of tduretz/HydroThermal3D in FP64:
 *   Dimensionless/kernels_HT3D.jl:73-154            (fluxes, residuals, updates)
 *   Dimensionless/HydroThermal3D_Alban_v6.jl:84-159 (closed-box BCs, Nu :345-348)
 *   Dimensional/kernels_HT3D_dimensional.jl:106-363 (DYREL, Gershgorin, Rayleigh)
 *   Dimensional/HydroThermal3D_v18.jl:424-538       (DYREL control flow)
 *   tools/Macros.jl:111,171-199                     (@dmul_ya, ResetA!, copies)
 *   tools/Weno5_Routines.jl  (v1)  and  tools/Weno5_Routines_v2.jl (v2)
 * with ParallelStencil.FiniteDifferences3D macro semantics (@d_xi, @av_yza, ...).
 *
 * Conventions (same as the reference):
 *   * arrays are column-major, x fastest; indices below are 1-based like Julia;
 *   * grid extents: c = (nx,ny,nz), E = (nx+2,ny+2,nz+2), X6 = (nx+6,ny+6,nz+6),
 *     Fx = (nx+1,ny,nz), Fy = (nx,ny+1,nz), Fz = (nx,ny,nz+1), V = (nx+1,ny+1,nz+1);
 *   * every expression keeps Julia's left-to-right evaluation order, and the file
 *     is compiled with -ffp-contract=off so that no multiply-add is fused
 *     (Julia's CPU code does not fuse either).
 *   * reductions (sum / mean / max) are deterministic: per z-plane partials in
 *     index order, then planes in order.  Julia's pairwise summation order is
 *     not reproduced (it only matters at the last bit).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ht3d_oracle_common.h"

/* ------------------------------------------------------------------------ */
/* Pointwise kernels: tools/Macros.jl:171-199, tools/Weno5_Routines.jl:271   */
/* ------------------------------------------------------------------------ */

/* ResetA!  tools/Macros.jl:171-177 (A and B of the same extents n1 x n2 x n3) */
EXPORT void o_ResetA(double *A, double *B, int n1, int n2, int n3) {
    size_t n = (size_t)n1 * n2 * n3;
    memset(A, 0, n * sizeof(double));
    memset(B, 0, n * sizeof(double));
}

/* Cpy_inn_to_all!  tools/Macros.jl:187-192: A (c) = inn(B (E)) */
EXPORT void o_Cpy_inn_to_all(double *A, const double *B, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Cc(A, i, j, k) = Ee(B, i + 1, j + 1, k + 1);
}

/* Cpy_all_to_inn!  tools/Macros.jl:194-199: inn(A (E)) = B (c) */
EXPORT void o_Cpy_all_to_inn(double *A, const double *B, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Ee(A, i + 1, j + 1, k + 1) = Cc(B, i, j, k);
}

/* ArrayEqualArray!  tools/Weno5_Routines.jl:271-276 */
EXPORT void o_ArrayEqualArray(double *F1, const double *F2, int n1, int n2, int n3) {
    memcpy(F1, F2, (size_t)n1 * n2 * n3 * sizeof(double));
}

/* Multiply  Dimensional/kernels_HT3D_dimensional.jl:106-109 */
EXPORT void o_Multiply(double *a, const double *b, const double *c, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Cc(a, i, j, k) = Cc(b, i, j, k) * Cc(c, i, j, k);
}

/* ------------------------------------------------------------------------ */
/* Coefficient set-up: kernels_HT3D.jl:73-107, Alban_v6:153-159               */
/* ------------------------------------------------------------------------ */

/* InitThermal!  Dimensionless/kernels_HT3D.jl:99-107 */
EXPORT void o_InitThermal(double *Rt, double *kx, double *ky, double *kz, const double *Tc_ex,
                          const double *ktv, double _dt, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Cc(Rt, i, j, k) = (-_dt) * Ee(Tc_ex, i + 1, j + 1, k + 1);
    PFOR_K(nz) FOR_JI(nx + 1, ny) FX(kx, i, j, k) = AV_YZA(ktv, nx + 1, ny + 1);
    PFOR_K(nz) FOR_JI(nx, ny + 1) FY(ky, i, j, k) = AV_XZA(ktv, nx + 1, ny + 1);
    PFOR_K(nz + 1) FOR_JI(nx, ny) FZ(kz, i, j, k) = AV_XYA(ktv, nx + 1, ny + 1);
}

/* InitConductivity!  Dimensional/kernels_HT3D_dimensional.jl:133-138 */
EXPORT void o_InitConductivity(double *kx, double *ky, double *kz, const double *ktv, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx + 1, ny) FX(kx, i, j, k) = AV_YZA(ktv, nx + 1, ny + 1);
    PFOR_K(nz) FOR_JI(nx, ny + 1) FY(ky, i, j, k) = AV_XZA(ktv, nx + 1, ny + 1);
    PFOR_K(nz + 1) FOR_JI(nx, ny) FZ(kz, i, j, k) = AV_XYA(ktv, nx + 1, ny + 1);
}

/* InitDarcy!  Dimensionless/kernels_HT3D.jl:82-90 */
EXPORT void o_InitDarcy(double *Ty, double *kx, double *ky, double *kz, const double *Tc_ex,
                        const double *kfv, double _dt, int nx, int ny, int nz) {
    (void)_dt;
    PFOR_K(nz) FOR_JI(nx, ny + 1) FY(Ty, i, j, k) = AV_YI(Tc_ex, nx + 2, ny + 2);
    PFOR_K(nz) FOR_JI(nx + 1, ny) FX(kx, i, j, k) = AV_YZA(kfv, nx + 1, ny + 1);
    PFOR_K(nz) FOR_JI(nx, ny + 1) FY(ky, i, j, k) = AV_XZA(kfv, nx + 1, ny + 1);
    PFOR_K(nz + 1) FOR_JI(nx, ny) FZ(kz, i, j, k) = AV_XYA(kfv, nx + 1, ny + 1);
}

/* Temperature-dependent viscosity mu = exp(-a T): face coefficients times
 * exp(a T_face), T_face = (T_left + T_right) * 0.5 with the wall ghosts.  (No
 * counterpart in the dimensionless kernels, which have constant viscosity.)  */
EXPORT void o_ApplyViscosity(double *kx, double *ky, double *kz, const double *Tc_ex, double a, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx + 1, ny)
        FX(kx, i, j, k) = FX(kx, i, j, k) * exp(a * ((Ee(Tc_ex, i, j + 1, k + 1) + Ee(Tc_ex, i + 1, j + 1, k + 1)) * 0.5));
    PFOR_K(nz) FOR_JI(nx, ny + 1)
        FY(ky, i, j, k) = FY(ky, i, j, k) * exp(a * ((Ee(Tc_ex, i + 1, j, k + 1) + Ee(Tc_ex, i + 1, j + 1, k + 1)) * 0.5));
    PFOR_K(nz + 1) FOR_JI(nx, ny)
        FZ(kz, i, j, k) = FZ(kz, i, j, k) * exp(a * ((Ee(Tc_ex, i + 1, j + 1, k) + Ee(Tc_ex, i + 1, j + 1, k + 1)) * 0.5));
}

/* KillNorthSouthFluxes!  Dimensionless/HydroThermal3D_Alban_v6.jl:153-159 */
EXPORT void o_KillNorthSouthFluxes(double *Ty, int nx, int ny, int nz) {
    PFOR_K(nz) for (int i = 1; i <= nx; ++i) {
        FY(Ty, i, 1, k) = 0.0;
        FY(Ty, i, ny + 1, k) = 0.0;
    }
}

/* Compute_Rp!  Dimensionless/kernels_HT3D.jl:92-97 with @dmul_ya tools/Macros.jl:111 */
EXPORT void o_Compute_Rp(double *Rp, const double *ky, const double *Ty, double Ra, double _dy,
                         int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny)
        Cc(Rp, i, j, k) = Ra * _dy * ((FY(ky, i, j + 1, k) * FY(Ty, i, j + 1, k)) - (FY(ky, i, j, k) * FY(Ty, i, j, k)));
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions (ghost fill)                                          */
/* ------------------------------------------------------------------------ */

/* Ghost fill of an E array.  type 0 = Neumann copy, 1 = Dirichlet
 * "2*val - interior".  Face ghosts (the only ghosts any hot-path stencil reads)
 * are exactly the reference's.  Edges and corners race in the reference kernel
 * (every thread writes its own point but reads ghosts written in the same
 * launch) and are never read downstream; here they are defined without any
 * dependence on earlier ghost values: x faces over the interior (j,k), then y
 * faces over all i and interior k, then z faces over all (i,j) -- every read is
 * of the interior or of a ghost written earlier in the same fill.            */
static void bc_fill6(double *A, int nx, int ny, int nz, const int type[6], const double val[6]) {
    const int s1 = nx + 2, s2 = ny + 2, s3 = nz + 2;
    for (int k = 2; k <= s3 - 1; ++k)
        for (int j = 2; j <= s2 - 1; ++j) {
            Ee(A, 1, j, k)  = type[0] ? 2 * val[0] - Ee(A, 2, j, k)      : Ee(A, 2, j, k);
            Ee(A, s1, j, k) = type[1] ? 2 * val[1] - Ee(A, s1 - 1, j, k) : Ee(A, s1 - 1, j, k);
        }
    for (int k = 2; k <= s3 - 1; ++k)
        for (int i = 1; i <= s1; ++i) {
            Ee(A, i, 1, k)  = type[2] ? 2 * val[2] - Ee(A, i, 2, k)      : Ee(A, i, 2, k);
            Ee(A, i, s2, k) = type[3] ? 2 * val[3] - Ee(A, i, s2 - 1, k) : Ee(A, i, s2 - 1, k);
        }
    for (int j = 1; j <= s2; ++j)
        for (int i = 1; i <= s1; ++i) {
            Ee(A, i, j, 1)  = type[4] ? 2 * val[4] - Ee(A, i, j, 2)      : Ee(A, i, j, 2);
            Ee(A, i, j, s3) = type[5] ? 2 * val[5] - Ee(A, i, j, s3 - 1) : Ee(A, i, j, s3 - 1);
        }
}

/* SetTemperatureBCs!  Alban_v6:108-118 (= SetPressureBCs! kernels_HT3D.jl:144-154):
 * Neumann in x and z, Dirichlet Pbot (y low) / Ptop (y high).                */
EXPORT void o_SetTemperatureBCs(double *Tc_ex, double Tbot, double Ttop, int nx, int ny, int nz) {
    const int type[6] = {0, 0, 1, 1, 0, 0};
    const double val[6] = {0.0, 0.0, Tbot, Ttop, 0.0, 0.0};
    bc_fill6(Tc_ex, nx, ny, nz, type, val);
}

/* SetPressureBCs!  Alban_v6:120-130: Neumann on all six faces (closed box). */
EXPORT void o_SetPressureBCs(double *Pc_ex, double Pbot, double Ptop, int nx, int ny, int nz) {
    (void)Pbot; (void)Ptop;
    const int type[6] = {0, 0, 0, 0, 0, 0};
    const double val[6] = {0, 0, 0, 0, 0, 0};
    bc_fill6(Pc_ex, nx, ny, nz, type, val);
}

/* v7 SetPressureBCs!  kernels_HT3D.jl:144-154 (Dirichlet in y for P). */
EXPORT void o_SetPressureBCs_v7(double *Pc_ex, double Pbot, double Ptop, int nx, int ny, int nz) {
    o_SetTemperatureBCs(Pc_ex, Pbot, Ptop, nx, ny, nz);
}

/* ------------------------------------------------------------------------ */
/* Fluxes, residuals and pseudo-transient updates                            */
/* ------------------------------------------------------------------------ */

/* ComputeFlux!  kernels_HT3D.jl:109-117 (= kernels_HT3D_dimensional.jl:140-146) */
EXPORT void o_ComputeFlux(double *qx, double *qy, double *qz, const double *kx, const double *ky,
                          const double *kz, const double *A, double _dx, double _dy, double _dz,
                          int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx + 1, ny) FX(qx, i, j, k) = -FX(kx, i, j, k) * (_dx * D_XI(A, nx + 2, ny + 2));
    PFOR_K(nz) FOR_JI(nx, ny + 1) FY(qy, i, j, k) = -FY(ky, i, j, k) * (_dy * D_YI(A, nx + 2, ny + 2));
    PFOR_K(nz + 1) FOR_JI(nx, ny) FZ(qz, i, j, k) = -FZ(kz, i, j, k) * (_dz * D_ZI(A, nx + 2, ny + 2));
}

#define DQX (FX(qx, i + 1, j, k) - FX(qx, i, j, k))
#define DQY (FY(qy, i, j + 1, k) - FY(qy, i, j, k))
#define DQZ (FZ(qz, i, j, k + 1) - FZ(qz, i, j, k))

/* ResidualDiffusion!  kernels_HT3D.jl:128-134 */
EXPORT void o_ResidualDiffusion(double *F, const double *R, const double *qx, const double *qy,
                                const double *qz, double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny)
        Cc(F, i, j, k) = Cc(R, i, j, k) + _dx * DQX + _dy * DQY + _dz * DQZ;
}

/* ResidualDiffusionTransient!  Alban_v6:84-90 */
EXPORT void o_ResidualDiffusionTransient(double *F, const double *T, const double *R, const double *qx,
                                         const double *qy, const double *qz, double _dt, double _dx,
                                         double _dy, double _dz, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny)
        Cc(F, i, j, k) = Cc(R, i, j, k) + _dt * Ee(T, i + 1, j + 1, k + 1) + _dx * DQX + _dy * DQY + _dz * DQZ;
}

/* Jacobi-preconditioned residuals used by the DYREL solves.  These compose the
 * dimensionless residuals above with the "F /= PC" step that closes every
 * residual of the dimensional DYREL path (kernels_HT3D_dimensional.jl:175,283). */
EXPORT void o_ResidualDiffusionPC(double *F, const double *R, const double *qx, const double *qy,
                                  const double *qz, const double *PC, double _dx, double _dy, double _dz,
                                  int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double f = Cc(R, i, j, k) + _dx * DQX + _dy * DQY + _dz * DQZ;
        Cc(F, i, j, k) = f / Cc(PC, i, j, k);
    }
}

EXPORT void o_ResidualDiffusionTransientPC(double *F, const double *T, const double *R, const double *qx,
                                           const double *qy, const double *qz, const double *PC, double _dt,
                                           double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double f = Cc(R, i, j, k) + _dt * Ee(T, i + 1, j + 1, k + 1) + _dx * DQX + _dy * DQY + _dz * DQZ;
        Cc(F, i, j, k) = f / Cc(PC, i, j, k);
    }
}

/* UpdateT!  kernels_HT3D.jl:119-126 (plain Richardson) */
EXPORT void o_UpdateT(double *F, double *T, const double *R, const double *qx, const double *qy,
                      const double *qz, double _dt, double dtau, double _dx, double _dy, double _dz,
                      int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        Cc(F, i, j, k) = Cc(R, i, j, k) + _dt * Ee(T, i + 1, j + 1, k + 1) + _dx * DQX + _dy * DQY + _dz * DQZ;
        Ee(T, i + 1, j + 1, k + 1) = Ee(T, i + 1, j + 1, k + 1) - dtau * Cc(F, i, j, k);
    }
}

/* UpdateP!  kernels_HT3D.jl:136-142 (= DampedUpdate! kernels_HT3D_dimensional.jl:316-320) */
EXPORT void o_UpdateP(double *F0, double *P, const double *F, double dampx, double dtau, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        Cc(F0, i, j, k) = Cc(F, i, j, k) + dampx * Cc(F0, i, j, k);
        Ee(P, i + 1, j + 1, k + 1) = Ee(P, i + 1, j + 1, k + 1) - dtau * Cc(F0, i, j, k);
    }
}

/* DYRELUpdate!  Dimensional/kernels_HT3D_dimensional.jl:322-326 */
EXPORT void o_DYRELUpdate(double *F0, double *X, const double *F, double h1, double h2, double dtau,
                          int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        Cc(F0, i, j, k) = h1 * Cc(F0, i, j, k) + h2 * Cc(F, i, j, k);
        Ee(X, i + 1, j + 1, k + 1) = Ee(X, i + 1, j + 1, k + 1) - dtau * Cc(F0, i, j, k);
    }
}

/* SwapDYREL!  kernels_HT3D_dimensional.jl:122-126: @all(Xc) = @inn(X0c_ex),
 * @all(Yc) = @all(Y0c).  The solve loops below pass Xc = NULL: the drivers
 * never read Xcit (HydroThermal3D_v18.jl:322,451,509), so only Fcit is kept. */
EXPORT void o_SwapDYREL(double *Xc, double *Yc, const double *X0c_ex, const double *Y0c, int nx, int ny, int nz) {
    if (Xc) {
        PFOR_K(nz) FOR_JI(nx, ny) Cc(Xc, i, j, k) = Ee(X0c_ex, i + 1, j + 1, k + 1);
    }
    memcpy(Yc, Y0c, (size_t)nx * ny * nz * sizeof(double));
}

/* GershgorinPoisson!  kernels_HT3D_dimensional.jl:207-221; the expression is the
 * stored SymPy output of notebooks/Gershgorin_TH3D.ipynb (cell 2).           */
EXPORT void o_GershgorinPoisson(double *G, double *D, const double *rhoC, const double *kx, const double *ky,
                                const double *kz, double transient, double dt, double dx, double dy, double dz,
                                int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double kW = FX(kx, i, j, k), kE = FX(kx, i + 1, j, k);
        double kS = FY(ky, i, j, k), kN = FY(ky, i, j + 1, k);
        double kF = FZ(kz, i, j, k + 1), kB = FZ(kz, i, j, k);
        double rc = Cc(rhoC, i, j, k);
        double diag = (kB / dz + kF / dz) / dz + (kN / dy + kS / dy) / dy + (kE / dx + kW / dx) / dx + rc * transient / dt;
        double g = fabs(kE / (dx * dx)) + fabs(kW / (dx * dx)) + fabs(kN / (dy * dy)) + fabs(kS / (dy * dy)) +
                   fabs(kB / (dz * dz)) + fabs(kF / (dz * dz)) + fabs(diag);
        Cc(D, i, j, k) = diag;
        Cc(G, i, j, k) = g / diag;
    }
}

/* GershgorinHydro!  kernels_HT3D_dimensional.jl:335-347 (no division by D) */
EXPORT void o_GershgorinHydro(double *G, const double *phidrhodP, const double *kx, const double *ky,
                              const double *kz, double transient, double dt, double dx, double dy, double dz,
                              int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double kW = FX(kx, i, j, k), kE = FX(kx, i + 1, j, k);
        double kS = FY(ky, i, j, k), kN = FY(ky, i, j + 1, k);
        double kF = FZ(kz, i, j, k + 1), kB = FZ(kz, i, j, k);
        double rc = Cc(phidrhodP, i, j, k);
        Cc(G, i, j, k) = fabs(kE / (dx * dx)) + fabs(kW / (dx * dx)) + fabs(kN / (dy * dy)) + fabs(kS / (dy * dy)) +
                         fabs(kB / (dz * dz)) + fabs(kF / (dz * dz)) +
                         fabs((kB / dz + kF / dz) / dz + (kN / dy + kS / dy) / dy + (kE / dx + kW / dx) / dx + rc * transient / dt);
    }
}

/* RayleighQuotientNumerator! / Denominator!  kernels_HT3D_dimensional.jl:349-363 */
EXPORT void o_RayleighQuotientNumerator(double *G, const double *Fc0, const double *Fcit, const double *Fc,
                                        double dtau, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double dx_ = dtau * Cc(Fc0, i, j, k);
        Cc(G, i, j, k) = (-dx_) * (Cc(Fc, i, j, k) - Cc(Fcit, i, j, k));
    }
}

EXPORT void o_RayleighQuotientDenominator(double *G, const double *Fc0, double dtau, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double dx_ = dtau * Cc(Fc0, i, j, k);
        Cc(G, i, j, k) = dx_ * dx_;
    }
}

/* ResidualTemperatureLinearised!  kernels_HT3D_dimensional.jl:162-181 */
EXPORT void o_ResidualTemperatureLinearised(double *F, const double *Tc_ex, const double *Tc_old, const double *rhoC_eff,
                                            const double *phc_ex, const double *PC, const double *qx, const double *qy,
                                            const double *qz, double Qt, double transient, double _dt, double _dx,
                                            double _dy, double _dz, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double ph = Ee(phc_ex, i + 1, j + 1, k + 1);
        if (fabs(ph) != 1.0) {
            double Q = 0.0;
            if (ph == 3.0) Q = Qt;
            double f = transient * Cc(rhoC_eff, i, j, k) * _dt * (Ee(Tc_ex, i + 1, j + 1, k + 1) - Cc(Tc_old, i, j, k));
            f += DQX * _dx;
            f += DQY * _dy;
            f += DQZ * _dz;
            f -= transient * Q;
            f /= Cc(PC, i, j, k);
            Cc(F, i, j, k) = f;
        } else {
            Cc(F, i, j, k) = 0.0;
        }
    }
}

/* ResidualFluidPressure!  kernels_HT3D_dimensional.jl:274-289 */
EXPORT void o_ResidualFluidPressure(double *F, const double *phc_ex, const double *phirhof, const double *phirhof0,
                                    const double *PC, const double *qx, const double *qy, const double *qz,
                                    double transient, double _dt, double _dx, double _dy, double _dz,
                                    int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double ph = Ee(phc_ex, i + 1, j + 1, k + 1);
        if (fabs(ph) != 1.0) {
            double f = transient * _dt * (Cc(phirhof, i, j, k) - Cc(phirhof0, i, j, k));
            f += DQX * _dx;
            f += DQY * _dy;
            f += DQZ * _dz;
            f /= Cc(PC, i, j, k);
            Cc(F, i, j, k) = f;
        } else {
            Cc(F, i, j, k) = 0.0;
        }
    }
}

/* ComputeDarcyFlux!  kernels_HT3D_dimensional.jl:266-272 */
EXPORT void o_ComputeDarcyFlux(double *qx, double *qy, double *qz, const double *rhofv, const double *kx,
                               const double *ky, const double *kz, const double *Pf, double g, double _dx,
                               double _dy, double _dz, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx + 1, ny) FX(qx, i, j, k) = -FX(kx, i, j, k) * (_dx * D_XI(Pf, nx + 2, ny + 2));
    PFOR_K(nz) FOR_JI(nx, ny + 1)
        FY(qy, i, j, k) = -FY(ky, i, j, k) * (_dy * D_YI(Pf, nx + 2, ny + 2) - g * AV_XZA(rhofv, nx + 1, ny + 1));
    PFOR_K(nz + 1) FOR_JI(nx, ny) FZ(qz, i, j, k) = -FZ(kz, i, j, k) * (_dz * D_ZI(Pf, nx + 2, ny + 2));
}

/* Init_vel!  Dimensionless/kernels_HT3D.jl:73-80 (note ((-k)*_dx)*dP order) */
EXPORT void o_Init_vel(double *Vx, double *Vy, double *Vz, const double *kx, const double *ky, const double *kz,
                       const double *Pc_ex, const double *Ty, double Ra, double _dx, double _dy, double _dz,
                       int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx + 1, ny) FX(Vx, i, j, k) = -FX(kx, i, j, k) * _dx * D_XI(Pc_ex, nx + 2, ny + 2);
    PFOR_K(nz) FOR_JI(nx, ny + 1)
        FY(Vy, i, j, k) = -FY(ky, i, j, k) * (_dy * D_YI(Pc_ex, nx + 2, ny + 2) - Ra * FY(Ty, i, j, k));
    PFOR_K(nz + 1) FOR_JI(nx, ny) FZ(Vz, i, j, k) = -FZ(kz, i, j, k) * _dz * D_ZI(Pc_ex, nx + 2, ny + 2);
}

/* ------------------------------------------------------------------------ */
/* WENO5 v1: tools/Weno5_Routines.jl                                          */
/* ------------------------------------------------------------------------ */

/* V?PlusMinus!  tools/Weno5_Routines.jl:5-42.  Vm/Vp are (nx,ny,nz); the
 * untouched branch (V == 0) keeps the value left by ResetA!.                  */
EXPORT void o_VxPlusMinus(double *Vxm, double *Vxp, const double *Vx, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double vl = FX(Vx, i, j, k), vr = FX(Vx, i + 1, j, k);
        if (vl < 0.0) Cc(Vxm, i, j, k) = vl;
        if (vl > 0.0) Cc(Vxm, i, j, k) = 0.0;
        if (vr < 0.0) Cc(Vxp, i, j, k) = 0.0;
        if (vr > 0.0) Cc(Vxp, i, j, k) = vr;
    }
}

EXPORT void o_VyPlusMinus(double *Vym, double *Vyp, const double *Vy, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double vl = FY(Vy, i, j, k), vr = FY(Vy, i, j + 1, k);
        if (vl < 0.0) Cc(Vym, i, j, k) = vl;
        if (vl > 0.0) Cc(Vym, i, j, k) = 0.0;
        if (vr < 0.0) Cc(Vyp, i, j, k) = 0.0;
        if (vr > 0.0) Cc(Vyp, i, j, k) = vr;
    }
}

EXPORT void o_VzPlusMinus(double *Vzm, double *Vzp, const double *Vz, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        double vl = FZ(Vz, i, j, k), vr = FZ(Vz, i, j, k + 1);
        if (vl < 0.0) Cc(Vzm, i, j, k) = vl;
        if (vl > 0.0) Cc(Vzm, i, j, k) = 0.0;
        if (vr < 0.0) Cc(Vzp, i, j, k) = 0.0;
        if (vr > 0.0) Cc(Vzp, i, j, k) = vr;
    }
}

/* dFdx_Weno5!  tools/Weno5_Routines.jl:44-68 (WENO5-JS).  The reference writes
 * eps = 10^(-99) + 1e-6*maxV; 10^(-99) is taken as the literal 1e-99.          */
static inline double weno5_js(double v1, double v2, double v3, double v4, double v5) {
    double p1 = v1 / 3.0 - 7.0 / 6.0 * v2 + 11.0 / 6.0 * v3;
    double p2 = -v2 / 6.0 + 5.0 / 6.0 * v3 + v4 / 3.0;
    double p3 = v3 / 3.0 + 5.0 / 6.0 * v4 - v5 / 6.0;
    double maxV = fmax(fmax(fmax(fmax(v1 * v1, v2 * v2), v3 * v3), v4 * v4), v5 * v5);
    double e = 1e-99 + 1e-6 * maxV;
    double a1 = v1 - 2.0 * v2 + v3, b1 = v1 - 4.0 * v2 + 3.0 * v3;
    double a2 = v2 - 2.0 * v3 + v4, b2 = v2 - v4;
    double a3 = v3 - 2.0 * v4 + v5, b3 = 3.0 * v3 - 4.0 * v4 + v5;
    double w1 = 13.0 / 12.0 * (a1 * a1) + 1.0 / 4.0 * (b1 * b1);
    double w2 = 13.0 / 12.0 * (a2 * a2) + 1.0 / 4.0 * (b2 * b2);
    double w3 = 13.0 / 12.0 * (a3 * a3) + 1.0 / 4.0 * (b3 * b3);
    w1 = 0.1 / ((w1 + e) * (w1 + e));
    w2 = 0.6 / ((w2 + e) * (w2 + e));
    w3 = 0.3 / ((w3 + e) * (w3 + e));
    double w = (w1 + w2 + w3);
    w1 = w1 / w;
    w2 = w2 / w;
    w3 = w3 / w;
    return w1 * p1 + w2 * p2 + w3 * p3;
}

EXPORT double o_weno5_js(double v1, double v2, double v3, double v4, double v5) { return weno5_js(v1, v2, v3, v4, v5); }

EXPORT void o_dFdx_Weno5(double *dFdxi, const double *V1, const double *V2, const double *V3, const double *V4,
                         const double *V5, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny)
        Cc(dFdxi, i, j, k) = weno5_js(Cc(V1, i, j, k), Cc(V2, i, j, k), Cc(V3, i, j, k), Cc(V4, i, j, k), Cc(V5, i, j, k));
}

/* Gradients_{minus,plus}_{x,y,z}_Weno5!  tools/Weno5_Routines.jl:70-128 with the
 * @in_xxx_* offsets of tools/Macros.jl:88-109 (centre at (i+3,j+3,k+3)).
 * axis: 1,2,3.  plus: 0 = minus (left-biased), 1 = plus (right-biased).       */
EXPORT void o_Gradients_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                              int axis, int plus, double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    const int ox = axis == 1, oy = axis == 2, oz = axis == 3;
    const double h = axis == 1 ? _dx : (axis == 2 ? _dy : _dz);
#define XS(o) X6(Fc_exxx, i + 3 + (o) * ox, j + 3 + (o) * oy, k + 3 + (o) * oz)
    PFOR_K(nz) FOR_JI(nx, ny) {
        if (!plus) {
            Cc(v1, i, j, k) = h * (XS(-2) - XS(-3));
            Cc(v2, i, j, k) = h * (XS(-1) - XS(-2));
            Cc(v3, i, j, k) = h * (XS(0) - XS(-1));
            Cc(v4, i, j, k) = h * (XS(1) - XS(0));
            Cc(v5, i, j, k) = h * (XS(2) - XS(1));
        } else {
            Cc(v1, i, j, k) = h * (XS(3) - XS(2));
            Cc(v2, i, j, k) = h * (XS(2) - XS(1));
            Cc(v3, i, j, k) = h * (XS(1) - XS(0));
            Cc(v4, i, j, k) = h * (XS(0) - XS(-1));
            Cc(v5, i, j, k) = h * (XS(-1) - XS(-2));
        }
    }
#undef XS
}

/* Boundaries_{x,y,z}_Weno5!  tools/Weno5_Routines.jl:130-255 (v1, source c array)
 * and tools/Weno5_Routines_v2.jl:221-347 (v2, source E array: src_ext = 1).
 * Copies the interior into X6 and fills the three ghost layers normal to
 * `axis` over the interior of the other two directions.
 * type: 0 Neumann (mirror), 1 Dirichlet (2*val - mirror), 2 periodic.        */
EXPORT void o_Boundaries_Weno5(double *Fc_exxx, const double *Fc, int axis, int src_ext, int type_lo, double val_lo,
                               int type_hi, double val_hi, int nx, int ny, int nz) {
    const int o = src_ext ? 1 : 0;
    const int s1 = src_ext ? nx + 2 : nx, s2 = src_ext ? ny + 2 : ny;
#define SRC(a, b, c) AT(Fc, s1, s2, (a) + o, (b) + o, (c) + o)
    PFOR_K(nz) FOR_JI(nx, ny) X6(Fc_exxx, i + 3, j + 3, k + 3) = SRC(i, j, k);
    const int n = axis == 1 ? nx : (axis == 2 ? ny : nz);
    /* ghost layer g = 1..3 on the low side sits at X6 index 4-g and mirrors
     * interior index g; on the high side at X6 index n+3+g mirrors n+1-g.
     * Periodic: low ghost g takes interior n+1-g... per the reference:
     * X6[1] = Fc[n-2], X6[2] = Fc[n-1], X6[3] = Fc[n];  X6[n+6] = Fc[3], ...  */
    for (int g = 1; g <= 3; ++g) {
        const int lo_dst = 4 - g, hi_dst = n + 3 + g;
        const int lo_mir = g, hi_mir = n + 1 - g;
        const int lo_per = n + 1 - g, hi_per = g;
        if (axis == 1) {
            for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) {
                double a = SRC(lo_mir, j, k), b = SRC(hi_mir, j, k);
                X6(Fc_exxx, lo_dst, j + 3, k + 3) = type_lo == 0 ? a : (type_lo == 1 ? 2 * val_lo - a : SRC(lo_per, j, k));
                X6(Fc_exxx, hi_dst, j + 3, k + 3) = type_hi == 0 ? b : (type_hi == 1 ? 2 * val_hi - b : SRC(hi_per, j, k));
            }
        } else if (axis == 2) {
            for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) {
                double a = SRC(i, lo_mir, k), b = SRC(i, hi_mir, k);
                X6(Fc_exxx, i + 3, lo_dst, k + 3) = type_lo == 0 ? a : (type_lo == 1 ? 2 * val_lo - a : SRC(i, lo_per, k));
                X6(Fc_exxx, i + 3, hi_dst, k + 3) = type_hi == 0 ? b : (type_hi == 1 ? 2 * val_hi - b : SRC(i, hi_per, k));
            }
        } else {
            for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) {
                double a = SRC(i, j, lo_mir), b = SRC(i, j, hi_mir);
                X6(Fc_exxx, i + 3, j + 3, lo_dst) = type_lo == 0 ? a : (type_lo == 1 ? 2 * val_lo - a : SRC(i, j, lo_per));
                X6(Fc_exxx, i + 3, j + 3, hi_dst) = type_hi == 0 ? b : (type_hi == 1 ? 2 * val_hi - b : SRC(i, j, hi_per));
            }
        }
    }
#undef SRC
}

/* Advect!  tools/Weno5_Routines.jl:257-262 */
EXPORT void o_Advect(double *Fc, const double *Vxp, const double *dTdxm, const double *Vxm, const double *dTdxp,
                     double dt, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny)
        Cc(Fc, i, j, k) = Cc(Fc, i, j, k) - dt * (Cc(Vxp, i, j, k) * Cc(dTdxm, i, j, k) + Cc(Vxm, i, j, k) * Cc(dTdxp, i, j, k));
}

/* TimeAveraging!  tools/Weno5_Routines.jl:264-269 (c arrays) */
EXPORT void o_TimeAveraging(double *Fc, const double *Fold, double order, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny)
        Cc(Fc, i, j, k) = (1.0 / order) * Cc(Fc, i, j, k) + (1.0 - 1.0 / order) * Cc(Fold, i, j, k);
}

/* TimeAveraging!  tools/Weno5_Routines_v2.jl:214-219 (Fc is E, Fold is c) */
EXPORT void o_TimeAveraging_v2(double *Fc, const double *Fold, double order, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny)
        Ee(Fc, i + 1, j + 1, k + 1) = (1.0 / order) * Ee(Fc, i + 1, j + 1, k + 1) + (1.0 - 1.0 / order) * Cc(Fold, i, j, k);
}

/* Advect!  tools/Weno5_Routines_v2.jl:55-212: upwind split, both WENO5 stencils
 * and the forward-Euler stage fused per cell.  The minus-side fifth difference
 * is X[I+2]-X[I-1] (Weno5_Routines_v2.jl:69,119,169), kept as in the reference. */
EXPORT void o_Advect_v2(double *Fc, const double *X, const double *V, int axis, double dt, double _dx, double _dy,
                        double _dz, int nx, int ny, int nz) {
    const int ox = axis == 1, oy = axis == 2, oz = axis == 3;
    const double h = axis == 1 ? _dx : (axis == 2 ? _dy : _dz);
    const int v1 = axis == 1 ? nx + 1 : nx, v2 = axis == 2 ? ny + 1 : ny;
#define XS(o) X6(X, i + 3 + (o) * ox, j + 3 + (o) * oy, k + 3 + (o) * oz)
    PFOR_K(nz) FOR_JI(nx, ny) {
        double vl = AT(V, v1, v2, i, j, k), vr = AT(V, v1, v2, i + ox, j + oy, k + oz);
        double Vm = vl < 0.0 ? vl : 0.0;
        double Vp = vr >= 0.0 ? vr : 0.0;
        double dm = weno5_js(h * (XS(-2) - XS(-3)), h * (XS(-1) - XS(-2)), h * (XS(0) - XS(-1)), h * (XS(1) - XS(0)),
                             h * (XS(2) - XS(-1)));
        double dp = weno5_js(h * (XS(3) - XS(2)), h * (XS(2) - XS(1)), h * (XS(1) - XS(0)), h * (XS(0) - XS(-1)),
                             h * (XS(-1) - XS(-2)));
        Ee(Fc, i + 1, j + 1, k + 1) = Ee(Fc, i + 1, j + 1, k + 1) - dt * (Vp * dm + Vm * dp);
    }
#undef XS
}

/* ------------------------------------------------------------------------ */
/* Reductions (maximum_g / minimum_g / mean_g / sum_g, tools/Macros.jl:20-25) */
/* ------------------------------------------------------------------------ */


/* Sums are accumulated in double-double (TwoSum, no FMA), so the rounded result
 * is the correctly rounded sum except in vanishingly rare ties, whatever the
 * order; maxima are exact.  (Julia's pairwise `sum` is a different rounding of
 * the same sum; the difference is at the last bit.)                          */
static inline void dd_add(double *hi, double *lo, double x) {
    double s = *hi + x, bp = s - *hi;
    double e = (*hi - (s - bp)) + (x - bp);
    *hi = s;
    *lo += e;
}

static inline void dd_combine(double *hi, double *lo, double h2, double l2) {
    double s = *hi + h2, bp = s - *hi;
    double e = (*hi - (s - bp)) + (h2 - bp);
    e += *lo + l2;
    *hi = s + e;
    *lo = e - (*hi - s);
}

EXPORT double o_reduce(const double *a, size_t n1, size_t n23, int op) {
    double *part = (double *)malloc(2 * n23 * sizeof(double));
    const int is_sum = (op == R_SUM || op == R_SUM_ABS);
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)n23; ++p) {
        const double *r = a + (size_t)p * n1;
        double hi = is_sum ? 0.0 : (op == R_MIN ? INFINITY : -INFINITY), lo = 0.0;
        for (size_t i = 0; i < n1; ++i) {
            double v = r[i];
            switch (op) {
            case R_SUM: dd_add(&hi, &lo, v); break;
            case R_SUM_ABS: dd_add(&hi, &lo, fabs(v)); break;
            case R_MAX: hi = (v > hi || v != v) ? v : hi; break;
            case R_MIN: hi = (v < hi || v != v) ? v : hi; break;
            default: hi = (fabs(v) > hi || v != v) ? fabs(v) : hi; break;
            }
        }
        part[2 * p] = hi;
        part[2 * p + 1] = lo;
    }
    double hi = is_sum ? 0.0 : (op == R_MIN ? INFINITY : -INFINITY), lo = 0.0;
    for (size_t p = 0; p < n23; ++p) {
        double v = part[2 * p];
        if (is_sum) dd_combine(&hi, &lo, v, part[2 * p + 1]);
        else if (op == R_MIN) hi = (v < hi || v != v) ? v : hi;
        else hi = (v > hi || v != v) ? v : hi;
    }
    free(part);
    return is_sum ? hi + lo : hi;
}

/* ------------------------------------------------------------------------ */
/* DYREL solve: Dimensional/HydroThermal3D_v18.jl:424-538                    */
/* ------------------------------------------------------------------------ */

/* o_dyrel_params and o_dyrel_report: ht3d_oracle_common.h */

typedef struct {
    int nx, ny, nz;
    double *X;             /* E: unknown with ghosts (Tc_ex or Pc_ex) */
    const double *R;       /* c: Rt or Rp */
    const double *kx, *ky, *kz;
    double *PC;            /* c: Gershgorin diagonal D (output of set-up) */
    double *G, *F, *F0, *Fcit, *dumc, *qx, *qy, *qz;
    int transient;         /* 1: thermal (R + _dt*T + div q), 0: Darcy (R + div q) */
    double rhoC;           /* rhoC for the Gershgorin diagonal */
    double dt, _dt, dx, dy, dz;
    int bc_type[6];
    double bc_val[6];
} o_dyrel_field;

static inline int is_lambda_iter(int iter, int nitout) { return iter <= 2 || iter % nitout == 0; }

EXPORT int o_dyrel_solve(o_dyrel_field *f, const o_dyrel_params *p, o_dyrel_report *rep) {
    const int nx = f->nx, ny = f->ny, nz = f->nz;
    const size_t N = (size_t)nx * ny * nz;
    const double _dx = 1.0 / f->dx, _dy = 1.0 / f->dy, _dz = 1.0 / f->dz;
    /* set-up: GershgorinPoisson! with a uniform rhoC (v18:432-439) */
    double *rc = f->dumc;
    for (size_t n = 0; n < N; ++n) rc[n] = f->rhoC;
    o_GershgorinPoisson(f->G, f->PC, rc, f->kx, f->ky, f->kz, (double)f->transient, f->dt, f->dx, f->dy, f->dz, nx, ny, nz);
    double lmax = o_reduce(f->G, (size_t)nx, (size_t)ny * nz, R_MAX);
    double lmin = lmax / p->lmin_div;
    double dtau = 2.0 / sqrt(lmax) * p->CFL;
    double c = 2.0 * sqrt(lmin) * p->cfact;
    double h1 = (2 - c * dtau) / (2 + c * dtau), h2 = 2 * dtau / (2 + c * dtau);
    o_ResetA(f->F, f->F0, nx, ny, nz);
    double nF_abs = 0.0, nF_rel = 0.0, nF_ini = 0.0;
    int status = 10, iter, n_checks = 0, n_lambda = 0;
    for (iter = 1; iter <= p->nitmax; ++iter) {
        const int lam = is_lambda_iter(iter, p->nitout);
        const int chk = lam || (p->ncheck > 0 && iter % p->ncheck == 0);
        if (lam) o_SwapDYREL(NULL, f->Fcit, f->X, f->F, nx, ny, nz);
        bc_fill6(f->X, nx, ny, nz, f->bc_type, f->bc_val);
        o_ComputeFlux(f->qx, f->qy, f->qz, f->kx, f->ky, f->kz, f->X, _dx, _dy, _dz, nx, ny, nz);
        if (f->transient)
            o_ResidualDiffusionTransientPC(f->F, f->X, f->R, f->qx, f->qy, f->qz, f->PC, f->_dt, _dx, _dy, _dz, nx, ny, nz);
        else
            o_ResidualDiffusionPC(f->F, f->R, f->qx, f->qy, f->qz, f->PC, _dx, _dy, _dz, nx, ny, nz);
        o_DYRELUpdate(f->F0, f->X, f->F, h1, h2, dtau, nx, ny, nz);
        if (chk) {
            ++n_checks;
            o_Multiply(f->dumc, f->F, f->PC, nx, ny, nz);
            if (p->norm == 0)
                nF_abs = o_reduce(f->dumc, (size_t)nx, (size_t)ny * nz, R_SUM_ABS) / (double)N / sqrt((double)N);
            else
                nF_abs = o_reduce(f->dumc, (size_t)nx, (size_t)ny * nz, R_MAX_ABS);
            if (iter == 1) nF_ini = nF_abs;
            nF_rel = nF_abs / nF_ini;
            if (isnan(nF_abs)) { status = 11; break; }
            if (nF_rel > 100) { status = 12; break; }
            if (lam && iter > 1) {
                o_RayleighQuotientNumerator(f->dumc, f->F0, f->Fcit, f->F, dtau, nx, ny, nz);
                lmin = fabs(o_reduce(f->dumc, (size_t)nx, (size_t)ny * nz, R_SUM));
                o_RayleighQuotientDenominator(f->dumc, f->F0, dtau, nx, ny, nz);
                lmin /= fabs(o_reduce(f->dumc, (size_t)nx, (size_t)ny * nz, R_SUM));
                c = 2.0 * sqrt(lmin) * p->cfact;
                h1 = (2 - c * dtau) / (2 + c * dtau);
                h2 = 2 * dtau / (2 + c * dtau);
                ++n_lambda;
            }
            if (nF_abs < p->tol || nF_rel < p->tol) { status = 0; break; }
        }
    }
    if (rep) {
        rep->iters = iter > p->nitmax ? p->nitmax : iter;
        rep->status = status;
        rep->n_checks = n_checks;
        rep->n_lambda = n_lambda;
        rep->lmax = lmax; rep->lmin = lmin; rep->dtau = dtau; rep->h1 = h1; rep->h2 = h2;
        rep->nF_abs = nF_abs; rep->nF_rel = nF_rel; rep->nF_ini = nF_ini;
    }
    return status;
}

/* ------------------------------------------------------------------------ */
/* WENO5 orchestration: kernels_HT3D.jl:156-233 (v1), Weno5_Routines_v2.jl:1-50 */
/* ------------------------------------------------------------------------ */

typedef struct {
    double *Tc, *Told, *Tc_exxx, *v1, *v2, *v3, *v4, *v5, *dTdxm, *dTdxp;
    double *Vxm, *Vxp, *Vym, *Vyp, *Vzm, *Vzp;
} o_weno_work;

/* closed-box WENO boundary types: x Neumann, y Dirichlet (Tbot, Ttop), z Neumann
 * (kernels_HT3D.jl:164-177, Weno5_Routines_v2.jl:8-21)                          */
static void weno_bc(int axis, double Tbot, double Ttop, int *tlo, double *vlo, int *thi, double *vhi) {
    if (axis == 2) { *tlo = 1; *vlo = Tbot; *thi = 1; *vhi = Ttop; }
    else { *tlo = 0; *vlo = 0.0; *thi = 0; *vhi = 0.0; }
}

/* AdvectWithWeno5 (v1) after Init_vel!: Tc_ex advanced by one SSP-RK2 step per direction. */
EXPORT void o_AdvectWithWeno5_v1(double *Tc_ex, const double *Vx, const double *Vy, const double *Vz,
                                 o_weno_work *w, double dt, double _dx, double _dy, double _dz, double Ttop,
                                 double Tbot, int nx, int ny, int nz) {
    const double order = 2.0;
    o_ResetA(w->Vxm, w->Vxp, nx, ny, nz); o_VxPlusMinus(w->Vxm, w->Vxp, Vx, nx, ny, nz);
    o_ResetA(w->Vym, w->Vyp, nx, ny, nz); o_VyPlusMinus(w->Vym, w->Vyp, Vy, nx, ny, nz);
    o_ResetA(w->Vzm, w->Vzp, nx, ny, nz); o_VzPlusMinus(w->Vzm, w->Vzp, Vz, nx, ny, nz);
    o_Cpy_inn_to_all(w->Tc, Tc_ex, nx, ny, nz);
    for (int axis = 1; axis <= 3; ++axis) {
        const double *Vp = axis == 1 ? w->Vxp : (axis == 2 ? w->Vyp : w->Vzp);
        const double *Vm = axis == 1 ? w->Vxm : (axis == 2 ? w->Vym : w->Vzm);
        int tlo, thi; double vlo, vhi;
        weno_bc(axis, Tbot, Ttop, &tlo, &vlo, &thi, &vhi);
        o_ArrayEqualArray(w->Told, w->Tc, nx, ny, nz);
        for (int io = 1; io <= 2; ++io) {
            o_Boundaries_Weno5(w->Tc_exxx, w->Tc, axis, 0, tlo, vlo, thi, vhi, nx, ny, nz);
            o_Gradients_Weno5(w->v1, w->v2, w->v3, w->v4, w->v5, w->Tc_exxx, axis, 0, _dx, _dy, _dz, nx, ny, nz);
            o_dFdx_Weno5(w->dTdxm, w->v1, w->v2, w->v3, w->v4, w->v5, nx, ny, nz);
            o_Gradients_Weno5(w->v1, w->v2, w->v3, w->v4, w->v5, w->Tc_exxx, axis, 1, _dx, _dy, _dz, nx, ny, nz);
            o_dFdx_Weno5(w->dTdxp, w->v1, w->v2, w->v3, w->v4, w->v5, nx, ny, nz);
            o_Advect(w->Tc, Vp, w->dTdxm, Vm, w->dTdxp, dt, nx, ny, nz);
        }
        o_TimeAveraging(w->Tc, w->Told, order, nx, ny, nz);
    }
    o_Cpy_all_to_inn(Tc_ex, w->Tc, nx, ny, nz);
}

/* AdvectWithWeno5_v2  tools/Weno5_Routines_v2.jl:1-50 */
EXPORT void o_AdvectWithWeno5_v2(double *Tc_ex, double *Tc_exxx, double *Told, const double *Vx, const double *Vy,
                                 const double *Vz, double dt, double _dx, double _dy, double _dz, double Ttop,
                                 double Tbot, int nx, int ny, int nz) {
    const double order = 2.0;
    for (int axis = 1; axis <= 3; ++axis) {
        const double *V = axis == 1 ? Vx : (axis == 2 ? Vy : Vz);
        int tlo, thi; double vlo, vhi;
        weno_bc(axis, Tbot, Ttop, &tlo, &vlo, &thi, &vhi);
        o_Cpy_inn_to_all(Told, Tc_ex, nx, ny, nz);
        for (int io = 1; io <= 2; ++io) {
            o_Boundaries_Weno5(Tc_exxx, Tc_ex, axis, 1, tlo, vlo, thi, vhi, nx, ny, nz);
            o_Advect_v2(Tc_ex, Tc_exxx, V, axis, dt, _dx, _dy, _dz, nx, ny, nz);
        }
        o_TimeAveraging_v2(Tc_ex, Told, order, nx, ny, nz);
    }
}

/* ------------------------------------------------------------------------ */
/* Closed-box Horton-Rogers-Lapwood physical step with DYREL solves            */
/*   step order: HydroThermal3D_Alban_v6.jl:322-466 / HydroThermal3D_v7.jl:167-222 */
/*   time step:  rule 0 HydroThermal3D_v18.jl:553-562 (advect with dt_adv from  */
/*               this step's V), 1 v18:563-570 (constant dt, sub-stepped),     */
/*               2 HydroThermal3D_v7.jl:215-221 (advect with dt, then dt_adv)  */
/*   solves:     HydroThermal3D_v18.jl:424-538                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz;
    double Lx, Ly, Lz, dx, dy, dz;
    double Ra, Tbot, Ttop, cfl, dt, time, Nu, mu_a;
    int weno_version, dt_rule, n_adv;
    /* fields */
    double *Tc_ex, *Pc_ex, *ktv, *kfv, *Ty, *kx, *ky, *kz, *qx, *qy, *qz, *Vx, *Vy, *Vz;
    double *Rt, *Rp, *F, *F0, *Fcit, *PC, *G, *dumc;
    o_weno_work w;
    o_dyrel_params pT, pP;
    o_dyrel_report rT, rP;
} o_hrl_model;

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

EXPORT o_hrl_model *o_hrl_create(int nx, int ny, int nz, double Lx, double Ly, double Lz, double Ra, double cfl) {
    o_hrl_model *m = (o_hrl_model *)calloc(1, sizeof(o_hrl_model));
    m->nx = nx; m->ny = ny; m->nz = nz;
    m->Lx = Lx; m->Ly = Ly; m->Lz = Lz;
    m->dx = Lx / nx; m->dy = Ly / ny; m->dz = Lz / nz;
    m->Ra = Ra; m->Tbot = 1.0; m->Ttop = 0.0; m->cfl = cfl;
    double h = fmin(fmin(m->dx, m->dy), m->dz);
    m->dt = h * h / 6.1; /* HydroThermal3D_v7.jl:68 */
    m->weno_version = 1;
    const size_t c = (size_t)nx * ny * nz, E = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t V = (size_t)(nx + 1) * (ny + 1) * (nz + 1), X = (size_t)(nx + 6) * (ny + 6) * (nz + 6);
    const size_t fx = (size_t)(nx + 1) * ny * nz, fy = (size_t)nx * (ny + 1) * nz, fz = (size_t)nx * ny * (nz + 1);
    m->Tc_ex = zalloc(E); m->Pc_ex = zalloc(E); m->ktv = zalloc(V); m->kfv = zalloc(V);
    for (size_t n = 0; n < V; ++n) { m->ktv[n] = 1.0; m->kfv[n] = 1.0; }
    m->Ty = zalloc(fy); m->kx = zalloc(fx); m->ky = zalloc(fy); m->kz = zalloc(fz);
    m->qx = zalloc(fx); m->qy = zalloc(fy); m->qz = zalloc(fz);
    m->Vx = zalloc(fx); m->Vy = zalloc(fy); m->Vz = zalloc(fz);
    m->Rt = zalloc(c); m->Rp = zalloc(c); m->F = zalloc(c); m->F0 = zalloc(c); m->Fcit = zalloc(c);
    m->PC = zalloc(c); m->G = zalloc(c); m->dumc = zalloc(c);
    m->w.Tc = zalloc(c); m->w.Told = zalloc(c); m->w.Tc_exxx = zalloc(X);
    m->w.v1 = zalloc(c); m->w.v2 = zalloc(c); m->w.v3 = zalloc(c); m->w.v4 = zalloc(c); m->w.v5 = zalloc(c);
    m->w.dTdxm = zalloc(c); m->w.dTdxp = zalloc(c);
    m->w.Vxm = zalloc(c); m->w.Vxp = zalloc(c); m->w.Vym = zalloc(c); m->w.Vyp = zalloc(c);
    m->w.Vzm = zalloc(c); m->w.Vzp = zalloc(c);
    o_dyrel_params d = {10000, 1000, 0, 0, 1e-8, 0.99, 0.5, 500.0};
    m->pT = d; m->pP = d;
    return m;
}

EXPORT void o_hrl_destroy(o_hrl_model *m) {
    if (!m) return;
    double *ptrs[] = {m->Tc_ex, m->Pc_ex, m->ktv, m->kfv, m->Ty, m->kx, m->ky, m->kz, m->qx, m->qy, m->qz, m->Vx,
                      m->Vy, m->Vz, m->Rt, m->Rp, m->F, m->F0, m->Fcit, m->PC, m->G, m->dumc, m->w.Tc, m->w.Told,
                      m->w.Tc_exxx, m->w.v1, m->w.v2, m->w.v3, m->w.v4, m->w.v5, m->w.dTdxm, m->w.dTdxp, m->w.Vxm,
                      m->w.Vxp, m->w.Vym, m->w.Vyp, m->w.Vzm, m->w.Vzp};
    for (size_t n = 0; n < sizeof(ptrs) / sizeof(ptrs[0]); ++n) free(ptrs[n]);
    free(m);
}

/* named field access for the Python wrapper */
EXPORT double *o_hrl_field(o_hrl_model *m, const char *name) {
#define F_(s) if (!strcmp(name, #s)) return m->s;
    F_(Tc_ex) F_(Pc_ex) F_(ktv) F_(kfv) F_(Ty) F_(kx) F_(ky) F_(kz) F_(qx) F_(qy) F_(qz) F_(Vx) F_(Vy) F_(Vz)
    F_(Rt) F_(Rp) F_(F) F_(F0) F_(Fcit) F_(PC) F_(G)
#undef F_
    return NULL;
}

EXPORT void o_hrl_set_params(o_hrl_model *m, const o_dyrel_params *pT, const o_dyrel_params *pP, int weno_version) {
    if (pT) m->pT = *pT;
    if (pP) m->pP = *pP;
    m->weno_version = weno_version;
}

EXPORT void o_hrl_get_reports(const o_hrl_model *m, o_dyrel_report *rT, o_dyrel_report *rP) {
    if (rT) *rT = m->rT;
    if (rP) *rP = m->rP;
}

EXPORT double o_hrl_scalar(const o_hrl_model *m, int which) {
    switch (which) {
    case 0: return m->dt;
    case 1: return m->time;
    case 2: return m->Nu;
    default: return NAN;
    }
}

EXPORT void o_hrl_set_dt(o_hrl_model *m, double dt) { m->dt = dt; }

EXPORT void o_hrl_set_viscosity(o_hrl_model *m, double a) { m->mu_a = a; }

/* OpenMP threads of the oracle's loops (small grids run faster on one) */
EXPORT void o_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

EXPORT int o_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

EXPORT void o_hrl_set_dt_rule(o_hrl_model *m, int rule) { m->dt_rule = rule; }

EXPORT int o_hrl_n_adv(const o_hrl_model *m) { return m->n_adv; }

/* One physical step.  Returns 0, or the first non-zero solver status. */
EXPORT int o_hrl_step(o_hrl_model *m) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const double _dx = 1.0 / m->dx, _dy = 1.0 / m->dy, _dz = 1.0 / m->dz;
    const double dt = m->dt, _dt = 1.0 / dt;
    o_dyrel_field f;
    memset(&f, 0, sizeof f);
    f.nx = nx; f.ny = ny; f.nz = nz;
    f.kx = m->kx; f.ky = m->ky; f.kz = m->kz;
    f.PC = m->PC; f.G = m->G; f.F = m->F; f.F0 = m->F0; f.Fcit = m->Fcit; f.dumc = m->dumc;
    f.qx = m->qx; f.qy = m->qy; f.qz = m->qz;
    f.dt = dt; f._dt = _dt; f.dx = m->dx; f.dy = m->dy; f.dz = m->dz;

    /* thermal solve (Alban_v6:327-342 with the DYREL loop of v18:432-478) */
    o_InitThermal(m->Rt, m->kx, m->ky, m->kz, m->Tc_ex, m->ktv, _dt, nx, ny, nz);
    f.X = m->Tc_ex; f.R = m->Rt; f.transient = 1; f.rhoC = 1.0;
    int tt[6] = {0, 0, 1, 1, 0, 0};
    double tv[6] = {0.0, 0.0, m->Tbot, m->Ttop, 0.0, 0.0};
    memcpy(f.bc_type, tt, sizeof tt); memcpy(f.bc_val, tv, sizeof tv);
    int st = o_dyrel_solve(&f, &m->pT, &m->rT);
    if (st == 11 || st == 12) return st;

    /* Nusselt number from the last thermal flux (Alban_v6:345-348) */
    {
        double s = 0.0, sl = 0.0;
        for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) dd_add(&s, &sl, FY(m->qy, i, ny + 1, k));
        m->Nu = (s + sl) / (m->Lx * m->Lz) / (double)(nx * nz);
    }

    /* Darcy solve (Alban_v6:353-370 with the DYREL loop of v18:489-537) */
    o_InitDarcy(m->Ty, m->kx, m->ky, m->kz, m->Tc_ex, m->kfv, _dt, nx, ny, nz);
    if (m->mu_a != 0.0) o_ApplyViscosity(m->kx, m->ky, m->kz, m->Tc_ex, m->mu_a, nx, ny, nz);
    o_KillNorthSouthFluxes(m->Ty, nx, ny, nz);
    o_Compute_Rp(m->Rp, m->ky, m->Ty, m->Ra, _dy, nx, ny, nz);
    f.X = m->Pc_ex; f.R = m->Rp; f.transient = 0; f.rhoC = 0.0;
    int pt[6] = {0, 0, 0, 0, 0, 0};
    double pv[6] = {0, 0, 0, 0, 0, 0};
    memcpy(f.bc_type, pt, sizeof pt); memcpy(f.bc_val, pv, sizeof pv);
    int sp = o_dyrel_solve(&f, &m->pP, &m->rP);
    if (sp == 11 || sp == 12) return sp;
    m->time += dt;

    /* velocity, then the advective time step from it (v18:553-561): the
       advection uses dt_adv and so does the next step (rule 0); the step's dt
       is split into ceil(dt/dt_adv) advection sub-steps (rule 1, v18:563-570);
       the advection uses dt and the next step dt_adv (rule 2, v7:215-221).
       Advection itself: Alban_v6:380-455.                                   */
    o_Init_vel(m->Vx, m->Vy, m->Vz, m->kx, m->ky, m->kz, m->Pc_ex, m->Ty, m->Ra, _dx, _dy, _dz, nx, ny, nz);
    double vmax = fmax(fmax(o_reduce(m->Vx, (size_t)(nx + 1), (size_t)ny * nz, R_MAX_ABS),
                            o_reduce(m->Vy, (size_t)nx, (size_t)(ny + 1) * nz, R_MAX_ABS)),
                       o_reduce(m->Vz, (size_t)nx, (size_t)ny * (nz + 1), R_MAX_ABS));
    double h = fmin(fmin(m->dx, m->dy), m->dz);
    double dt_adv = m->cfl * h / vmax, dt_a = dt_adv;
    int n_adv = 1;
    if (m->dt_rule == 1) {
        n_adv = (int)fmax(ceil(dt / dt_adv), 1.0);
        dt_a = dt / n_adv;
    } else {
        if (m->dt_rule == 2) dt_a = dt;
        m->dt = dt_adv;
    }
    for (int s = 0; s < n_adv; ++s) {
        if (m->weno_version == 2)
            o_AdvectWithWeno5_v2(m->Tc_ex, m->w.Tc_exxx, m->w.Told, m->Vx, m->Vy, m->Vz, dt_a, _dx, _dy, _dz, m->Ttop,
                                 m->Tbot, nx, ny, nz);
        else
            o_AdvectWithWeno5_v1(m->Tc_ex, m->Vx, m->Vy, m->Vz, &m->w, dt_a, _dx, _dy, _dz, m->Ttop, m->Tbot, nx, ny,
                                 nz);
    }
    m->n_adv = n_adv;
    return (st == 10 || sp == 10) ? 10 : 0;
}

/* ------------------------------------------------------------------------ */
/* Synthetic permeability: lognormal field + planar fault (vertex array)      */
/* ------------------------------------------------------------------------ */

/* kfv(x) = 10^(sigma * G(x)), G(x) = sqrt(2/M) sum_m cos(K_m . x + phi_m)
 * (spectral randomisation; K_m drawn by the caller from the exponential-
 * covariance spectrum), times fault_mult where |n.(x - x0)| <= fault_half.
 * Vertex coordinates: x = x0 + (i-1)*dx.                                    */
EXPORT void o_lognormal_field(double *kfv, int nx, int ny, int nz, double xmin, double ymin, double zmin, double dx,
                              double dy, double dz, const double *K, const double *phi, int M, double sigma,
                              double fault_mult, double fault_half, const double *fault_n, const double *fault_x0) {
    const double amp = sqrt(2.0 / (double)M);
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz + 1; ++k)
        for (int j = 1; j <= ny + 1; ++j)
            for (int i = 1; i <= nx + 1; ++i) {
                const double x = xmin + (i - 1) * dx, y = ymin + (j - 1) * dy, z = zmin + (k - 1) * dz;
                double s = 0.0;
                for (int m = 0; m < M; ++m) s += cos(K[3 * m] * x + K[3 * m + 1] * y + K[3 * m + 2] * z + phi[m]);
                double v = pow(10.0, sigma * (amp * s));
                if (fault_mult != 1.0) {
                    double d = fault_n[0] * (x - fault_x0[0]) + fault_n[1] * (y - fault_x0[1]) + fault_n[2] * (z - fault_x0[2]);
                    if (fabs(d) <= fault_half) v *= fault_mult;
                }
                VV(kfv, i, j, k) = v;
            }
}
