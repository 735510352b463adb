/*
 * ht3d.h -- C ABI of libht3d.so, the MI355X (gfx950) implementation of the
 * HydroThermal3D solver hot path.
 *
 * The library replaces ParallelStencil's `@parallel Kernel!(args...)` launches of
 * tduretz/HydroThermal3D for the pseudo-transient (PT) Darcy / heat solves and
 * the WENO5 advection.  Three layers of entry points:
 *
 *   1. runtime:   context, device memory, copies, reductions;
 *   2. kernels:   one entry per reference kernel, same name (without "!") and
 *                 argument order, arrays as device pointers, the cell-grid
 *                 extents (ncx,ncy,ncz) appended; every staggered extent is
 *                 derived from them exactly as in the reference:
 *                   c  = (ncx,ncy,ncz)            E  = (ncx+2,ncy+2,ncz+2)
 *                   X6 = (ncx+6,ncy+6,ncz+6)      V  = (ncx+1,ncy+1,ncz+1)
 *                   Fx = (ncx+1,ncy,ncz)  Fy = (ncx,ncy+1,ncz)  Fz = (ncx,ncy,ncz+1)
 *                 all column-major, x fastest, FP64;
 *   3. fused:     a whole DYREL solve, a whole WENO5 advection step, a whole
 *                 closed-box physical step, and the mixed-precision (FP32 inner
 *                 sweeps + FP64 defect correction) solve.
 *
 * Conventions: every function returns 0 (HT3D_OK) or an error code; nothing
 * throws across the ABI (each entry catches C++ exceptions and returns
 * HT3D_ERR_NOMEM for a failed host allocation, HT3D_ERR_INTERNAL otherwise).  Buffers are owned by the caller (ht3d_alloc/ht3d_free);
 * the library never frees caller memory.  One host thread per context.  Kernels
 * run on the library stream; the per-kernel entries synchronise before returning
 * (ParallelStencil launches are synchronous) unless ht3d_set_sync(0) was called.
 */
#ifndef HT3D_H
#define HT3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    HT3D_OK = 0,
    HT3D_ERR_ARG = 1,        /* bad argument (null pointer, extent < 3, ...)   */
    HT3D_ERR_HIP = 2,        /* HIP runtime error (see ht3d_last_error)         */
    HT3D_ERR_NOMEM = 3,
    HT3D_ERR_STATE = 4,      /* library not initialised                         */
    HT3D_ERR_INTERNAL = 5,   /* a C++ exception caught at the ABI boundary      */
    HT3D_NOT_CONVERGED = 10, /* solve hit nitmax (silent in the reference)      */
    HT3D_NAN = 11,           /* v18:464 error("Nan ...")                        */
    HT3D_DIVERGED = 12       /* v18:465 error("Diverged ...")                   */
};

/* ---------------------------------------------------------------- runtime */
/* CUDA.device!(GPU_ID), HydroThermal3D_v7.jl:9 */
int ht3d_init(int device);
int ht3d_finalize(void);
const char *ht3d_strerror(int code);
const char *ht3d_last_error(void);
int ht3d_set_sync(int synchronous);
int ht3d_sync(void);
int ht3d_device_name(char *buf, size_t len);
/* Test hook, no reference counterpart: the next `count` calls of guarded
 * entries fail as if a host allocation had thrown (HT3D_ERR_NOMEM), so callers
 * can exercise their error path without exhausting memory.  0 disarms it. */
int ht3d_debug_inject_host_failure(int count);

/* @zeros / Data.Array(host) / Array(device) (HydroThermal3D_v7.jl:87-121,153-157) */
int ht3d_alloc(size_t bytes, void **dptr);
int ht3d_free(void *dptr);
int ht3d_h2d(void *dst, const void *src, size_t bytes);
int ht3d_d2h(void *dst, const void *src, size_t bytes);
int ht3d_d2d(void *dst, const void *src, size_t bytes);
int ht3d_memset(void *dptr, int value, size_t bytes);

/* Live timing of the fused pseudo-iteration kernel: when enabled, the DYREL
 * drivers bracket each run of back-to-back iteration launches with HIP events
 * on the library stream and accumulate the kernel time, the launch count and
 * the algorithmic bytes (8 words per cell per iteration: read X, F0, kx, ky,
 * kz, R; write X, F0).  No reference counterpart (the reference prints a T_eff
 * template it never fills, HydroThermal3D_v7.jl:257-258).                   */
typedef struct {
    double ms;             /* summed event time of the iteration kernels    */
    long long launches;    /* iteration-kernel launches timed               */
    double bytes;          /* algorithmic bytes of those launches           */
    double ms_fp32;        /* the same for the FP32 inner sweeps            */
    long long launches_fp32;
    double bytes_fp32;
} ht3d_timing_stats;
int ht3d_timing_enable(int on);
int ht3d_timing_get(ht3d_timing_stats *out);
int ht3d_timing_reset(void);

/* Tile autotuner of the fused iteration kernel.  A configuration is (rows of 64
 * lanes per block, z cells marched per thread, x cells per lane: 1, or 2 as
 * one vector load, offered only for an even ncx) from a fixed table plus a
 * variant: bit 0 reads the stored Gershgorin diagonal instead of recomputing it
 * from the face coefficients (bit-identical either way), bit 1 writes the
 * updated unknown with a non-temporal store, bit 2 lets the face-coefficient
 * loads be cached; bit 3 forms the face coefficients from the solve's vertex
 * field (ht3d_dyrel_field.kv; tiles with 2 cells per lane, and 4 in FP32;
 * there bit 0 applies in FP64 only and bit 2 means a z march unrolled by two, in FP32
 * only); a vertex bit the solve cannot honour is dropped.  With autotuning on
 * (default) the
 * first solve of each (grid, operator, precision) times every configuration on
 * its own inputs and caches the fastest; force_tile >= 0 pins a table entry and
 * variant (a 2-cell entry on an odd ncx runs as its 1-cell twin).
 * ht3d_tune_last_tile also reports the x cells per lane.  No reference counterpart (ParallelStencil takes explicit
 * cublocks/cuthreads, HydroThermal3D_v6.jl:264-265).                        */
int ht3d_tune_set(int autotune, int force_tile, int force_variant);
int ht3d_tune_clear(void);
int ht3d_tune_last(int *rows_per_block, int *z_per_thread, int *variant);
int ht3d_tune_last_tile(int *rows_per_block, int *z_per_thread, int *x_per_thread, int *variant);
int ht3d_tile_count(void);
/* Two pseudo-iterations per launch for runs of plain iterations (overlapped
 * tiles, the intermediate iterate kept on chip; bit-identical to two single
 * launches).  mode 0: never, 1: when the tuner times them faster per iteration
 * (default), 2: always.  ht3d_tune_last_pair reports the last solve's pair
 * tile (0 rows when none was used).                                          */
int ht3d_tune_pairs(int mode);
/* The check-interval tuner's rule (used when ncheck < 0): the divisor L of
 * nitout closest (in ratio) to sqrt(2 iters t_check / t_iter), which minimises
 * iters*t_iter + (iters/L)*t_check + (L/2)*t_iter; without history (any
 * argument <= 0) the largest divisor of nitout up to 50.  Host only.          */
int ht3d_tune_check_interval(double iters, double t_iter_us, double t_check_us, int nitout);
int ht3d_tune_last_pair(int *rows_per_block, int *z_per_block, int *variant);
/* The pair tile table also has tiles with 2 or 4 x-cells per lane (for an ncx
 * divisible by that number; FP32 only for 4).  ht3d_tune_last_pair_tile also
 * reports the cells per lane (0 when no pair tile was used).
 * ht3d_tune_force_pair pins pair tile `tile` (< 0: none) and variant (bit 0
 * stored diagonal, bit 1 cached coefficients; on the band tiles of the FP32
 * sweeps bit 3: coefficients from the vertex field, when the solve has one)
 * for the plain runs of every solve, whatever the timings; a solve whose grid
 * the tile does not fit runs without pairs, and pairs mode 0 still turns pairs
 * off.  No reference counterpart.                                            */
int ht3d_tune_last_pair_tile(int *rows_per_block, int *z_per_block, int *x_per_thread, int *variant);
int ht3d_tune_force_pair(int tile, int variant);
int ht3d_pair_tile_count(void);
/* Sweep direction of the fused iteration kernels: with alternate = 1 (default)
 * every other launch walks each XCD's slab of tiles backwards, starting on the
 * planes the previous launch touched last (still in the 256 MB MALL).  Results
 * do not depend on it; 0 sweeps forward always.  No reference counterpart.   */
int ht3d_tune_sweep(int alternate);

/* maximum_g / minimum_g / sum_g / mean_g of abs.(A) (tools/Macros.jl:20-25);
 * the result lands in host memory.  Deterministic order.                    */
enum { HT3D_SUM = 0, HT3D_SUM_ABS = 1, HT3D_MAX = 2, HT3D_MIN = 3, HT3D_MAX_ABS = 4 };
int ht3d_reduce(double *out_host, const double *a, size_t n, int op);

/* ---------------------------------------------------------------- kernels */
/* tools/Macros.jl:171-199, tools/Weno5_Routines.jl:271-276 */
int ht3d_ResetA(double *A, double *B, int n1, int n2, int n3);
int ht3d_Cpy_inn_to_all(double *A, const double *B, int ncx, int ncy, int ncz);
int ht3d_Cpy_all_to_inn(double *A, const double *B, int ncx, int ncy, int ncz);
int ht3d_ArrayEqualArray(double *F1, const double *F2, int n1, int n2, int n3);
/* kernels_HT3D_dimensional.jl:106-109 */
int ht3d_Multiply(double *a, const double *b, const double *c, int ncx, int ncy, int ncz);

/* Dimensionless/kernels_HT3D.jl:99-107 */
int ht3d_InitThermal(double *Rt, double *kx, double *ky, double *kz, const double *Tc_ex, const double *ktv,
                     double _dt, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:133-138 */
int ht3d_InitConductivity(double *kx, double *ky, double *kz, const double *ktv, int ncx, int ncy, int ncz);
/* Dimensionless/kernels_HT3D.jl:82-90 */
int ht3d_InitDarcy(double *Ty, double *kx, double *ky, double *kz, const double *Tc_ex, const double *kfv,
                   double _dt, int ncx, int ncy, int ncz);
/* Dimensionless/HydroThermal3D_Alban_v6.jl:153-159 */
int ht3d_KillNorthSouthFluxes(double *Ty, int ncx, int ncy, int ncz);
/* Dimensionless/kernels_HT3D.jl:92-97 (+ @dmul_ya, tools/Macros.jl:111) */
int ht3d_Compute_Rp(double *Rp, const double *ky, const double *Ty, double Ra, double _dy, int ncx, int ncy,
                    int ncz);

/* Ghost fills on E arrays.  Face ghosts as in the reference.  Edge and corner
 * ghosts race in the reference and are never read by the hot path; here they
 * are deterministic (x faces over the interior, then y faces, then z faces).
 * Alban_v6:108-118 (= kernels_HT3D.jl:144-154): Neumann x,z; Dirichlet y.    */
int ht3d_SetTemperatureBCs(double *Tc_ex, double Tbot, double Ttop, int ncx, int ncy, int ncz);
/* Alban_v6:120-130: Neumann on all faces (closed box). */
int ht3d_SetPressureBCs(double *Pc_ex, double Pbot, double Ptop, int ncx, int ncy, int ncz);
/* kernels_HT3D.jl:144-154: the v7 pressure BC (Dirichlet in y). */
int ht3d_SetPressureBCs_v7(double *Pc_ex, double Pbot, double Ptop, int ncx, int ncy, int ncz);

/* kernels_HT3D.jl:109-117 (= kernels_HT3D_dimensional.jl:140-146) */
int ht3d_ComputeFlux(double *qx, double *qy, double *qz, const double *kx, const double *ky, const double *kz,
                     const double *A, double _dx, double _dy, double _dz, int ncx, int ncy, int ncz);
/* kernels_HT3D.jl:119-126 */
int ht3d_UpdateT(double *F, double *T, const double *R, const double *qx, const double *qy, const double *qz,
                 double _dt, double dtau, double _dx, double _dy, double _dz, int ncx, int ncy, int ncz);
/* kernels_HT3D.jl:128-134 */
int ht3d_ResidualDiffusion(double *F, const double *R, const double *qx, const double *qy, const double *qz,
                           double _dx, double _dy, double _dz, int ncx, int ncy, int ncz);
/* Alban_v6:84-90 */
int ht3d_ResidualDiffusionTransient(double *F, const double *T, const double *R, const double *qx,
                                    const double *qy, const double *qz, double _dt, double _dx, double _dy,
                                    double _dz, int ncx, int ncy, int ncz);
/* The two residuals above followed by the "F /= PC" (Jacobi) step that ends every
 * DYREL residual of kernels_HT3D_dimensional.jl:175,283.                      */
int ht3d_ResidualDiffusionPC(double *F, const double *R, const double *qx, const double *qy, const double *qz,
                             const double *PC, double _dx, double _dy, double _dz, int ncx, int ncy, int ncz);
int ht3d_ResidualDiffusionTransientPC(double *F, const double *T, const double *R, const double *qx,
                                      const double *qy, const double *qz, const double *PC, double _dt,
                                      double _dx, double _dy, double _dz, int ncx, int ncy, int ncz);
/* kernels_HT3D.jl:136-142 (= DampedUpdate!, kernels_HT3D_dimensional.jl:316-320) */
int ht3d_UpdateP(double *F0, double *P, const double *F, double dampx, double dtau, int ncx, int ncy, int ncz);
int ht3d_DampedUpdate(double *F0, double *X, const double *F, double dampx, double dtau, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:322-326 */
int ht3d_DYRELUpdate(double *F0, double *X, const double *F, double h1, double h2, double dtau, int ncx, int ncy,
                     int ncz);
/* kernels_HT3D_dimensional.jl:122-126 (only the residual copy, as in the reference) */
int ht3d_SwapDYREL(double *Xc, double *Yc, const double *X0c_ex, const double *Y0c, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:207-221 */
int ht3d_GershgorinPoisson(double *G, double *D, const double *rhoC, const double *kx, const double *ky,
                           const double *kz, double transient, double dt, double dx, double dy, double dz, int ncx,
                           int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:335-347 */
int ht3d_GershgorinHydro(double *G, const double *phidrhodP, const double *kx, const double *ky, const double *kz,
                         double transient, double dt, double dx, double dy, double dz, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:349-363 */
int ht3d_RayleighQuotientNumerator(double *G, const double *Fc0, const double *Fcit, const double *Fc, double dtau,
                                   int ncx, int ncy, int ncz);
int ht3d_RayleighQuotientDenominator(double *G, const double *Fc0, double dtau, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:162-181 */
int ht3d_ResidualTemperatureLinearised(double *F, const double *Tc_ex, const double *Tc_old, const double *rhoC_eff,
                                       const double *phc_ex, const double *PC, const double *qx, const double *qy,
                                       const double *qz, double Qt, double transient, double _dt, double _dx,
                                       double _dy, double _dz, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:274-289 */
int ht3d_ResidualFluidPressure(double *F, const double *phc_ex, const double *phirhof, const double *phirhof0,
                               const double *PC, const double *qx, const double *qy, const double *qz,
                               double transient, double _dt, double _dx, double _dy, double _dz, int ncx, int ncy,
                               int ncz);
/* kernels_HT3D_dimensional.jl:266-272 */
int ht3d_ComputeDarcyFlux(double *qx, double *qy, double *qz, const double *rhofv, const double *kx,
                          const double *ky, const double *kz, const double *Pf, double g, double _dx, double _dy,
                          double _dz, int ncx, int ncy, int ncz);
/* Dimensionless/kernels_HT3D.jl:73-80 */
int ht3d_Init_vel(double *Vx, double *Vy, double *Vz, const double *kx, const double *ky, const double *kz,
                  const double *Pc_ex, const double *Ty, double Ra, double _dx, double _dy, double _dz, int ncx,
                  int ncy, int ncz);

/* ---- Dimensional (v18) property kernels and immersed-surface ghost fills.
 * They evaluate the closures of Dimensional/HydroThermal3D_v18.jl:29-39 (rock and
 * fluid conductivity, heat capacity, density, the fluid-density polynomial and
 * its pressure derivative, viscosity, depth-dependent permeability, porosity
 * per phase) in the reference's order; only 10^x and exp() may differ from a
 * CPU evaluation, by an ulp.  Phases: 1 air, -1 immersed surface cell, 2 fault,
 * 3 intrusion.  wN is the (ncx+2, ncz+2) table of surface weights.           */
/* kernels_HT3D_dimensional.jl:40-49 (vertex array from extended centroids) */
int ht3d_ComputeThermalConductivity(double *ktv, const double *Tce, const double *phv, double phi0, double scale_kt,
                                    double scale_T, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:51-83; mode 1: rho_f k/mu, 2: k/mu, 3: k */
int ht3d_ComputeHydroConductivity(double *kfv, const double *Tce, const double *Pce, const double *phv, double ymin,
                                  double dy, double k_fact, double delta, double scale_sigma, double scale_T,
                                  double scale_L, double scale_t, double scale_eta, int mode, int ncx, int ncy,
                                  int ncz);
/* kernels_HT3D_dimensional.jl:85-99 */
int ht3d_ComputeFluidDensity(double *rhofv, const double *Tce, const double *Pce, const double *phv,
                             double scale_sigma, double scale_T, double scale_rho, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:110-113 (cells = @av of vertices) */
int ht3d_SmoothConductivityV2C(double *ktc, const double *ktv, int ncx, int ncy, int ncz);
/* kernels_HT3D_dimensional.jl:115-118 (inner vertices = @av of cells) */
int ht3d_SmoothConductivityC2V(double *ktv, const double *ktc, int ncx, int ncy, int ncz);
/* ComputeρCeffective!  kernels_HT3D_dimensional.jl:148-160 (air cells untouched) */
int ht3d_ComputeRhoCeffective(double *rhoC_eff, const double *Tc_ex, const double *Pc_ex, const double *phc_ex,
                              double phi_i, double sca_T, double sca_sigma, double sca_C, double sca_rho, int ncx,
                              int ncy, int ncz);
/* ϕρf  kernels_HT3D_dimensional.jl:225-235 (phase-1 cells untouched) */
int ht3d_PhiRhof(double *phirhofc, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double phi_i,
                 double scale_rho, double scale_T, double scale_sigma, int ncx, int ncy, int ncz);
/* ϕρf_old  kernels_HT3D_dimensional.jl:237-247 (T, P as cell arrays) */
int ht3d_PhiRhof_old(double *phirhofc, const double *Tc0, const double *Pc0, const double *phc_ex, double phi_i,
                     double scale_rho, double scale_T, double scale_sigma, int ncx, int ncy, int ncz);
/* ϕdρdP!  kernels_HT3D_dimensional.jl:249-259 */
int ht3d_PhiDrhoDP(double *phidrhodP, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double phi_i,
                   double scale_rho, double scale_T, double scale_sigma, int ncx, int ncy, int ncz);
/* Init_vel!  kernels_HT3D_dimensional.jl:328-332 (V = q / face-averaged rho_f;
 * named dim_ because Dimensionless/kernels_HT3D.jl has another Init_vel!) */
int ht3d_dim_Init_vel(double *Vx, double *Vy, double *Vz, const double *qx, const double *qy, const double *qz,
                      const double *rhofv, int ncx, int ncy, int ncz);
/* SetTemperatureBCs!  kernels_HT3D_dimensional.jl:183-205: air = TN, Neumann x/z,
 * basal heat flux qyS (bottom), immersed-surface Dirichlet TN.  Every argument of
 * the reference is kept; kS, TS, y_plateau, a2, b2, dTdy, dx, dy, sticky_air,
 * sca_T, sca_kT and phi_i are unused there and here.  A rule reads its source
 * cell after the source's own rule, so the result does not depend on thread order. */
int ht3d_dim_SetTemperatureBCs(double *Tc_ex, const double *ky, const double *wN, const double *phc_ex, double qyS,
                               double _dy, double kS, double TS, double TN, double y_plateau, double a2, double b2,
                               double dTdy, double dx, double dy, double sticky_air, double sca_T, double sca_kT,
                               double phi_i, int ncx, int ncy, int ncz);
/* SetPressureBCs!  kernels_HT3D_dimensional.jl:291-314: air = PN, Neumann x/z and
 * bottom, immersed-surface Dirichlet PN (PS, y_plateau, a2, b2, rhof, g, dx, dy,
 * sticky_air unused, as in the reference) */
int ht3d_dim_SetPressureBCs(double *Pc_ex, const double *wN, const double *phc_ex, double PS, double PN,
                            double y_plateau, double a2, double b2, double rhof, double g, double dx, double dy,
                            double sticky_air, int ncx, int ncy, int ncz);

/* ---- Output fields of Dimensional/kernels_HT3D_postprocessing.jl (the VTK
 * fields of HydroThermal3D_v18.jl:663-687): cell arrays in physical units,
 * NaN where the phase is 1 (air).                                             */
/* Phase!  :1-10 */
int ht3d_Phase(double *X, const double *phc_ex, int ncx, int ncy, int ncz);
/* Pressure!  :12-21 and Temperature!  :23-32 (field * scale) */
int ht3d_Pressure(double *X, const double *Pc_ex, const double *phc_ex, double scale_sigma, int ncx, int ncy,
                  int ncz);
int ht3d_Temperature(double *X, const double *Tc_ex, const double *phc_ex, double scale_T, int ncx, int ncy, int ncz);
/* Velocity!  :34-45 (cell averages of the face velocities; in air only X is set) */
int ht3d_Velocity(double *X, double *Y, double *Z, const double *Vx, const double *Vy, const double *Vz,
                  const double *phc_ex, double scale_V, int ncx, int ncy, int ncz);
/* EffectiveThermalConductivity!  :47-57 */
int ht3d_EffectiveThermalConductivity(double *X, const double *Tc_ex, double phi0, const double *phc_ex,
                                      double scale_T, int ncx, int ncy, int ncz);
/* EffectiveHeatCapacity!  :60-70 (the solid's C; air cells untouched) */
int ht3d_EffectiveHeatCapacity(double *Ceff, const double *Tc_ex, const double *Pc_ex, const double *phc_ex,
                               double phi_i, double sca_T, double sca_sigma, double sca_C, double sca_rho, int ncx,
                               int ncy, int ncz);
/* Permeability!  :72-85 */
int ht3d_Permeability(double *X, double k_fact, double ymin, double dy, const double *phc_ex, double delta,
                      double scale_L, int ncx, int ncy, int ncz);
/* FluidDensity!  :87-98 */
int ht3d_FluidDensity(double *X, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double scale_T,
                      double scale_sigma, int ncx, int ncy, int ncz);
/* Viscosity!  :100-109 */
int ht3d_Viscosity(double *X, const double *Tc_ex, const double *phc_ex, double scale_T, int ncx, int ncy, int ncz);
/* Peclet!  :111-128 */
int ht3d_Peclet(double *X, const double *Tc_ex, const double *Pc_ex, const double *Vx, const double *Vy,
                const double *Vz, const double *phc_ex, double phi0, double scale_T, double scale_sigma,
                double scale_V, double scale_L, int ncx, int ncy, int ncz);

/* Temperature-dependent viscosity mu = exp(-a T): face coefficients *= exp(a T_face),
 * T_face the mean of the two adjacent cells (ghosts at walls).  No counterpart in
 * the dimensionless kernels (constant viscosity there); the dimensionless analogue
 * of the T-dependent conductivities of kernels_HT3D_dimensional.jl.             */
int ht3d_ApplyViscosity(double *kx, double *ky, double *kz, const double *Tc_ex, double a, int ncx, int ncy, int ncz);

/* tools/Weno5_Routines.jl:5-42 */
int ht3d_VxPlusMinus(double *Vxm, double *Vxp, const double *Vx, int ncx, int ncy, int ncz);
int ht3d_VyPlusMinus(double *Vym, double *Vyp, const double *Vy, int ncx, int ncy, int ncz);
int ht3d_VzPlusMinus(double *Vzm, double *Vzp, const double *Vz, int ncx, int ncy, int ncz);
/* tools/Weno5_Routines.jl:44-68 */
int ht3d_dFdx_Weno5(double *dFdxi, const double *V1, const double *V2, const double *V3, const double *V4,
                    const double *V5, int ncx, int ncy, int ncz);
/* tools/Weno5_Routines.jl:70-128; axis 1..3, plus 0 (minus) / 1 (plus) */
int ht3d_Gradients_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                         int axis, int plus, double _dx, double _dy, double _dz, int ncx, int ncy, int ncz);
/* tools/Weno5_Routines.jl:130-255 (src_ext=0, Fc is c) and
 * tools/Weno5_Routines_v2.jl:221-347 (src_ext=1, Fc is E); type 0/1/2 =
 * Neumann / Dirichlet / periodic                                             */
int ht3d_Boundaries_Weno5(double *Fc_exxx, const double *Fc, int axis, int src_ext, int type_lo, double val_lo,
                          int type_hi, double val_hi, int ncx, int ncy, int ncz);
/* The same kernels under the reference's names and argument orders:
 * Gradients_{minus,plus}_{x,y,z}_Weno5!  tools/Weno5_Routines.jl:70-128, and
 * Boundaries_{x,y,z}_Weno5!  tools/Weno5_Routines.jl:130-255 (Fc a cell array;
 * the _v2 entries take Fc extended, tools/Weno5_Routines_v2.jl:221-347).     */
int ht3d_Gradients_minus_x_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5,
                              const double *Fc_exxx, double _dx, double _dy, double _dz, int ncx, int ncy,
                              int ncz);
int ht3d_Gradients_plus_x_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5,
                              const double *Fc_exxx, double _dx, double _dy, double _dz, int ncx, int ncy,
                              int ncz);
int ht3d_Gradients_minus_y_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5,
                              const double *Fc_exxx, double _dx, double _dy, double _dz, int ncx, int ncy,
                              int ncz);
int ht3d_Gradients_plus_y_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5,
                              const double *Fc_exxx, double _dx, double _dy, double _dz, int ncx, int ncy,
                              int ncz);
int ht3d_Gradients_minus_z_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5,
                              const double *Fc_exxx, double _dx, double _dy, double _dz, int ncx, int ncy,
                              int ncz);
int ht3d_Gradients_plus_z_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5,
                              const double *Fc_exxx, double _dx, double _dy, double _dz, int ncx, int ncy,
                              int ncz);
int ht3d_Boundaries_x_Weno5(double *Fc_exxx, const double *Fc, int type_W, double val_W,
                              int type_E, double val_E, int ncx, int ncy, int ncz);
int ht3d_Boundaries_x_Weno5_v2(double *Fc_exxx, const double *Fc, int type_W, double val_W,
                              int type_E, double val_E, int ncx, int ncy, int ncz);
int ht3d_Boundaries_y_Weno5(double *Fc_exxx, const double *Fc, int type_S, double val_S,
                              int type_N, double val_N, int ncx, int ncy, int ncz);
int ht3d_Boundaries_y_Weno5_v2(double *Fc_exxx, const double *Fc, int type_S, double val_S,
                              int type_N, double val_N, int ncx, int ncy, int ncz);
int ht3d_Boundaries_z_Weno5(double *Fc_exxx, const double *Fc, int type_B, double val_B,
                              int type_F, double val_F, int ncx, int ncy, int ncz);
int ht3d_Boundaries_z_Weno5_v2(double *Fc_exxx, const double *Fc, int type_B, double val_B,
                              int type_F, double val_F, int ncx, int ncy, int ncz);
/* tools/Weno5_Routines.jl:257-269 */
int ht3d_Advect(double *Fc, const double *Vxp, const double *dTdxm, const double *Vxm, const double *dTdxp,
                double dt, int ncx, int ncy, int ncz);
int ht3d_TimeAveraging(double *Fc, const double *Fold, double order, int ncx, int ncy, int ncz);
/* tools/Weno5_Routines_v2.jl:55-219 */
int ht3d_Advect_v2(double *Fc, const double *X, const double *V, int axis, double dt, double _dx, double _dy,
                   double _dz, int ncx, int ncy, int ncz);
int ht3d_TimeAveraging_v2(double *Fc, const double *Fold, double order, int ncx, int ncy, int ncz);

/* ---------------------------------------------------------------- fused */

/* One DYREL solve (HydroThermal3D_v18.jl:424-478 / :481-538) on the
 * dimensionless operators of kernels_HT3D.jl: Gershgorin set-up, then up to
 * nitmax fused pseudo-iterations (ghost fill + flux + residual/D + DYREL update
 * in one kernel), Rayleigh-quotient lambda_min updates on the reference schedule
 * and convergence checks.                                                     */
typedef struct {
    int nitmax;        /* v18:309 (1e4)                                              */
    int nitout;        /* lambda_min + reference check schedule: iter<=2 || iter%nitout==0 */
    int ncheck;        /* >0: also test convergence every ncheck iterations; <0: every L
                          iterations, L chosen by the check-interval tuner (divisor of nitout) */
    int norm;          /* 0: mean|F*D|/sqrt(N) (v18:459-460); 1: max|F*D|               */
    double tol;        /* converged if norm < tol or norm/norm(iter 1) < tol (v18:476)   */
    double CFL;        /* v18:435 (0.99)                                                 */
    double cfact;      /* v18:436 (0.5)                                                  */
    double lmin_div;   /* v18:434 (500)                                                  */
} ht3d_dyrel_params;

typedef struct {
    int iters, status, n_checks, n_lambda;
    double lmax, lmin, dtau, h1, h2;
    double nF_abs, nF_rel, nF_ini;
    int ncheck;        /* check interval used (the tuner's choice when ncheck < 0) */
} ht3d_dyrel_report;

typedef struct {
    int ncx, ncy, ncz;
    double dx, dy, dz;
    double *X;                          /* E, in/out                                  */
    const double *R;                    /* c: Rt (heat) or Rp (Darcy)                 */
    const double *kx, *ky, *kz;         /* Fx, Fy, Fz                                 */
    int transient;                      /* 1: F = R + X/dt + div q; 0: F = R + div q  */
    double rhoC;                        /* uniform rhoC of the Gershgorin diagonal    */
    double dt;
    int bc_type[6];                     /* x-,x+,y-,y+,z-,z+: 0 Neumann, 1 Dirichlet  */
    double bc_val[6];
    const double *kv;                   /* V, or NULL: the vertex field whose face
                                           averages kx, ky, kz are (@av_yza/@av_xza/@av_xya,
                                           as InitThermal!/InitDarcy! form them,
                                           kernels_HT3D.jl:82-107).  When given, the FP64
                                           solve may read it instead of the three face
                                           arrays (6 instead of 8 words per cell and
                                           iteration); the results are the same bit for bit
                                           when kx, ky, kz are those averages.  The
                                           defect-correction solve then forms its FP32
                                           faces as FP32 averages of the FP32 vertex field
                                           (not the FP64 faces rounded), so its FP32 sweeps
                                           differ from a solve without kv at FP32 rounding;
                                           its FP64 residuals, and so its tolerance, do not. */
} ht3d_dyrel_field;

int ht3d_dyrel_solve(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep);

/* ht3d_dyrel_solve without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  Every device buffer of the solve comes
 * from a static device array of the library (about 105 MB per GPU, held until
 * the process ends), so the grid is limited to 2^21 cells and 2^22 entries of
 * its extended array (HT3D_ERR_ARG beyond); no tuning pass runs (the tile the
 * tuner chose for this grid earlier in the process, else the default tile:
 * the results are the same bit for bit).  Calls that overlap on one GPU race
 * on that array; each call clears what must start at zero. */
int ht3d_dyrel_solve_noalloc(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep);

/* ---- Fused DYREL solves of the dimensional (v18) model: the heat loop of
 * Dimensional/HydroThermal3D_v18.jl:432-478 (SetTemperatureBCs!, ComputeFlux!,
 * ResidualTemperatureLinearised!, DYRELUpdate!) and the Darcy loop of :489-537
 * (ϕρf, SetPressureBCs!, ComputeDarcyFlux!, ResidualFluidPressure!,
 * DYRELUpdate!), from GershgorinPoisson! to convergence, one launch per
 * pseudo-iteration.  The caller runs the property kernels before (as the v18
 * driver does: conductivities, rho*C or phi*drho/dP, ϕρf_old, fluid density).
 * Same iterations, lambda updates and fields as the unfused kernel sequence;
 * memory is left as that sequence leaves it (see ht3d_v18.hip).             */
typedef struct {
    int ncx, ncy, ncz;
    double dx, dy, dz, dt, transient;
    double *X;                    /* Tc_ex (heat) or Pc_ex (Darcy): E, in/out                       */
    const double *kx, *ky, *kz;   /* face coefficients (InitConductivity!)                          */
    const double *phc_ex;         /* E: centroid phases (1 air, -1 immersed, 2 fault, 3 intrusion)  */
    const double *wN;             /* (ncx+2, ncz+2): immersed-surface weights                       */
    double *rhoC;                 /* c: heat: rho*C (in); Darcy: phi*drho/dP (in), phi*rho_f (out)   */
    double *PC, *G;               /* c: Gershgorin diagonal and bound (out; NULL: internal)         */
    double top;                   /* TN or PN, the surface value                                    */
    double norm_scale;            /* dimensional scale of the norm (v18:460 / :517)                 */
    const double *Tc_old;         /* heat: c, temperature of the previous step (Tc0)                */
    double qyS, Qt;               /* heat: basal heat flux; heat source of phase-3 cells            */
    const double *Tc_ex;          /* Darcy: E, temperature of the fluid-density closure            */
    const double *phirhof0;       /* Darcy: c, phi*rho_f of the previous step (ϕρf_old, Xc0)         */
    const double *rhofv;          /* Darcy: V, fluid density (ComputeFluidDensity)                  */
    double g, phi_i, scale_rho, scale_T, scale_sigma;   /* Darcy: gravity, porosity, scales        */
    const double *kv;             /* V, or NULL: the vertex field whose face averages kx, ky, kz are
                                     (InitConductivity!); when given, the tiles that read it instead
                                     of the three face arrays are offered (same results bit for bit) */
} ht3d_v18_field;
int ht3d_v18_thermal_solve(const ht3d_v18_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep);
int ht3d_v18_darcy_solve(const ht3d_v18_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep);
/* tile of the v18 iteration kernel: force_tile >= 0 forces one of the
 * ht3d_v18_tile_count() candidates (6 with one x-cell per lane, then 4 with
 * two, which need an even ncx and otherwise run as their one-cell twin, then
 * the same 4 forming the face coefficients from the field's vertex field,
 * which without one run as their face-array twin), -1
 * lets the autotuner choose; *last_tile (may be NULL) returns the tile of the
 * last solve.  Every tile gives bit-identical results.  Analogue of
 * cublocks/cuthreads.                                                        */
int ht3d_v18_tune(int force_tile, int *last_tile);
int ht3d_v18_tile_count(void);

/* FP32 versions of the inner-sweep kernels for a defect-correction driver (the
 * fused ht3d_dyrel_solve_dc runs the same arithmetic).  Same argument order as
 * the FP64 entries, float arrays and scalars, expressions evaluated in float in
 * the reference's order.  No reference counterpart (the reference is FP64 only,
 * HydroThermal3D_v7.jl:13).  ht3d_to_f32 converts n doubles; ht3d_add_f32 adds
 * an FP32 correction to the interior of an FP64 E array (X += e).           */
int ht3d_SetTemperatureBCs_f32(float *Tc_ex, float Tbot, float Ttop, int ncx, int ncy, int ncz);
int ht3d_SetPressureBCs_f32(float *Pc_ex, float Pbot, float Ptop, int ncx, int ncy, int ncz);
int ht3d_ComputeFlux_f32(float *qx, float *qy, float *qz, const float *kx, const float *ky, const float *kz,
                         const float *A, float _dx, float _dy, float _dz, int ncx, int ncy, int ncz);
int ht3d_ResidualDiffusion_f32(float *F, const float *R, const float *qx, const float *qy, const float *qz, float _dx,
                               float _dy, float _dz, int ncx, int ncy, int ncz);
int ht3d_ResidualDiffusionTransient_f32(float *F, const float *T, const float *R, const float *qx, const float *qy,
                                        const float *qz, float _dt, float _dx, float _dy, float _dz, int ncx, int ncy,
                                        int ncz);
int ht3d_ResidualDiffusionPC_f32(float *F, const float *R, const float *qx, const float *qy, const float *qz,
                                 const float *PC, float _dx, float _dy, float _dz, int ncx, int ncy, int ncz);
int ht3d_ResidualDiffusionTransientPC_f32(float *F, const float *T, const float *R, const float *qx, const float *qy,
                                          const float *qz, const float *PC, float _dt, float _dx, float _dy,
                                          float _dz, int ncx, int ncy, int ncz);
int ht3d_DYRELUpdate_f32(float *F0, float *X, const float *F, float h1, float h2, float dtau, int ncx, int ncy,
                         int ncz);
int ht3d_GershgorinPoisson_f32(float *G, float *D, const float *rhoC, const float *kx, const float *ky,
                               const float *kz, float transient, float dt, float dx, float dy, float dz, int ncx,
                               int ncy, int ncz);
int ht3d_to_f32(float *out, const double *in, size_t n);
int ht3d_add_f32(double *X, const float *e, int ncx, int ncy, int ncz);

/* The same solve with FP32 DYREL sweeps on the correction equation A e = r inside
 * an FP64 defect-correction loop (r = b - A x and x += e in FP64).  No reference
 * counterpart: the reference is FP64 only (HydroThermal3D_v7.jl:13).         */
int ht3d_dyrel_solve_dc(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep,
                        int *outer_iters, int *inner_iters);

/* One whole WENO5 advection step on Tc_ex (AdvectWithWeno5, kernels_HT3D.jl:156-233,
 * version 1; AdvectWithWeno5_v2, tools/Weno5_Routines_v2.jl:1-50, version 2) with
 * the closed-box boundary types, fused into two passes per direction.       */
int ht3d_weno5_advect(double *Tc_ex, const double *Vx, const double *Vy, const double *Vz, double dt, double _dx,
                      double _dy, double _dz, double Ttop, double Tbot, int version, int ncx, int ncy, int ncz);

/* Init_vel! (kernels_HT3D.jl:73-80), the advective time step and the Nusselt
 * number in one pass and one read-back: out[0] = max|V| over Vx, Vy, Vz,
 * out[1] = cfl*min(dx,dy,dz)/max|V| (HydroThermal3D_v18.jl:554, v7:220),
 * out[2] = Nu = sum_top(-kty dT/dy)/(Lx*Lz)/(ncx*ncz) of Tc_ex with filled
 * ghosts (HydroThermal3D_Alban_v6.jl:345-346; 0 when Tc_ex is null).
 * V is bit-identical to ht3d_Init_vel.                                      */
int ht3d_velocity_dt_nu(double *Vx, double *Vy, double *Vz, const double *kx, const double *ky, const double *kz,
                        const double *Pc_ex, const double *Ty, double Ra, double dx, double dy, double dz, double cfl,
                        const double *Tc_ex, const double *kty, double Lx, double Lz, double *out, int ncx, int ncy,
                        int ncz);

/* Closed-box Horton-Rogers-Lapwood porous convection (HydroThermal3D_Alban_v6.jl)
 * stepped with DYREL solves and WENO5 advection, all on the device.  Time step
 * rules, with dt_adv = cfl*min(dx,dy,dz)/max|V| from this step's velocity:
 *   HT3D_DT_V18           advect with dt_adv; the next step's solves use it too
 *                         (HydroThermal3D_v18.jl:553-562, dt_constant = false);
 *   HT3D_DT_V18_CONSTANT  keep dt; advect in ceil(dt/dt_adv) equal sub-steps
 *                         (HydroThermal3D_v18.jl:563-570, dt_constant = true);
 *   HT3D_DT_V7            advect with this step's dt, then dt = dt_adv for the
 *                         next step (HydroThermal3D_v7.jl:215-221).          */
typedef struct ht3d_hrl ht3d_hrl;
typedef struct {
    ht3d_dyrel_report thermal, darcy;
    double dt_used, dt_next, time, Nu, vmax;
    int status;
    int dc_outer_thermal, dc_outer_darcy;   /* defect-correction outer iterations */
    int inner_thermal, inner_darcy;         /* FP32 inner sweeps in total          */
    int n_adv;                              /* advection sub-steps                 */
} ht3d_step_report;

enum { HT3D_FP64 = 0, HT3D_FP32_DC = 1 };

int ht3d_hrl_create(int ncx, int ncy, int ncz, double Lx, double Ly, double Lz, double Ra, double cfl,
                    ht3d_hrl **out);
int ht3d_hrl_destroy(ht3d_hrl *m);
/* names: Tc_ex, Pc_ex, ktv, kfv, Ty, kx, ky, kz, Vx, Vy, Vz, Rt, Rp */
int ht3d_hrl_field(ht3d_hrl *m, const char *name, double **dptr, size_t *count);
int ht3d_hrl_set_params(ht3d_hrl *m, const ht3d_dyrel_params *pT, const ht3d_dyrel_params *pP, int weno_version,
                        int precision);
int ht3d_hrl_set_dt(ht3d_hrl *m, double dt);
enum { HT3D_DT_V18 = 0, HT3D_DT_V18_CONSTANT = 1, HT3D_DT_V7 = 2 };
/* time step rule (default HT3D_DT_V18) */
int ht3d_hrl_set_dt_rule(ht3d_hrl *m, int rule);
/* mu = exp(-a T) in the Darcy solve (a = 0: constant viscosity, the default) */
int ht3d_hrl_set_viscosity(ht3d_hrl *m, double a);
int ht3d_hrl_step(ht3d_hrl *m, ht3d_step_report *rep);

/* Many independent problems stepped together (no reference counterpart: the
 * reference steps one problem per process).  A batch groups ht3d_hrl models of
 * any sizes, Rayleigh numbers, permeabilities, viscosities and time step rules
 * (the models stay owned by the caller) and steps them in lockstep.  The DYREL
 * solves of all problems run as one persistent launch per pseudo-iteration
 * over a packed (problem, tile) work list; the DYREL control runs on the
 * device, so each run of pseudo-iterations between two checks is replayed as
 * one HIP graph (use_graphs = 1) or launched kernel by kernel (0; also
 * whenever a rocprofiler tool is configured, ROCP_TOOL_LIBRARIES set, unless
 * HT3D_BATCH_GRAPHS_UNDER_PROFILER=1: a kernel trace of the graph form ended in
 * a host-side fault inside the runtime's graph launch).  The DYREL
 * parameters and the precision are the batch's (all problems share the check
 * schedule); results equal those of stepping each model alone.              */
typedef struct ht3d_batch ht3d_batch;
typedef struct {
    double ms_solve;            /* batched solve loops (device time, first to last block) */
    double bytes_solve;         /* their algorithmic bytes: 8 N s per pseudo-iteration,
                                   6 N x 8 per defect-correction anchor                   */
    double ms_step;             /* whole batched steps (host wall clock)                  */
    long long steps, blocks, graph_launches, graphs, repacks, iterations;
    int tile_rows[4], tile_z[4];/* chosen tiles: thermal FP64, Darcy FP64, thermal FP32, Darcy FP32 */
    int tile_cpl[4];            /* their x cells per lane                                 */
    int grid_mode[4];           /* their launch grids: 0 persistent, 1 one block per item */
    int pair_rows[4], pair_z[4];/* pair tiles (two iterations per launch) in the same order;
                                   0: the solve kind runs single iterations only          */
    int pair_cpl[4];            /* their x cells per lane                                 */
    int vertex[4];              /* 1: the single-iteration launches form the face
                                   coefficients from the vertex field (ht3d_batch_tune_vertex) */
} ht3d_batch_stats;

int ht3d_batch_create(ht3d_hrl *const *models, int nprob, ht3d_batch **out);
int ht3d_batch_destroy(ht3d_batch *b);
int ht3d_batch_set_params(ht3d_batch *b, const ht3d_dyrel_params *pT, const ht3d_dyrel_params *pP, int precision,
                          int use_graphs);
/* one physical step of every problem; reps: one report per problem */
int ht3d_batch_step(ht3d_batch *b, ht3d_step_report *reps);
int ht3d_batch_stats_get(const ht3d_batch *b, ht3d_batch_stats *s);
int ht3d_batch_stats_reset(ht3d_batch *b);
/* The batch's own tile table: (rows per block, z cells per thread, x cells per
 * lane), the 2- and 4-cell tiles offered only when every problem's ncx is a
 * multiple of the count (4 only for the FP32 sweeps).  By default each solve
 * kind times every tile the batch can run on its first step;
 * ht3d_batch_tune_force(t) pins entry t for batches created or re-tuned
 * afterwards (a tile a batch cannot run falls back to the default), -1
 * restores timing.  No reference counterpart (the reference has no batch).   */
int ht3d_batch_tune_force(int tile);
int ht3d_batch_tile_count(void);
/* The batch's launch grid, also chosen by timing (ht3d_batch_stats.grid_mode):
 * 0 persistent (a few blocks per CU walk the work list), 1 one block per work
 * item.  Pins it for batches tuned afterwards (-1: time both).  Results are
 * the same bit for bit.  No reference counterpart.                           */
int ht3d_batch_tune_force_grid(int mode);
/* The vertex-coefficient form of the batched single-iteration launches: for
 * the problems whose solve has a vertex field (ht3d_dyrel_field.kv, the model's
 * ktv, and kfv without viscosity), the face coefficients are formed from it in
 * registers (6 instead of 8 words per cell and iteration, the FP64 anchor
 * residual of the defect correction 3 instead of 5 words read).  mode -1
 * (default): timed against the face-array form per solve kind and precision;
 * 0: never; 1: always, on tiles with 2 cells per lane (and 4 in FP32).
 * Results are the same bit for bit.  Acts on batches tuned afterwards.  No
 * reference counterpart.                                                     */
int ht3d_batch_tune_vertex(int mode);
/* Two pseudo-iterations per launch for the runs of plain iterations (no check,
 * lambda update or stored F): bands of whole rows of a problem, 64*rows/ncx
 * rows each, overlapping by one row, with the intermediate iterate in LDS.  A
 * pair tile fits a batch when every problem has at least three rows per band.
 * mode -1 (default): used when the timing of the fastest pair tile beats the
 * single-iteration launches; 0: never; 1: always, when a tile fits.
 * ht3d_batch_tune_force_pair(t) pins pair tile t (-1: time them); results are
 * the same bit for bit either way.  Both act on batches tuned afterwards.  No
 * reference counterpart.                                                     */
int ht3d_batch_tune_pairs(int mode);
int ht3d_batch_tune_force_pair(int tile);
int ht3d_batch_pair_tile_count(void);
/* Pair tile `tile` of the batch's table: rows of 64 lanes per block, planes per
 * item and x cells per lane (4 cells: the FP32 sweeps only).  No reference
 * counterpart.                                                               */
int ht3d_batch_pair_tile_info(int tile, int *rows, int *planes, int *cells_per_lane);

/* Lognormal permeability on vertices (spectral randomisation) plus an optional
 * planar fault; K (3*M) and phi (M) in host memory.  Set-up only.            */
int ht3d_lognormal_field(double *kfv, int ncx, int ncy, int ncz, double xmin, double ymin, double zmin, double dx,
                         double dy, double dz, const double *K, const double *phi, int M, double sigma,
                         double fault_mult, double fault_half, const double *fault_n, const double *fault_x0);

#ifdef __cplusplus
}
#endif
#endif /* HT3D_H */
