/*
 * ht3d_oracle_dim.c -- CPU oracle for the dimensional (v18) property kernels.
 *
 * TEST INFRASTRUCTURE ONLY, like ht3d_oracle.c: the oracle the HIP library's
 * dimensional entries are compared with.  The product path never loads it.
 *
 * Restates, one loop nest per reference kernel and in FP64:
 *   Dimensional/HydroThermal3D_v18.jl:29-39           (EOS and property closures)
 *   Dimensional/kernels_HT3D_dimensional.jl:40-104    (vertex conductivities, fluid density)
 *   Dimensional/kernels_HT3D_dimensional.jl:110-119   (vertex <-> cell smoothing)
 *   Dimensional/kernels_HT3D_dimensional.jl:148-160   (effective rho*C)
 *   Dimensional/kernels_HT3D_dimensional.jl:183-205   (immersed-surface temperature BCs)
 *   Dimensional/kernels_HT3D_dimensional.jl:225-259   (phi*rho_f, phi*drho/dP)
 *   Dimensional/kernels_HT3D_dimensional.jl:291-314   (immersed-surface pressure BCs)
 *   Dimensional/kernels_HT3D_dimensional.jl:328-332   (Init_vel!, dimensional)
 *
 * Evaluation-order policy (the same in the HIP library):
 *   * sums and products fold left to right, as Julia evaluates `a+b+c` and
 *     `a*b*c`; integer literals are promoted to double first;
 *   * x^2 = x*x and x^3 = (x*x)*x (Julia's literal powers); x^4 is Julia's
 *     compensated power by squaring for Float64 (each square carries its exact
 *     rounding error, obtained with fma), which returns x^4 rounded once in
 *     all but rare ties;
 *   * 10^x and exp(x) call the C library; they are the only operations that
 *     can differ from the device by an ulp, so the GPU parity tests give the
 *     kernels that use them (mu_f, k_F) a relative tolerance of 1e-14.
 * The file is compiled with -ffp-contract=off: fma appears only where written.
 *
 * Ghost-fill semantics of the immersed-surface BC kernels.  In the reference
 * every rule of SetTemperatureBCs! / SetPressureBCs! runs in every thread at
 * once, and a rule's source cell may be rewritten by another rule in the same
 * launch.  The oracle and the HIP kernel define the result without a race: a
 * rule reads its source's value after the source's own rule has been
 * applied (the order of dependencies).  Where sources are not rewritten -- air
 * cells already at the surface value, no immersed cell next to a ghost layer,
 * which is every state the v18 driver creates -- this equals the reference's
 * result in any execution order, except on the ghost lines shared by an x face
 * and a z face (i = 1 or ncx+2 with k = ncz+2 in the temperature fill), whose
 * value depends on the thread order in the reference and which no stencil on
 * the path reads.  tests/test_oracle_dim.py checks this against a literal
 * one-thread execution of the reference kernel.
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
/* EOS and property closures: Dimensional/HydroThermal3D_v18.jl:29-39        */
/* ------------------------------------------------------------------------ */

/* Julia's Float64 x^4 (Base.pow_body with n = 4): two squarings, each with its
   error term carried in a second word. */
static inline double pow4(double x) {
    const double x2 = x * x;
    const double e2 = fma(x, x, -x2);
    const double x4 = x2 * x2;
    const double e4 = fma(x2, x2, -x4);
    const double lo = e4 + (x2 * 2.0) * e2;
    return x4 + lo;
}

static inline double kTs(double T) { return 3.1138 - 0.0023 * T; }                 /* v18:29 */
static inline double kTf(double T) { return -0.069 + 0.0012 * T; }                 /* v18:30 */
static inline double Cf(double T) { return 1000.0 + 7.5 * T; }                     /* v18:31 */
static inline double Cs(double T) { return 0.5915 * T + 636.14; }                  /* v18:32 */
static inline double rhos(double T) { return 2800.0 * (1.0 - (0.000024 * (T - 293.0))); }  /* v18:33 */

/* v18:35, fifteen terms summed left to right */
static inline double rhof_C(double Tc, double p) {
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

/* v18:36, ten terms left to right */
static inline double drhodP_C(double Tc, double p) {
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

static inline double muf(double T) { return 2.414e-5 * pow(10.0, 247.8 / (T - 140.0)); }  /* v18:37 */
static inline double kF(double y, double delta) { return 5e-16 * exp(y / delta); }     /* v18:38 */
static inline double phi_of(double phase, double phi0) { return phase == 2.0 ? 3.0 * phi0 : phi0; }  /* v18:39 */

/* scalar entries for the closed-form checks */
EXPORT double o_eos(int which, double a, double b) {
    switch (which) {
        case 0: return kTs(a);
        case 1: return kTf(a);
        case 2: return Cf(a);
        case 3: return Cs(a);
        case 4: return rhos(a);
        case 5: return rhof_C(a, b);
        case 6: return drhodP_C(a, b);
        case 7: return muf(a);
        case 8: return kF(a, b);
        case 9: return phi_of(a, b);
        case 10: return pow4(a);
        default: return NAN;
    }
}

/* Temperature (or pressure) interpolated from the eight extended centroids
   around vertex (i,j,k): kernels_HT3D_dimensional.jl:42-43 and :51-52. */
#define VTX_AVG(A)                                                                                      \
    (0.125 * (((Ee(A, i + 1, j + 1, k + 1) + Ee(A, i + 1, j, k + 1)) + Ee(A, i, j + 1, k + 1)) + Ee(A, i, j, k + 1)) + \
     0.125 * (((Ee(A, i + 1, j + 1, k) + Ee(A, i + 1, j, k)) + Ee(A, i, j + 1, k)) + Ee(A, i, j, k)))

/* ComputeThermalConductivity  kernels_HT3D_dimensional.jl:40-49 */
EXPORT void o_ComputeThermalConductivity(double *ktv, const double *Tce, const double *phv, double phi0,
                                         double scale_kt, double scale_T, int nx, int ny, int nz) {
    PFOR_K(nz + 1) FOR_JI(nx + 1, ny + 1) {
        double Tv = VTX_AVG(Tce);
        Tv *= scale_T;
        const double ph = VV(phv, i, j, k);
        VV(ktv, i, j, k) = ((1.0 - phi_of(ph, phi0)) * kTs(Tv) + phi_of(ph, phi0) * kTf(Tv)) / scale_kt;
    }
}

/* ComputeHydroConductivity  kernels_HT3D_dimensional.jl:51-83 */
EXPORT void o_ComputeHydroConductivity(double *kfv, const double *Tce, const double *Pce, const double *phv,
                                       double ymin, double dy, double k_fact, double delta, double scale_sigma,
                                       double scale_T, double scale_L, double scale_t, double scale_eta, int mode,
                                       int nx, int ny, int nz) {
    PFOR_K(nz + 1) FOR_JI(nx + 1, ny + 1) {
        double Tv = VTX_AVG(Tce);
        Tv *= scale_T;
        if (Tv < 293.15) Tv = 293.15;
        double Pv = VTX_AVG(Pce);
        Pv *= scale_sigma;
        const double y = ymin + (double)(j - 1) * dy;
        const double rk_mu = rhof_C(Tv - 273.15, Pv) * kF(y * scale_L, delta * scale_L) / muf(Tv);
        double *out = &VV(kfv, i, j, k);
        if (mode == 1)
            *out = rk_mu / scale_t;
        else if (mode == 2)
            *out = (kF(y * scale_L, delta * scale_L) / muf(Tv)) / (scale_L * scale_L / scale_eta);
        else if (mode == 3)
            *out = kF(y * scale_L, delta * scale_L) / (scale_L * scale_L);
        if (VV(phv, i, j, k) == 2.0) *out *= k_fact;
    }
}

/* ComputeFluidDensity  kernels_HT3D_dimensional.jl:85-99 */
EXPORT void o_ComputeFluidDensity(double *rhofv, const double *Tce, const double *Pce, const double *phv,
                                  double scale_sigma, double scale_T, double scale_rho, int nx, int ny, int nz) {
    PFOR_K(nz + 1) FOR_JI(nx + 1, ny + 1) {
        double Tv = VTX_AVG(Tce);
        Tv *= scale_T;
        double Pv = VTX_AVG(Pce);
        Pv *= scale_sigma;
        VV(rhofv, i, j, k) = rhof_C(Tv - 273.15, Pv) / scale_rho;
        if (VV(phv, i, j, k) == 1.0) VV(rhofv, i, j, k) = 0.0 / scale_rho;
    }
}

/* @av: eight-point cube mean, terms in x, then y, then z order, times 0.125 */
#define AV8(A, n1, n2)                                                                                 \
    ((AT(A, n1, n2, i, j, k) + AT(A, n1, n2, i + 1, j, k) + AT(A, n1, n2, i, j + 1, k) +                \
      AT(A, n1, n2, i + 1, j + 1, k) + AT(A, n1, n2, i, j, k + 1) + AT(A, n1, n2, i + 1, j, k + 1) +    \
      AT(A, n1, n2, i, j + 1, k + 1) + AT(A, n1, n2, i + 1, j + 1, k + 1)) * 0.125)

/* SmoothConductivityV2C  kernels_HT3D_dimensional.jl:110-113: @all(ktc) = @av(ktv) */
EXPORT void o_SmoothConductivityV2C(double *ktc, const double *ktv, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Cc(ktc, i, j, k) = AV8(ktv, nx + 1, ny + 1);
}

/* SmoothConductivityC2V  kernels_HT3D_dimensional.jl:115-118: @inn(ktv) = @av(ktc) */
EXPORT void o_SmoothConductivityC2V(double *ktv, const double *ktc, int nx, int ny, int nz) {
    PFOR_K(nz - 1) FOR_JI(nx - 1, ny - 1) VV(ktv, i + 1, j + 1, k + 1) = AV8(ktc, nx, ny);
}

/* ComputeρCeffective!  kernels_HT3D_dimensional.jl:148-160 (air cells are left unchanged) */
EXPORT void o_ComputeRhoCeffective(double *rhoC_eff, const double *Tc_ex, const double *Pc_ex,
                                   const double *phc_ex, double phi_i, double sca_T, double sca_sigma,
                                   double sca_C, double sca_rho, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        const double ph = Ee(phc_ex, i + 1, j + 1, k + 1);
        if (fabs(ph) != 1.0) {
            const double Tsca = Ee(Tc_ex, i + 1, j + 1, k + 1) * sca_T;
            const double Psca = Ee(Pc_ex, i + 1, j + 1, k + 1) * sca_sigma;
            const double phi = phi_of(ph, phi_i);
            const double rho_eff = phi * rhof_C(Tsca - 273.15, Psca) + (1.0 - phi) * rhos(Tsca);
            const double Ceff = phi * Cf(Tsca) + (1.0 - phi) * Cs(Tsca);
            Cc(rhoC_eff, i, j, k) = (rho_eff * Ceff) / sca_rho / sca_C;
        }
    }
}

/* ϕρf / ϕρf_old  kernels_HT3D_dimensional.jl:225-247 (cells with phase 1 unchanged).
   Temperatures and pressures come from E arrays (ϕρf) or cell arrays (ϕρf_old). */
static void phirhof_impl(double *out, const double *T, const double *P, int ext, const double *phc_ex,
                         double phi_i, double scale_rho, double scale_T, double scale_sigma, int deriv,
                         int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        const double ph = Ee(phc_ex, i + 1, j + 1, k + 1);
        if (ph != 1.0) {
            const double Tk = ext ? Ee(T, i + 1, j + 1, k + 1) : Cc(T, i, j, k);
            const double Pk = ext ? Ee(P, i + 1, j + 1, k + 1) : Cc(P, i, j, k);
            const double Tc = Tk * scale_T - 273.15;
            const double Pp = Pk * scale_sigma;
            if (deriv)
                Cc(out, i, j, k) = phi_of(ph, phi_i) * drhodP_C(Tc, Pp) / (scale_rho / scale_sigma);
            else
                Cc(out, i, j, k) = phi_of(ph, phi_i) * rhof_C(Tc, Pp) / scale_rho;
        }
    }
}

EXPORT void o_PhiRhof(double *out, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double phi_i,
                      double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    phirhof_impl(out, Tc_ex, Pc_ex, 1, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, 0, nx, ny, nz);
}

EXPORT void o_PhiRhof_old(double *out, const double *Tc0, const double *Pc0, const double *phc_ex, double phi_i,
                          double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    phirhof_impl(out, Tc0, Pc0, 0, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, 0, nx, ny, nz);
}

/* ϕdρdP!  kernels_HT3D_dimensional.jl:249-259 */
EXPORT void o_PhiDrhoDP(double *out, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double phi_i,
                        double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    phirhof_impl(out, Tc_ex, Pc_ex, 1, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, 1, nx, ny, nz);
}

/* Init_vel!  kernels_HT3D_dimensional.jl:328-332: V = q / (rho_f averaged to the face) */
EXPORT void o_dim_Init_vel(double *Vx, double *Vy, double *Vz, const double *qx, const double *qy,
                           const double *qz, const double *rhofv, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx + 1, ny) FX(Vx, i, j, k) = FX(qx, i, j, k) / AV_YZA(rhofv, nx + 1, ny + 1);
    PFOR_K(nz) FOR_JI(nx, ny + 1) FY(Vy, i, j, k) = FY(qy, i, j, k) / AV_XZA(rhofv, nx + 1, ny + 1);
    PFOR_K(nz + 1) FOR_JI(nx, ny) FZ(Vz, i, j, k) = FZ(qz, i, j, k) / AV_XYA(rhofv, nx + 1, ny + 1);
}

/* ------------------------------------------------------------------------ */
/* Immersed-surface ghost fills                                              */
/* ------------------------------------------------------------------------ */

/* Which rule of the reference kernel gives cell (i,j,k) of an E array its final
   value: the later statements of the kernel body override the earlier ones. */
enum { R_KEEP, R_AIR, R_XLO, R_XHI, R_BOT, R_IMM, R_ZLO, R_ZHI };

typedef struct {
    int nx, ny, nz;
    int temperature;       /* 1: SetTemperatureBCs!, 0: SetPressureBCs! */
    const double *A;       /* values before the launch */
    const double *ky, *wN, *ph;
    double top, qyS_dy;    /* TN or PN; qyS/_dy */
} bc_ctx;

static int bc_rule(const bc_ctx *c, int i, int j, int k) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const int iE = nx + 2, jE = ny + 2, kE = nz + 2;
    if (i > 1 && i <= iE - 1 && j > 1 && j <= jE - 1) {
        if (k == 1) return R_ZLO;
        if (k == kE) return R_ZHI;
    }
    const double ph = Ee(c->ph, i, j, k);
    if (ph == -1.0 && j > 1) return R_IMM;   /* the reference reads the cell below; none below j = 1 */
    const int kx_hi = c->temperature ? kE : kE - 1;   /* the temperature x fill runs to k = ncz+2 */
    if (j > 1 && j <= jE - 1 && k > 1 && k <= kx_hi) {
        if (i == 1) return R_XLO;
        if (i == iE) return R_XHI;
    }
    if (j == 1 && i > 1 && i <= iE - 1 && k > 1 && k <= kE - 1) return R_BOT;
    if (ph == 1.0) return R_AIR;
    return R_KEEP;
}

static double bc_value(const bc_ctx *c, int i, int j, int k, int depth) {
    const int nx = c->nx, ny = c->ny;
    if (depth > 4) return Ee(c->A, i, j, k);
    switch (bc_rule(c, i, j, k)) {
        case R_ZLO: return bc_value(c, i, j, 2, depth + 1);
        case R_ZHI: return bc_value(c, i, j, c->nz + 1, depth + 1);
        case R_IMM: {
            const double w = AT(c->wN, nx + 2, c->nz + 2, i, k, 1);   /* wN: (ncx+2) x (ncz+2) */
            return (c->top - (1.0 - w) * bc_value(c, i, j - 1, k, depth + 1)) / w;
        }
        case R_XLO: return bc_value(c, 2, j, k, depth + 1);
        case R_XHI: return bc_value(c, nx + 1, j, k, depth + 1);
        case R_BOT:
            if (c->temperature) {
                const double kS = FY(c->ky, i - 1, 1, k - 1);
                return c->qyS_dy / kS + bc_value(c, i, 2, k, depth + 1);
            }
            return bc_value(c, i, 2, k, depth + 1);
        case R_AIR: return c->top;
        default: return Ee(c->A, i, j, k);
    }
}

static void bc_apply(bc_ctx *c, double *A) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t n = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    double *old = (double *)malloc(n * sizeof(double));
    memcpy(old, A, n * sizeof(double));
    c->A = old;
    PFOR_K(nz + 2) FOR_JI(nx + 2, ny + 2) {
        if (bc_rule(c, i, j, k) != R_KEEP) Ee(A, i, j, k) = bc_value(c, i, j, k, 0);
    }
    free(old);
}

/* SetTemperatureBCs!  kernels_HT3D_dimensional.jl:183-205.  kS, TS, y_plateau,
   a2, b2, dTdy, dx, dy, sticky_air, sca_T, sca_kT and phi_i are accepted for
   the reference's argument order and unused, as in the reference. */
EXPORT void o_dim_SetTemperatureBCs(double *Tc_ex, const double *ky, const double *wN, const double *phc_ex,
                                    double qyS, double _dy, double kS, double TS, double TN, double y_plateau,
                                    double a2, double b2, double dTdy, double dx, double dy, double sticky_air,
                                    double sca_T, double sca_kT, double phi_i, int nx, int ny, int nz) {
    bc_ctx c = {nx, ny, nz, 1, NULL, ky, wN, phc_ex, TN, qyS / _dy};
    bc_apply(&c, Tc_ex);
}

/* SetPressureBCs!  kernels_HT3D_dimensional.jl:291-314 (bottom: Neumann copy). */
EXPORT void o_dim_SetPressureBCs(double *Pc_ex, const double *wN, const double *phc_ex, double PS, double PN,
                                 double y_plateau, double a2, double b2, double rhof, double g, double dx,
                                 double dy, double sticky_air, int nx, int ny, int nz) {
    bc_ctx c = {nx, ny, nz, 0, NULL, NULL, wN, phc_ex, PN, 0.0};
    bc_apply(&c, Pc_ex);
}

/* ------------------------------------------------------------------------ */
/* v18 DYREL solves, unfused: Dimensional/HydroThermal3D_v18.jl:432-478 (heat) */
/* and :489-537 (Darcy), from the Gershgorin set-up to convergence.           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, dt, transient;
    double *Tc_ex;                 /* E, in/out                                          */
    const double *Tc0;             /* c: temperature of the previous step (InitThermal!) */
    const double *rhoC;            /* c: rho*C (ComputeρCeffective!)                       */
    const double *kx, *ky, *kz;    /* face conductivities (InitConductivity!)            */
    const double *phc_ex, *wN;
    double qyS, TN, Qt, norm_scale;
    double *G, *PC, *F, *F0, *Fcit, *dumc, *qx, *qy, *qz;   /* left as the reference leaves them */
} o_v18_thermal;

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, dt, transient;
    double *Pc_ex;                 /* E, in/out                                              */
    const double *Tc_ex;           /* E: temperature of the fluid-density closure            */
    double *phirho;                /* c: phi*drho/dP on entry, phi*rho_f of the last iteration on exit */
    const double *phirhof0;        /* c: phi*rho_f of the previous step (ϕρf_old)            */
    const double *rhofv;           /* V: fluid density (ComputeFluidDensity)                 */
    const double *kx, *ky, *kz, *phc_ex, *wN;
    double PN, g, phi_i, scale_rho, scale_T, scale_sigma, norm_scale;
    double *G, *PC, *F, *F0, *Fcit, *dumc, *qx, *qy, *qz;
} o_v18_darcy;

/* The DYREL control flow shared by both solves (v18:432-478): `sweep` runs the
   reference's per-iteration kernels on the unknown X. */
typedef void (*sweep_fn)(void *ctx);

static int v18_control(int nx, int ny, int nz, const double *rhoC, const double *kx, const double *ky,
                       const double *kz, double transient, double dt, double dx, double dy, double dz, double *X,
                       double *G, double *PC, double *F, double *F0, double *Fcit, double *dumc, double norm_scale,
                       const o_dyrel_params *p, o_dyrel_report *rep, sweep_fn sweep, void *ctx) {
    const size_t N = (size_t)nx * ny * nz;
    o_GershgorinPoisson(G, PC, rhoC, kx, ky, kz, transient, dt, dx, dy, dz, nx, ny, nz);
    double lmax = o_reduce(G, (size_t)nx, (size_t)ny * nz, R_MAX);
    double lmin = lmax / p->lmin_div;
    const double dtau = 2.0 / sqrt(lmax) * p->CFL;
    double c = 2.0 * sqrt(lmin) * p->cfact;
    double h1 = (2 - c * dtau) / (2 + c * dtau), h2 = 2 * dtau / (2 + c * dtau);
    o_ResetA(F, F0, nx, ny, nz);
    double nF_abs = 0.0, nF_rel = 0.0, nF_ini = 0.0;
    int status = 10, iter, n_checks = 0, n_lambda = 0;
    for (iter = 1; iter <= p->nitmax; ++iter) {
        const int lam = iter <= 2 || iter % p->nitout == 0;
        const int chk = lam || (p->ncheck > 0 && iter % p->ncheck == 0);
        if (lam) o_SwapDYREL(NULL, Fcit, X, F, nx, ny, nz);
        sweep(ctx);
        o_DYRELUpdate(F0, X, F, h1, h2, dtau, nx, ny, nz);
        if (chk) {
            ++n_checks;
            o_Multiply(dumc, F, PC, nx, ny, nz);
            if (p->norm == 0)
                nF_abs = o_reduce(dumc, (size_t)nx, (size_t)ny * nz, R_SUM_ABS) / (double)N / sqrt((double)N) * norm_scale;
            else
                nF_abs = o_reduce(dumc, (size_t)nx, (size_t)ny * nz, R_MAX_ABS) * norm_scale;
            if (iter == 1) nF_ini = nF_abs;
            nF_rel = nF_abs / nF_ini;
            if (isnan(nF_abs)) { status = 11; break; }
            if (nF_rel > 100) { status = 12; break; }
            if (lam && iter > 1) {
                o_RayleighQuotientNumerator(dumc, F0, Fcit, F, dtau, nx, ny, nz);
                lmin = fabs(o_reduce(dumc, (size_t)nx, (size_t)ny * nz, R_SUM));
                o_RayleighQuotientDenominator(dumc, F0, dtau, nx, ny, nz);
                lmin /= fabs(o_reduce(dumc, (size_t)nx, (size_t)ny * nz, R_SUM));
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

static void thermal_sweep(void *ctx) {   /* v18:449-452 */
    o_v18_thermal *f = (o_v18_thermal *)ctx;
    const int nx = f->nx, ny = f->ny, nz = f->nz;
    const double _dx = 1.0 / f->dx, _dy = 1.0 / f->dy, _dz = 1.0 / f->dz, _dt = 1.0 / f->dt;
    o_dim_SetTemperatureBCs(f->Tc_ex, f->ky, f->wN, f->phc_ex, f->qyS, _dy, 0, 0, f->TN, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                            0, nx, ny, nz);
    o_ComputeFlux(f->qx, f->qy, f->qz, f->kx, f->ky, f->kz, f->Tc_ex, _dx, _dy, _dz, nx, ny, nz);
    o_ResidualTemperatureLinearised(f->F, f->Tc_ex, f->Tc0, f->rhoC, f->phc_ex, f->PC, f->qx, f->qy, f->qz, f->Qt,
                                    f->transient, _dt, _dx, _dy, _dz, nx, ny, nz);
}

static void darcy_sweep(void *ctx) {     /* v18:508-512 */
    o_v18_darcy *f = (o_v18_darcy *)ctx;
    const int nx = f->nx, ny = f->ny, nz = f->nz;
    const double _dx = 1.0 / f->dx, _dy = 1.0 / f->dy, _dz = 1.0 / f->dz, _dt = 1.0 / f->dt;
    o_PhiRhof(f->phirho, f->Tc_ex, f->Pc_ex, f->phc_ex, f->phi_i, f->scale_rho, f->scale_T, f->scale_sigma, nx, ny, nz);
    o_dim_SetPressureBCs(f->Pc_ex, f->wN, f->phc_ex, 0, f->PN, 0, 0, 0, 0, 0, 0, 0, 0, nx, ny, nz);
    o_ComputeDarcyFlux(f->qx, f->qy, f->qz, f->rhofv, f->kx, f->ky, f->kz, f->Pc_ex, f->g, _dx, _dy, _dz, nx, ny, nz);
    o_ResidualFluidPressure(f->F, f->phc_ex, f->phirho, f->phirhof0, f->PC, f->qx, f->qy, f->qz, f->transient, _dt,
                            _dx, _dy, _dz, nx, ny, nz);
}

EXPORT int o_v18_thermal_solve(o_v18_thermal *f, const o_dyrel_params *p, o_dyrel_report *rep) {
    return v18_control(f->nx, f->ny, f->nz, f->rhoC, f->kx, f->ky, f->kz, f->transient, f->dt, f->dx, f->dy, f->dz,
                       f->Tc_ex, f->G, f->PC, f->F, f->F0, f->Fcit, f->dumc, f->norm_scale, p, rep, thermal_sweep, f);
}

EXPORT int o_v18_darcy_solve(o_v18_darcy *f, const o_dyrel_params *p, o_dyrel_report *rep) {
    return v18_control(f->nx, f->ny, f->nz, f->phirho, f->kx, f->ky, f->kz, f->transient, f->dt, f->dx, f->dy, f->dz,
                       f->Pc_ex, f->G, f->PC, f->F, f->F0, f->Fcit, f->dumc, f->norm_scale, p, rep, darcy_sweep, f);
}

/* ------------------------------------------------------------------------ */
/* Output fields: Dimensional/kernels_HT3D_postprocessing.jl (cell arrays,   */
/* NaN in air unless noted)                                                  */
/* ------------------------------------------------------------------------ */

#define PH_C (Ee(phc_ex, i + 1, j + 1, k + 1))

/* Phase!  :1-10 */
EXPORT void o_Phase(double *X, const double *phc_ex, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Cc(X, i, j, k) = PH_C != 1.0 ? PH_C : NAN;
}

/* Pressure! / Temperature!  :12-32 (field * scale) */
EXPORT void o_Pressure(double *X, const double *Pc_ex, const double *phc_ex, double scale, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Cc(X, i, j, k) = PH_C != 1.0 ? Ee(Pc_ex, i + 1, j + 1, k + 1) * scale : NAN;
}

/* Velocity!  :34-45 (face velocities averaged to the cell; only X is set to NaN in air) */
EXPORT void o_Velocity(double *X, double *Y, double *Z, const double *Vx, const double *Vy, const double *Vz,
                       const double *phc_ex, double scale_V, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        if (PH_C != 1.0) {
            Cc(X, i, j, k) = 0.5 * (FX(Vx, i, j, k) + FX(Vx, i + 1, j, k)) * scale_V;
            Cc(Y, i, j, k) = 0.5 * (FY(Vy, i, j, k) + FY(Vy, i, j + 1, k)) * scale_V;
            Cc(Z, i, j, k) = 0.5 * (FZ(Vz, i, j, k) + FZ(Vz, i, j, k + 1)) * scale_V;
        } else {
            Cc(X, i, j, k) = NAN;
        }
    }
}

/* EffectiveThermalConductivity!  :47-57 */
EXPORT void o_EffectiveThermalConductivity(double *X, const double *Tc_ex, double phi0, const double *phc_ex,
                                           double scale_T, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        if (PH_C != 1.0) {
            const double T = Ee(Tc_ex, i + 1, j + 1, k + 1) * scale_T;
            Cc(X, i, j, k) = ((1.0 - phi_of(PH_C, phi0)) * kTs(T) + phi_of(PH_C, phi0) * kTf(T));
        } else {
            Cc(X, i, j, k) = NAN;
        }
    }
}

/* EffectiveHeatCapacity!  :60-70 (the solid's C; air cells untouched) */
EXPORT void o_EffectiveHeatCapacity(double *Ceff, const double *Tc_ex, const double *Pc_ex, const double *phc_ex,
                                    double phi_i, double sca_T, double sca_sigma, double sca_C, double sca_rho,
                                    int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        if (PH_C != 1.0) Cc(Ceff, i, j, k) = Cs(Ee(Tc_ex, i + 1, j + 1, k + 1) * sca_T);
    }
}

/* Permeability!  :72-85 (cell-centre depth) */
EXPORT void o_Permeability(double *X, double k_fact, double ymin, double dy, const double *phc_ex, double delta,
                           double scale_L, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        if (PH_C != 1.0) {
            const double y = ymin + (double)(j - 1) * dy + dy / 2.0;
            double v = kF(y * scale_L, delta * scale_L);
            if (PH_C == 2.0) v *= k_fact;
            Cc(X, i, j, k) = v;
        } else {
            Cc(X, i, j, k) = NAN;
        }
    }
}

/* FluidDensity!  :87-98 (kg/m^3) */
EXPORT void o_FluidDensity(double *X, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double scale_T,
                           double scale_sigma, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        Cc(X, i, j, k) = PH_C != 1.0 ? rhof_C(Ee(Tc_ex, i + 1, j + 1, k + 1) * scale_T - 273.15,
                                              Ee(Pc_ex, i + 1, j + 1, k + 1) * scale_sigma)
                                     : NAN;
    }
}

/* Viscosity!  :100-109 (Pa s) */
EXPORT void o_Viscosity(double *X, const double *Tc_ex, const double *phc_ex, double scale_T, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) Cc(X, i, j, k) = PH_C != 1.0 ? muf(Ee(Tc_ex, i + 1, j + 1, k + 1) * scale_T) : NAN;
}

/* Peclet!  :111-128: L rho_f C_f |V| / k_eff */
EXPORT void o_Peclet(double *X, const double *Tc_ex, const double *Pc_ex, const double *Vx, const double *Vy,
                     const double *Vz, const double *phc_ex, double phi0, double scale_T, double scale_sigma,
                     double scale_V, double scale_L, int nx, int ny, int nz) {
    PFOR_K(nz) FOR_JI(nx, ny) {
        if (PH_C != 1.0) {
            const double x = 0.5 * (FX(Vx, i, j, k) + FX(Vx, i + 1, j, k)) * scale_V;
            const double y = 0.5 * (FY(Vy, i, j, k) + FY(Vy, i, j + 1, k)) * scale_V;
            const double z = 0.5 * (FZ(Vz, i, j, k) + FZ(Vz, i, j, k + 1)) * scale_V;
            const double V = sqrt(x * x + y * y + z * z);
            const double T = Ee(Tc_ex, i + 1, j + 1, k + 1) * scale_T;
            const double P = Ee(Pc_ex, i + 1, j + 1, k + 1) * scale_sigma;
            const double rho = rhof_C(T - 273.15, P);
            const double C = Cf(T);
            const double keff = ((1.0 - phi_of(PH_C, phi0)) * kTs(T) + phi_of(PH_C, phi0) * kTf(T));
            Cc(X, i, j, k) = scale_L * rho * C * V / keff;
        } else {
            Cc(X, i, j, k) = NAN;
        }
    }
}
