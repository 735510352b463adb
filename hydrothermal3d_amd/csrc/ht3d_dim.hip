// ht3d_dim.hip -- the dimensional (v18) property kernels and immersed-surface
// ghost fills, one HIP kernel per reference kernel.
//
// Entries replace the `@parallel` launches of
// Dimensional/kernels_HT3D_dimensional.jl:40-119,148-160,183-205,225-259,291-314,328-332
// with the reference's argument order (cell extents appended).  They evaluate
// the closures of Dimensional/HydroThermal3D_v18.jl:29-39 (fluid and rock
// properties) in the reference's order:
//   * sums and products fold left to right; x^2 = x*x, x^3 = (x*x)*x;
//   * x^4 is Julia's Float64 power by squaring with each square's rounding
//     error carried exactly (fma), i.e. x^4 rounded once in all but rare ties;
//   * 10^x and exp(x) use the device math library (within an ulp of the host's).
// Built with -ffp-contract=off like the rest of the library, so every kernel
// without 10^x or exp() equals an FP64 CPU evaluation bit for bit.
#include <cmath>

#include "ht3d_internal.hpp"
#include "ht3d_launch.hpp"

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
    double *A;
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

// Final value of cell (i,j,k): each rule reads its source cell's final value
// (the order of dependencies), so no cell is read while another thread of the
// launch rewrites it.  Chains longer than five rules (stacked immersed cells)
// read the stored value, as the CPU oracle does.
template <int D>
__device__ double bc_value(const BcArgs &a, int i, int j, int k) {
    if constexpr (D > 4) {
        return a.A[a.g.e(i, j, k)];
    } else {
        switch (bc_rule(a, i, j, k)) {
            case R_ZLO: return bc_value<D + 1>(a, i, j, 1);
            case R_ZHI: return bc_value<D + 1>(a, i, j, a.g.nz);
            case R_IMM: {
                const double w = a.wN[(size_t)k * (a.g.nx + 2) + i];
                return (a.top - (1.0 - w) * bc_value<D + 1>(a, i, j - 1, k)) / w;
            }
            case R_XLO: return bc_value<D + 1>(a, 1, j, k);
            case R_XHI: return bc_value<D + 1>(a, a.g.nx, j, k);
            case R_BOT:
                if (a.temperature) {
                    const double kS = a.ky[a.g.fy(i - 1, 0, k - 1)];
                    return a.qyS_dy / kS + bc_value<D + 1>(a, i, 1, k);
                }
                return bc_value<D + 1>(a, i, 1, k);
            case R_AIR: return a.top;
            default: return a.A[a.g.e(i, j, k)];
        }
    }
}

static int bc_launch(const char *name, const BcArgs &a) {
    const Grid g = a.g;
    return launch3d(name, g.nx + 2, g.ny + 2, g.nz + 2, [=] __device__(int i, int j, int k) {
        if (bc_rule(a, i, j, k) == R_KEEP) return;
        const double v = bc_value<0>(a, i, j, k);
        a.A[g.e(i, j, k)] = v;
    });
}

}  // namespace dim
}  // namespace ht3d

using namespace ht3d;
using namespace ht3d::dim;

extern "C" {

int ht3d_ComputeThermalConductivity(double *ktv, const double *Tce, const double *phv, double phi0, double scale_kt,
                                    double scale_T, int nx, int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(ktv); CHECK_PTR(Tce); CHECK_PTR(phv);
    Grid g{nx, ny, nz};
    return launch3d("ComputeThermalConductivity", nx + 1, ny + 1, nz + 1, [=] __device__(int i, int j, int k) {
        double Tv = vertex_mean(Tce, g, i, j, k);
        Tv *= scale_T;
        const double ph = phv[g.v(i, j, k)];
        ktv[g.v(i, j, k)] = ((1.0 - phi_of(ph, phi0)) * kTs(Tv) + phi_of(ph, phi0) * kTf(Tv)) / scale_kt;
    });
}

int ht3d_ComputeHydroConductivity(double *kfv, const double *Tce, const double *Pce, const double *phv, double ymin,
                                  double dy, double k_fact, double delta, double scale_sigma, double scale_T,
                                  double scale_L, double scale_t, double scale_eta, int mode, int nx, int ny,
                                  int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(kfv); CHECK_PTR(Tce); CHECK_PTR(Pce); CHECK_PTR(phv);
    Grid g{nx, ny, nz};
    return launch3d("ComputeHydroConductivity", nx + 1, ny + 1, nz + 1, [=] __device__(int i, int j, int k) {
        double Tv = vertex_mean(Tce, g, i, j, k);
        Tv *= scale_T;
        if (Tv < 293.15) Tv = 293.15;
        double Pv = vertex_mean(Pce, g, i, j, k);
        Pv *= scale_sigma;
        const double y = ymin + (double)j * dy;
        const size_t n = g.v(i, j, k);
        double out = kfv[n];
        if (mode == 1)
            out = rhof_C(Tv - 273.15, Pv) * kF(y * scale_L, delta * scale_L) / muf(Tv) / scale_t;
        else if (mode == 2)
            out = (kF(y * scale_L, delta * scale_L) / muf(Tv)) / (scale_L * scale_L / scale_eta);
        else if (mode == 3)
            out = kF(y * scale_L, delta * scale_L) / (scale_L * scale_L);
        if (phv[n] == 2.0) out *= k_fact;
        kfv[n] = out;
    });
}

int ht3d_ComputeFluidDensity(double *rhofv, const double *Tce, const double *Pce, const double *phv,
                             double scale_sigma, double scale_T, double scale_rho, int nx, int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(rhofv); CHECK_PTR(Tce); CHECK_PTR(Pce); CHECK_PTR(phv);
    Grid g{nx, ny, nz};
    return launch3d("ComputeFluidDensity", nx + 1, ny + 1, nz + 1, [=] __device__(int i, int j, int k) {
        double Tv = vertex_mean(Tce, g, i, j, k);
        Tv *= scale_T;
        double Pv = vertex_mean(Pce, g, i, j, k);
        Pv *= scale_sigma;
        const size_t n = g.v(i, j, k);
        rhofv[n] = phv[n] == 1.0 ? 0.0 / scale_rho : rhof_C(Tv - 273.15, Pv) / scale_rho;
    });
}

int ht3d_SmoothConductivityV2C(double *ktc, const double *ktv, int nx, int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(ktc); CHECK_PTR(ktv);
    Grid g{nx, ny, nz};
    return launch3d("SmoothConductivityV2C", nx, ny, nz, [=] __device__(int i, int j, int k) {
        ktc[g.c(i, j, k)] = cube_mean(ktv, [=](int a, int b, int c) { return g.v(a, b, c); }, i, j, k);
    });
}

int ht3d_SmoothConductivityC2V(double *ktv, const double *ktc, int nx, int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(ktc); CHECK_PTR(ktv);
    Grid g{nx, ny, nz};
    return launch3d("SmoothConductivityC2V", nx - 1, ny - 1, nz - 1, [=] __device__(int i, int j, int k) {
        ktv[g.v(i + 1, j + 1, k + 1)] = cube_mean(ktc, [=](int a, int b, int c) { return g.c(a, b, c); }, i, j, k);
    });
}

int ht3d_ComputeRhoCeffective(double *rhoC_eff, const double *Tc_ex, const double *Pc_ex, const double *phc_ex,
                              double phi_i, double sca_T, double sca_sigma, double sca_C, double sca_rho, int nx,
                              int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(rhoC_eff); CHECK_PTR(Tc_ex); CHECK_PTR(Pc_ex); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("ComputeRhoCeffective", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        const double ph = phc_ex[e];
        if (fabs(ph) == 1.0) return;
        const double Tsca = Tc_ex[e] * sca_T;
        const double Psca = Pc_ex[e] * sca_sigma;
        const double phi = phi_of(ph, phi_i);
        const double rho_eff = phi * rhof_C(Tsca - 273.15, Psca) + (1.0 - phi) * rhos(Tsca);
        const double Ceff = phi * Cf(Tsca) + (1.0 - phi) * Cs(Tsca);
        rhoC_eff[g.c(i, j, k)] = (rho_eff * Ceff) / sca_rho / sca_C;
    });
}

// ϕρf, ϕρf_old and ϕdρdP!: cells whose phase is 1 are left unchanged.
static int phirhof(const char *name, double *out, const double *T, const double *P, bool ext,
                   const double *phc_ex, double phi_i, double scale_rho, double scale_T, double scale_sigma,
                   bool deriv, int nx, int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(out); CHECK_PTR(T); CHECK_PTR(P); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d(name, nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1), c = g.c(i, j, k);
        const double ph = phc_ex[e];
        if (ph == 1.0) return;
        const double Tc = (ext ? T[e] : T[c]) * scale_T - 273.15;
        const double Pp = (ext ? P[e] : P[c]) * scale_sigma;
        out[c] = deriv ? phi_of(ph, phi_i) * drhodP_C(Tc, Pp) / (scale_rho / scale_sigma)
                       : phi_of(ph, phi_i) * rhof_C(Tc, Pp) / scale_rho;
    });
}

int ht3d_PhiRhof(double *phirhofc, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double phi_i,
                 double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    return phirhof("PhiRhof", phirhofc, Tc_ex, Pc_ex, true, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, false,
                   nx, ny, nz);
}

int ht3d_PhiRhof_old(double *phirhofc, const double *Tc0, const double *Pc0, const double *phc_ex, double phi_i,
                     double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    return phirhof("PhiRhof_old", phirhofc, Tc0, Pc0, false, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, false,
                   nx, ny, nz);
}

int ht3d_PhiDrhoDP(double *phidrhodP, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double phi_i,
                   double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    return phirhof("PhiDrhoDP", phidrhodP, Tc_ex, Pc_ex, true, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, true,
                   nx, ny, nz);
}

int ht3d_dim_Init_vel(double *Vx, double *Vy, double *Vz, const double *qx, const double *qy, const double *qz,
                      const double *rhofv, int nx, int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Vx); CHECK_PTR(Vy); CHECK_PTR(Vz); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(rhofv);
    Grid g{nx, ny, nz};
    int rc = launch3d("dim_Init_vel", nx + 1, ny, nz, [=] __device__(int i, int j, int k) {   // @av_yza
        const double r = (rhofv[g.v(i, j, k)] + rhofv[g.v(i, j + 1, k)] + rhofv[g.v(i, j, k + 1)] +
                          rhofv[g.v(i, j + 1, k + 1)]) * 0.25;
        Vx[g.fx(i, j, k)] = qx[g.fx(i, j, k)] / r;
    });
    if (rc) return rc;
    rc = launch3d("dim_Init_vel", nx, ny + 1, nz, [=] __device__(int i, int j, int k) {       // @av_xza
        const double r = (rhofv[g.v(i, j, k)] + rhofv[g.v(i + 1, j, k)] + rhofv[g.v(i, j, k + 1)] +
                          rhofv[g.v(i + 1, j, k + 1)]) * 0.25;
        Vy[g.fy(i, j, k)] = qy[g.fy(i, j, k)] / r;
    });
    if (rc) return rc;
    return launch3d("dim_Init_vel", nx, ny, nz + 1, [=] __device__(int i, int j, int k) {     // @av_xya
        const double r = (rhofv[g.v(i, j, k)] + rhofv[g.v(i + 1, j, k)] + rhofv[g.v(i, j + 1, k)] +
                          rhofv[g.v(i + 1, j + 1, k)]) * 0.25;
        Vz[g.fz(i, j, k)] = qz[g.fz(i, j, k)] / r;
    });
}

int ht3d_dim_SetTemperatureBCs(double *Tc_ex, const double *ky, const double *wN, const double *phc_ex, double qyS,
                               double _dy, double kS, double TS, double TN, double y_plateau, double a2, double b2,
                               double dTdy, double dx, double dy, double sticky_air, double sca_T, double sca_kT,
                               double phi_i, int nx, int ny, int nz) {
    (void)kS; (void)TS; (void)y_plateau; (void)a2; (void)b2; (void)dTdy; (void)dx; (void)dy; (void)sticky_air;
    (void)sca_T; (void)sca_kT; (void)phi_i;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Tc_ex); CHECK_PTR(ky); CHECK_PTR(wN); CHECK_PTR(phc_ex);
    BcArgs a{Tc_ex, ky, wN, phc_ex, TN, qyS / _dy, Grid{nx, ny, nz}, 1};
    return bc_launch("dim_SetTemperatureBCs", a);
}

int ht3d_dim_SetPressureBCs(double *Pc_ex, const double *wN, const double *phc_ex, double PS, double PN,
                            double y_plateau, double a2, double b2, double rhof, double g_, double dx, double dy,
                            double sticky_air, int nx, int ny, int nz) {
    (void)PS; (void)y_plateau; (void)a2; (void)b2; (void)rhof; (void)g_; (void)dx; (void)dy; (void)sticky_air;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Pc_ex); CHECK_PTR(wN); CHECK_PTR(phc_ex);
    BcArgs a{Pc_ex, nullptr, wN, phc_ex, PN, 0.0, Grid{nx, ny, nz}, 0};
    return bc_launch("dim_SetPressureBCs", a);
}

}  // extern "C"
