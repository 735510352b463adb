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

#include "ht3d_dim.hpp"

namespace ht3d {
namespace dim {

// The fill is done in place: every rule's sources are cells that no rule
// rewrites (bc_value), so reading a.A while writing `out` (the same array) is
// free of races.
static int bc_launch(const char *name, const BcArgs &a, double *out) {
    const Grid g = a.g;
    return launch3d(name, g.nx + 2, g.ny + 2, g.nz + 2, [=] __device__(int i, int j, int k) {
        if (bc_rule(a, i, j, k) == R_KEEP) return;
        const double v = bc_value(a, i, j, k);
        out[g.e(i, j, k)] = v;
    });
}

}  // namespace dim
}  // namespace ht3d

using namespace ht3d;
using namespace ht3d::dim;

extern "C" {

int ht3d_ComputeThermalConductivity(double *ktv, const double *Tce, const double *phv, double phi0, double scale_kt,
                                    double scale_T, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(ktv); CHECK_PTR(Tce); CHECK_PTR(phv);
    Grid g{nx, ny, nz};
    return launch3d("ComputeThermalConductivity", nx + 1, ny + 1, nz + 1, [=] __device__(int i, int j, int k) {
        double Tv = vertex_mean(Tce, g, i, j, k);
        Tv *= scale_T;
        const double ph = phv[g.v(i, j, k)];
        ktv[g.v(i, j, k)] = ((1.0 - phi_of(ph, phi0)) * kTs(Tv) + phi_of(ph, phi0) * kTf(Tv)) / scale_kt;
    });
    HT3D_API_END
}

int ht3d_ComputeHydroConductivity(double *kfv, const double *Tce, const double *Pce, const double *phv, double ymin,
                                  double dy, double k_fact, double delta, double scale_sigma, double scale_T,
                                  double scale_L, double scale_t, double scale_eta, int mode, int nx, int ny,
                                  int nz) {
    HT3D_API_BEGIN
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
    HT3D_API_END
}

int ht3d_ComputeFluidDensity(double *rhofv, const double *Tce, const double *Pce, const double *phv,
                             double scale_sigma, double scale_T, double scale_rho, int nx, int ny, int nz) {
    HT3D_API_BEGIN
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
    HT3D_API_END
}

int ht3d_SmoothConductivityV2C(double *ktc, const double *ktv, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(ktc); CHECK_PTR(ktv);
    Grid g{nx, ny, nz};
    return launch3d("SmoothConductivityV2C", nx, ny, nz, [=] __device__(int i, int j, int k) {
        ktc[g.c(i, j, k)] = cube_mean(ktv, [=](int a, int b, int c) { return g.v(a, b, c); }, i, j, k);
    });
    HT3D_API_END
}

int ht3d_SmoothConductivityC2V(double *ktv, const double *ktc, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(ktc); CHECK_PTR(ktv);
    Grid g{nx, ny, nz};
    return launch3d("SmoothConductivityC2V", nx - 1, ny - 1, nz - 1, [=] __device__(int i, int j, int k) {
        ktv[g.v(i + 1, j + 1, k + 1)] = cube_mean(ktc, [=](int a, int b, int c) { return g.c(a, b, c); }, i, j, k);
    });
    HT3D_API_END
}

int ht3d_ComputeRhoCeffective(double *rhoC_eff, const double *Tc_ex, const double *Pc_ex, const double *phc_ex,
                              double phi_i, double sca_T, double sca_sigma, double sca_C, double sca_rho, int nx,
                              int ny, int nz) {
    HT3D_API_BEGIN
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
    HT3D_API_END
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
    HT3D_API_BEGIN
    return phirhof("PhiRhof", phirhofc, Tc_ex, Pc_ex, true, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, false,
                   nx, ny, nz);
    HT3D_API_END
}

int ht3d_PhiRhof_old(double *phirhofc, const double *Tc0, const double *Pc0, const double *phc_ex, double phi_i,
                     double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return phirhof("PhiRhof_old", phirhofc, Tc0, Pc0, false, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, false,
                   nx, ny, nz);
    HT3D_API_END
}

int ht3d_PhiDrhoDP(double *phidrhodP, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double phi_i,
                   double scale_rho, double scale_T, double scale_sigma, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return phirhof("PhiDrhoDP", phidrhodP, Tc_ex, Pc_ex, true, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, true,
                   nx, ny, nz);
    HT3D_API_END
}

int ht3d_dim_Init_vel(double *Vx, double *Vy, double *Vz, const double *qx, const double *qy, const double *qz,
                      const double *rhofv, int nx, int ny, int nz) {
    HT3D_API_BEGIN
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
    HT3D_API_END
}

int ht3d_dim_SetTemperatureBCs(double *Tc_ex, const double *ky, const double *wN, const double *phc_ex, double qyS,
                               double _dy, double kS, double TS, double TN, double y_plateau, double a2, double b2,
                               double dTdy, double dx, double dy, double sticky_air, double sca_T, double sca_kT,
                               double phi_i, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    (void)kS; (void)TS; (void)y_plateau; (void)a2; (void)b2; (void)dTdy; (void)dx; (void)dy; (void)sticky_air;
    (void)sca_T; (void)sca_kT; (void)phi_i;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Tc_ex); CHECK_PTR(ky); CHECK_PTR(wN); CHECK_PTR(phc_ex);
    BcArgs a{Tc_ex, ky, wN, phc_ex, TN, qyS / _dy, Grid{nx, ny, nz}, 1};
    return bc_launch("dim_SetTemperatureBCs", a, Tc_ex);
    HT3D_API_END
}

int ht3d_dim_SetPressureBCs(double *Pc_ex, const double *wN, const double *phc_ex, double PS, double PN,
                            double y_plateau, double a2, double b2, double rhof, double g_, double dx, double dy,
                            double sticky_air, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    (void)PS; (void)y_plateau; (void)a2; (void)b2; (void)rhof; (void)g_; (void)dx; (void)dy; (void)sticky_air;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Pc_ex); CHECK_PTR(wN); CHECK_PTR(phc_ex);
    BcArgs a{Pc_ex, nullptr, wN, phc_ex, PN, 0.0, Grid{nx, ny, nz}, 0};
    return bc_launch("dim_SetPressureBCs", a, Pc_ex);
    HT3D_API_END
}

// ---------------------------------------------------------------- output fields
// Dimensional/kernels_HT3D_postprocessing.jl: cell arrays for the VTK output,
// NaN in air unless noted.

int ht3d_Phase(double *X, const double *phc_ex, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("Phase", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const double ph = phc_ex[g.e(i + 1, j + 1, k + 1)];
        X[g.c(i, j, k)] = ph != 1.0 ? ph : NAN;
    });
    HT3D_API_END
}

static int scaled_field(const char *name, double *X, const double *F, const double *phc_ex, double scale, int nx,
                        int ny, int nz) {
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(F); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d(name, nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        X[g.c(i, j, k)] = phc_ex[e] != 1.0 ? F[e] * scale : NAN;
    });
}

int ht3d_Pressure(double *X, const double *Pc_ex, const double *phc_ex, double scale_sigma, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return scaled_field("Pressure", X, Pc_ex, phc_ex, scale_sigma, nx, ny, nz);
    HT3D_API_END
}

int ht3d_Temperature(double *X, const double *Tc_ex, const double *phc_ex, double scale_T, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return scaled_field("Temperature", X, Tc_ex, phc_ex, scale_T, nx, ny, nz);
    HT3D_API_END
}

int ht3d_Velocity(double *X, double *Y, double *Z, const double *Vx, const double *Vy, const double *Vz,
                  const double *phc_ex, double scale_V, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(Y); CHECK_PTR(Z); CHECK_PTR(Vx); CHECK_PTR(Vy); CHECK_PTR(Vz); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("Velocity", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        if (phc_ex[g.e(i + 1, j + 1, k + 1)] != 1.0) {
            X[c] = 0.5 * (Vx[g.fx(i, j, k)] + Vx[g.fx(i + 1, j, k)]) * scale_V;
            Y[c] = 0.5 * (Vy[g.fy(i, j, k)] + Vy[g.fy(i, j + 1, k)]) * scale_V;
            Z[c] = 0.5 * (Vz[g.fz(i, j, k)] + Vz[g.fz(i, j, k + 1)]) * scale_V;
        } else {
            X[c] = NAN;   // Y and Z are left as they are, as in the reference
        }
    });
    HT3D_API_END
}

int ht3d_EffectiveThermalConductivity(double *X, const double *Tc_ex, double phi0, const double *phc_ex,
                                      double scale_T, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(Tc_ex); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("EffectiveThermalConductivity", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        const double ph = phc_ex[e];
        if (ph != 1.0) {
            const double T = Tc_ex[e] * scale_T;
            X[g.c(i, j, k)] = ((1.0 - phi_of(ph, phi0)) * kTs(T) + phi_of(ph, phi0) * kTf(T));
        } else {
            X[g.c(i, j, k)] = NAN;
        }
    });
    HT3D_API_END
}

int ht3d_EffectiveHeatCapacity(double *Ceff, const double *Tc_ex, const double *Pc_ex, const double *phc_ex,
                               double phi_i, double sca_T, double sca_sigma, double sca_C, double sca_rho, int nx,
                               int ny, int nz) {
    HT3D_API_BEGIN
    (void)phi_i; (void)sca_sigma; (void)sca_C; (void)sca_rho;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Ceff); CHECK_PTR(Tc_ex); CHECK_PTR(Pc_ex); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("EffectiveHeatCapacity", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        if (phc_ex[e] != 1.0) Ceff[g.c(i, j, k)] = Cs(Tc_ex[e] * sca_T);
    });
    HT3D_API_END
}

int ht3d_Permeability(double *X, double k_fact, double ymin, double dy, const double *phc_ex, double delta,
                      double scale_L, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("Permeability", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const double ph = phc_ex[g.e(i + 1, j + 1, k + 1)];
        if (ph != 1.0) {
            const double y = ymin + (double)j * dy + dy / 2.0;
            double v = kF(y * scale_L, delta * scale_L);
            if (ph == 2.0) v *= k_fact;
            X[g.c(i, j, k)] = v;
        } else {
            X[g.c(i, j, k)] = NAN;
        }
    });
    HT3D_API_END
}

int ht3d_FluidDensity(double *X, const double *Tc_ex, const double *Pc_ex, const double *phc_ex, double scale_T,
                      double scale_sigma, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(Tc_ex); CHECK_PTR(Pc_ex); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("FluidDensity", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        X[g.c(i, j, k)] = phc_ex[e] != 1.0 ? rhof_C(Tc_ex[e] * scale_T - 273.15, Pc_ex[e] * scale_sigma) : NAN;
    });
    HT3D_API_END
}

int ht3d_Viscosity(double *X, const double *Tc_ex, const double *phc_ex, double scale_T, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(Tc_ex); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("Viscosity", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        X[g.c(i, j, k)] = phc_ex[e] != 1.0 ? muf(Tc_ex[e] * scale_T) : NAN;
    });
    HT3D_API_END
}

int ht3d_Peclet(double *X, const double *Tc_ex, const double *Pc_ex, const double *Vx, const double *Vy,
                const double *Vz, const double *phc_ex, double phi0, double scale_T, double scale_sigma,
                double scale_V, double scale_L, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(Tc_ex); CHECK_PTR(Pc_ex); CHECK_PTR(Vx); CHECK_PTR(Vy); CHECK_PTR(Vz); CHECK_PTR(phc_ex);
    Grid g{nx, ny, nz};
    return launch3d("Peclet", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        const double ph = phc_ex[e];
        if (ph != 1.0) {
            const double x = 0.5 * (Vx[g.fx(i, j, k)] + Vx[g.fx(i + 1, j, k)]) * scale_V;
            const double y = 0.5 * (Vy[g.fy(i, j, k)] + Vy[g.fy(i, j + 1, k)]) * scale_V;
            const double z = 0.5 * (Vz[g.fz(i, j, k)] + Vz[g.fz(i, j, k + 1)]) * scale_V;
            const double V = sqrt(x * x + y * y + z * z);
            const double T = Tc_ex[e] * scale_T;
            const double P = Pc_ex[e] * scale_sigma;
            const double rho = rhof_C(T - 273.15, P);
            const double C = Cf(T);
            const double keff = ((1.0 - phi_of(ph, phi0)) * kTs(T) + phi_of(ph, phi0) * kTf(T));
            X[g.c(i, j, k)] = scale_L * rho * C * V / keff;
        } else {
            X[g.c(i, j, k)] = NAN;
        }
    });
    HT3D_API_END
}

}  // extern "C"
