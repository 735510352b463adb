// ht3d_kernels.hip -- one HIP kernel per reference kernel (the drop-in layer).
//
// Each extern "C" entry replaces one `@parallel Kernel!(args...)` launch of
// tduretz/HydroThermal3D with the same argument order; the cell extents
// (ncx,ncy,ncz) are appended.  Arithmetic keeps the reference's evaluation order
// and the library is built with -ffp-contract=off, so every element equals the
// FP64 CPU result bit for bit.  These kernels are one thread per output point,
// x fastest (coalesced); the tuned path is the fused solver in ht3d_dyrel.hip.
#include <cmath>

#include "ht3d_internal.hpp"
#include "ht3d_launch.hpp"

namespace ht3d {

int to_float(float *out, const double *in, size_t n);   // ht3d_dc.hip

// ---------------------------------------------------------------- WENO5-JS
// tools/Weno5_Routines.jl:44-68 (same expression order; eps uses 1e-99).
__device__ __forceinline__ double weno5_js(double v1, double v2, double v3, double v4, double v5) {
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

}  // namespace ht3d

using namespace ht3d;

extern "C" {

// ------------------------------------------------------------------ pointwise
int ht3d_ResetA(double *A, double *B, int n1, int n2, int n3) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    CHECK_PTR(A);
    CHECK_PTR(B);
    return launch3d("ResetA", n1, n2, n3, [=] __device__(int i, int j, int k) {
        size_t n = ((size_t)k * n2 + j) * n1 + i;
        A[n] = 0.0;
        B[n] = 0.0;
    });
    HT3D_API_END
}

int ht3d_Cpy_inn_to_all(double *A, const double *B, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(A);
    CHECK_PTR(B);
    Grid g{nx, ny, nz};
    return launch3d("Cpy_inn_to_all", nx, ny, nz,
                    [=] __device__(int i, int j, int k) { A[g.c(i, j, k)] = B[g.e(i + 1, j + 1, k + 1)]; });
    HT3D_API_END
}

int ht3d_Cpy_all_to_inn(double *A, const double *B, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(A);
    CHECK_PTR(B);
    Grid g{nx, ny, nz};
    return launch3d("Cpy_all_to_inn", nx, ny, nz,
                    [=] __device__(int i, int j, int k) { A[g.e(i + 1, j + 1, k + 1)] = B[g.c(i, j, k)]; });
    HT3D_API_END
}

int ht3d_ArrayEqualArray(double *F1, const double *F2, int n1, int n2, int n3) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    CHECK_PTR(F1);
    CHECK_PTR(F2);
    return launch3d("ArrayEqualArray", n1, n2, n3, [=] __device__(int i, int j, int k) {
        size_t n = ((size_t)k * n2 + j) * n1 + i;
        F1[n] = F2[n];
    });
    HT3D_API_END
}

int ht3d_Multiply(double *a, const double *b, const double *c, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(a);
    CHECK_PTR(b);
    CHECK_PTR(c);
    Grid g{nx, ny, nz};
    return launch3d("Multiply", nx, ny, nz, [=] __device__(int i, int j, int k) {
        size_t n = g.c(i, j, k);
        a[n] = b[n] * c[n];
    });
    HT3D_API_END
}

// ------------------------------------------------------------------ coefficients
static int face_averages(const char *name, double *kx, double *ky, double *kz, const double *kv, int nx, int ny,
                         int nz) {
    Grid g{nx, ny, nz};
    int rc = launch3d(name, nx + 1, ny, nz, [=] __device__(int i, int j, int k) {  // @av_yza
        kx[g.fx(i, j, k)] = (kv[g.v(i, j, k)] + kv[g.v(i, j + 1, k)] + kv[g.v(i, j, k + 1)] + kv[g.v(i, j + 1, k + 1)]) * 0.25;
    });
    if (rc) return rc;
    rc = launch3d(name, nx, ny + 1, nz, [=] __device__(int i, int j, int k) {  // @av_xza
        ky[g.fy(i, j, k)] = (kv[g.v(i, j, k)] + kv[g.v(i + 1, j, k)] + kv[g.v(i, j, k + 1)] + kv[g.v(i + 1, j, k + 1)]) * 0.25;
    });
    if (rc) return rc;
    return launch3d(name, nx, ny, nz + 1, [=] __device__(int i, int j, int k) {  // @av_xya
        kz[g.fz(i, j, k)] = (kv[g.v(i, j, k)] + kv[g.v(i + 1, j, k)] + kv[g.v(i, j + 1, k)] + kv[g.v(i + 1, j + 1, k)]) * 0.25;
    });
}

int ht3d_InitThermal(double *Rt, double *kx, double *ky, double *kz, const double *Tc_ex, const double *ktv,
                     double _dt, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Rt); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz); CHECK_PTR(Tc_ex); CHECK_PTR(ktv);
    Grid g{nx, ny, nz};
    int rc = launch3d("InitThermal", nx, ny, nz,
                      [=] __device__(int i, int j, int k) { Rt[g.c(i, j, k)] = (-_dt) * Tc_ex[g.e(i + 1, j + 1, k + 1)]; });
    if (rc) return rc;
    return face_averages("InitThermal", kx, ky, kz, ktv, nx, ny, nz);
    HT3D_API_END
}

int ht3d_InitConductivity(double *kx, double *ky, double *kz, const double *ktv, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz); CHECK_PTR(ktv);
    return face_averages("InitConductivity", kx, ky, kz, ktv, nx, ny, nz);
    HT3D_API_END
}

int ht3d_InitDarcy(double *Ty, double *kx, double *ky, double *kz, const double *Tc_ex, const double *kfv,
                   double _dt, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    (void)_dt;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Ty); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz); CHECK_PTR(Tc_ex); CHECK_PTR(kfv);
    Grid g{nx, ny, nz};
    int rc = launch3d("InitDarcy", nx, ny + 1, nz, [=] __device__(int i, int j, int k) {  // @av_yi
        Ty[g.fy(i, j, k)] = (Tc_ex[g.e(i + 1, j, k + 1)] + Tc_ex[g.e(i + 1, j + 1, k + 1)]) * 0.5;
    });
    if (rc) return rc;
    return face_averages("InitDarcy", kx, ky, kz, kfv, nx, ny, nz);
    HT3D_API_END
}

int ht3d_KillNorthSouthFluxes(double *Ty, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Ty);
    Grid g{nx, ny, nz};
    return launch3d("KillNorthSouthFluxes", nx, 2, nz, [=] __device__(int i, int j, int k) {
        Ty[g.fy(i, j == 0 ? 0 : ny, k)] = 0.0;
    });
    HT3D_API_END
}

int ht3d_Compute_Rp(double *Rp, const double *ky, const double *Ty, double Ra, double _dy, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Rp); CHECK_PTR(ky); CHECK_PTR(Ty);
    Grid g{nx, ny, nz};
    return launch3d("Compute_Rp", nx, ny, nz, [=] __device__(int i, int j, int k) {
        Rp[g.c(i, j, k)] = Ra * _dy * ((ky[g.fy(i, j + 1, k)] * Ty[g.fy(i, j + 1, k)]) - (ky[g.fy(i, j, k)] * Ty[g.fy(i, j, k)]));
    });
    HT3D_API_END
}

// ------------------------------------------------------------------ ghost fills
// Face-parallel ghost fill of an E array in three launches: x faces over the
// interior (j,k), y faces over all i and interior k, z faces over all (i,j).
// Every read is of the interior or of a ghost written by an earlier launch, so
// edges and corners (which race in the reference and are never read by the
// hot path) are deterministic and independent of earlier ghost values.
}  // extern "C"
template <typename T>
static int bc_fill6(const char *name, T *A, int nx, int ny, int nz, const int type[6], const double val[6]) {
    Grid g{nx, ny, nz};
    const int s1 = nx + 2, s2 = ny + 2, s3 = nz + 2;
    const int t0 = type[0], t1 = type[1], t2 = type[2], t3 = type[3], t4 = type[4], t5 = type[5];
    const T v0 = (T)val[0], v1 = (T)val[1], v2 = (T)val[2], v3 = (T)val[3], v4 = (T)val[4], v5 = (T)val[5];
    int rc = launch3d(name, 2, ny, nz, [=] __device__(int f, int j0, int k0) {
        const int j = j0 + 1, k = k0 + 1;
        if (f == 0) A[g.e(0, j, k)] = t0 ? 2 * v0 - A[g.e(1, j, k)] : A[g.e(1, j, k)];
        else A[g.e(s1 - 1, j, k)] = t1 ? 2 * v1 - A[g.e(s1 - 2, j, k)] : A[g.e(s1 - 2, j, k)];
    });
    if (rc) return rc;
    rc = launch3d(name, s1, 2, nz, [=] __device__(int i, int f, int k0) {
        const int k = k0 + 1;
        if (f == 0) A[g.e(i, 0, k)] = t2 ? 2 * v2 - A[g.e(i, 1, k)] : A[g.e(i, 1, k)];
        else A[g.e(i, s2 - 1, k)] = t3 ? 2 * v3 - A[g.e(i, s2 - 2, k)] : A[g.e(i, s2 - 2, k)];
    });
    if (rc) return rc;
    return launch3d(name, s1, s2, 2, [=] __device__(int i, int j, int f) {
        if (f == 0) A[g.e(i, j, 0)] = t4 ? 2 * v4 - A[g.e(i, j, 1)] : A[g.e(i, j, 1)];
        else A[g.e(i, j, s3 - 1)] = t5 ? 2 * v5 - A[g.e(i, j, s3 - 2)] : A[g.e(i, j, s3 - 2)];
    });
}

namespace ht3d {
int bc_fill_table(double *A, int nx, int ny, int nz, const int type[6], const double val[6]) {
    return bc_fill6("bc_fill", A, nx, ny, nz, type, val);
}
}  // namespace ht3d

extern "C" {

int ht3d_SetTemperatureBCs(double *Tc_ex, double Tbot, double Ttop, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Tc_ex);
    const int t[6] = {0, 0, 1, 1, 0, 0};
    const double v[6] = {0.0, 0.0, Tbot, Ttop, 0.0, 0.0};
    return bc_fill6("SetTemperatureBCs", Tc_ex, nx, ny, nz, t, v);
    HT3D_API_END
}

int ht3d_SetPressureBCs(double *Pc_ex, double Pbot, double Ptop, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    (void)Pbot; (void)Ptop;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Pc_ex);
    const int t[6] = {0, 0, 0, 0, 0, 0};
    const double v[6] = {0, 0, 0, 0, 0, 0};
    return bc_fill6("SetPressureBCs", Pc_ex, nx, ny, nz, t, v);
    HT3D_API_END
}

int ht3d_SetPressureBCs_v7(double *Pc_ex, double Pbot, double Ptop, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Pc_ex);
    const int t[6] = {0, 0, 1, 1, 0, 0};
    const double v[6] = {0.0, 0.0, Pbot, Ptop, 0.0, 0.0};
    return bc_fill6("SetPressureBCs_v7", Pc_ex, nx, ny, nz, t, v);
    HT3D_API_END
}

// ------------------------------------------------------------------ fluxes / residuals / updates
int ht3d_ComputeFlux(double *qx, double *qy, double *qz, const double *kx, const double *ky, const double *kz,
                     const double *A, double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz); CHECK_PTR(A);
    Grid g{nx, ny, nz};
    int rc = launch3d("ComputeFlux", nx + 1, ny, nz, [=] __device__(int i, int j, int k) {
        qx[g.fx(i, j, k)] = -kx[g.fx(i, j, k)] * (_dx * (A[g.e(i + 1, j + 1, k + 1)] - A[g.e(i, j + 1, k + 1)]));
    });
    if (rc) return rc;
    rc = launch3d("ComputeFlux", nx, ny + 1, nz, [=] __device__(int i, int j, int k) {
        qy[g.fy(i, j, k)] = -ky[g.fy(i, j, k)] * (_dy * (A[g.e(i + 1, j + 1, k + 1)] - A[g.e(i + 1, j, k + 1)]));
    });
    if (rc) return rc;
    return launch3d("ComputeFlux", nx, ny, nz + 1, [=] __device__(int i, int j, int k) {
        qz[g.fz(i, j, k)] = -kz[g.fz(i, j, k)] * (_dz * (A[g.e(i + 1, j + 1, k + 1)] - A[g.e(i + 1, j + 1, k)]));
    });
    HT3D_API_END
}

#define DQX (qx[g.fx(i + 1, j, k)] - qx[g.fx(i, j, k)])
#define DQY (qy[g.fy(i, j + 1, k)] - qy[g.fy(i, j, k)])
#define DQZ (qz[g.fz(i, j, k + 1)] - qz[g.fz(i, j, k)])

int ht3d_UpdateT(double *F, double *T, const double *R, const double *qx, const double *qy, const double *qz,
                 double _dt, double dtau, double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(T); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz);
    Grid g{nx, ny, nz};
    return launch3d("UpdateT", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k), e = g.e(i + 1, j + 1, k + 1);
        double f = R[c] + _dt * T[e] + _dx * DQX + _dy * DQY + _dz * DQZ;
        F[c] = f;
        T[e] = T[e] - dtau * f;
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusion(double *F, const double *R, const double *qx, const double *qy, const double *qz,
                           double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusion", nx, ny, nz, [=] __device__(int i, int j, int k) {
        F[g.c(i, j, k)] = R[g.c(i, j, k)] + _dx * DQX + _dy * DQY + _dz * DQZ;
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusionTransient(double *F, const double *T, const double *R, const double *qx,
                                    const double *qy, const double *qz, double _dt, double _dx, double _dy,
                                    double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(T); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusionTransient", nx, ny, nz, [=] __device__(int i, int j, int k) {
        F[g.c(i, j, k)] = R[g.c(i, j, k)] + _dt * T[g.e(i + 1, j + 1, k + 1)] + _dx * DQX + _dy * DQY + _dz * DQZ;
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusionPC(double *F, const double *R, const double *qx, const double *qy, const double *qz,
                             const double *PC, double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(PC);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusionPC", nx, ny, nz, [=] __device__(int i, int j, int k) {
        double f = R[g.c(i, j, k)] + _dx * DQX + _dy * DQY + _dz * DQZ;
        F[g.c(i, j, k)] = f / PC[g.c(i, j, k)];
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusionTransientPC(double *F, const double *T, const double *R, const double *qx,
                                      const double *qy, const double *qz, const double *PC, double _dt,
                                      double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(T); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(PC);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusionTransientPC", nx, ny, nz, [=] __device__(int i, int j, int k) {
        double f = R[g.c(i, j, k)] + _dt * T[g.e(i + 1, j + 1, k + 1)] + _dx * DQX + _dy * DQY + _dz * DQZ;
        F[g.c(i, j, k)] = f / PC[g.c(i, j, k)];
    });
    HT3D_API_END
}

int ht3d_UpdateP(double *F0, double *P, const double *F, double dampx, double dtau, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F0); CHECK_PTR(P); CHECK_PTR(F);
    Grid g{nx, ny, nz};
    return launch3d("UpdateP", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k), e = g.e(i + 1, j + 1, k + 1);
        double f0 = F[c] + dampx * F0[c];
        F0[c] = f0;
        P[e] = P[e] - dtau * f0;
    });
    HT3D_API_END
}

int ht3d_DampedUpdate(double *F0, double *X, const double *F, double dampx, double dtau, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_UpdateP(F0, X, F, dampx, dtau, nx, ny, nz);
    HT3D_API_END
}

int ht3d_DYRELUpdate(double *F0, double *X, const double *F, double h1, double h2, double dtau, int nx, int ny,
                     int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F0); CHECK_PTR(X); CHECK_PTR(F);
    Grid g{nx, ny, nz};
    return launch3d("DYRELUpdate", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k), e = g.e(i + 1, j + 1, k + 1);
        double f0 = h1 * F0[c] + h2 * F[c];
        F0[c] = f0;
        X[e] = X[e] - dtau * f0;
    });
    HT3D_API_END
}

// SwapDYREL!  kernels_HT3D_dimensional.jl:122-126: Xc = inn(X0c_ex), Yc = Y0c
int ht3d_SwapDYREL(double *Xc, double *Yc, const double *X0c_ex, const double *Y0c, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Xc); CHECK_PTR(Yc); CHECK_PTR(X0c_ex); CHECK_PTR(Y0c);
    Grid g{nx, ny, nz};
    return launch3d("SwapDYREL", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        Xc[c] = X0c_ex[g.e(i + 1, j + 1, k + 1)];
        Yc[c] = Y0c[c];
    });
    HT3D_API_END
}

// the Gershgorin row sum of notebooks/Gershgorin_TH3D.ipynb (cell 2 output)
__device__ __forceinline__ double gersh_diag(double kW, double kE, double kS, double kN, double kB, double kF,
                                             double rc_tr_dt, double dx, double dy, double dz) {
    return (kB / dz + kF / dz) / dz + (kN / dy + kS / dy) / dy + (kE / dx + kW / dx) / dx + rc_tr_dt;
}
__device__ __forceinline__ double gersh_offsum(double kW, double kE, double kS, double kN, double kB, double kF,
                                               double dx, double dy, double dz) {
    return fabs(kE / (dx * dx)) + fabs(kW / (dx * dx)) + fabs(kN / (dy * dy)) + fabs(kS / (dy * dy)) +
           fabs(kB / (dz * dz)) + fabs(kF / (dz * dz));
}

int ht3d_GershgorinPoisson(double *G, double *D, const double *rhoC, const double *kx, const double *ky,
                           const double *kz, double transient, double dt, double dx, double dy, double dz, int nx,
                           int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(G); CHECK_PTR(D); CHECK_PTR(rhoC); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz);
    Grid g{nx, ny, nz};
    return launch3d("GershgorinPoisson", nx, ny, nz, [=] __device__(int i, int j, int k) {
        double kW = kx[g.fx(i, j, k)], kE = kx[g.fx(i + 1, j, k)];
        double kS = ky[g.fy(i, j, k)], kN = ky[g.fy(i, j + 1, k)];
        double kF = kz[g.fz(i, j, k + 1)], kB = kz[g.fz(i, j, k)];
        double diag = gersh_diag(kW, kE, kS, kN, kB, kF, rhoC[g.c(i, j, k)] * transient / dt, dx, dy, dz);
        double gg = gersh_offsum(kW, kE, kS, kN, kB, kF, dx, dy, dz) + fabs(diag);
        D[g.c(i, j, k)] = diag;
        G[g.c(i, j, k)] = gg / diag;
    });
    HT3D_API_END
}

int ht3d_GershgorinHydro(double *G, const double *phidrhodP, const double *kx, const double *ky, const double *kz,
                         double transient, double dt, double dx, double dy, double dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(G); CHECK_PTR(phidrhodP); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz);
    Grid g{nx, ny, nz};
    return launch3d("GershgorinHydro", nx, ny, nz, [=] __device__(int i, int j, int k) {
        double kW = kx[g.fx(i, j, k)], kE = kx[g.fx(i + 1, j, k)];
        double kS = ky[g.fy(i, j, k)], kN = ky[g.fy(i, j + 1, k)];
        double kF = kz[g.fz(i, j, k + 1)], kB = kz[g.fz(i, j, k)];
        double diag = gersh_diag(kW, kE, kS, kN, kB, kF, phidrhodP[g.c(i, j, k)] * transient / dt, dx, dy, dz);
        G[g.c(i, j, k)] = gersh_offsum(kW, kE, kS, kN, kB, kF, dx, dy, dz) + fabs(diag);
    });
    HT3D_API_END
}

int ht3d_RayleighQuotientNumerator(double *G, const double *Fc0, const double *Fcit, const double *Fc, double dtau,
                                   int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(G); CHECK_PTR(Fc0); CHECK_PTR(Fcit); CHECK_PTR(Fc);
    Grid g{nx, ny, nz};
    return launch3d("RayleighQuotientNumerator", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        double d = dtau * Fc0[c];
        G[c] = (-d) * (Fc[c] - Fcit[c]);
    });
    HT3D_API_END
}

int ht3d_RayleighQuotientDenominator(double *G, const double *Fc0, double dtau, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(G); CHECK_PTR(Fc0);
    Grid g{nx, ny, nz};
    return launch3d("RayleighQuotientDenominator", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        double d = dtau * Fc0[c];
        G[c] = d * d;
    });
    HT3D_API_END
}

int ht3d_ResidualTemperatureLinearised(double *F, const double *Tc_ex, const double *Tc_old, const double *rhoC_eff,
                                       const double *phc_ex, const double *PC, const double *qx, const double *qy,
                                       const double *qz, double Qt, double transient, double _dt, double _dx,
                                       double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(Tc_ex); CHECK_PTR(Tc_old); CHECK_PTR(rhoC_eff); CHECK_PTR(phc_ex); CHECK_PTR(PC);
    CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz);
    Grid g{nx, ny, nz};
    return launch3d("ResidualTemperatureLinearised", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k), e = g.e(i + 1, j + 1, k + 1);
        double ph = phc_ex[e];
        if (fabs(ph) != 1.0) {
            double Q = 0.0;
            if (ph == 3.0) Q = Qt;
            double f = transient * rhoC_eff[c] * _dt * (Tc_ex[e] - Tc_old[c]);
            f += DQX * _dx;
            f += DQY * _dy;
            f += DQZ * _dz;
            f -= transient * Q;
            f /= PC[c];
            F[c] = f;
        } else {
            F[c] = 0.0;
        }
    });
    HT3D_API_END
}

int ht3d_ResidualFluidPressure(double *F, const double *phc_ex, const double *phirhof, const double *phirhof0,
                               const double *PC, const double *qx, const double *qy, const double *qz,
                               double transient, double _dt, double _dx, double _dy, double _dz, int nx, int ny,
                               int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(phc_ex); CHECK_PTR(phirhof); CHECK_PTR(phirhof0); CHECK_PTR(PC);
    CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz);
    Grid g{nx, ny, nz};
    return launch3d("ResidualFluidPressure", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        double ph = phc_ex[g.e(i + 1, j + 1, k + 1)];
        if (fabs(ph) != 1.0) {
            double f = transient * _dt * (phirhof[c] - phirhof0[c]);
            f += DQX * _dx;
            f += DQY * _dy;
            f += DQZ * _dz;
            f /= PC[c];
            F[c] = f;
        } else {
            F[c] = 0.0;
        }
    });
    HT3D_API_END
}

int ht3d_ComputeDarcyFlux(double *qx, double *qy, double *qz, const double *rhofv, const double *kx,
                          const double *ky, const double *kz, const double *Pf, double gr, double _dx, double _dy,
                          double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(rhofv); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz);
    CHECK_PTR(Pf);
    Grid g{nx, ny, nz};
    int rc = launch3d("ComputeDarcyFlux", nx + 1, ny, nz, [=] __device__(int i, int j, int k) {
        qx[g.fx(i, j, k)] = -kx[g.fx(i, j, k)] * (_dx * (Pf[g.e(i + 1, j + 1, k + 1)] - Pf[g.e(i, j + 1, k + 1)]));
    });
    if (rc) return rc;
    rc = launch3d("ComputeDarcyFlux", nx, ny + 1, nz, [=] __device__(int i, int j, int k) {
        double av = (rhofv[g.v(i, j, k)] + rhofv[g.v(i + 1, j, k)] + rhofv[g.v(i, j, k + 1)] + rhofv[g.v(i + 1, j, k + 1)]) * 0.25;
        qy[g.fy(i, j, k)] = -ky[g.fy(i, j, k)] * (_dy * (Pf[g.e(i + 1, j + 1, k + 1)] - Pf[g.e(i + 1, j, k + 1)]) - gr * av);
    });
    if (rc) return rc;
    return launch3d("ComputeDarcyFlux", nx, ny, nz + 1, [=] __device__(int i, int j, int k) {
        qz[g.fz(i, j, k)] = -kz[g.fz(i, j, k)] * (_dz * (Pf[g.e(i + 1, j + 1, k + 1)] - Pf[g.e(i + 1, j + 1, k)]));
    });
    HT3D_API_END
}

int ht3d_Init_vel(double *Vx, double *Vy, double *Vz, const double *kx, const double *ky, const double *kz,
                  const double *Pc_ex, const double *Ty, double Ra, double _dx, double _dy, double _dz, int nx,
                  int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Vx); CHECK_PTR(Vy); CHECK_PTR(Vz); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz); CHECK_PTR(Pc_ex);
    CHECK_PTR(Ty);
    Grid g{nx, ny, nz};
    int rc = launch3d("Init_vel", nx + 1, ny, nz, [=] __device__(int i, int j, int k) {
        Vx[g.fx(i, j, k)] = -kx[g.fx(i, j, k)] * _dx * (Pc_ex[g.e(i + 1, j + 1, k + 1)] - Pc_ex[g.e(i, j + 1, k + 1)]);
    });
    if (rc) return rc;
    rc = launch3d("Init_vel", nx, ny + 1, nz, [=] __device__(int i, int j, int k) {
        Vy[g.fy(i, j, k)] = -ky[g.fy(i, j, k)] * (_dy * (Pc_ex[g.e(i + 1, j + 1, k + 1)] - Pc_ex[g.e(i + 1, j, k + 1)]) - Ra * Ty[g.fy(i, j, k)]);
    });
    if (rc) return rc;
    return launch3d("Init_vel", nx, ny, nz + 1, [=] __device__(int i, int j, int k) {
        Vz[g.fz(i, j, k)] = -kz[g.fz(i, j, k)] * _dz * (Pc_ex[g.e(i + 1, j + 1, k + 1)] - Pc_ex[g.e(i + 1, j + 1, k)]);
    });
    HT3D_API_END
}

// Temperature-dependent viscosity mu = exp(-a T): face coefficients become the
// mobility k/mu(T_face) = k exp(a T_face) with T_face = (T_left + T_right) * 0.5
// (ghost values at the walls, i.e. the boundary condition).  The dimensionless
// kernels of the reference have constant viscosity; this is the dimensionless
// analogue of the T-dependent hydraulic conductivity of
// Dimensional/kernels_HT3D_dimensional.jl (ComputeHydroConductivity), applied
// once per Darcy solve after InitDarcy!.
int ht3d_ApplyViscosity(double *kx, double *ky, double *kz, const double *Tc_ex, double a, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz); CHECK_PTR(Tc_ex);
    Grid g{nx, ny, nz};
    int rc = launch3d("ApplyViscosity", nx + 1, ny, nz, [=] __device__(int i, int j, int k) {
        kx[g.fx(i, j, k)] = kx[g.fx(i, j, k)] * exp(a * ((Tc_ex[g.e(i, j + 1, k + 1)] + Tc_ex[g.e(i + 1, j + 1, k + 1)]) * 0.5));
    });
    if (rc) return rc;
    rc = launch3d("ApplyViscosity", nx, ny + 1, nz, [=] __device__(int i, int j, int k) {
        ky[g.fy(i, j, k)] = ky[g.fy(i, j, k)] * exp(a * ((Tc_ex[g.e(i + 1, j, k + 1)] + Tc_ex[g.e(i + 1, j + 1, k + 1)]) * 0.5));
    });
    if (rc) return rc;
    return launch3d("ApplyViscosity", nx, ny, nz + 1, [=] __device__(int i, int j, int k) {
        kz[g.fz(i, j, k)] = kz[g.fz(i, j, k)] * exp(a * ((Tc_ex[g.e(i + 1, j + 1, k)] + Tc_ex[g.e(i + 1, j + 1, k + 1)]) * 0.5));
    });
    HT3D_API_END
}

// ------------------------------------------------------------------ WENO5 (v1 kernels)
static int plus_minus(const char *name, double *Vm, double *Vp, const double *V, int axis, int nx, int ny, int nz) {
    Grid g{nx, ny, nz};
    return launch3d(name, nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        double vl, vr;
        if (axis == 1) { vl = V[g.fx(i, j, k)]; vr = V[g.fx(i + 1, j, k)]; }
        else if (axis == 2) { vl = V[g.fy(i, j, k)]; vr = V[g.fy(i, j + 1, k)]; }
        else { vl = V[g.fz(i, j, k)]; vr = V[g.fz(i, j, k + 1)]; }
        if (vl < 0.0) Vm[c] = vl;
        if (vl > 0.0) Vm[c] = 0.0;
        if (vr < 0.0) Vp[c] = 0.0;
        if (vr > 0.0) Vp[c] = vr;
    });
}

int ht3d_VxPlusMinus(double *Vxm, double *Vxp, const double *Vx, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Vxm); CHECK_PTR(Vxp); CHECK_PTR(Vx);
    return plus_minus("VxPlusMinus", Vxm, Vxp, Vx, 1, nx, ny, nz);
    HT3D_API_END
}
int ht3d_VyPlusMinus(double *Vym, double *Vyp, const double *Vy, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Vym); CHECK_PTR(Vyp); CHECK_PTR(Vy);
    return plus_minus("VyPlusMinus", Vym, Vyp, Vy, 2, nx, ny, nz);
    HT3D_API_END
}
int ht3d_VzPlusMinus(double *Vzm, double *Vzp, const double *Vz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Vzm); CHECK_PTR(Vzp); CHECK_PTR(Vz);
    return plus_minus("VzPlusMinus", Vzm, Vzp, Vz, 3, nx, ny, nz);
    HT3D_API_END
}

int ht3d_dFdx_Weno5(double *dFdxi, const double *V1, const double *V2, const double *V3, const double *V4,
                    const double *V5, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(dFdxi); CHECK_PTR(V1); CHECK_PTR(V2); CHECK_PTR(V3); CHECK_PTR(V4); CHECK_PTR(V5);
    Grid g{nx, ny, nz};
    return launch3d("dFdx_Weno5", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        dFdxi[c] = weno5_js(V1[c], V2[c], V3[c], V4[c], V5[c]);
    });
    HT3D_API_END
}

int ht3d_Gradients_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *X,
                         int axis, int plus, double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(v1); CHECK_PTR(v2); CHECK_PTR(v3); CHECK_PTR(v4); CHECK_PTR(v5); CHECK_PTR(X);
    if (axis < 1 || axis > 3) return arg_fail("Gradients_Weno5: axis must be 1..3"), HT3D_ERR_ARG;
    Grid g{nx, ny, nz};
    const int ox = axis == 1, oy = axis == 2, oz = axis == 3;
    const double h = axis == 1 ? _dx : (axis == 2 ? _dy : _dz);
    return launch3d("Gradients_Weno5", nx, ny, nz, [=] __device__(int i, int j, int k) {
        auto xs = [&](int o) { return X[g.x6(i + 3 + o * ox, j + 3 + o * oy, k + 3 + o * oz)]; };
        const size_t c = g.c(i, j, k);
        if (!plus) {
            v1[c] = h * (xs(-2) - xs(-3));
            v2[c] = h * (xs(-1) - xs(-2));
            v3[c] = h * (xs(0) - xs(-1));
            v4[c] = h * (xs(1) - xs(0));
            v5[c] = h * (xs(2) - xs(1));
        } else {
            v1[c] = h * (xs(3) - xs(2));
            v2[c] = h * (xs(2) - xs(1));
            v3[c] = h * (xs(1) - xs(0));
            v4[c] = h * (xs(0) - xs(-1));
            v5[c] = h * (xs(-1) - xs(-2));
        }
    });
    HT3D_API_END
}

int ht3d_Boundaries_Weno5(double *Fc_exxx, const double *Fc, int axis, int src_ext, int type_lo, double val_lo,
                          int type_hi, double val_hi, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Fc_exxx); CHECK_PTR(Fc);
    if (axis < 1 || axis > 3) return arg_fail("Boundaries_Weno5: axis must be 1..3"), HT3D_ERR_ARG;
    Grid g{nx, ny, nz};
    const int o = src_ext ? 1 : 0;
    const int s1 = src_ext ? nx + 2 : nx, s2 = src_ext ? ny + 2 : ny;
    auto src = [=] __device__(int a, int b, int c) { return Fc[((size_t)(c + o) * s2 + (b + o)) * s1 + (a + o)]; };
    int rc = launch3d("Boundaries_Weno5", nx, ny, nz,
                      [=] __device__(int i, int j, int k) { Fc_exxx[g.x6(i + 3, j + 3, k + 3)] = src(i, j, k); });
    if (rc) return rc;
    const int n = axis == 1 ? nx : (axis == 2 ? ny : nz);
    // ghost layer g (1..3): low side X6 index 3-g mirrors interior g-1; high side
    // X6 index n+2+g mirrors n-g; periodic: low takes n-g, high takes g-1 (0-based)
    auto ghost = [=] __device__(int gl, int hi, int a, int b) {
        // (a,b) index the two transverse directions
        int mir = hi ? n - gl : gl - 1, per = hi ? gl - 1 : n - gl;
        int t = hi ? type_hi : type_lo;
        double val = hi ? val_hi : val_lo;
        int d = hi ? n + 2 + gl : 3 - gl;
        double sm, sp;
        if (axis == 1) { sm = src(mir, a, b); sp = src(per, a, b); }
        else if (axis == 2) { sm = src(a, mir, b); sp = src(a, per, b); }
        else { sm = src(a, b, mir); sp = src(a, b, per); }
        double v = t == 0 ? sm : (t == 1 ? 2 * val - sm : sp);
        if (axis == 1) Fc_exxx[g.x6(d, a + 3, b + 3)] = v;
        else if (axis == 2) Fc_exxx[g.x6(a + 3, d, b + 3)] = v;
        else Fc_exxx[g.x6(a + 3, b + 3, d)] = v;
    };
    const int na = axis == 1 ? ny : nx, nb = axis == 3 ? ny : nz;
    return launch3d("Boundaries_Weno5", na, nb, 6, [=] __device__(int a, int b, int s) { ghost(s % 3 + 1, s / 3, a, b); });
    HT3D_API_END
}

// Per-name entries of tools/Weno5_Routines.jl:70-255 with the reference's
// argument orders: the six one-sided difference kernels and the three ghost
// fills of a cell array (Fc is c); the _v2 fills take the extended array (Fc
// is E, tools/Weno5_Routines_v2.jl:221-347).  Each forwards to the
// parameterised kernel above.
int ht3d_Gradients_minus_x_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                              double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Gradients_Weno5(v1, v2, v3, v4, v5, Fc_exxx, 1, 0, _dx, _dy, _dz, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Gradients_plus_x_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                              double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Gradients_Weno5(v1, v2, v3, v4, v5, Fc_exxx, 1, 1, _dx, _dy, _dz, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Gradients_minus_y_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                              double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Gradients_Weno5(v1, v2, v3, v4, v5, Fc_exxx, 2, 0, _dx, _dy, _dz, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Gradients_plus_y_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                              double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Gradients_Weno5(v1, v2, v3, v4, v5, Fc_exxx, 2, 1, _dx, _dy, _dz, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Gradients_minus_z_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                              double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Gradients_Weno5(v1, v2, v3, v4, v5, Fc_exxx, 3, 0, _dx, _dy, _dz, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Gradients_plus_z_Weno5(double *v1, double *v2, double *v3, double *v4, double *v5, const double *Fc_exxx,
                              double _dx, double _dy, double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Gradients_Weno5(v1, v2, v3, v4, v5, Fc_exxx, 3, 1, _dx, _dy, _dz, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Boundaries_x_Weno5(double *Fc_exxx, const double *Fc, int type_W, double val_W, int type_E,
                              double val_E, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Boundaries_Weno5(Fc_exxx, Fc, 1, 0, type_W, val_W, type_E, val_E, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Boundaries_x_Weno5_v2(double *Fc_exxx, const double *Fc, int type_W, double val_W, int type_E,
                              double val_E, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Boundaries_Weno5(Fc_exxx, Fc, 1, 1, type_W, val_W, type_E, val_E, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Boundaries_y_Weno5(double *Fc_exxx, const double *Fc, int type_S, double val_S, int type_N,
                              double val_N, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Boundaries_Weno5(Fc_exxx, Fc, 2, 0, type_S, val_S, type_N, val_N, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Boundaries_y_Weno5_v2(double *Fc_exxx, const double *Fc, int type_S, double val_S, int type_N,
                              double val_N, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Boundaries_Weno5(Fc_exxx, Fc, 2, 1, type_S, val_S, type_N, val_N, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Boundaries_z_Weno5(double *Fc_exxx, const double *Fc, int type_B, double val_B, int type_F,
                              double val_F, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Boundaries_Weno5(Fc_exxx, Fc, 3, 0, type_B, val_B, type_F, val_F, nx, ny, nz);
    HT3D_API_END
}
int ht3d_Boundaries_z_Weno5_v2(double *Fc_exxx, const double *Fc, int type_B, double val_B, int type_F,
                              double val_F, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    return ht3d_Boundaries_Weno5(Fc_exxx, Fc, 3, 1, type_B, val_B, type_F, val_F, nx, ny, nz);
    HT3D_API_END
}

int ht3d_Advect(double *Fc, const double *Vxp, const double *dTdxm, const double *Vxm, const double *dTdxp,
                double dt, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Fc); CHECK_PTR(Vxp); CHECK_PTR(dTdxm); CHECK_PTR(Vxm); CHECK_PTR(dTdxp);
    Grid g{nx, ny, nz};
    return launch3d("Advect", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        Fc[c] = Fc[c] - dt * (Vxp[c] * dTdxm[c] + Vxm[c] * dTdxp[c]);
    });
    HT3D_API_END
}

int ht3d_TimeAveraging(double *Fc, const double *Fold, double order, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Fc); CHECK_PTR(Fold);
    Grid g{nx, ny, nz};
    return launch3d("TimeAveraging", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k);
        Fc[c] = (1.0 / order) * Fc[c] + (1.0 - 1.0 / order) * Fold[c];
    });
    HT3D_API_END
}

int ht3d_TimeAveraging_v2(double *Fc, const double *Fold, double order, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Fc); CHECK_PTR(Fold);
    Grid g{nx, ny, nz};
    return launch3d("TimeAveraging_v2", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t e = g.e(i + 1, j + 1, k + 1);
        Fc[e] = (1.0 / order) * Fc[e] + (1.0 - 1.0 / order) * Fold[g.c(i, j, k)];
    });
    HT3D_API_END
}

int ht3d_Advect_v2(double *Fc, const double *X, const double *V, int axis, double dt, double _dx, double _dy,
                   double _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Fc); CHECK_PTR(X); CHECK_PTR(V);
    if (axis < 1 || axis > 3) return arg_fail("Advect_v2: axis must be 1..3"), HT3D_ERR_ARG;
    Grid g{nx, ny, nz};
    const int ox = axis == 1, oy = axis == 2, oz = axis == 3;
    const double h = axis == 1 ? _dx : (axis == 2 ? _dy : _dz);
    const int w1 = axis == 1 ? nx + 1 : nx, w2 = axis == 2 ? ny + 1 : ny;
    return launch3d("Advect_v2", nx, ny, nz, [=] __device__(int i, int j, int k) {
        auto xs = [&](int o) { return X[g.x6(i + 3 + o * ox, j + 3 + o * oy, k + 3 + o * oz)]; };
        double vl = V[((size_t)k * w2 + j) * w1 + i];
        double vr = V[((size_t)(k + oz) * w2 + (j + oy)) * w1 + (i + ox)];
        double Vm = vl < 0.0 ? vl : 0.0;
        double Vp = vr >= 0.0 ? vr : 0.0;
        double dm = weno5_js(h * (xs(-2) - xs(-3)), h * (xs(-1) - xs(-2)), h * (xs(0) - xs(-1)), h * (xs(1) - xs(0)),
                             h * (xs(2) - xs(-1)));
        double dp = weno5_js(h * (xs(3) - xs(2)), h * (xs(2) - xs(1)), h * (xs(1) - xs(0)), h * (xs(0) - xs(-1)),
                             h * (xs(-1) - xs(-2)));
        const size_t e = g.e(i + 1, j + 1, k + 1);
        Fc[e] = Fc[e] - dt * (Vp * dm + Vm * dp);
    });
    HT3D_API_END
}


// ---------------------------------------------------------------- FP32 sweeps
// FP32 versions of the kernels a defect-correction driver runs in its inner
// sweeps (the correction equation carries homogeneous boundary values), with
// the FP64 entries' expressions and order evaluated in float, as the fused FP32
// sweep does.  No reference counterpart: the reference is FP64 only
// (HydroThermal3D_v7.jl:13).  ht3d_to_f32 and ht3d_add_f32 move data between the
// FP64 anchor and the FP32 correction.
int ht3d_SetTemperatureBCs_f32(float *Tc_ex, float Tbot, float Ttop, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Tc_ex);
    const int t[6] = {0, 0, 1, 1, 0, 0};
    const double v[6] = {0.0, 0.0, (double)Tbot, (double)Ttop, 0.0, 0.0};
    return bc_fill6("SetTemperatureBCs_f32", Tc_ex, nx, ny, nz, t, v);
    HT3D_API_END
}

int ht3d_SetPressureBCs_f32(float *Pc_ex, float Pbot, float Ptop, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    (void)Pbot; (void)Ptop;
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(Pc_ex);
    const int t[6] = {0, 0, 0, 0, 0, 0};
    const double v[6] = {0, 0, 0, 0, 0, 0};
    return bc_fill6("SetPressureBCs_f32", Pc_ex, nx, ny, nz, t, v);
    HT3D_API_END
}

int ht3d_ComputeFlux_f32(float *qx, float *qy, float *qz, const float *kx, const float *ky, const float *kz,
                         const float *A, float _dx, float _dy, float _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz); CHECK_PTR(A);
    Grid g{nx, ny, nz};
    int rc = launch3d("ComputeFlux_f32", nx + 1, ny, nz, [=] __device__(int i, int j, int k) {
        qx[g.fx(i, j, k)] = -kx[g.fx(i, j, k)] * (_dx * (A[g.e(i + 1, j + 1, k + 1)] - A[g.e(i, j + 1, k + 1)]));
    });
    if (rc) return rc;
    rc = launch3d("ComputeFlux_f32", nx, ny + 1, nz, [=] __device__(int i, int j, int k) {
        qy[g.fy(i, j, k)] = -ky[g.fy(i, j, k)] * (_dy * (A[g.e(i + 1, j + 1, k + 1)] - A[g.e(i + 1, j, k + 1)]));
    });
    if (rc) return rc;
    return launch3d("ComputeFlux_f32", nx, ny, nz + 1, [=] __device__(int i, int j, int k) {
        qz[g.fz(i, j, k)] = -kz[g.fz(i, j, k)] * (_dz * (A[g.e(i + 1, j + 1, k + 1)] - A[g.e(i + 1, j + 1, k)]));
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusion_f32(float *F, const float *R, const float *qx, const float *qy, const float *qz, float _dx,
                               float _dy, float _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusion_f32", nx, ny, nz, [=] __device__(int i, int j, int k) {
        F[g.c(i, j, k)] = R[g.c(i, j, k)] + _dx * DQX + _dy * DQY + _dz * DQZ;
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusionTransient_f32(float *F, const float *T, const float *R, const float *qx, const float *qy,
                                        const float *qz, float _dt, float _dx, float _dy, float _dz, int nx, int ny,
                                        int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(T); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusionTransient_f32", nx, ny, nz, [=] __device__(int i, int j, int k) {
        F[g.c(i, j, k)] = R[g.c(i, j, k)] + _dt * T[g.e(i + 1, j + 1, k + 1)] + _dx * DQX + _dy * DQY + _dz * DQZ;
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusionPC_f32(float *F, const float *R, const float *qx, const float *qy, const float *qz,
                                 const float *PC, float _dx, float _dy, float _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(PC);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusionPC_f32", nx, ny, nz, [=] __device__(int i, int j, int k) {
        float f = R[g.c(i, j, k)] + _dx * DQX + _dy * DQY + _dz * DQZ;
        F[g.c(i, j, k)] = f / PC[g.c(i, j, k)];
    });
    HT3D_API_END
}

int ht3d_ResidualDiffusionTransientPC_f32(float *F, const float *T, const float *R, const float *qx, const float *qy,
                                          const float *qz, const float *PC, float _dt, float _dx, float _dy,
                                          float _dz, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F); CHECK_PTR(T); CHECK_PTR(R); CHECK_PTR(qx); CHECK_PTR(qy); CHECK_PTR(qz); CHECK_PTR(PC);
    Grid g{nx, ny, nz};
    return launch3d("ResidualDiffusionTransientPC_f32", nx, ny, nz, [=] __device__(int i, int j, int k) {
        float f = R[g.c(i, j, k)] + _dt * T[g.e(i + 1, j + 1, k + 1)] + _dx * DQX + _dy * DQY + _dz * DQZ;
        F[g.c(i, j, k)] = f / PC[g.c(i, j, k)];
    });
    HT3D_API_END
}

int ht3d_DYRELUpdate_f32(float *F0, float *X, const float *F, float h1, float h2, float dtau, int nx, int ny,
                         int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(F0); CHECK_PTR(X); CHECK_PTR(F);
    Grid g{nx, ny, nz};
    return launch3d("DYRELUpdate_f32", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.c(i, j, k), e = g.e(i + 1, j + 1, k + 1);
        float f0 = h1 * F0[c] + h2 * F[c];
        F0[c] = f0;
        X[e] = X[e] - dtau * f0;
    });
    HT3D_API_END
}

int ht3d_GershgorinPoisson_f32(float *G, float *D, const float *rhoC, const float *kx, const float *ky,
                               const float *kz, float transient, float dt, float dx, float dy, float dz, int nx,
                               int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(G); CHECK_PTR(D); CHECK_PTR(rhoC); CHECK_PTR(kx); CHECK_PTR(ky); CHECK_PTR(kz);
    Grid g{nx, ny, nz};
    return launch3d("GershgorinPoisson_f32", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const float kW = kx[g.fx(i, j, k)], kE = kx[g.fx(i + 1, j, k)];
        const float kS = ky[g.fy(i, j, k)], kN = ky[g.fy(i, j + 1, k)];
        const float kF = kz[g.fz(i, j, k + 1)], kB = kz[g.fz(i, j, k)];
        const float rc = rhoC[g.c(i, j, k)] * transient / dt;
        const float diag = (kB / dz + kF / dz) / dz + (kN / dy + kS / dy) / dy + (kE / dx + kW / dx) / dx + rc;
        const float off = fabsf(kE / (dx * dx)) + fabsf(kW / (dx * dx)) + fabsf(kN / (dy * dy)) +
                          fabsf(kS / (dy * dy)) + fabsf(kB / (dz * dz)) + fabsf(kF / (dz * dz));
        D[g.c(i, j, k)] = diag;
        G[g.c(i, j, k)] = (off + fabsf(diag)) / diag;
    });
    HT3D_API_END
}

int ht3d_to_f32(float *out, const double *in, size_t n) {
    HT3D_API_BEGIN
    CHECK_PTR(out); CHECK_PTR(in);
    if (!n) return HT3D_OK;
    int rc = to_float(out, in, n);
    if (rc) return rc;
    return after_launch("to_f32");
    HT3D_API_END
}

int ht3d_add_f32(double *X, const float *e, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    CHECK_GRID(nx, ny, nz);
    CHECK_PTR(X); CHECK_PTR(e);
    Grid g{nx, ny, nz};
    return launch3d("add_f32", nx, ny, nz, [=] __device__(int i, int j, int k) {
        const size_t c = g.e(i + 1, j + 1, k + 1);
        X[c] += (double)e[c];
    });
    HT3D_API_END
}
}  // extern "C"
