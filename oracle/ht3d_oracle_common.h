/*
 * ht3d_oracle_common.h -- index macros shared by the CPU oracle's C files.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle; see ht3d_oracle.c).
 * 1-based, column-major (x fastest) access with the reference's staggered
 * extents, and the ParallelStencil.FiniteDifferences3D index macros.
 */
#ifndef HT3D_ORACLE_COMMON_H
#define HT3D_ORACLE_COMMON_H
#include <stddef.h>

#define EXPORT __attribute__((visibility("default")))

/* 1-based, column-major element access for an array of leading extents (n1,n2).
 * The checked build (-DHT3D_ORACLE_BOUNDS, `make sanitize`) aborts on an index
 * outside 1..n1, 1..n2 or below 1 in the last dimension; AddressSanitizer in the
 * same build catches accesses past the end of an allocation. */
#ifdef HT3D_ORACLE_BOUNDS
#include <stdio.h>
#include <stdlib.h>
static inline size_t o_index_checked(size_t n1, size_t n2, long i, long j, long k, const char *file, int line) {
    if (i < 1 || i > (long)n1 || j < 1 || j > (long)n2 || k < 1) {
        fprintf(stderr, "%s:%d: index (%ld, %ld, %ld) outside the extents (%zu, %zu, .)\n", file, line, i, j, k, n1,
                n2);
        abort();
    }
    return ((size_t)(k - 1) * n2 + (size_t)(j - 1)) * n1 + (size_t)(i - 1);
}
#define AT(a, n1, n2, i, j, k) \
    ((a)[o_index_checked((size_t)(n1), (size_t)(n2), (long)(i), (long)(j), (long)(k), __FILE__, __LINE__)])
#else
#define AT(a, n1, n2, i, j, k) \
    ((a)[(((size_t)(k) - 1) * (size_t)(n2) + ((size_t)(j) - 1)) * (size_t)(n1) + ((size_t)(i) - 1)])
#endif

#define Cc(a, i, j, k)  AT(a, nx, ny, i, j, k)
#define Ee(a, i, j, k)  AT(a, nx + 2, ny + 2, i, j, k)
#define X6(a, i, j, k)  AT(a, nx + 6, ny + 6, i, j, k)
#define FX(a, i, j, k)  AT(a, nx + 1, ny, i, j, k)
#define FY(a, i, j, k)  AT(a, nx, ny + 1, i, j, k)
#define FZ(a, i, j, k)  AT(a, nx, ny, i, j, k)
#define VV(a, i, j, k)  AT(a, nx + 1, ny + 1, i, j, k)

#define PFOR_K(n3) _Pragma("omp parallel for schedule(static)") for (int k = 1; k <= (n3); ++k)
#define FOR_JI(n1, n2) for (int j = 1; j <= (n2); ++j) for (int i = 1; i <= (n1); ++i)

/* ParallelStencil.FiniteDifferences3D index macros on an array of extents (n1,n2,.) */
#define D_XI(A, n1, n2) (AT(A, n1, n2, i + 1, j + 1, k + 1) - AT(A, n1, n2, i, j + 1, k + 1))
#define D_YI(A, n1, n2) (AT(A, n1, n2, i + 1, j + 1, k + 1) - AT(A, n1, n2, i + 1, j, k + 1))
#define D_ZI(A, n1, n2) (AT(A, n1, n2, i + 1, j + 1, k + 1) - AT(A, n1, n2, i + 1, j + 1, k))
#define AV_YZA(A, n1, n2) ((AT(A, n1, n2, i, j, k) + AT(A, n1, n2, i, j + 1, k) + AT(A, n1, n2, i, j, k + 1) + AT(A, n1, n2, i, j + 1, k + 1)) * 0.25)
#define AV_XZA(A, n1, n2) ((AT(A, n1, n2, i, j, k) + AT(A, n1, n2, i + 1, j, k) + AT(A, n1, n2, i, j, k + 1) + AT(A, n1, n2, i + 1, j, k + 1)) * 0.25)
#define AV_XYA(A, n1, n2) ((AT(A, n1, n2, i, j, k) + AT(A, n1, n2, i + 1, j, k) + AT(A, n1, n2, i, j + 1, k) + AT(A, n1, n2, i + 1, j + 1, k)) * 0.25)
#define AV_YI(A, n1, n2) ((AT(A, n1, n2, i + 1, j, k + 1) + AT(A, n1, n2, i + 1, j + 1, k + 1)) * 0.5)

/* DYREL parameters and report (Dimensional/HydroThermal3D_v18.jl:424-538) */
typedef struct {
    int nitmax;      /* maximum pseudo-iterations (v18: 1e4)                          */
    int nitout;      /* lambda_min / reference-norm schedule: iter<=2 || iter%nitout==0  */
    int ncheck;      /* >0: also test convergence every ncheck iterations               */
    int norm;        /* 0: mean|F*D|/sqrt(N) (v18:459-460); 1: max|F*D|                 */
    double tol;      /* converged if norm < tol or norm/norm_ini < tol (v18:476)         */
    double CFL;      /* 0.99 (v18:435)                                                  */
    double cfact;    /* 0.5  (v18:436)                                                  */
    double lmin_div; /* 500  (v18:434)                                                  */
} o_dyrel_params;

typedef struct {
    int iters;
    int status;      /* 0 converged, 10 nitmax, 11 NaN, 12 diverged (norm_rel > 100) */
    int n_checks;
    int n_lambda;
    double lmax, lmin, dtau, h1, h2;
    double nF_abs, nF_rel, nF_ini;
} o_dyrel_report;

enum { R_SUM = 0, R_SUM_ABS = 1, R_MAX = 2, R_MIN = 3, R_MAX_ABS = 4 };

/* kernels of ht3d_oracle.c used by the v18 solves of ht3d_oracle_dim.c */
double o_reduce(const double *a, size_t n1, size_t n23, int op);
void o_ResetA(double *A, double *B, int n1, int n2, int n3);
void o_SwapDYREL(double *Xc, double *Yc, const double *X0c_ex, const double *Y0c, int nx, int ny, int nz);
void o_ComputeFlux(double *qx, double *qy, double *qz, const double *kx, const double *ky, const double *kz,
                   const double *A, double _dx, double _dy, double _dz, int nx, int ny, int nz);
void o_ComputeDarcyFlux(double *qx, double *qy, double *qz, const double *rhofv, const double *kx, const double *ky,
                        const double *kz, const double *Pf, double g, double _dx, double _dy, double _dz, int nx,
                        int ny, int nz);
void o_ResidualTemperatureLinearised(double *F, const double *Tc_ex, const double *Tc_old, const double *rhoC_eff,
                                     const double *phc_ex, const double *PC, const double *qx, const double *qy,
                                     const double *qz, double Qt, double transient, double _dt, double _dx,
                                     double _dy, double _dz, int nx, int ny, int nz);
void o_ResidualFluidPressure(double *F, const double *phc_ex, const double *phirhof, const double *phirhof0,
                             const double *PC, const double *qx, const double *qy, const double *qz,
                             double transient, double _dt, double _dx, double _dy, double _dz, int nx, int ny,
                             int nz);
void o_DYRELUpdate(double *F0, double *X, const double *F, double h1, double h2, double dtau, int nx, int ny, int nz);
void o_Multiply(double *a, const double *b, const double *c, int nx, int ny, int nz);
void o_GershgorinPoisson(double *G, double *D, const double *rhoC, const double *kx, const double *ky,
                         const double *kz, double transient, double dt, double dx, double dy, double dz, int nx,
                         int ny, int nz);
void o_RayleighQuotientNumerator(double *G, const double *Fc0, const double *Fcit, const double *Fc, double dtau,
                                 int nx, int ny, int nz);
void o_RayleighQuotientDenominator(double *G, const double *Fc0, double dtau, int nx, int ny, int nz);

#endif
