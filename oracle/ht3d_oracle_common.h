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

/* 1-based, column-major element access for an array of leading extents (n1,n2). */
#define AT(a, n1, n2, i, j, k) \
    ((a)[(((size_t)(k) - 1) * (size_t)(n2) + ((size_t)(j) - 1)) * (size_t)(n1) + ((size_t)(i) - 1)])

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

#endif
