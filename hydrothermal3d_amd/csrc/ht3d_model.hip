// ht3d_model.hip -- fused WENO5 advection, the closed-box physical step, and
// the synthetic permeability generator.
//
// WENO5: AdvectWithWeno5 (Dimensionless/kernels_HT3D.jl:156-233, stencils of
// tools/Weno5_Routines.jl) and AdvectWithWeno5_v2 (tools/Weno5_Routines_v2.jl:1-50)
// become two passes per direction.  Pass 1 reads T and writes T1 = T - dt L(T);
// pass 2 reads T1 and writes T = 1/2 (T1 - dt L(T1)) + 1/2 T in place (each
// thread reads the old T only at its own cell).  The three-layer halo copy of
// Boundaries_?_Weno5! is replaced by index mirroring inside the gather, so the
// (ncx+6)^3 scratch array and the ten cell-sized temporaries of v1 are gone.
//
// Physical step: HydroThermal3D_Alban_v6.jl:322-466 order (thermal solve, Nu,
// Darcy solve, velocity, advection, dt) with the DYREL solves of
// HydroThermal3D_v18.jl:424-538.
#include <cmath>
#include <cstring>
#include <new>

#include "ht3d_hrl.hpp"

namespace ht3d {

int dc_solve_impl(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep,
                  double *nu_out, int *outer, int *inner);

__device__ __forceinline__ double weno5_js_m(double v1, double v2, double v3, double v4, double v5) {
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

struct WenoArgs {
    const double *__restrict__ Tin;   // E
    double *Tout;                     // E (may alias Told in pass 2)
    const double *Told;               // E, pass 2 only
    const double *__restrict__ V;     // face array normal to the axis
    int nx, ny, nz;
    double h, dt;
    int tlo, thi;
    double vlo, vhi;
};

// One WENO5 pass along AXIS (0,1,2) at cell (i, j, k).  V2 selects the v2
// minus-side stencil quirk and its >= 0 upwind test; STAGE2 fuses
// TimeAveraging!(T, Told, 2).
template <int AXIS, bool V2, bool STAGE2>
__device__ __forceinline__ void weno_cell(const WenoArgs &a, const int i, const int j, const int k) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (i >= nx || j >= ny || k >= nz) return;
    const size_t sE1 = (size_t)nx + 2, sE2 = sE1 * (ny + 2);
    const size_t ec = (size_t)(k + 1) * sE2 + (size_t)(j + 1) * sE1 + (size_t)(i + 1);
    const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const int p0 = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const size_t st = AXIS == 0 ? 1 : (AXIS == 1 ? sE1 : sE2);
    double u[7];
#pragma unroll
    for (int o = -3; o <= 3; ++o) {
        const int p = p0 + o;
        if (p < 0) {            // low ghost layer -p: mirror of interior -p-1 (Boundaries_?_Weno5!)
            const double m = a.Tin[ec + (ptrdiff_t)(-p - 1 - p0) * (ptrdiff_t)st];
            u[o + 3] = a.tlo == 1 ? 2 * a.vlo - m : (a.tlo == 2 ? a.Tin[ec + (ptrdiff_t)(p + n - p0) * (ptrdiff_t)st] : m);
        } else if (p >= n) {    // high ghost: mirror of 2n-1-p
            const double m = a.Tin[ec + (ptrdiff_t)(2 * n - 1 - p - p0) * (ptrdiff_t)st];
            u[o + 3] = a.thi == 1 ? 2 * a.vhi - m : (a.thi == 2 ? a.Tin[ec + (ptrdiff_t)(p - n - p0) * (ptrdiff_t)st] : m);
        } else {
            u[o + 3] = a.Tin[ec + (ptrdiff_t)o * (ptrdiff_t)st];
        }
    }
    // face velocities: left face = cell index, right face = +1 along the axis
    size_t fl;
    if (AXIS == 0) fl = ((size_t)k * ny + j) * (nx + 1) + i;
    else if (AXIS == 1) fl = ((size_t)k * (ny + 1) + j) * nx + i;
    else fl = ((size_t)k * ny + j) * nx + i;
    const size_t fs = AXIS == 0 ? 1 : (AXIS == 1 ? (size_t)nx : (size_t)nx * ny);
    const double vl = a.V[fl], vr = a.V[fl + fs];
    double Vm, Vp;
    if (V2) {
        Vm = vl < 0.0 ? vl : 0.0;
        Vp = vr >= 0.0 ? vr : 0.0;
    } else {  // ResetA! then V?PlusMinus! (tools/Weno5_Routines.jl:5-42)
        Vm = vl < 0.0 ? vl : 0.0;
        Vp = vr > 0.0 ? vr : 0.0;
    }
    const double h = a.h;
    const double dm = weno5_js_m(h * (u[1] - u[0]), h * (u[2] - u[1]), h * (u[3] - u[2]), h * (u[4] - u[3]),
                                 V2 ? h * (u[5] - u[2]) : h * (u[5] - u[4]));
    const double dp = weno5_js_m(h * (u[6] - u[5]), h * (u[5] - u[4]), h * (u[4] - u[3]), h * (u[3] - u[2]),
                                 h * (u[2] - u[1]));
    double tn = u[3] - a.dt * (Vp * dm + Vm * dp);
    if (STAGE2) tn = (1.0 / 2.0) * tn + (1.0 - 1.0 / 2.0) * a.Told[ec];
    a.Tout[ec] = tn;
}

template <int AXIS, bool V2, bool STAGE2>
__global__ __launch_bounds__(256) void weno_pass_kernel(WenoArgs a) {
    weno_cell<AXIS, V2, STAGE2>(a, blockIdx.x * 64 + (threadIdx.x & 63), blockIdx.y * 4 + (threadIdx.x >> 6),
                                blockIdx.z);
}

// The same pass over every model of a batch in one launch: a 1-D grid whose
// blocks are dealt to the models in item order, each model's share laid out as
// weno_pass_kernel's grid (64 x 4 cells per block).  Pass 1 writes W from T,
// pass 2 writes T from W and the old T, as weno_axis does per model.
template <int AXIS, bool STAGE2>
__global__ __launch_bounds__(256) void weno_pass_batch_kernel(const WenoItem *__restrict__ items, int n) {
    const unsigned b = blockIdx.x;
    int lo = 0, hi = n - 1;   // the last item whose first block is <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].b0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const WenoItem &it = items[lo];
    const unsigned lb = b - it.b0, bx = ((unsigned)it.nx + 63) / 64, by = ((unsigned)it.ny + 3) / 4;
    WenoArgs a;
    a.Tin = STAGE2 ? it.W : it.T;
    a.Tout = STAGE2 ? it.T : it.W;
    a.Told = it.T;
    a.V = it.V[AXIS];
    a.nx = it.nx;
    a.ny = it.ny;
    a.nz = it.nz;
    a.h = it.h[AXIS];
    a.dt = it.dt;
    a.tlo = a.thi = AXIS == 1 ? 1 : 0;   // y: Dirichlet Tbot / Ttop; x, z: Neumann
    a.vlo = AXIS == 1 ? it.Tbot : 0.0;
    a.vhi = AXIS == 1 ? it.Ttop : 0.0;
    const int i = (int)(lb % bx) * 64 + (threadIdx.x & 63);
    const int j = (int)((lb / bx) % by) * 4 + (threadIdx.x >> 6);
    const int k = (int)(lb / (bx * by));
    if (it.v2) weno_cell<AXIS, true, STAGE2>(a, i, j, k);
    else weno_cell<AXIS, false, STAGE2>(a, i, j, k);
}

template <int AXIS, bool V2, bool STAGE2>
static int launch_weno(const WenoArgs &a) {
    dim3 blk(256), grd(div_up(a.nx, 64), div_up(a.ny, 4), a.nz);
    hipLaunchKernelGGL((weno_pass_kernel<AXIS, V2, STAGE2>), grd, blk, 0, ctx().stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "weno_pass_kernel");
    return HT3D_OK;
}

template <int AXIS, bool V2>
static int weno_axis(double *T, double *W, const double *V, int nx, int ny, int nz, double h, double dt, int tlo,
                     double vlo, int thi, double vhi) {
    WenoArgs a{T, W, nullptr, V, nx, ny, nz, h, dt, tlo, thi, vlo, vhi};
    int rc = launch_weno<AXIS, V2, false>(a);
    if (rc) return rc;
    WenoArgs b{W, T, T, V, nx, ny, nz, h, dt, tlo, thi, vlo, vhi};
    return launch_weno<AXIS, V2, true>(b);
}

// Fused AdvectWithWeno5 / _v2 with the closed-box boundary types
// (x Neumann, y Dirichlet Tbot/Ttop, z Neumann).  W: E-sized workspace.
int weno_advect(double *Tc_ex, double *W, const double *Vx, const double *Vy, const double *Vz, double dt, double _dx,
                double _dy, double _dz, double Ttop, double Tbot, int version, int nx, int ny, int nz) {
    int rc;
    if (version == 2) {
        if ((rc = weno_axis<0, true>(Tc_ex, W, Vx, nx, ny, nz, _dx, dt, 0, 0.0, 0, 0.0))) return rc;
        if ((rc = weno_axis<1, true>(Tc_ex, W, Vy, nx, ny, nz, _dy, dt, 1, Tbot, 1, Ttop))) return rc;
        return weno_axis<2, true>(Tc_ex, W, Vz, nx, ny, nz, _dz, dt, 0, 0.0, 0, 0.0);
    }
    if ((rc = weno_axis<0, false>(Tc_ex, W, Vx, nx, ny, nz, _dx, dt, 0, 0.0, 0, 0.0))) return rc;
    if ((rc = weno_axis<1, false>(Tc_ex, W, Vy, nx, ny, nz, _dy, dt, 1, Tbot, 1, Ttop))) return rc;
    return weno_axis<2, false>(Tc_ex, W, Vz, nx, ny, nz, _dz, dt, 0, 0.0, 0, 0.0);
}

unsigned weno_blocks(int nx, int ny, int nz) { return div_up(nx, 64) * div_up(ny, 4) * (unsigned)nz; }

template <int AXIS, bool STAGE2>
static int launch_weno_batch(const WenoItem *d_items, int n, unsigned nblocks) {
    hipLaunchKernelGGL((weno_pass_batch_kernel<AXIS, STAGE2>), dim3(nblocks), dim3(256), 0, ctx().stream, d_items, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "weno_pass_batch_kernel");
    return HT3D_OK;
}

int weno_advect_batch(const WenoItem *d_items, int n, unsigned nblocks) {
    if (n < 1 || nblocks == 0) return HT3D_OK;
    int rc;
    if ((rc = launch_weno_batch<0, false>(d_items, n, nblocks)) || (rc = launch_weno_batch<0, true>(d_items, n, nblocks)) ||
        (rc = launch_weno_batch<1, false>(d_items, n, nblocks)) || (rc = launch_weno_batch<1, true>(d_items, n, nblocks)) ||
        (rc = launch_weno_batch<2, false>(d_items, n, nblocks)))
        return rc;
    return launch_weno_batch<2, true>(d_items, n, nblocks);
}

// ------------------------------------------------------------------ random field
__global__ __launch_bounds__(256) void lognormal_kernel(double *__restrict__ kfv, int nx, int ny, int nz, double xmin,
                                                        double ymin, double zmin, double dx, double dy, double dz,
                                                        const double *__restrict__ K, const double *__restrict__ phi,
                                                        int M, double sigma, double fault_mult, double fault_half,
                                                        double n0, double n1, double n2, double x0, double y0,
                                                        double z0) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const double x = xmin + i * dx, y = ymin + j * dy, z = zmin + k * dz;
    const double amp = sqrt(2.0 / (double)M);
    double s = 0.0;
    for (int m = 0; m < M; ++m) s += cos(K[3 * m] * x + K[3 * m + 1] * y + K[3 * m + 2] * z + phi[m]);
    double v = pow(10.0, sigma * (amp * s));
    if (fault_mult != 1.0) {
        const double d = n0 * (x - x0) + n1 * (y - y0) + n2 * (z - z0);
        if (fabs(d) <= fault_half) v *= fault_mult;
    }
    kfv[((size_t)k * (ny + 1) + j) * (nx + 1) + i] = v;
}

// Init_vel! (kernels_HT3D.jl:73-80) over the three face arrays in one pass over
// the vertex-shaped index space, fused with max|V| (for the advective time step)
// and, when T is given, the top-face Nusselt sum of HydroThermal3D_Alban_v6.jl:345-346
// (flux -ky (T(ny+1) - T(ny)) / dy of an E array whose ghosts are filled).
// Slots: 0 max|V|, 1 Nu sum (double-double).
__global__ __launch_bounds__(256) void velocity_vmax_kernel(
    double *__restrict__ Vx, double *__restrict__ Vy, double *__restrict__ Vz, const double *__restrict__ kx,
    const double *__restrict__ ky, const double *__restrict__ kz, const double *__restrict__ P,
    const double *__restrict__ Ty, double Ra, double _dx, double _dy, double _dz, const double *__restrict__ T,
    const double *__restrict__ kt_y, int nx, int ny, int nz, double *__restrict__ partials) {
    const Grid g{nx, ny, nz};
    const size_t n1 = (size_t)nx + 1, n12 = n1 * (ny + 1), n = n12 * (nz + 1);
    double vmax = 0.0, s = 0.0, sl = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % n1), j = (int)((t / n1) % (ny + 1)), k = (int)(t / n12);
        if (j < ny && k < nz) {
            const double v = -kx[g.fx(i, j, k)] * _dx * (P[g.e(i + 1, j + 1, k + 1)] - P[g.e(i, j + 1, k + 1)]);
            Vx[g.fx(i, j, k)] = v;
            vmax = (fabs(v) > vmax || v != v) ? fabs(v) : vmax;
        }
        if (i < nx && k < nz) {
            const double v = -ky[g.fy(i, j, k)] * (_dy * (P[g.e(i + 1, j + 1, k + 1)] - P[g.e(i + 1, j, k + 1)]) -
                                                   Ra * Ty[g.fy(i, j, k)]);
            Vy[g.fy(i, j, k)] = v;
            vmax = (fabs(v) > vmax || v != v) ? fabs(v) : vmax;
            if (T && j == ny) {
                const double a = T[g.e(i + 1, ny + 1, k + 1)], b = T[g.e(i + 1, ny, k + 1)];
                dd_add(s, sl, -kt_y[g.fy(i, ny, k)] * (_dy * (a - b)));
            }
        }
        if (i < nx && j < ny) {
            const double v = -kz[g.fz(i, j, k)] * _dz * (P[g.e(i + 1, j + 1, k + 1)] - P[g.e(i + 1, j + 1, k)]);
            Vz[g.fz(i, j, k)] = v;
            vmax = (fabs(v) > vmax || v != v) ? fabs(v) : vmax;
        }
    }
    const double h4[4] = {vmax, s, 0.0, 0.0}, l4[4] = {0.0, sl, 0.0, 0.0};
    block_reduce4(h4, l4, 0x1u, partials, blockIdx.x);
}

// V, max|V| and (T given) the Nu sum in one pass and one read-back
// the velocity kernel's block partials: at most this many blocks, so one
// workspace of this size serves any grid (and launches queued back to back
// for several grids never see it reallocated)
static constexpr unsigned kVelBlocks = 4096;

// velocity, max|V| (and the Nu sum) into res_dev[0..1] on the device, no read-back
int velocity_vmax_dev(double *Vx, double *Vy, double *Vz, const double *kx, const double *ky, const double *kz,
                      const double *P, const double *Ty, double Ra, double _dx, double _dy, double _dz, const double *T,
                      const double *kt_y, int nx, int ny, int nz, double *res_dev) {
    const size_t n = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    unsigned nb = div_up(n, 256);
    if (nb > kVelBlocks) nb = kVelBlocks;
    void *part = nullptr;
    int rc;
    if ((rc = workspace(WS_MISC, (size_t)kVelBlocks * 8 * sizeof(double), &part))) return rc;
    hipLaunchKernelGGL(velocity_vmax_kernel, dim3(nb), dim3(256), 0, ctx().stream, Vx, Vy, Vz, kx, ky, kz, P, Ty, Ra,
                       _dx, _dy, _dz, T, kt_y, nx, ny, nz, (double *)part);
    HIP_TRY(hipGetLastError());
    return finish_partials((double *)part, nb, 0x1u, res_dev);
}

int velocity_vmax(double *Vx, double *Vy, double *Vz, const double *kx, const double *ky, const double *kz,
                  const double *P, const double *Ty, double Ra, double _dx, double _dy, double _dz, const double *T,
                  const double *kt_y, int nx, int ny, int nz, double *vmax, double *nusum) {
    void *res = nullptr;
    int rc;
    if ((rc = workspace(WS_RESULT, 64 * sizeof(double), &res))) return rc;
    if ((rc = velocity_vmax_dev(Vx, Vy, Vz, kx, ky, kz, P, Ty, Ra, _dx, _dy, _dz, T, kt_y, nx, ny, nz, (double *)res)))
        return rc;
    double r4[4];
    if ((rc = read_result4((double *)res, r4))) return rc;
    *vmax = r4[0];
    if (nusum) *nusum = r4[1];
    return HT3D_OK;
}

__global__ __launch_bounds__(256) void fill_kernel(double *__restrict__ a, size_t n, double v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

}  // namespace ht3d

using namespace ht3d;

namespace ht3d {

int hrl_pre_thermal(ht3d_hrl *m, ht3d_dyrel_field *f) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    int rc;
    if ((rc = ht3d_InitThermal(m->Rt, m->ktx, m->kty, m->ktz, m->Tc_ex, m->ktv, 1.0 / m->dt, nx, ny, nz))) return rc;
    std::memset(f, 0, sizeof *f);
    f->ncx = nx; f->ncy = ny; f->ncz = nz; f->dx = m->dx; f->dy = m->dy; f->dz = m->dz; f->dt = m->dt;
    f->X = m->Tc_ex; f->R = m->Rt; f->kx = m->ktx; f->ky = m->kty; f->kz = m->ktz;
    f->kv = m->ktv;   // ktx, kty, ktz are its face averages
    f->transient = 1; f->rhoC = 1.0;
    const int tt[6] = {0, 0, 1, 1, 0, 0};
    const double tv[6] = {0.0, 0.0, m->Tbot, m->Ttop, 0.0, 0.0};
    std::memcpy(f->bc_type, tt, sizeof tt);
    std::memcpy(f->bc_val, tv, sizeof tv);
    return HT3D_OK;
}

int hrl_pre_darcy(ht3d_hrl *m, ht3d_dyrel_field *f) {
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    int rc;
    if ((rc = ht3d_InitDarcy(m->Ty, m->kx, m->ky, m->kz, m->Tc_ex, m->kfv, 1.0 / m->dt, nx, ny, nz))) return rc;
    if (m->mu_a != 0.0 && (rc = ht3d_ApplyViscosity(m->kx, m->ky, m->kz, m->Tc_ex, m->mu_a, nx, ny, nz))) return rc;
    if ((rc = ht3d_KillNorthSouthFluxes(m->Ty, nx, ny, nz))) return rc;
    if ((rc = ht3d_Compute_Rp(m->Rp, m->ky, m->Ty, m->Ra, 1.0 / m->dy, nx, ny, nz))) return rc;
    std::memset(f, 0, sizeof *f);
    f->ncx = nx; f->ncy = ny; f->ncz = nz; f->dx = m->dx; f->dy = m->dy; f->dz = m->dz; f->dt = m->dt;
    f->X = m->Pc_ex; f->R = m->Rp; f->kx = m->kx; f->ky = m->ky; f->kz = m->kz;
    f->kv = m->mu_a == 0.0 ? m->kfv : nullptr;   // the face averages of kfv, unless divided by mu(T)
    f->transient = 0; f->rhoC = 0.0;
    return HT3D_OK;
}

int hrl_velocity_dev(ht3d_hrl *m, double *res_dev) {
    return velocity_vmax_dev(m->Vx, m->Vy, m->Vz, m->kx, m->ky, m->kz, m->Pc_ex, m->Ty, m->Ra, 1.0 / m->dx,
                             1.0 / m->dy, 1.0 / m->dz, nullptr, nullptr, m->nx, m->ny, m->nz, res_dev);
}

int hrl_advect(ht3d_hrl *m, ht3d_step_report *r) {
    // velocity, advective time step from this velocity, advection
    // (HydroThermal3D_v18.jl:551-570, HydroThermal3D_v7.jl:215-221)
    double vmax = 0.0;
    int rc;
    if ((rc = velocity_vmax(m->Vx, m->Vy, m->Vz, m->kx, m->ky, m->kz, m->Pc_ex, m->Ty, m->Ra, 1.0 / m->dx,
                            1.0 / m->dy, 1.0 / m->dz, nullptr, nullptr, m->nx, m->ny, m->nz, &vmax, nullptr)))
        return rc;
    return hrl_advect_vmax(m, vmax, r);
}

int hrl_advect_plan(ht3d_hrl *m, double vmax, ht3d_step_report *r, int *n_adv_out, double *dt_a_out) {
    const double dt = m->dt;
    m->time += dt;
    const double h = fmin(fmin(m->dx, m->dy), m->dz);
    const double dt_adv = m->cfl * h / vmax;
    int n_adv = 1;
    double dt_a = dt_adv;
    if (m->dt_rule == HT3D_DT_V18_CONSTANT) {
        const double q = std::fmax(std::ceil(dt / dt_adv), 1.0);  // 1 when V = 0
        if (!(q <= 1e6)) return arg_fail("ht3d_hrl_step: advection sub-step count out of range");
        n_adv = (int)q;
        dt_a = dt / n_adv;
    } else {
        if (!(dt_adv > 0.0) || !std::isfinite(dt_adv)) return arg_fail("ht3d_hrl_step: zero velocity, no CFL time step");
        if (m->dt_rule == HT3D_DT_V7) dt_a = dt;
        m->dt = dt_adv;
    }
    r->n_adv = n_adv;
    r->dt_next = m->dt;
    r->time = m->time;
    r->vmax = vmax;
    *n_adv_out = n_adv;
    *dt_a_out = dt_a;
    return HT3D_OK;
}

int hrl_weno(ht3d_hrl *m, double dt_a) {
    return weno_advect(m->Tc_ex, m->W, m->Vx, m->Vy, m->Vz, dt_a, 1.0 / m->dx, 1.0 / m->dy, 1.0 / m->dz, m->Ttop,
                       m->Tbot, m->weno_version, m->nx, m->ny, m->nz);
}

WenoItem hrl_weno_item(const ht3d_hrl *m, double dt_a, unsigned b0) {
    WenoItem it;
    it.T = m->Tc_ex;
    it.W = m->W;
    it.V[0] = m->Vx;
    it.V[1] = m->Vy;
    it.V[2] = m->Vz;
    it.h[0] = 1.0 / m->dx;
    it.h[1] = 1.0 / m->dy;
    it.h[2] = 1.0 / m->dz;
    it.dt = dt_a;
    it.Tbot = m->Tbot;
    it.Ttop = m->Ttop;
    it.nx = m->nx;
    it.ny = m->ny;
    it.nz = m->nz;
    it.v2 = m->weno_version == 2;
    it.b0 = b0;
    return it;
}

int hrl_advect_vmax(ht3d_hrl *m, double vmax, ht3d_step_report *r) {
    int n_adv, rc;
    double dt_a;
    if ((rc = hrl_advect_plan(m, vmax, r, &n_adv, &dt_a))) return rc;
    for (int s = 0; s < n_adv; ++s)
        if ((rc = hrl_weno(m, dt_a))) return rc;
    return HT3D_OK;
}

}  // namespace ht3d

extern "C" {

int ht3d_weno5_advect(double *Tc_ex, const double *Vx, const double *Vy, const double *Vz, double dt, double _dx,
                      double _dy, double _dz, double Ttop, double Tbot, int version, int nx, int ny, int nz) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!grid_ok(nx, ny, nz)) return arg_fail("ht3d_weno5_advect: extents must be >= 3");
    if (!Tc_ex || !Vx || !Vy || !Vz) return arg_fail("ht3d_weno5_advect: null pointer");
    void *w = nullptr;
    const size_t NE = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    int rc = workspace(WS_XB, NE * sizeof(double), &w);
    if (rc) return rc;
    rc = weno_advect(Tc_ex, (double *)w, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot, version, nx, ny, nz);
    if (rc) return rc;
    return after_launch("ht3d_weno5_advect");
    HT3D_API_END
}

int ht3d_lognormal_field(double *kfv, int nx, int ny, int nz, double xmin, double ymin, double zmin, double dx,
                         double dy, double dz, const double *K, const double *phi, int M, double sigma,
                         double fault_mult, double fault_half, const double *fault_n, const double *fault_x0) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!kfv || !K || !phi || M < 1 || nx < 1 || ny < 1 || nz < 1) return arg_fail("ht3d_lognormal_field: bad argument");
    double *dK = nullptr, *dphi = nullptr;
    HIP_TRY(hipMalloc(&dK, (size_t)3 * M * sizeof(double)));
    HIP_TRY(hipMalloc(&dphi, (size_t)M * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(dK, K, (size_t)3 * M * sizeof(double), hipMemcpyHostToDevice, ctx().stream));
    HIP_TRY(hipMemcpyAsync(dphi, phi, (size_t)M * sizeof(double), hipMemcpyHostToDevice, ctx().stream));
    const double n0 = fault_n ? fault_n[0] : 0.0, n1 = fault_n ? fault_n[1] : 0.0, n2 = fault_n ? fault_n[2] : 0.0;
    const double x0 = fault_x0 ? fault_x0[0] : 0.0, y0 = fault_x0 ? fault_x0[1] : 0.0, z0 = fault_x0 ? fault_x0[2] : 0.0;
    dim3 grd(div_up(nx + 1, 64), div_up(ny + 1, 4), nz + 1);
    hipLaunchKernelGGL(lognormal_kernel, grd, dim3(256), 0, ctx().stream, kfv, nx, ny, nz, xmin, ymin, zmin, dx, dy,
                       dz, dK, dphi, M, sigma, fault_mult, fault_half, n0, n1, n2, x0, y0, z0);
    int rc = after_launch("ht3d_lognormal_field", true);
    (void)hipFree(dK);
    (void)hipFree(dphi);
    return rc;
    HT3D_API_END
}

int ht3d_velocity_dt_nu(double *Vx, double *Vy, double *Vz, const double *kx, const double *ky, const double *kz,
                        const double *Pc_ex, const double *Ty, double Ra, double dx, double dy, double dz, double cfl,
                        const double *Tc_ex, const double *kty, double Lx, double Lz, double *out, int nx, int ny,
                        int nz) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!grid_ok(nx, ny, nz)) return arg_fail("ht3d_velocity_dt_nu: extents must be >= 3");
    if (!Vx || !Vy || !Vz || !kx || !ky || !kz || !Pc_ex || !Ty || !out || (Tc_ex && !kty))
        return arg_fail("ht3d_velocity_dt_nu: null pointer");
    if (!(dx > 0) || !(dy > 0) || !(dz > 0) || !(cfl > 0)) return arg_fail("ht3d_velocity_dt_nu: bad spacing or cfl");
    double vmax = 0.0, nusum = 0.0;
    int rc = velocity_vmax(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, 1.0 / dx, 1.0 / dy, 1.0 / dz, Tc_ex, kty, nx, ny, nz,
                           &vmax, &nusum);
    if (rc) return rc;
    out[0] = vmax;
    out[1] = cfl * fmin(fmin(dx, dy), dz) / vmax;                       // HydroThermal3D_v18.jl:554 rule
    out[2] = Tc_ex ? nusum / (Lx * Lz) / (double)(nx * nz) : 0.0;      // HydroThermal3D_Alban_v6.jl:346
    return after_launch("ht3d_velocity_dt_nu");
    HT3D_API_END
}

int ht3d_hrl_create(int nx, int ny, int nz, double Lx, double Ly, double Lz, double Ra, double cfl, ht3d_hrl **out) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!out || !grid_ok(nx, ny, nz) || !(Lx > 0) || !(Ly > 0) || !(Lz > 0) || !(cfl > 0))
        return arg_fail("ht3d_hrl_create: bad argument");
    ht3d_hrl *m = new (std::nothrow) ht3d_hrl();
    if (!m) return HT3D_ERR_NOMEM;
    std::memset(m, 0, sizeof *m);
    m->nx = nx; m->ny = ny; m->nz = nz;
    m->Lx = Lx; m->Ly = Ly; m->Lz = Lz;
    m->dx = Lx / nx; m->dy = Ly / ny; m->dz = Lz / nz;
    m->Ra = Ra; m->Tbot = 1.0; m->Ttop = 0.0; m->cfl = cfl;
    const double h = fmin(fmin(m->dx, m->dy), m->dz);
    m->dt = h * h / 6.1;  // HydroThermal3D_v7.jl:68
    m->weno_version = 1;
    m->precision = HT3D_FP64;
    ht3d_dyrel_params d = {10000, 1000, 0, 0, 1e-8, 0.99, 0.5, 500.0};
    m->pT = d; m->pP = d;
    const size_t c = (size_t)nx * ny * nz, E = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t V = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    const size_t fx = (size_t)(nx + 1) * ny * nz, fy = (size_t)nx * (ny + 1) * nz, fz = (size_t)nx * ny * (nz + 1);
    struct { double **p; size_t n; } arr[] = {
        {&m->Tc_ex, E}, {&m->Pc_ex, E}, {&m->ktv, V}, {&m->kfv, V}, {&m->Ty, fy}, {&m->kx, fx}, {&m->ky, fy},
        {&m->kz, fz}, {&m->ktx, fx}, {&m->kty, fy}, {&m->ktz, fz}, {&m->Vx, fx}, {&m->Vy, fy}, {&m->Vz, fz},
        {&m->Rt, c}, {&m->Rp, c}, {&m->W, E}};
    for (auto &a : arr) {
        hipError_t e = hipMalloc(a.p, a.n * sizeof(double));
        if (e != hipSuccess) { ht3d_hrl_destroy(m); return hip_fail(e, "hipMalloc(hrl)"); }
        e = hipMemsetAsync(*a.p, 0, a.n * sizeof(double), ctx().stream);
        if (e != hipSuccess) { ht3d_hrl_destroy(m); return hip_fail(e, "hipMemset(hrl)"); }
    }
    // ktv and kfv start as @ones (HydroThermal3D_Alban_v6.jl:246-247)
    for (double *p : {m->ktv, m->kfv}) {
        unsigned nb = div_up(V, 256);
        if (nb > 4096) nb = 4096;
        hipLaunchKernelGGL(fill_kernel, dim3(nb), dim3(256), 0, ctx().stream, p, V, 1.0);
    }
    *out = m;
    return after_launch("ht3d_hrl_create", true);
    HT3D_API_END
}

int ht3d_hrl_destroy(ht3d_hrl *m) {
    HT3D_API_BEGIN
    if (!m) return HT3D_OK;
    if (ctx().initialised) (void)hipStreamSynchronize(ctx().stream);
    double *ptrs[] = {m->Tc_ex, m->Pc_ex, m->ktv, m->kfv, m->Ty, m->kx, m->ky, m->kz, m->ktx, m->kty,
                      m->ktz, m->Vx, m->Vy, m->Vz, m->Rt, m->Rp, m->W};
    for (double *p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_hrl_field(ht3d_hrl *m, const char *name, double **dptr, size_t *count) {
    HT3D_API_BEGIN
    if (!m || !name || !dptr) return arg_fail("ht3d_hrl_field: bad argument");
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const size_t c = (size_t)nx * ny * nz, E = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t V = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    const size_t fx = (size_t)(nx + 1) * ny * nz, fy = (size_t)nx * (ny + 1) * nz, fz = (size_t)nx * ny * (nz + 1);
    struct { const char *n; double *p; size_t c; } tab[] = {
        {"Tc_ex", m->Tc_ex, E}, {"Pc_ex", m->Pc_ex, E}, {"ktv", m->ktv, V}, {"kfv", m->kfv, V}, {"Ty", m->Ty, fy},
        {"kx", m->kx, fx}, {"ky", m->ky, fy}, {"kz", m->kz, fz}, {"ktx", m->ktx, fx}, {"kty", m->kty, fy},
        {"ktz", m->ktz, fz}, {"Vx", m->Vx, fx}, {"Vy", m->Vy, fy}, {"Vz", m->Vz, fz}, {"Rt", m->Rt, c},
        {"Rp", m->Rp, c}};
    for (auto &t : tab)
        if (!std::strcmp(t.n, name)) {
            *dptr = t.p;
            if (count) *count = t.c;
            return HT3D_OK;
        }
    return arg_fail("ht3d_hrl_field: unknown field name");
    HT3D_API_END
}

int ht3d_hrl_set_params(ht3d_hrl *m, const ht3d_dyrel_params *pT, const ht3d_dyrel_params *pP, int weno_version,
                        int precision) {
    HT3D_API_BEGIN
    if (!m) return arg_fail("ht3d_hrl_set_params: null model");
    if (weno_version != 1 && weno_version != 2) return arg_fail("ht3d_hrl_set_params: weno_version must be 1 or 2");
    if (precision != HT3D_FP64 && precision != HT3D_FP32_DC) return arg_fail("ht3d_hrl_set_params: bad precision");
    if (pT) m->pT = *pT;
    if (pP) m->pP = *pP;
    m->weno_version = weno_version;
    m->precision = precision;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_hrl_set_viscosity(ht3d_hrl *m, double a) {
    HT3D_API_BEGIN
    if (!m || !(a == a)) return arg_fail("ht3d_hrl_set_viscosity: bad argument");
    m->mu_a = a;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_hrl_set_dt_rule(ht3d_hrl *m, int rule) {
    HT3D_API_BEGIN
    if (!m || rule < HT3D_DT_V18 || rule > HT3D_DT_V7) return arg_fail("ht3d_hrl_set_dt_rule: bad argument");
    m->dt_rule = rule;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_hrl_set_dt(ht3d_hrl *m, double dt) {
    HT3D_API_BEGIN
    if (!m || !(dt > 0)) return arg_fail("ht3d_hrl_set_dt: bad argument");
    m->dt = dt;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_hrl_step(ht3d_hrl *m, ht3d_step_report *rep) {
    HT3D_API_BEGIN
    HT3D_CHECK_INIT();
    if (!m) return arg_fail("ht3d_hrl_step: null model");
    SyncGuard guard;
    ht3d_step_report r;
    std::memset(&r, 0, sizeof r);
    r.dt_used = m->dt;
    int rc;
    ht3d_dyrel_field f;

    // thermal solve
    if ((rc = hrl_pre_thermal(m, &f))) return rc;
    double nusum = 0.0;
    int st;
    if (m->precision == HT3D_FP32_DC)
        st = dc_solve_impl(&f, &m->pT, &r.thermal, &nusum, &r.dc_outer_thermal, &r.inner_thermal);
    else
        st = dyrel_solve_impl(&f, &m->pT, &r.thermal, &nusum, 1.0);
    if (st != HT3D_OK && st != HT3D_NOT_CONVERGED) { r.status = st; if (rep) *rep = r; return st; }
    hrl_set_nu(m, nusum);

    // Darcy solve
    if ((rc = hrl_pre_darcy(m, &f))) return rc;
    int sp;
    if (m->precision == HT3D_FP32_DC)
        sp = dc_solve_impl(&f, &m->pP, &r.darcy, nullptr, &r.dc_outer_darcy, &r.inner_darcy);
    else
        sp = dyrel_solve_impl(&f, &m->pP, &r.darcy, nullptr, 1.0);
    if (sp != HT3D_OK && sp != HT3D_NOT_CONVERGED) { r.status = sp; if (rep) *rep = r; return sp; }

    // velocity, time step, advection
    if ((rc = hrl_advect(m, &r))) return rc;
    r.Nu = m->Nu;
    r.status = (st == HT3D_NOT_CONVERGED || sp == HT3D_NOT_CONVERGED) ? HT3D_NOT_CONVERGED : HT3D_OK;
    if (rep) *rep = r;
    rc = after_launch("ht3d_hrl_step", true);
    if (rc) return rc;
    return r.status;
    HT3D_API_END
}

}  // extern "C"
