// ht3d_dyrel.hip -- fused DYREL pseudo-transient solve (the hot loop).
//
// One kernel launch = one pseudo-iteration of HydroThermal3D_v18.jl:448-478
// (:506-537) on the dimensionless operators of Dimensionless/kernels_HT3D.jl:
//
//   SetTemperatureBCs!/SetPressureBCs!  -> ghost values derived on the fly
//   ComputeFlux!                         -> q on the six faces of the cell
//   ResidualDiffusion(Transient)! / PC   -> F = (R [+ X/dt] + div q) / D
//   DYRELUpdate!                         -> F0 = h1 F0 + h2 F ; X -= dtau F0
//   Multiply + mean/max, Rayleigh sums   -> per-block partials (check iterations)
//
// D (GershgorinPoisson!, kernels_HT3D_dimensional.jl:207-221) is recomputed from
// the six face coefficients the flux already loaded, so an iteration moves the
// algorithmic minimum 8 N words: read X, F0, kx, ky, kz, R; write X', F0.
// X is double-buffered (read Xin, write Xout) because neighbours must see the
// pre-update values, exactly like the reference's separate flux kernel.
//
// Layout: E arrays (ncx+2)(ncy+2)(ncz+2), x fastest.  A block is 64 x TY threads
// over an (x,y) tile; each thread marches KZ cells in z with a 3-deep register
// queue for X and a 2-deep one for kz.  Tiles are numbered so that the blocks
// the hardware sends to one XCD (block id mod 8) own a contiguous slab, keeping
// halo rows inside one L2.  Arithmetic follows the reference's evaluation order
// with no FMA contraction, so FP64 iterations equal the CPU ones bit for bit.
#include <cmath>
#include <cstring>
#include <vector>

#include "ht3d_iter.hpp"

namespace ht3d {
namespace v18 {
void clear_tiles();   // ht3d_v18.hip
}

// Gershgorin bound (GershgorinPoisson!, kernels_HT3D_dimensional.jl:207-221): max of G
// (optionally storing the diagonal D for the DSTORED iteration variants)
__global__ __launch_bounds__(256) void gershgorin_max_kernel(const double *__restrict__ kx,
                                                             const double *__restrict__ ky,
                                                             const double *__restrict__ kz, int nx, int ny, int nz,
                                                             double dx, double dy, double dz, double rc_tr_dt,
                                                             double *__restrict__ partials, double *__restrict__ Dout) {
    const size_t N = (size_t)nx * ny * nz;
    double m = -INFINITY;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((size_t)nx * ny));
        const double kW = kx[((size_t)k * ny + j) * (nx + 1) + i], kE = kx[((size_t)k * ny + j) * (nx + 1) + i + 1];
        const double kS = ky[((size_t)k * (ny + 1) + j) * nx + i], kN = ky[((size_t)k * (ny + 1) + j + 1) * nx + i];
        const double kB = kz[c], kF = kz[c + (size_t)nx * ny];
        const double diag = (kB / dz + kF / dz) / dz + (kN / dy + kS / dy) / dy + (kE / dx + kW / dx) / dx + rc_tr_dt;
        const double g = fabs(kE / (dx * dx)) + fabs(kW / (dx * dx)) + fabs(kN / (dy * dy)) + fabs(kS / (dy * dy)) +
                         fabs(kB / (dz * dz)) + fabs(kF / (dz * dz)) + fabs(diag);
        const double G = g / diag;
        if (Dout) Dout[c] = diag;
        m = (G > m || G != G) ? G : m;
    }
    const double h4[4] = {m, 0.0, 0.0, 0.0}, l4[4] = {0.0, 0.0, 0.0, 0.0};
    block_reduce4(h4, l4, 1u, partials, blockIdx.x);
}

// copy the ghost layer of src into dst (E arrays)
__global__ __launch_bounds__(256) void copy_ghosts_kernel(double *__restrict__ dst, const double *__restrict__ src,
                                                          int nx, int ny, int nz) {
    const int s1 = nx + 2, s2 = ny + 2, s3 = nz + 2;
    const size_t nf = (size_t)2 * s2 * s3 + (size_t)2 * s1 * s3 + (size_t)2 * s1 * s2;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nf; t += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        size_t r = t;
        if (r < (size_t)2 * s2 * s3) {
            i = (r & 1) ? s1 - 1 : 0; r >>= 1; j = (int)(r % s2); k = (int)(r / s2);
        } else if ((r -= (size_t)2 * s2 * s3) < (size_t)2 * s1 * s3) {
            j = (r & 1) ? s2 - 1 : 0; r >>= 1; i = (int)(r % s1); k = (int)(r / s1);
        } else {
            r -= (size_t)2 * s1 * s3;
            k = (r & 1) ? s3 - 1 : 0; r >>= 1; i = (int)(r % s1); j = (int)(r / s1);
        }
        const size_t e = ((size_t)k * s2 + j) * s1 + i;
        dst[e] = src[e];
    }
}

// Nusselt sum: top-face thermal flux qy(:, ncy+1, :) of HydroThermal3D_Alban_v6.jl:345-346,
// evaluated on the pre-update temperature whose ghosts have been filled.
__global__ __launch_bounds__(256) void nu_sum_kernel(const double *__restrict__ T, const double *__restrict__ ky,
                                                     int nx, int ny, int nz, double _dy, double *__restrict__ partials) {
    double s = 0.0, sl = 0.0;
    const size_t n = (size_t)nx * nz;
    const size_t sE1 = (size_t)nx + 2, sE2 = sE1 * (ny + 2);
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), k = (int)(t / nx);
        const double kN = ky[((size_t)k * (ny + 1) + ny) * nx + i];
        const double a = T[(size_t)(k + 1) * sE2 + (size_t)(ny + 1) * sE1 + i + 1];
        const double b = T[(size_t)(k + 1) * sE2 + (size_t)ny * sE1 + i + 1];
        dd_add(s, sl, -kN * (_dy * (a - b)));
    }
    const double h4[4] = {s, 0.0, 0.0, 0.0}, l4[4] = {sl, 0.0, 0.0, 0.0};
    block_reduce4(h4, l4, 0u, partials, blockIdx.x);
}

// shared state of the autotuner (declared in ht3d_iter.hpp)
std::vector<TuneEntry> g_tune;
bool g_autotune = true;
int g_force_tile = -1, g_force_variant = 0;
int g_pairs = 1;
int g_force_pair_tile = -1;
int g_force_pair_variant = 0;
int g_alternate = 1;

IterCfg &last_cfg() {
    static IterCfg c;
    return c;
}

PairCfg &last_pair() {
    static PairCfg c;
    return c;
}

static inline bool lambda_iter(int iter, int nitout) { return iter <= 2 || (nitout > 0 && iter % nitout == 0); }

// face-by-face ghost fill with an explicit BC table (same order as the per-kernel entry)
int bc_fill_table(double *A, int nx, int ny, int nz, const int type[6], const double val[6]);

unsigned gershgorin_blocks(size_t N) {
    const unsigned nb = div_up(N, 256);
    return nb > 4096 ? 4096 : nb;
}

int launch_gershgorin(const double *kx, const double *ky, const double *kz, int nx, int ny, int nz, double dx,
                      double dy, double dz, double rc_tr_dt, double *partials, unsigned nb, double *Dout) {
    hipLaunchKernelGGL(gershgorin_max_kernel, dim3(nb), dim3(256), 0, ctx().stream, kx, ky, kz, nx, ny, nz, dx, dy,
                       dz, rc_tr_dt, partials, Dout);
    HIP_TRY(hipGetLastError());
    return HT3D_OK;
}

int gershgorin_lmax(const double *kx, const double *ky, const double *kz, int nx, int ny, int nz, double dx,
                    double dy, double dz, double rc_tr_dt, double *lmax, double *Dout) {
    void *part = nullptr, *res = nullptr;
    const unsigned nb = gershgorin_blocks((size_t)nx * ny * nz);
    int rc = workspace(WS_MISC, (size_t)nb * 8 * sizeof(double), &part);
    if (rc) return rc;
    rc = workspace(WS_RESULT, 64 * sizeof(double), &res);
    if (rc) return rc;
    if ((rc = launch_gershgorin(kx, ky, kz, nx, ny, nz, dx, dy, dz, rc_tr_dt, (double *)part, nb, Dout))) return rc;
    rc = finish_partials((double *)part, nb, 1u, (double *)res);
    if (rc) return rc;
    double r4[4];
    rc = read_result4((double *)res, r4);
    if (rc) return rc;
    *lmax = r4[0];
    return HT3D_OK;
}

unsigned nu_blocks(int nx, int nz) {
    const unsigned nb = div_up((size_t)nx * nz, 256);
    return nb > 1024 ? 1024 : nb;
}

int launch_nu_partials(const double *T, const double *ky, int nx, int ny, int nz, double _dy, double *partials,
                       unsigned nb) {
    hipLaunchKernelGGL(nu_sum_kernel, dim3(nb), dim3(256), 0, ctx().stream, T, ky, nx, ny, nz, _dy, partials);
    HIP_TRY(hipGetLastError());
    return HT3D_OK;
}

// Sum of the top-face flux qy(:, ncy+1, :) of an E array whose ghosts are filled.
int nu_sum(const double *T, const double *ky, int nx, int ny, int nz, double _dy, double *sum) {
    const unsigned nb = nu_blocks(nx, nz);
    void *npart = nullptr, *res = nullptr;
    int rc;
    if ((rc = workspace(WS_MISC, (size_t)nb * 8 * sizeof(double), &npart))) return rc;
    if ((rc = workspace(WS_RESULT, 64 * sizeof(double), &res))) return rc;
    if ((rc = launch_nu_partials(T, ky, nx, ny, nz, _dy, (double *)npart, nb))) return rc;
    if ((rc = finish_partials((double *)npart, nb, 0u, (double *)res))) return rc;
    double r4[4];
    if ((rc = read_result4((double *)res, r4))) return rc;
    *sum = r4[0];
    return HT3D_OK;
}

// copy the ghost layer of src into dst (E arrays)
int copy_ghosts(double *dst, const double *src, int nx, int ny, int nz) {
    const size_t nf = (size_t)2 * ((ny + 2) * (size_t)(nz + 2) + (nx + 2) * (size_t)(nz + 2) + (nx + 2) * (size_t)(ny + 2));
    unsigned nb = div_up(nf, 256);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(copy_ghosts_kernel, dim3(nb), dim3(256), 0, ctx().stream, dst, src, nx, ny, nz);
    HIP_TRY(hipGetLastError());
    return HT3D_OK;
}

// The DYREL solve.  If nu_out is given (heat solve), the Nusselt number of the
// last pseudo-iteration is returned through it (sum / nu_div).  With `given`
// the device buffers come from the caller and nothing is allocated: no tuning
// pass (the cached tile of the grid, else the default tile) and F0, Fst
// cleared by a fill kernel.
int dyrel_solve_impl(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep,
                     double *nu_out, double nu_div, const SolveBuffers *given) {
    HT3D_CHECK_INIT();
    if (!f || !p) return arg_fail("ht3d_dyrel_solve: null argument");
    const int nx = f->ncx, ny = f->ncy, nz = f->ncz;
    if (!grid_ok(nx, ny, nz)) return arg_fail("ht3d_dyrel_solve: extents must be >= 3");
    if (!f->X || !f->R || !f->kx || !f->ky || !f->kz) return arg_fail("ht3d_dyrel_solve: null field pointer");
    if (p->nitmax < 1 || p->nitout < 1 || !(p->tol >= 0.0)) return arg_fail("ht3d_dyrel_solve: bad parameters");
    Context &cx = ctx();
    const size_t N = (size_t)nx * ny * nz, NE = (size_t)(nx + 2) * (ny + 2) * (nz + 2);

    void *xb = nullptr, *f0 = nullptr, *f0b = nullptr, *fst = nullptr, *part = nullptr, *res = nullptr, *dst = nullptr;
    int rc;
    IterArgs<double> a;
    std::memset(&a, 0, sizeof a);
    a.nx = nx; a.ny = ny; a.nz = nz;
    const double rc_tr_dt = f->rhoC * (double)f->transient / f->dt;
    double lmax = 0.0;
    if (given) {
        xb = given->xb; f0 = given->f0; f0b = given->f0b; fst = given->fst; dst = given->dst;
        part = given->partials; res = given->result;
        // set-up (v18:432-439) on the given partials; the diagonal D is kept for the DSTORED variants
        const unsigned nb = gershgorin_blocks(N);
        if ((rc = launch_gershgorin(f->kx, f->ky, f->kz, nx, ny, nz, f->dx, f->dy, f->dz, rc_tr_dt, (double *)part, nb,
                                    (double *)dst)) ||
            (rc = finish_partials((double *)part, nb, 1u, (double *)res)))
            return rc;
        double r4[4];
        if ((rc = read_result4((double *)res, r4))) return rc;
        lmax = r4[0];
    } else {
        if ((rc = workspace(WS_XB, NE * sizeof(double), &xb))) return rc;
        if ((rc = workspace(WS_F0, N * sizeof(double), &f0))) return rc;
        if ((rc = workspace(WS_F0B, N * sizeof(double), &f0b))) return rc;
        if ((rc = workspace(WS_FCIT, N * sizeof(double), &fst))) return rc;
        if ((rc = workspace(WS_DIAG, N * sizeof(double), &dst))) return rc;
        // most blocks of any tile configuration (1 row, or 2 planes, per block at the least)
        const unsigned max_blocks = div_up(nx, 64) * div_up(ny, 1) * div_up(nz, 2);
        if ((rc = workspace(WS_PARTIALS, (size_t)max_blocks * 8 * sizeof(double), &part))) return rc;
        if ((rc = workspace(WS_RESULT, 64 * sizeof(double), &res))) return rc;
        // set-up (v18:432-439); the diagonal D is kept for the DSTORED variants
        if ((rc = gershgorin_lmax(f->kx, f->ky, f->kz, nx, ny, nz, f->dx, f->dy, f->dz, rc_tr_dt, &lmax, (double *)dst)))
            return rc;
    }
    double lmin = lmax / p->lmin_div;
    const double dtau = 2.0 / sqrt(lmax) * p->CFL;
    double c = 2.0 * sqrt(lmin) * p->cfact;
    double h1 = (2 - c * dtau) / (2 + c * dtau), h2 = 2 * dtau / (2 + c * dtau);

    a.R = f->R; a.kx = f->kx; a.ky = f->ky; a.kz = f->kz; a.kv = f->kv; a.Dst = (const double *)dst;
    a.F0 = (double *)f0; a.Fst = (double *)fst; a.partials = (double *)part;
    a._dx = 1.0 / f->dx; a._dy = 1.0 / f->dy; a._dz = 1.0 / f->dz; a._dt = 1.0 / f->dt;
    a.dx = f->dx; a.dy = f->dy; a.dz = f->dz; a.rc_tr_dt = rc_tr_dt;
    set_inv_h2(a, f->dx, f->dy, f->dz);
    a.dtau = dtau; a.h1 = h1; a.h2 = h2;
    for (int q = 0; q < 6; ++q) { a.bct[q] = f->bc_type[q]; a.bcv[q] = f->bc_val[q]; }
    a.Xin = f->X;
    a.Xout = (double *)xb;
    IterCfg cfg;
    PairCfg pcfg;
    TuneEntry *te = tune_find(TuneKey{nx, ny, nz, f->transient != 0 ? 1 : 0, 0, f->kv ? 1 : 0});
    if (given) {   // no tuning pass: the grid's cached choice, else the default tile
        if (te) {
            cfg = te->cfg;
            kv_normalize(a, cfg);
            if (te->pair.tile >= 0 && (g_pairs == 2 || (g_pairs == 1 && te->pair_us < te->us))) pcfg = te->pair;
        } else {
            cfg = IterCfg{tile_for_grid(kDefaultTile, nx), 6};
        }
        set_tiling(a, cfg.tile);
        if (given->max_blocks < (size_t)a.nbx * a.nby * a.nbz)
            return arg_fail("ht3d_dyrel_solve_noalloc: the grid needs more block partials than the scratch holds");
        if (pcfg.tile >= 0) {
            IterArgs<double> t = a;
            set_pair_tiling(t, pcfg.tile);
            if (given->max_blocks < (size_t)t.nbx * t.nby * t.nbz) pcfg.tile = -1;
        }
    } else if ((rc = choose_cfg(a, f->transient != 0, true, 0, (double *)f0b, &cfg, &pcfg))) {
        return rc;
    }
    const unsigned nblocks = (unsigned)a.nbx * a.nby * a.nbz;
    int ncheck = p->ncheck;
    if (ncheck < 0)   // the check-interval tuner
        ncheck = tuned_check_interval(te ? te->last_iters : -1.0,
                                      te ? (pcfg.tile >= 0 && te->pair_us < te->us ? te->pair_us : te->us) : -1.0,
                                      te ? te->check_us : -1.0, p->nitout);
    IterArgs<double> ap = a;   // the same arguments with the pair tiling
    if (pcfg.tile >= 0) set_pair_tiling(ap, pcfg.tile);

    if (given) {   // ResetA!(Fc, Fc0) with a fill kernel
        if ((rc = fill_zero(f0, N * sizeof(double))) || (rc = fill_zero(fst, N * sizeof(double)))) return rc;
    } else {
        HIP_TRY(hipMemsetAsync(f0, 0, N * sizeof(double), cx.stream));   // ResetA!(Fc, Fc0)
        HIP_TRY(hipMemsetAsync(fst, 0, N * sizeof(double), cx.stream));
    }

    double *buf[2] = {f->X, (double *)xb};
    double *f0buf[2] = {(double *)f0, (double *)f0b};
    int cur = 0, fc = 0;
    bool rev = false;   // sweep direction of the next launch
    SegmentTimer timer;
    last_cfg() = cfg;
    last_pair() = pcfg;
    // algorithmic bytes per iteration of the reference's DYREL iteration: X, F0
    // read and written, R and the face coefficients kx, ky, kz read (8 words
    // per cell), whatever the launch reads (pairs cross HBM once per two
    // iterations; the vertex-coefficient kernels read kv instead of the faces)
    const double bytes_per_iter = 8.0 * (double)N * sizeof(double);
    double nF_abs = 0.0, nF_rel = 0.0, nF_ini = 0.0;
    int status = HT3D_NOT_CONVERGED, iter, n_checks = 0, n_lambda = 0, done_iters = 0;
    auto plain = [&](int it) {   // no check, no stored F, no lambda update
        return !lambda_iter(it, p->nitout) && !(ncheck > 0 && it % ncheck == 0) &&
               !(it < p->nitmax && lambda_iter(it + 1, p->nitout));
    };
    for (iter = 1; iter <= p->nitmax; ++iter) {
        if (pcfg.tile >= 0 && iter + 1 < p->nitmax && plain(iter) && plain(iter + 1)) {
            // iterations iter and iter+1 in one launch (the last iteration of a
            // solve is always a single launch, so `prev` below is its input)
            ap.Xin = buf[cur];
            ap.Xout = buf[cur ^ 1];
            ap.F0 = f0buf[fc ^ 1];
            ap.h1 = h1; ap.h2 = h2;
            ap.flags = rev ? FL_REVERSE : 0u;
            rev = g_alternate && !rev;
            timer.before_launch();
            if ((rc = launch_pair(ap, (const double *)f0buf[fc], f->transient != 0, pcfg))) return rc;
            timer.count += 2;
            cur ^= 1;
            fc ^= 1;
            done_iters = ++iter;
            continue;
        }
        const bool lam = lambda_iter(iter, p->nitout);
        const bool chk = lam || (ncheck > 0 && iter % ncheck == 0);
        const bool store = iter < p->nitmax && lambda_iter(iter + 1, p->nitout);
        a.Xin = buf[cur];
        a.Xout = buf[cur ^ 1];
        a.F0 = f0buf[fc];
        a.h1 = h1; a.h2 = h2;
        a.flags = (chk ? FL_CHECK : 0u) | (lam ? FL_LAMBDA : 0u) | (store ? FL_STORE_F : 0u) |
                  (rev ? FL_REVERSE : 0u);
        rev = g_alternate && !rev;
        timer.before_launch();
        rc = launch_iter(a, f->transient != 0, cfg);
        if (rc) return rc;
        ++timer.count;
        cur ^= 1;
        done_iters = iter;
        if (chk || iter == p->nitmax) timer.close();
        if (chk) {
            ++n_checks;
            if ((rc = finish_partials((double *)part, nblocks, 0x2u, (double *)res))) return rc;
            double r4[4];
            if ((rc = read_result4((double *)res, r4))) return rc;
            timer.collect(bytes_per_iter, false);
            nF_abs = p->norm == 0 ? r4[0] / (double)N / sqrt((double)N) : r4[1];
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
    // Finalise: ghosts of the result = BC applied to the pre-update state, as the
    // reference leaves them (its BC kernel runs before the last update).
    double *res_buf = buf[cur], *prev = buf[cur ^ 1];
    if ((rc = bc_fill_table(prev, nx, ny, nz, f->bc_type, f->bc_val))) return rc;
    if ((rc = copy_ghosts(res_buf, prev, nx, ny, nz))) return rc;
    if (nu_out) {
        double sum = 0.0;
        if ((rc = nu_sum(prev, f->ky, nx, ny, nz, 1.0 / f->dy, &sum))) return rc;
        *nu_out = sum / nu_div;
    }
    if (res_buf != f->X) HIP_TRY(hipMemcpyAsync(f->X, res_buf, NE * sizeof(double), hipMemcpyDeviceToDevice, cx.stream));
    if (te) te->last_iters = done_iters;
    if (rep) {
        rep->ncheck = ncheck;
        rep->iters = done_iters;
        rep->status = status;
        rep->n_checks = n_checks;
        rep->n_lambda = n_lambda;
        rep->lmax = lmax; rep->lmin = lmin; rep->dtau = dtau; rep->h1 = h1; rep->h2 = h2;
        rep->nF_abs = nF_abs; rep->nF_rel = nF_rel; rep->nF_ini = nF_ini;
    }
    rc = after_launch("ht3d_dyrel_solve", true);
    if (rc) return rc;
    return status;
}

}  // namespace ht3d

extern "C" {

int ht3d_dyrel_solve(const ht3d_dyrel_field *f, const ht3d_dyrel_params *p, ht3d_dyrel_report *rep) {
    HT3D_API_BEGIN
    return ht3d::dyrel_solve_impl(f, p, rep, nullptr, 1.0);
    HT3D_API_END
}

int ht3d_tune_set(int autotune, int force_tile, int force_variant) {
    HT3D_API_BEGIN
    if (force_tile >= ht3d::kNumTiles || force_variant < 0 || force_variant > 15)
        return ht3d::arg_fail("ht3d_tune_set: no such tile configuration or variant");
    ht3d::g_autotune = autotune != 0;
    ht3d::g_force_tile = force_tile;
    ht3d::g_force_variant = force_variant;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_tune_clear(void) {
    HT3D_API_BEGIN
    ht3d::g_tune.clear();
    ht3d::v18::clear_tiles();
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_tune_last(int *rows_per_block, int *z_per_thread, int *variant) {
    HT3D_API_BEGIN
    const ht3d::IterCfg &c = ht3d::last_cfg();
    if (rows_per_block) *rows_per_block = ht3d::kTiles[c.tile].ty;
    if (z_per_thread) *z_per_thread = ht3d::kTiles[c.tile].kz;
    if (variant) *variant = c.variant;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_tune_last_tile(int *rows_per_block, int *z_per_thread, int *x_per_thread, int *variant) {
    HT3D_API_BEGIN
    const ht3d::IterCfg &c = ht3d::last_cfg();
    if (x_per_thread) *x_per_thread = ht3d::kTiles[c.tile].cpl;
    return ht3d_tune_last(rows_per_block, z_per_thread, variant);
    HT3D_API_END
}

int ht3d_tile_count(void) { return ht3d::kNumTiles; }

int ht3d_tune_check_interval(double iters, double t_iter_us, double t_check_us, int nitout) {
    HT3D_API_BEGIN
    if (nitout < 1) return 0;
    return ht3d::tuned_check_interval(iters, t_iter_us, t_check_us, nitout);
    HT3D_API_END
}

int ht3d_tune_sweep(int alternate) {
    HT3D_API_BEGIN
    ht3d::g_alternate = alternate ? 1 : 0;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_tune_pairs(int mode) {
    HT3D_API_BEGIN
    if (mode < 0 || mode > 2) return ht3d::arg_fail("ht3d_tune_pairs: mode must be 0, 1 or 2");
    ht3d::g_pairs = mode;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_tune_last_pair(int *rows_per_block, int *z_per_block, int *variant) {
    HT3D_API_BEGIN
    const ht3d::PairCfg &c = ht3d::last_pair();
    if (rows_per_block) *rows_per_block = c.tile >= 0 ? ht3d::kPairTiles[c.tile].ty : 0;
    if (z_per_block) *z_per_block = c.tile >= 0 ? ht3d::kPairTiles[c.tile].kz : 0;
    if (variant) *variant = c.tile >= 0 ? c.variant : -1;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_tune_last_pair_tile(int *rows_per_block, int *z_per_block, int *x_per_thread, int *variant) {
    HT3D_API_BEGIN
    const ht3d::PairCfg &c = ht3d::last_pair();
    if (x_per_thread) *x_per_thread = c.tile >= 0 ? ht3d::kPairTiles[c.tile].cpl : 0;
    return ht3d_tune_last_pair(rows_per_block, z_per_block, variant);
    HT3D_API_END
}

int ht3d_tune_force_pair(int tile, int variant) {
    HT3D_API_BEGIN
    if (tile >= ht3d::kNumPairTiles || variant < 0 || variant > 15)
        return ht3d::arg_fail("ht3d_tune_force_pair: no such pair tile or variant");
    ht3d::g_force_pair_tile = tile;
    ht3d::g_force_pair_variant = variant;
    return HT3D_OK;
    HT3D_API_END
}

int ht3d_pair_tile_count(void) { return ht3d::kNumPairTiles; }

}  // extern "C"
