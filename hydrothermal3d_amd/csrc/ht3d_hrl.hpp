// ht3d_hrl.hpp -- the closed-box HRL model object and the phases of its
// physical step, shared by the single-model step (ht3d_model.hip) and the
// batched step (ht3d_batch.hip).
#pragma once

#include "ht3d_internal.hpp"

struct ht3d_hrl {
    int nx, ny, nz;
    double Lx, Ly, Lz, dx, dy, dz;
    double Ra, Tbot, Ttop, cfl, dt, time, Nu, mu_a;
    int weno_version, precision, dt_rule;
    ht3d_dyrel_params pT, pP;
    double *Tc_ex, *Pc_ex, *ktv, *kfv, *Ty, *kx, *ky, *kz, *ktx, *kty, *ktz, *Vx, *Vy, *Vz, *Rt, *Rp, *W;
};

namespace ht3d {

// the model's internal launches do not synchronise one by one
struct SyncGuard {
    bool saved;
    SyncGuard() : saved(ctx().sync) { ctx().sync = false; }
    ~SyncGuard() { ctx().sync = saved; }
};

// InitThermal! (HydroThermal3D_Alban_v6.jl:328) and the thermal solve's field
int hrl_pre_thermal(ht3d_hrl *m, ht3d_dyrel_field *f);
// InitDarcy!, the T-dependent mobility, KillNorthSouthFluxes!, Compute_Rp!
// (Alban_v6:354-356) and the Darcy solve's field
int hrl_pre_darcy(ht3d_hrl *m, ht3d_dyrel_field *f);
// Init_vel!, the advective time step and WENO5 advection (HydroThermal3D_v18.jl:548-570);
// fills dt_next, vmax and n_adv of the report and advances m->time by the step's dt
int hrl_advect(ht3d_hrl *m, ht3d_step_report *r);
// hrl_advect in two parts: the velocity and max|V| into res_dev[0] (queued, no
// read-back), then the time step from max|V| and the advection
int hrl_velocity_dev(ht3d_hrl *m, double *res_dev);
int hrl_advect_vmax(ht3d_hrl *m, double vmax, ht3d_step_report *r);
// hrl_advect_vmax without the advection: advances m->time, applies the time
// step rule, fills the report and returns the sub-step count and sub-step dt
int hrl_advect_plan(ht3d_hrl *m, double vmax, ht3d_step_report *r, int *n_adv, double *dt_a);
// one advection sub-step of the model (the six passes of weno_advect)
int hrl_weno(ht3d_hrl *m, double dt_a);

// One model of a batched advection sub-step: its fields, grid spacings, sub-step
// dt and boundary values, and the first block of its share of the launch.
struct WenoItem {
    double *T, *W;
    const double *V[3];
    double h[3];   // 1/dx, 1/dy, 1/dz
    double dt, Tbot, Ttop;
    int nx, ny, nz, v2;
    unsigned b0;
};
// blocks of one model in a batched pass
unsigned weno_blocks(int nx, int ny, int nz);
WenoItem hrl_weno_item(const ht3d_hrl *m, double dt_a, unsigned b0);
// one advection sub-step of every item (device table, sorted by b0; nblocks =
// the last item's b0 + its weno_blocks): six launches in all
int weno_advect_batch(const WenoItem *d_items, int n, unsigned nblocks);
// Nu from the thermal flux sum (Alban_v6:345-348)
inline void hrl_set_nu(ht3d_hrl *m, double nusum) { m->Nu = nusum / (m->Lx * m->Lz) / (double)(m->nx * m->nz); }

}  // namespace ht3d
