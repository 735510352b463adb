"""Second, independent restatement of the hot path in NumPy array slicing.

TEST INFRASTRUCTURE ONLY (it verifies the C oracle).  Written from the
reference's Julia kernels independently of ht3d_oracle.c: whole-array slices
with 0-based offsets instead of 1-based point loops.  Every expression keeps the
reference's left-to-right evaluation order and NumPy never fuses multiply-adds,
so per-kernel results must equal the C oracle bit for bit; a transcription slip
in either restatement shows up as a mismatch.

Arrays are NumPy arrays indexed [i, j, k] (x, y, z) with the reference's
staggered shapes; sources are cited per function.
"""
from __future__ import annotations

import math

import numpy as np

I = slice(1, -1)  # interior of an extended array


def _f(a):
    return np.asfortranarray(a)


# ------------------------------------------------------------------ coefficients
def av_yza(v):   # @av_yza on a vertex array -> x faces
    return (v[:, :-1, :-1] + v[:, 1:, :-1] + v[:, :-1, 1:] + v[:, 1:, 1:]) * 0.25


def av_xza(v):   # @av_xza -> y faces
    return (v[:-1, :, :-1] + v[1:, :, :-1] + v[:-1, :, 1:] + v[1:, :, 1:]) * 0.25


def av_xya(v):   # @av_xya -> z faces
    return (v[:-1, :-1, :] + v[1:, :-1, :] + v[:-1, 1:, :] + v[1:, 1:, :]) * 0.25


def InitThermal(T, ktv, _dt):
    """kernels_HT3D.jl:99-107 -> (Rt, kx, ky, kz)"""
    return (-_dt) * T[I, I, I], av_yza(ktv), av_xza(ktv), av_xya(ktv)


def InitDarcy(T, kfv):
    """kernels_HT3D.jl:82-90 -> (Ty, kx, ky, kz); Ty = @av_yi(Tc_ex)"""
    Ty = (T[I, :-1, I] + T[I, 1:, I]) * 0.5
    return Ty, av_yza(kfv), av_xza(kfv), av_xya(kfv)


def ApplyViscosity(kx, ky, kz, T, a):
    """mu = exp(-a T): face coefficients times exp(a (T_left + T_right) * 0.5)."""
    return (kx * np.exp(a * ((T[:-1, I, I] + T[1:, I, I]) * 0.5)),
            ky * np.exp(a * ((T[I, :-1, I] + T[I, 1:, I]) * 0.5)),
            kz * np.exp(a * ((T[I, I, :-1] + T[I, I, 1:]) * 0.5)))


def KillNorthSouthFluxes(Ty):
    """HydroThermal3D_Alban_v6.jl:153-159"""
    Ty = Ty.copy()
    Ty[:, 0, :] = 0.0
    Ty[:, -1, :] = 0.0
    return Ty


def Compute_Rp(ky, Ty, Ra, _dy):
    """kernels_HT3D.jl:92-97 with @dmul_ya (tools/Macros.jl:111)"""
    return Ra * _dy * ((ky[:, 1:, :] * Ty[:, 1:, :]) - (ky[:, :-1, :] * Ty[:, :-1, :]))


# ------------------------------------------------------------------ ghosts
def fill_ghosts(A, types, vals):
    """Face ghosts: 0 Neumann copy, 1 Dirichlet 2*val - interior (Alban_v6:108-130).
    Edges/corners follow the same deterministic order as the library: x faces over
    the interior, y faces over all x, z faces over everything."""
    A = A.copy()
    t, v = types, vals
    A[0, I, I] = 2 * v[0] - A[1, I, I] if t[0] else A[1, I, I]
    A[-1, I, I] = 2 * v[1] - A[-2, I, I] if t[1] else A[-2, I, I]
    A[:, 0, I] = 2 * v[2] - A[:, 1, I] if t[2] else A[:, 1, I]
    A[:, -1, I] = 2 * v[3] - A[:, -2, I] if t[3] else A[:, -2, I]
    A[:, :, 0] = 2 * v[4] - A[:, :, 1] if t[4] else A[:, :, 1]
    A[:, :, -1] = 2 * v[5] - A[:, :, -2] if t[5] else A[:, :, -2]
    return A


def temperature_bc(A, Tbot, Ttop):
    return fill_ghosts(A, (0, 0, 1, 1, 0, 0), (0.0, 0.0, Tbot, Ttop, 0.0, 0.0))


def pressure_bc(A):
    return fill_ghosts(A, (0,) * 6, (0.0,) * 6)


# ------------------------------------------------------------------ fluxes and residuals
def ComputeFlux(kx, ky, kz, A, _dx, _dy, _dz):
    """kernels_HT3D.jl:109-117 -> (qx, qy, qz); q = (-k) * (_dh * @d_hi(A))"""
    qx = -kx * (_dx * (A[1:, I, I] - A[:-1, I, I]))
    qy = -ky * (_dy * (A[I, 1:, I] - A[I, :-1, I]))
    qz = -kz * (_dz * (A[I, I, 1:] - A[I, I, :-1]))
    return qx, qy, qz


def _div(qx, qy, qz, _dx, _dy, _dz):
    return _dx * (qx[1:] - qx[:-1]), _dy * (qy[:, 1:] - qy[:, :-1]), _dz * (qz[:, :, 1:] - qz[:, :, :-1])


def ResidualDiffusion(R, qx, qy, qz, _dx, _dy, _dz):
    """kernels_HT3D.jl:128-134"""
    a, b, c = _div(qx, qy, qz, _dx, _dy, _dz)
    return R + a + b + c


def ResidualDiffusionTransient(T, R, qx, qy, qz, _dt, _dx, _dy, _dz):
    """Alban_v6:84-90"""
    a, b, c = _div(qx, qy, qz, _dx, _dy, _dz)
    return R + _dt * T[I, I, I] + a + b + c


def GershgorinPoisson(rhoC, kx, ky, kz, transient, dt, dx, dy, dz):
    """kernels_HT3D_dimensional.jl:207-221 -> (G, D)"""
    kW, kE = kx[:-1], kx[1:]
    kS, kN = ky[:, :-1], ky[:, 1:]
    kB, kF = kz[:, :, :-1], kz[:, :, 1:]
    D = (kB / dz + kF / dz) / dz + (kN / dy + kS / dy) / dy + (kE / dx + kW / dx) / dx + rhoC * transient / dt
    G = (np.abs(kE / (dx * dx)) + np.abs(kW / (dx * dx)) + np.abs(kN / (dy * dy)) + np.abs(kS / (dy * dy)) +
         np.abs(kB / (dz * dz)) + np.abs(kF / (dz * dz)) + np.abs(D))
    return G / D, D


def DYRELUpdate(F0, X, F, h1, h2, dtau):
    """kernels_HT3D_dimensional.jl:322-326 -> (F0, X)"""
    F0 = h1 * F0 + h2 * F
    X = X.copy()
    X[I, I, I] = X[I, I, I] - dtau * F0
    return F0, X


def Init_vel(kx, ky, kz, P, Ty, Ra, _dx, _dy, _dz):
    """kernels_HT3D.jl:73-80 (note ((-k)*_dh)*dP for x and z)"""
    Vx = -kx * _dx * (P[1:, I, I] - P[:-1, I, I])
    Vy = -ky * (_dy * (P[I, 1:, I] - P[I, :-1, I]) - Ra * Ty)
    Vz = -kz * _dz * (P[I, I, 1:] - P[I, I, :-1])
    return Vx, Vy, Vz


# ------------------------------------------------------------------ WENO5 (tools/Weno5_Routines*.jl)
def weno5(v1, v2, v3, v4, v5):
    """dFdx_Weno5! tools/Weno5_Routines.jl:44-68 (eps term 10^-99 taken as 1e-99)"""
    p1 = v1 / 3.0 - 7.0 / 6.0 * v2 + 11.0 / 6.0 * v3
    p2 = -v2 / 6.0 + 5.0 / 6.0 * v3 + v4 / 3.0
    p3 = v3 / 3.0 + 5.0 / 6.0 * v4 - v5 / 6.0
    maxV = np.maximum(np.maximum(np.maximum(np.maximum(v1 * v1, v2 * v2), v3 * v3), v4 * v4), v5 * v5)
    e = 1e-99 + 1e-6 * maxV
    s1 = v1 - 2.0 * v2 + v3
    t1 = v1 - 4.0 * v2 + 3.0 * v3
    s2 = v2 - 2.0 * v3 + v4
    t2 = v2 - v4
    s3 = v3 - 2.0 * v4 + v5
    t3 = 3.0 * v3 - 4.0 * v4 + v5
    w1 = 13.0 / 12.0 * (s1 * s1) + 1.0 / 4.0 * (t1 * t1)
    w2 = 13.0 / 12.0 * (s2 * s2) + 1.0 / 4.0 * (t2 * t2)
    w3 = 13.0 / 12.0 * (s3 * s3) + 1.0 / 4.0 * (t3 * t3)
    w1 = 0.1 / ((w1 + e) * (w1 + e))
    w2 = 0.6 / ((w2 + e) * (w2 + e))
    w3 = 0.3 / ((w3 + e) * (w3 + e))
    w = w1 + w2 + w3
    w1 = w1 / w
    w2 = w2 / w
    w3 = w3 / w
    return w1 * p1 + w2 * p2 + w3 * p3


def halo3(Tc, axis, tlo, vlo, thi, vhi):
    """Boundaries_?_Weno5! (tools/Weno5_Routines.jl:130-255): Tc padded by three
    layers along `axis` (0,1,2); 0 mirror, 1 2*val - mirror, 2 periodic."""
    Tm = np.moveaxis(Tc, axis, 0)
    n = Tm.shape[0]
    lo_m = Tm[[2, 1, 0]]               # ghosts -3,-2,-1 <- interior 2,1,0
    hi_m = Tm[[n - 1, n - 2, n - 3]]   # ghosts n,n+1,n+2 <- interior n-1,n-2,n-3
    lo = lo_m if tlo == 0 else (2 * vlo - lo_m if tlo == 1 else Tm[n - 3:])
    hi = hi_m if thi == 0 else (2 * vhi - hi_m if thi == 1 else Tm[:3])
    return np.moveaxis(np.concatenate([lo, Tm, hi], axis=0), 0, axis)


def weno_derivatives(Tc, axis, h, tlo, vlo, thi, vhi, v2_quirk=False):
    """(dTdx minus, dTdx plus) of Gradients_{minus,plus}_*_Weno5! + dFdx_Weno5!;
    v2_quirk uses X[I+2]-X[I-1] for the fifth minus difference (Weno5_Routines_v2.jl:69)."""
    X = halo3(Tc, axis, tlo, vlo, thi, vhi)
    n = Tc.shape[axis]

    def s(o):  # X at offset o from the centre, over the interior
        sl = [slice(None)] * 3
        sl[axis] = slice(3 + o, 3 + o + n)
        return X[tuple(sl)]

    v5m = h * (s(2) - s(-1)) if v2_quirk else h * (s(2) - s(1))
    dm = weno5(h * (s(-2) - s(-3)), h * (s(-1) - s(-2)), h * (s(0) - s(-1)), h * (s(1) - s(0)), v5m)
    dp = weno5(h * (s(3) - s(2)), h * (s(2) - s(1)), h * (s(1) - s(0)), h * (s(0) - s(-1)), h * (s(-1) - s(-2)))
    return dm, dp


def upwind(V, axis, v2=False):
    """V?PlusMinus! (tools/Weno5_Routines.jl:5-42) after ResetA!, or the v2 split."""
    sl_l = [slice(None)] * 3
    sl_r = [slice(None)] * 3
    sl_l[axis] = slice(None, -1)
    sl_r[axis] = slice(1, None)
    vl, vr = V[tuple(sl_l)], V[tuple(sl_r)]
    Vm = np.where(vl < 0.0, vl, 0.0)
    Vp = np.where(vr >= 0.0, vr, 0.0) if v2 else np.where(vr > 0.0, vr, 0.0)
    return Vm, Vp


def advect_weno5(Tc_ex, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot, version=1):
    """AdvectWithWeno5 (kernels_HT3D.jl:156-233) / _v2 (Weno5_Routines_v2.jl:1-50)
    on the interior of Tc_ex; closed-box types x: Neumann, y: Dirichlet, z: Neumann."""
    Tc = Tc_ex[I, I, I].copy()
    bcs = [(0, 0.0, 0, 0.0), (1, Tbot, 1, Ttop), (0, 0.0, 0, 0.0)]
    for axis, (V, h) in enumerate(((Vx, _dx), (Vy, _dy), (Vz, _dz))):
        Vm, Vp = upwind(V, axis, version == 2)
        Told = Tc.copy()
        for _ in range(2):
            dm, dp = weno_derivatives(Tc, axis, h, *bcs[axis], v2_quirk=(version == 2))
            Tc = Tc - dt * (Vp * dm + Vm * dp)
        Tc = (1.0 / 2.0) * Tc + (1.0 - 1.0 / 2.0) * Told
    out = Tc_ex.copy()
    out[I, I, I] = Tc
    return out


# ------------------------------------------------------------------ DYREL (HydroThermal3D_v18.jl:424-538)
def dyrel_solve(X, R, kx, ky, kz, *, dx, dy, dz, dt, transient, rhoC, bc_type, bc_val, nitmax=10000, nitout=1000,
                ncheck=0, norm=0, tol=1e-8, CFL=0.99, cfact=0.5, lmin_div=500.0):
    """Unfused DYREL solve; returns (X, report, last thermal flux qy)."""
    _dx, _dy, _dz, _dt = 1.0 / dx, 1.0 / dy, 1.0 / dz, 1.0 / dt
    shape = R.shape
    G, D = GershgorinPoisson(np.full(shape, float(rhoC)), kx, ky, kz, float(transient), dt, dx, dy, dz)
    lmax = float(G.max())
    lmin = lmax / lmin_div
    dtau = 2.0 / math.sqrt(lmax) * CFL
    c = 2.0 * math.sqrt(lmin) * cfact
    h1, h2 = (2 - c * dtau) / (2 + c * dtau), 2 * dtau / (2 + c * dtau)
    F = np.zeros(shape)
    F0 = np.zeros(shape)
    Fcit = np.zeros(shape)
    N = R.size
    nF_abs = nF_rel = nF_ini = 0.0
    status, it, n_lambda = 10, 0, 0
    qy = None
    for it in range(1, nitmax + 1):
        lam = it <= 2 or it % nitout == 0
        chk = lam or (ncheck > 0 and it % ncheck == 0)
        if lam:
            Fcit = F.copy()
        X = fill_ghosts(X, bc_type, bc_val)
        qx, qy, qz = ComputeFlux(kx, ky, kz, X, _dx, _dy, _dz)
        Fr = ResidualDiffusionTransient(X, R, qx, qy, qz, _dt, _dx, _dy, _dz) if transient else \
            ResidualDiffusion(R, qx, qy, qz, _dx, _dy, _dz)
        F = Fr / D
        F0, X = DYRELUpdate(F0, X, F, h1, h2, dtau)
        if chk:
            FD = F * D
            nF_abs = math.fsum(np.abs(FD).ravel()) / N / math.sqrt(N) if norm == 0 else float(np.abs(FD).max())
            if it == 1:
                nF_ini = nF_abs
            nF_rel = nF_abs / nF_ini
            if math.isnan(nF_abs):
                status = 11
                break
            if nF_rel > 100:
                status = 12
                break
            if lam and it > 1:
                d = dtau * F0
                lmin = abs(math.fsum(((-d) * (F - Fcit)).ravel()))
                lmin /= abs(math.fsum((d * d).ravel()))
                c = 2.0 * math.sqrt(lmin) * cfact
                h1, h2 = (2 - c * dtau) / (2 + c * dtau), 2 * dtau / (2 + c * dtau)
                n_lambda += 1
            if nF_abs < tol or nF_rel < tol:
                status = 0
                break
    rep = dict(iters=it, status=status, n_lambda=n_lambda, lmax=lmax, lmin=lmin, dtau=dtau, h1=h1, h2=h2,
               nF_abs=nF_abs, nF_rel=nF_rel, nF_ini=nF_ini)
    return X, rep, qy


# ------------------------------------------------------------------ physical step
class HRL:
    """Closed-box HRL convection stepped like HydroThermal3D_Alban_v6.jl:322-466
    with DYREL solves and the advective time step rules of HydroThermal3D_v18.jl:551-570
    (0, 1) and HydroThermal3D_v7.jl:215-221 (2);
    same step order as the C oracle's o_hrl_step."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, Ra, cfl=0.5):
        self.n = (nx, ny, nz)
        self.L = (Lx, Ly, Lz)
        self.h = (Lx / nx, Ly / ny, Lz / nz)
        self.Ra, self.cfl = Ra, cfl
        hm = min(self.h)
        self.dt = hm * hm / 6.1
        self.T = np.zeros((nx + 2, ny + 2, nz + 2))
        self.P = np.zeros((nx + 2, ny + 2, nz + 2))
        self.ktv = np.ones((nx + 1, ny + 1, nz + 1))
        self.kfv = np.ones((nx + 1, ny + 1, nz + 1))
        self.prm = dict(nitmax=10000, nitout=1000, ncheck=0, norm=0, tol=1e-8)
        self.weno = 1
        self.time = 0.0
        self.mu_a = 0.0
        self.dt_rule = 0          # 0: v18, 1: v18 constant dt + sub-steps, 2: v7

    def step(self):
        nx, ny, nz = self.n
        dx, dy, dz = self.h
        _dx, _dy, _dz = 1.0 / dx, 1.0 / dy, 1.0 / dz
        dt = self.dt
        _dt = 1.0 / dt
        Rt, kx, ky, kz = InitThermal(self.T, self.ktv, _dt)
        self.T, repT, qy = dyrel_solve(self.T, Rt, kx, ky, kz, dx=dx, dy=dy, dz=dz, dt=dt, transient=1, rhoC=1.0,
                                       bc_type=(0, 0, 1, 1, 0, 0), bc_val=(0.0, 0.0, 1.0, 0.0, 0.0, 0.0), **self.prm)
        Nu = math.fsum(qy[:, -1, :].ravel()) / (self.L[0] * self.L[2]) / (nx * nz)
        Ty, kx, ky, kz = InitDarcy(self.T, self.kfv)
        if self.mu_a != 0.0:
            kx, ky, kz = ApplyViscosity(kx, ky, kz, self.T, self.mu_a)
        Ty = KillNorthSouthFluxes(Ty)
        Rp = Compute_Rp(ky, Ty, self.Ra, _dy)
        self.P, repP, _ = dyrel_solve(self.P, Rp, kx, ky, kz, dx=dx, dy=dy, dz=dz, dt=dt, transient=0, rhoC=0.0,
                                      bc_type=(0,) * 6, bc_val=(0.0,) * 6, **self.prm)
        self.time += dt
        self.V = Init_vel(kx, ky, kz, self.P, Ty, self.Ra, _dx, _dy, _dz)
        vmax = max(float(np.abs(v).max()) for v in self.V)
        dt_adv = self.cfl * min(self.h) / vmax
        if self.dt_rule == 1:
            n_adv = max(math.ceil(dt / dt_adv), 1)
            dt_a = dt / n_adv
        else:
            n_adv, dt_a = 1, (dt if self.dt_rule == 2 else dt_adv)
            self.dt = dt_adv
        for _ in range(n_adv):
            self.T = advect_weno5(self.T, *self.V, dt_a, _dx, _dy, _dz, 0.0, 1.0, self.weno)
        return dict(thermal=repT, darcy=repP, Nu=Nu, dt_next=self.dt, n_adv=n_adv)
