#!/usr/bin/env python3
"""Benchmark: closed-box porous convection physical steps on one MI355X.

A step is one physical time step of the 256x128x256 (2x1x2), Ra=500,
lognormal-permeability workload (thermal DYREL solve, Nu, Darcy DYREL solve,
velocity, WENO5 advection, CFL time step), inputs resident in HBM.

Prints one JSON line:
  value         T_eff (GB/s) of the fused DYREL pseudo-iterations over the timed
                steps: algorithmic bytes of the reference's iteration (8 words
                per cell: X and F0 read and written, R, kx, ky, kz read; the
                ParallelStencil A_eff) / summed kernel time from HIP events on
                the library stream.  The kernel that runs may read fewer: with
                the vertex field it reads kv instead of kx, ky, kz (6 words, 7
                with the stored diagonal), reported as roofline.kernel_words and
                roofline.kernel_frac;
  ms_per_step   wall time per physical step (all kernels, reductions and host
                convergence checks);
  roofline      the iteration kernel against the 8 TB/s HBM peak;
  cpu_baseline  the CPU oracle (unfused reference kernel sequence, FP64, OpenMP)
                on a bounded sample of the same workload, as T_eff.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "T_eff GB/s (% HBM peak) + s/physical step, FP32+DC vs FP64 at 1e-6 rel err"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _ensure_built() -> None:
    from hydrothermal3d_amd import _lib
    if not _lib.LIBPATH.exists():
        import __graft_entry__
        __graft_entry__.build()


def timing_stats():
    import ctypes as C
    from hydrothermal3d_amd import _lib

    class Stats(C.Structure):
        _fields_ = [("ms", C.c_double), ("launches", C.c_longlong), ("bytes", C.c_double),
                    ("ms_fp32", C.c_double), ("launches_fp32", C.c_longlong), ("bytes_fp32", C.c_double)]

    s = Stats()
    _lib.call("ht3d_timing_get", C.byref(s))
    return {n: getattr(s, n) for n, _ in Stats._fields_}


def rel(a, b):
    import numpy as np
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run_gpu(w, steps: int, warmup: int, precision: int, prm):
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import _lib, inputs

    g = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    g["Tc_ex"].set(inputs.initial_temperature(w))
    inputs.permeability_device(w, g["kfv"])
    g.set_params(prm, prm, 1, precision)
    g.set_viscosity(w.mu_a)
    _lib.sync()
    for s in range(warmup):
        r = g.step()
        log(f"warmup step {s}: iters T {r['thermal']['iters']} P {r['darcy']['iters']} Nu {r['Nu']:.5f}")
    _lib.call("ht3d_timing_reset")
    _lib.call("ht3d_timing_enable", 1)
    _lib.sync()
    t0 = time.perf_counter()
    reports = []
    for s in range(steps):
        r = g.step()
        reports.append(r)
        log(f"step {s}: iters T {r['thermal']['iters']} P {r['darcy']['iters']} Nu {r['Nu']:.6f} "
            f"dt {r['dt_used']:.3e} norms T {r['thermal']['nF_abs']:.3e} P {r['darcy']['nF_abs']:.3e} "
            f"({time.perf_counter() - t0:.2f} s)")
    _lib.sync()
    wall = time.perf_counter() - t0
    _lib.call("ht3d_timing_enable", 0)
    st = timing_stats()
    import ctypes as C
    ty, kz, xc, var = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _lib.call("ht3d_tune_last_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
    st["tile"] = (ty.value, kz.value, var.value, xc.value)
    _lib.call("ht3d_tune_last_pair_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
    st["pair"] = (ty.value, kz.value, var.value, xc.value) if ty.value > 0 else None
    kfv = g["kfv"].get()
    I = (slice(1, -1),) * 3
    fields = {"T": g["Tc_ex"].get()[I], "P": g["Pc_ex"].get()[I]}
    g.close()
    return wall, st, reports, kfv, fields


def _cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo), for the CPU baseline."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def tile_label(tile) -> str:
    """rows x z-planes of an iteration tile (rows, z, variant[, x cells per
    lane]), with " x2" / " x4" when each lane holds that many x-cells."""
    return f"{tile[0]}x{tile[1]}" + (f" x{tile[3]}" if len(tile) > 3 and tile[3] > 1 else "")


def pmc_traffic(workload: str, tile) -> float | None:
    """HBM bytes per launch of the single-iteration kernel for this tile and
    variant, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the
    workload (calibrated against a copy kernel there); None when no summary
    covers it."""
    path = ROOT / "profiles" / f"dyrel_iter_{workload}_pmc.md"
    if not path.exists():
        return None
    want = f"| single | {tile_label(tile)} | {tile[2]} |"
    for line in path.read_text().splitlines():
        if line.startswith(want):
            return float(line.split("|")[9])
    return None


def _batch_models(ws, prm, precision):
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import inputs
    ms = []
    for w in ws:
        m = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        m["Tc_ex"].set(inputs.initial_temperature(w))
        inputs.permeability_device(w, m["kfv"])
        m.set_params(prm, prm, 1, precision)
        ms.append(m)
    return ms


def _batch_state(ms, reps):
    I = (slice(1, -1),) * 3
    out = []
    for m, r in zip(ms, reps):
        P = m["Pc_ex"].get()[I]
        out.append({"T": m["Tc_ex"].get()[I], "P": P - P.mean(), "Nu": r["Nu"]})
    return out


def run_batch(n: int, steps: int, warmup: int, prm, graphs: bool = True,
              modes=("plain_fp64", "batched_fp64", "batched_dc")):
    """Many-problem workload: plain FP64 (each model stepped alone), then the
    batched FP64 and FP32+DC forms with graph-replayed iteration blocks (kernel
    by kernel when `graphs` is false), all on the same inputs and the same
    number of steps; `modes` picks a subset of the three."""
    import numpy as np
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import _lib, inputs

    ws = inputs.batch_workloads(n, seed=13000, steps=steps)
    cells = sum(w.ncell for w in ws)
    res = {}
    states = {}
    for mode in modes:
        precision = _lib.FP32_DC if mode == "batched_dc" else _lib.FP64
        ms = _batch_models(ws, prm, precision)
        b = None if mode == "plain_fp64" else ht.HydroThermal3DBatch(ms, prm, prm, precision=precision,
                                                                     graphs=graphs)
        reps = None
        for s in range(warmup + steps):
            if s == warmup:
                _lib.sync()
                if b is not None:
                    b.reset_stats()
                t0 = time.perf_counter()
            ts = time.perf_counter()
            reps = [m.step() for m in ms] if b is None else b.step()
            it = [(r["thermal"]["iters"], r["darcy"]["iters"]) for r in reps]
            log(f"batch {mode} step {s}: max iters T {max(i[0] for i in it)} P {max(i[1] for i in it)} "
                f"({time.perf_counter() - ts:.2f} s)")
        _lib.sync()
        wall = time.perf_counter() - t0
        res[mode] = {"s_per_step": round(wall / steps, 4)}
        if b is not None:
            st = b.stats()
            res[mode].update({"t_eff_solves_gbs": round(st["bytes_solve"] / (st["ms_solve"] / 1e3) / 1e9, 2)
                              if st["ms_solve"] > 0 else None,
                              "solve_share": round(st["ms_solve"] / 1e3 / wall, 3),
                              "graph_launches": st["graph_launches"], "graphs": st["graphs"],
                              "repacks": st["repacks"],
                              "tiles": [f"{r}x{z}" + (f" x{c}" if c > 1 else "")
                                        for r, z, c in zip(st["tile_rows"], st["tile_z"], st["tile_cpl"])],
                              "grid": ["per-item" if g else "persistent" for g in st["grid_mode"]],
                              "pairs": [f"{r}x{z}" + (f" x{c}" if c > 1 else "") if r else None
                                        for r, z, c in zip(st["pair_rows"], st["pair_z"], st["pair_cpl"])],
                              "vertex": list(st["vertex"])})
            b.close()
        states[mode] = _batch_state(ms, reps)
        for m in ms:
            m.close()

    def errs(a, c):
        return {"T": max(rel(x["T"], y["T"]) for x, y in zip(a, c)),
                "P": max(rel(x["P"], y["P"]) for x, y in zip(a, c)),
                "Nu": max(abs(x["Nu"] - y["Nu"]) / abs(y["Nu"]) for x, y in zip(a, c))}

    out = {"workload": f"{n} problems, edge n in {{32,48,64,96,128}}, n^3 or n x n/2 x n, Ra log-uniform "
                       f"[50,2000], sigma(log10 k) U[0,1], seeds 13000+p",
           "problems": n, "cells": cells, "steps": steps, "warmup": warmup, "graphs": graphs, **res}
    if "plain_fp64" in res:
        plain = res["plain_fp64"]["s_per_step"]
        for mode in ("batched_fp64", "batched_dc"):
            if mode in res:
                tag = "dc" if mode == "batched_dc" else "fp64"
                out[f"speedup_batched_{tag}_vs_plain" + ("_fp64" if tag == "dc" else "")] = \
                    round(plain / res[mode]["s_per_step"], 3)
                out[f"rel_err_{mode}_vs_plain" + ("_fp64" if tag == "dc" else "")] = \
                    errs(states[mode], states["plain_fp64"])
    if "batched_fp64" in res and "batched_dc" in res:
        out["speedup_batched_dc_vs_batched_fp64"] = round(res["batched_fp64"]["s_per_step"] /
                                                          res["batched_dc"]["s_per_step"], 3)
    return out


def calibration_copies(copies: int) -> None:
    """`copies` plain copies of a 256x128x256 cell array (ArrayEqualArray: N words
    read, N written), whose counter readings calibrate FETCH_SIZE/WRITE_SIZE."""
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import _lib
    a = ht.DeviceArray.zeros((256, 128, 256))
    b = ht.DeviceArray.zeros((256, 128, 256))
    for _ in range(copies):
        ht.kernels.ArrayEqualArray(a, b)
    _lib.sync()
    log(f"{copies} calibration copies of {8 * 256 * 128 * 256} bytes")


def probe(w, iters: int, copies: int) -> int:
    """Counter probe for rocprofv3 --pmc: `iters` fused Darcy pseudo-iterations
    of the workload with the autotuned tile (no check iterations after the first
    two), then `copies` plain copies of a cell array (ArrayEqualArray, N words
    read and N written) whose counter readings calibrate FETCH_SIZE/WRITE_SIZE."""
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import _lib, inputs
    g = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    inputs.permeability_device(w, g["kfv"])
    g["Tc_ex"].set(inputs.initial_temperature(w))
    k = ht.kernels
    k.InitDarcy(g["Ty"], g["kx"], g["ky"], g["kz"], g["Tc_ex"], g["kfv"], 1.0)
    k.KillNorthSouthFluxes(g["Ty"])
    k.Compute_Rp(g["Rp"], g["ky"], g["Ty"], w.Ra, 1.0 / w.h[1])
    ht.dyrel_solve(g["Pc_ex"], g["Rp"], g["kx"], g["ky"], g["kz"], dx=w.h[0], dy=w.h[1], dz=w.h[2], dt=1.0,
                   transient=0, rhoC=0.0, params=ht.DyrelParams(nitmax=iters, nitout=10**6, tol=0.0),
                   kv=g["kfv"] if w.mu_a == 0.0 else None)
    a = ht.DeviceArray.zeros((w.ncx, w.ncy, w.ncz))
    b = ht.DeviceArray.zeros((w.ncx, w.ncy, w.ncz))
    for _ in range(copies):
        k.ArrayEqualArray(a, b)
    _lib.sync()
    import ctypes as C
    ty, kz, xc, var = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _lib.call("ht3d_tune_last_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
    single = f"{tile_label((ty.value, kz.value, var.value, xc.value))} variant {var.value}"
    _lib.call("ht3d_tune_last_pair_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
    pair = f"{tile_label((ty.value, kz.value, var.value, xc.value))} variant {var.value}" if ty.value > 0 else "none"
    log(f"probe: {iters} pseudo-iterations on {w.name} (single tile {single}, pair tile {pair}), "
        f"{copies} copies of {8 * w.ncell} bytes")
    g.close()
    return 0


def cpu_baseline(w, iters: int, kfv=None):
    """The oracle's unfused DYREL Darcy solve at full size for a fixed number of
    pseudo-iterations (tol 0), timed on the host; reported as T_eff."""
    import numpy as np
    from oracle import oracle as orc
    from hydrothermal3d_amd import inputs

    nx, ny, nz = w.ncx, w.ncy, w.ncz
    if kfv is None:   # the same field the GPU run generated, else the oracle's generator
        kfv = np.ones((nx + 1, ny + 1, nz + 1), order="F")
        if w.sigma > 0 or w.fault:
            orc.lognormal_field(kfv, *inputs.field_args(w))
    T = inputs.initial_temperature(w)
    Ty = np.zeros((nx, ny + 1, nz), order="F")
    kx = np.zeros((nx + 1, ny, nz), order="F")
    ky = np.zeros((nx, ny + 1, nz), order="F")
    kz = np.zeros((nx, ny, nz + 1), order="F")
    Rp = np.zeros((nx, ny, nz), order="F")
    dx, dy, dz = w.h
    orc.InitDarcy(Ty, kx, ky, kz, T, kfv, 1.0)
    orc.KillNorthSouthFluxes(Ty)
    orc.Compute_Rp(Rp, ky, Ty, w.Ra, 1.0 / dy)
    P = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    p = orc.params(nitmax=iters, nitout=10**9, ncheck=0, norm=0, tol=0.0)
    t0 = time.perf_counter()
    orc.dyrel_solve(P, Rp, kx, ky, kz, dx=dx, dy=dy, dz=dz, dt=1.0, transient=0, rhoC=0.0, p=p)
    dt = time.perf_counter() - t0
    nbytes = 8.0 * w.ncell * 8 * iters
    return nbytes / dt / 1e9, dt


def run_config0(steps: int, tol: float):
    """The reference's CPU-runnable configuration in full: 32^3 cells, Ra=200,
    k=1, `steps` physical steps, max-norm tolerance `tol`, on the CPU oracle
    (unfused reference kernel sequence, OpenMP) and on the GPU, with the GPU
    result checked against the CPU one."""
    import numpy as np
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import _lib, inputs
    from oracle import oracle as orc
    w = inputs.WORKLOADS["hrl32"]
    T0 = inputs.initial_temperature(w)
    prm = ht.production_params(tol=tol)
    g = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    g["Tc_ex"].set(T0)
    g.set_params(prm, prm, 1, _lib.FP64)
    _lib.sync()
    t0 = time.perf_counter()
    rg = [g.step() for _ in range(steps)]
    _lib.sync()
    gpu_s = time.perf_counter() - t0
    o = orc.HRLModel(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    o.field("Tc_ex")[...] = T0
    po = orc.params(prm.nitmax, prm.nitout, prm.ncheck, prm.norm, prm.tol)
    o.set_params(po, po, 1)
    t0 = time.perf_counter()
    ro = [o.step() for _ in range(steps)]
    cpu_s = time.perf_counter() - t0
    I = (slice(1, -1),) * 3
    Tg, To = g["Tc_ex"].get()[I], o.field("Tc_ex")[I]
    out = {"workload": w.name, "steps": steps, "tol": tol,
           "cpu_s_per_step": round(cpu_s / steps, 4), "gpu_s_per_step": round(gpu_s / steps, 5),
           "gpu_vs_cpu_speedup": round(cpu_s / gpu_s, 1),
           "iters_equal": all(a["thermal"]["iters"] == c["thermal"]["iters"] and
                              a["darcy"]["iters"] == c["darcy"]["iters"] for a, c in zip(rg, ro)),
           "rel_err_T_gpu_vs_cpu": rel(Tg, To),
           "rel_err_Nu_gpu_vs_cpu": max(abs(a["Nu"] - c["Nu"]) / abs(c["Nu"]) for a, c in zip(rg, ro)),
           "Nu_last": rg[-1]["Nu"]}
    g.close()
    o.close()
    return out


def run_v18(iters: int, unfused_iters: int):
    """The dimensional (v18) DYREL loops (HydroThermal3D_v18.jl:432-478, :489-537)
    on a 256x128x256 synthetic v18 geometry after the v18 property set-up:
    the fused solves (one launch per pseudo-iteration; kernel time from HIP
    events, algorithmic bytes 73 B/cell heat, 89 B/cell Darcy) against the
    reference's unfused launch sequence made of the library's per-kernel
    entries on the same GPU (4 launches per heat iteration, 5 per Darcy one;
    launched asynchronously, wall time over `unfused_iters` iterations)."""
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import _lib, kernels as K
    from hydrothermal3d_amd.device import DeviceArray
    from hydrothermal3d_amd.v18 import V18Model, v18_geometry
    nx, ny, nz = 256, 128, 256
    N = nx * ny * nz
    geo = v18_geometry(nx, ny, nz, seed=18)
    prm = ht.DyrelParams(nitmax=iters, nitout=100, tol=0.0)
    m = V18Model(nx, ny, nz, geo, prm, prm)
    a, sc, ph = m.a, m.sc, m.phys
    K.Cpy_inn_to_all(a["Tc0"], a["Tc_ex"])
    K.Cpy_inn_to_all(a["Pc0"], a["Pc_ex"])
    out = {"grid": f"{nx}x{ny}x{nz}", "iters": iters, "unfused_iters": unfused_iters}
    z = lambda s: DeviceArray.zeros(s)
    F, F0 = z((nx, ny, nz)), z((nx, ny, nz))
    qx, qy, qz = z((nx + 1, ny, nz)), z((nx, ny + 1, nz)), z((nx, ny, nz + 1))
    _d = (1.0 / m.dx, 1.0 / m.dy, 1.0 / m.dz)
    for kind in ("heat", "darcy"):
        m.thermal(1.0) if kind == "heat" else None            # untimed: tile choice, state
        _lib.call("ht3d_timing_enable", 1)
        _lib.call("ht3d_timing_reset")
        rep = m.thermal(1.0) if kind == "heat" else m.darcy(1.0)
        _lib.sync()
        st = timing_stats()
        _lib.call("ht3d_timing_enable", 0)
        us = st["ms"] / max(st["launches"], 1) * 1e3
        teff = st["bytes"] / (st["ms"] / 1e3) / 1e9 if st["ms"] > 0 else 0.0
        h1, h2, dtau = rep["h1"], rep["h2"], rep["dtau"]
        _lib.call("ht3d_set_sync", 0)
        _lib.sync()
        t0 = time.perf_counter()
        for _ in range(unfused_iters):
            if kind == "heat":
                K.dim_SetTemperatureBCs(a["Tc_ex"], a["ky"], a["wN"], a["phc_ex"], ph.qyS / (sc.W / sc.L ** 2), _d[1],
                                        0.0, 0.0, ph.Ttop / sc.T, *([0.0] * 10))
                K.ComputeFlux(qx, qy, qz, a["kx"], a["ky"], a["kz"], a["Tc_ex"], *_d)
                K.ResidualTemperatureLinearised(F, a["Tc_ex"], a["Tc0"], a["rhoC_phirho"], a["phc_ex"], a["PC"], qx,
                                                qy, qz, 0.0, 1.0, 1.0 / m.dt, *_d)
                K.DYRELUpdate(F0, a["Tc_ex"], F, h1, h2, dtau)
            else:
                K.PhiRhof(a["rhoC_phirho"], a["Tc_ex"], a["Pc_ex"], a["phc_ex"], ph.phi_i, sc.rho, sc.T, sc.sigma)
                K.dim_SetPressureBCs(a["Pc_ex"], a["wN"], a["phc_ex"], 0.0, ph.Ptop / sc.sigma, *([0.0] * 8))
                K.ComputeDarcyFlux(qx, qy, qz, a["k_rhof"], a["kx"], a["ky"], a["kz"], a["Pc_ex"],
                                   ph.g / (sc.L / sc.t ** 2), *_d)
                K.ResidualFluidPressure(F, a["phc_ex"], a["rhoC_phirho"], a["Xc0"], a["PC"], qx, qy, qz, 1.0,
                                        1.0 / m.dt, *_d)
                K.DYRELUpdate(F0, a["Pc_ex"], F, h1, h2, dtau)
        _lib.sync()
        un_us = (time.perf_counter() - t0) / unfused_iters * 1e6
        _lib.call("ht3d_set_sync", 1)
        import ctypes as C
        tile = C.c_int()
        _lib.call("ht3d_v18_tune", -1, C.byref(tile))
        out[kind] = {"fused_us_per_iter": round(us, 2), "t_eff_GBs": round(teff, 1),
                     "hbm_frac": round(teff / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_iter": int(st["bytes"] / max(st["launches"], 1)),
                     "tile": tile.value, "iters": rep["iters"], "status": rep["status"],
                     "unfused_us_per_iter": round(un_us, 2), "fusion_speedup": round(un_us / us, 2)}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hrl256")
    ap.add_argument("--precision", choices=["fp64", "dc", "both"], default="both",
                    help="fp64: plain FP64 path; dc: FP32 inner sweeps + FP64 defect correction; both: run the two "
                         "on the same inputs and report the DC speed-up and its error against FP64")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--check", default="50", help="check interval: an integer, or 'auto' for the check-interval tuner")
    ap.add_argument("--ref-tol", type=float, default=0.0,
                    help="also run FP64 at this tighter tolerance on the same inputs and steps, and report the T, P "
                         "and Nu errors of both forms against it (0: skip)")
    ap.add_argument("--ref-schedule-steps", type=int, default=2,
                    help="also time this many FP64 steps (after one warm-up) on the reference's solver schedule: "
                         "lambda_min refresh and mean-norm convergence test at iter<=2 or every 1000 iterations, "
                         "nitmax 10000 (HydroThermal3D_v18.jl:309-310,450-476); 0: skip")
    ap.add_argument("--cpu-iters", type=int, default=3000,
                    help="pseudo-iterations of the CPU-baseline sample (about 10 s on 16 host threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config0-steps", type=int, default=20,
                    help="physical steps of the 32^3 configuration run in full on CPU and GPU (0: skip)")
    ap.add_argument("--batch", type=int, default=96, help="problems of the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="timed physical steps of the batched leg (the 96-problem workload is defined over 20)")
    ap.add_argument("--batch-tol", type=float, default=1e-8,
                    help="DYREL tolerance of the batched leg (the configurations' 1e-8 by default)")
    ap.add_argument("--batch-warmup", type=int, default=1)
    ap.add_argument("--batch-graphs", type=int, choices=[0, 1], default=1,
                    help="replay the batch's pseudo-iteration blocks as HIP graphs (1) or launch them kernel by "
                         "kernel (0; the form to run under a profiler)")
    ap.add_argument("--batch-pairs", type=int, choices=[-1, 0, 1], default=-1,
                    help="two pseudo-iterations per launch in the batch: -1 when timed faster, 0 never, 1 always")
    ap.add_argument("--batch-modes", default="plain_fp64,batched_fp64,batched_dc",
                    help="comma list of the batched leg's forms: plain_fp64, batched_fp64, batched_dc")
    ap.add_argument("--forward-only", action="store_true",
                    help="sweep every launch forwards (default: alternate the direction between launches)")
    ap.add_argument("--v18-only", action="store_true", help="run only the v18 leg and print its JSON")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched leg and print its JSON")
    ap.add_argument("--v18-iters", type=int, default=400,
                    help="pseudo-iterations per fused v18 solve in the v18 leg (0: skip it)")
    ap.add_argument("--calib-copies", type=int, default=0,
                    help="after the run, copy a 256x128x256 cell array this many times (ArrayEqualArray: 64 MiB read "
                         "and 64 MiB written each), the yardstick of rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    ap.add_argument("--probe", action="store_true",
                    help="counter probe only (fused iterations + calibration copies), no JSON line")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU halo exchange is a design "
              "note, not built.", file=sys.stderr)
        return 2

    _ensure_built()
    import hydrothermal3d_amd as ht
    from hydrothermal3d_amd import _lib, inputs

    w = inputs.WORKLOADS[args.workload]
    _lib.call("ht3d_batch_tune_pairs", args.batch_pairs)
    if args.forward_only:
        _lib.call("ht3d_tune_sweep", 0)
    if args.probe:
        return probe(w, 20, 5)
    if args.v18_only:
        print(json.dumps({"v18": run_v18(max(args.v18_iters, 1), 100)}), flush=True)
        return 0
    ncheck = -1 if args.check == "auto" else int(args.check)
    prm = ht.production_params(tol=args.tol, ncheck=ncheck)
    if args.batch_only:
        log(f"batched leg: {max(args.batch, 1)} problems")
        print(json.dumps({"batch": run_batch(max(args.batch, 1), args.batch_steps, args.batch_warmup,
                                             ht.production_params(tol=args.batch_tol, ncheck=ncheck),
                                             bool(args.batch_graphs), tuple(args.batch_modes.split(",")))}),
              flush=True)
        if args.calib_copies > 0:
            calibration_copies(args.calib_copies)
        return 0
    precision = _lib.FP32_DC if args.precision == "dc" else _lib.FP64
    log(f"device: {_lib.device_name()}; workload {w.name}; precision {args.precision}")
    wall, st, reports, kfv, fields = run_gpu(w, args.steps, args.warmup, precision, prm)
    ms_step = wall / args.steps * 1e3
    dc = None
    if args.precision == "both":
        log("FP32 inner sweeps + FP64 defect correction on the same inputs")
        wall_dc, st_dc, rep_dc, _, f_dc = run_gpu(w, args.steps, args.warmup, _lib.FP32_DC, prm)
        ms_dc = wall_dc / args.steps * 1e3
        import numpy as np
        Pg, Pd = fields["P"] - fields["P"].mean(), f_dc["P"] - f_dc["P"].mean()
        nu = np.array([r["Nu"] for r in reports])
        nu_dc = np.array([r["Nu"] for r in rep_dc])
        dc = {"ms_per_step": round(ms_dc, 3), "speedup_vs_fp64": round(ms_step / ms_dc, 3),
              "rel_err_T": rel(f_dc["T"], fields["T"]), "rel_err_P": rel(Pd, Pg),
              "rel_err_Nu": float(np.max(np.abs(nu_dc - nu) / np.abs(nu))),
              "t_eff_fp32_inner": round(st_dc["bytes_fp32"] / (st_dc["ms_fp32"] / 1e3) / 1e9, 2)
              if st_dc["ms_fp32"] > 0 else None,
              "inner_iters_per_step": [(r["inner_thermal"], r["inner_darcy"]) for r in rep_dc],
              "outer_iters_per_step": [(r["dc_outer_thermal"], r["dc_outer_darcy"]) for r in rep_dc]}
    ref = None
    if args.ref_tol > 0:
        import numpy as np
        log(f"FP64 reference at tol {args.ref_tol:g} on the same inputs")
        _, _, rep_ref, _, f_ref = run_gpu(w, args.steps, args.warmup, _lib.FP64,
                                          ht.production_params(tol=args.ref_tol, ncheck=ncheck))
        Pr = f_ref["P"] - f_ref["P"].mean()
        nu_ref = np.array([r["Nu"] for r in rep_ref])

        def vs_ref(f, reps):
            nu = np.array([r["Nu"] for r in reps])
            return {"T": rel(f["T"], f_ref["T"]), "P": rel(f["P"] - f["P"].mean(), Pr),
                    "Nu": float(np.max(np.abs(nu - nu_ref) / np.abs(nu_ref)))}
        ref = {"tol": args.ref_tol, "iters_per_step": [(r["thermal"]["iters"], r["darcy"]["iters"]) for r in rep_ref],
               "rel_err_fp64": vs_ref(fields, reports)}
        if args.precision == "both":
            ref["rel_err_dc"] = vs_ref(f_dc, rep_dc)
    refsched = None
    if args.ref_schedule_steps > 0:
        log("FP64 on the reference's solver schedule (checks and lambda refresh every 1000 iterations)")
        wall_r, st_r, rep_r, _, _ = run_gpu(w, args.ref_schedule_steps, 1, _lib.FP64, ht.reference_params(tol=args.tol))
        refsched = {"ms_per_step": round(wall_r / args.ref_schedule_steps * 1e3, 3),
                    "iters_per_step": [(r["thermal"]["iters"], r["darcy"]["iters"]) for r in rep_r],
                    "status_per_step": [(r["thermal"]["status"], r["darcy"]["status"]) for r in rep_r],
                    "t_eff": round(st_r["bytes"] / (st_r["ms"] / 1e3) / 1e9, 2) if st_r["ms"] > 0 else None,
                    "schedule": "nitout 1000, convergence test at iter<=2 or iter%1000==0, mean|F*D|/sqrt(N), "
                                "nitmax 10000"}
    teff = st["bytes"] / (st["ms"] / 1e3) / 1e9 if st["ms"] > 0 else 0.0
    # words per cell and iteration that the kernel which ran must move (variant
    # bit 8: face coefficients formed from the vertex field, bit 1: stored
    # diagonal read; pairs read the faces, once per two iterations)
    kwords = (7 if st["tile"][2] & 1 else 6) if st["tile"][2] & 8 else (9 if st["tile"][2] & 1 else 8)
    if st["pair"] is not None:
        kwords = 4
    teff32 = st["bytes_fp32"] / (st["ms_fp32"] / 1e3) / 1e9 if st["ms_fp32"] > 0 else None
    avg_us = st["ms"] / max(st["launches"], 1) * 1e3
    iters = [(r["thermal"]["iters"], r["darcy"]["iters"]) for r in reports]

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU baseline: oracle Darcy DYREL, {args.cpu_iters} iterations at full size, {threads} threads")
        cpu_teff, cpu_s = cpu_baseline(w, args.cpu_iters, kfv)
        # per physical step: the sample's time per pseudo-iteration times the
        # pseudo-iterations the timed GPU steps took (same algorithm, same counts)
        mean_iters = sum(a + b for a, b in iters) / max(len(iters), 1)
        cpu = {"value": round(cpu_teff, 3), "unit": "GB/s", "cores": threads, "kind": "port", "cpu_model": _cpu_model(),
               "sample": f"{w.name}: CPU oracle (unfused reference kernel sequence, FP64) Darcy DYREL solve, "
                         f"{args.cpu_iters} pseudo-iterations at full size incl. one Gershgorin set-up pass, "
                         f"{cpu_s:.2f} s",
               "est_s_per_step": round(cpu_s / args.cpu_iters * mean_iters, 3),
               "est_s_per_step_basis": f"sample time per pseudo-iteration x {mean_iters:.0f} pseudo-iterations "
                                       f"(thermal + Darcy, mean of the timed GPU steps); advection not included",
               "gpu_fp64_speedup_est": round(cpu_s / args.cpu_iters * mean_iters / (ms_step / 1e3), 1)}

    out = {
        "metric": METRIC,
        "value": round(teff, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != "dc" else "f32+f64dc",
        "data": "synthetic",
        "config": {"workload": w.name, "ncx": w.ncx, "ncy": w.ncy, "ncz": w.ncz, "Ra": w.Ra, "mu_a": w.mu_a,
                   "precision": args.precision, "tol": args.tol, "norm": "max|F*D|",
                   "check_interval": prm.ncheck, "lambda_interval": prm.nitout,
                   "sweep": "forward" if args.forward_only else "alternating"},
        "s_per_step": round(ms_step / 1e3, 4),
        "hbm_frac": round(teff / HBM_PEAK_GBS, 4),
        "iters_per_step": iters,
        "iteration_kernel_avg_us": round(avg_us, 2),
        "roofline": {"bound": "hbm", "achieved": round(teff, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(teff / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.workload, st["tile"]),
                     "algorithmic_bytes": 8 * 8 * w.ncell,
                     "kernel_words": kwords,
                     "kernel_frac": round(teff * kwords / 8 / HBM_PEAK_GBS, 4),
                     "tile": f"{tile_label(st['tile'])} variant {st['tile'][2]}",
                     "pair_tile": None if st["pair"] is None else
                     f"{tile_label(st['pair'])} variant {st['pair'][2]}"},
        "cpu_baseline": cpu,
        "device": _lib.device_name(),
    }
    if not args.no_cpu and args.config0_steps > 0:
        log(f"32^3 configuration in full on the CPU oracle and the GPU ({args.config0_steps} steps)")
        c0 = run_config0(args.config0_steps, args.tol)
        c0["cores"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["config_32cube"] = c0
    if args.batch > 0:
        log(f"batched leg: {args.batch} problems")
        out["batch"] = run_batch(args.batch, args.batch_steps, args.batch_warmup,
                                 ht.production_params(tol=args.batch_tol, ncheck=ncheck), bool(args.batch_graphs),
                                 tuple(args.batch_modes.split(",")))
        out["batch"]["tol"] = args.batch_tol
    if args.v18_iters > 0:
        log("v18 leg: fused dimensional DYREL loops vs the unfused launch sequence")
        out["v18"] = run_v18(args.v18_iters, 100)
    if teff32 is not None:
        out["t_eff_fp32_inner"] = round(teff32, 2)
    if dc is not None:
        out["fp32_dc"] = dc
    if ref is not None:
        out["vs_converged_fp64"] = ref
    if refsched is not None:
        out["reference_schedule_fp64"] = refsched
    print(json.dumps(out), flush=True)
    if args.calib_copies > 0:
        calibration_copies(args.calib_copies)
    return 0


if __name__ == "__main__":
    sys.exit(main())
