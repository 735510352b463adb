# HT3D_ccalls.jl -- generated from include/ht3d.h by gen_ccalls.py; do not edit.
# One method per C entry point, each a ccall with the entry's literal signature.
# Device-array parameters go through devptr (device arrays or raw pointers only).

ht3d_init(device) =
    ccall((:ht3d_init, libht3d), Cint, (Cint,), device)
ht3d_finalize() =
    ccall((:ht3d_finalize, libht3d), Cint, ())
ht3d_strerror(code) =
    ccall((:ht3d_strerror, libht3d), Cstring, (Cint,), code)
ht3d_last_error() =
    ccall((:ht3d_last_error, libht3d), Cstring, ())
ht3d_set_sync(synchronous) =
    ccall((:ht3d_set_sync, libht3d), Cint, (Cint,), synchronous)
ht3d_sync() =
    ccall((:ht3d_sync, libht3d), Cint, ())
ht3d_device_name(buf, len) =
    GC.@preserve buf ccall((:ht3d_device_name, libht3d), Cint, (Ptr{UInt8}, Csize_t), buf, len)
ht3d_debug_inject_host_failure(count) =
    ccall((:ht3d_debug_inject_host_failure, libht3d), Cint, (Cint,), count)
ht3d_alloc(bytes, dptr) =
    GC.@preserve dptr ccall((:ht3d_alloc, libht3d), Cint, (Csize_t, Ptr{Cvoid}), bytes, dptr)
ht3d_free(dptr) =
    GC.@preserve dptr ccall((:ht3d_free, libht3d), Cint, (Ptr{Cvoid},), dptr)
ht3d_h2d(dst, src, bytes) =
    GC.@preserve dst src ccall((:ht3d_h2d, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), dst, src, bytes)
ht3d_d2h(dst, src, bytes) =
    GC.@preserve dst src ccall((:ht3d_d2h, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), dst, src, bytes)
ht3d_d2d(dst, src, bytes) =
    GC.@preserve dst src ccall((:ht3d_d2d, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), dst, src, bytes)
ht3d_memset(dptr, value, bytes) =
    GC.@preserve dptr ccall((:ht3d_memset, libht3d), Cint, (Ptr{Cvoid}, Cint, Csize_t), dptr, value, bytes)
ht3d_timing_enable(on) =
    ccall((:ht3d_timing_enable, libht3d), Cint, (Cint,), on)
ht3d_timing_get(out) =
    GC.@preserve out ccall((:ht3d_timing_get, libht3d), Cint, (Ptr{Cvoid},), out)
ht3d_timing_reset() =
    ccall((:ht3d_timing_reset, libht3d), Cint, ())
ht3d_tune_set(autotune, force_tile, force_variant) =
    ccall((:ht3d_tune_set, libht3d), Cint, (Cint, Cint, Cint), autotune, force_tile, force_variant)
ht3d_tune_clear() =
    ccall((:ht3d_tune_clear, libht3d), Cint, ())
ht3d_tune_last(rows_per_block, z_per_thread, variant) =
    GC.@preserve rows_per_block z_per_thread variant ccall((:ht3d_tune_last, libht3d), Cint, (Ptr{Cint}, Ptr{Cint}, Ptr{Cint}), rows_per_block, z_per_thread, variant)
ht3d_tune_last_tile(rows_per_block, z_per_thread, x_per_thread, variant) =
    GC.@preserve rows_per_block z_per_thread x_per_thread variant ccall((:ht3d_tune_last_tile, libht3d), Cint, (Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}), rows_per_block, z_per_thread, x_per_thread, variant)
ht3d_tile_count() =
    ccall((:ht3d_tile_count, libht3d), Cint, ())
ht3d_tune_pairs(mode) =
    ccall((:ht3d_tune_pairs, libht3d), Cint, (Cint,), mode)
ht3d_tune_check_interval(iters, t_iter_us, t_check_us, nitout) =
    ccall((:ht3d_tune_check_interval, libht3d), Cint, (Cdouble, Cdouble, Cdouble, Cint), iters, t_iter_us, t_check_us, nitout)
ht3d_tune_last_pair(rows_per_block, z_per_block, variant) =
    GC.@preserve rows_per_block z_per_block variant ccall((:ht3d_tune_last_pair, libht3d), Cint, (Ptr{Cint}, Ptr{Cint}, Ptr{Cint}), rows_per_block, z_per_block, variant)
ht3d_tune_last_pair_tile(rows_per_block, z_per_block, x_per_thread, variant) =
    GC.@preserve rows_per_block z_per_block x_per_thread variant ccall((:ht3d_tune_last_pair_tile, libht3d), Cint, (Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}), rows_per_block, z_per_block, x_per_thread, variant)
ht3d_tune_force_pair(tile, variant) =
    ccall((:ht3d_tune_force_pair, libht3d), Cint, (Cint, Cint), tile, variant)
ht3d_pair_tile_count() =
    ccall((:ht3d_pair_tile_count, libht3d), Cint, ())
ht3d_tune_sweep(alternate) =
    ccall((:ht3d_tune_sweep, libht3d), Cint, (Cint,), alternate)
ht3d_reduce(out_host, a, n, op) =
    GC.@preserve out_host a ccall((:ht3d_reduce, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Csize_t, Cint), out_host, devptr(Cdouble, a), n, op)
ht3d_ResetA(A, B, n1, n2, n3) =
    GC.@preserve A B ccall((:ht3d_ResetA, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, A), devptr(Cdouble, B), n1, n2, n3)
ht3d_Cpy_inn_to_all(A, B, ncx, ncy, ncz) =
    GC.@preserve A B ccall((:ht3d_Cpy_inn_to_all, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, A), devptr(Cdouble, B), ncx, ncy, ncz)
ht3d_Cpy_all_to_inn(A, B, ncx, ncy, ncz) =
    GC.@preserve A B ccall((:ht3d_Cpy_all_to_inn, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, A), devptr(Cdouble, B), ncx, ncy, ncz)
ht3d_ArrayEqualArray(F1, F2, n1, n2, n3) =
    GC.@preserve F1 F2 ccall((:ht3d_ArrayEqualArray, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, F1), devptr(Cdouble, F2), n1, n2, n3)
ht3d_Multiply(a, b, c, ncx, ncy, ncz) =
    GC.@preserve a b c ccall((:ht3d_Multiply, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, a), devptr(Cdouble, b), devptr(Cdouble, c), ncx, ncy, ncz)
ht3d_InitThermal(Rt, kx, ky, kz, Tc_ex, ktv, _dt, ncx, ncy, ncz) =
    GC.@preserve Rt kx ky kz Tc_ex ktv ccall((:ht3d_InitThermal, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Rt), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, Tc_ex), devptr(Cdouble, ktv), _dt, ncx, ncy, ncz)
ht3d_InitConductivity(kx, ky, kz, ktv, ncx, ncy, ncz) =
    GC.@preserve kx ky kz ktv ccall((:ht3d_InitConductivity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, ktv), ncx, ncy, ncz)
ht3d_InitDarcy(Ty, kx, ky, kz, Tc_ex, kfv, _dt, ncx, ncy, ncz) =
    GC.@preserve Ty kx ky kz Tc_ex kfv ccall((:ht3d_InitDarcy, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Ty), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, Tc_ex), devptr(Cdouble, kfv), _dt, ncx, ncy, ncz)
ht3d_KillNorthSouthFluxes(Ty, ncx, ncy, ncz) =
    GC.@preserve Ty ccall((:ht3d_KillNorthSouthFluxes, libht3d), Cint, (Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, Ty), ncx, ncy, ncz)
ht3d_Compute_Rp(Rp, ky, Ty, Ra, _dy, ncx, ncy, ncz) =
    GC.@preserve Rp ky Ty ccall((:ht3d_Compute_Rp, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Rp), devptr(Cdouble, ky), devptr(Cdouble, Ty), Ra, _dy, ncx, ncy, ncz)
ht3d_SetTemperatureBCs(Tc_ex, Tbot, Ttop, ncx, ncy, ncz) =
    GC.@preserve Tc_ex ccall((:ht3d_SetTemperatureBCs, libht3d), Cint, (Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Tc_ex), Tbot, Ttop, ncx, ncy, ncz)
ht3d_SetPressureBCs(Pc_ex, Pbot, Ptop, ncx, ncy, ncz) =
    GC.@preserve Pc_ex ccall((:ht3d_SetPressureBCs, libht3d), Cint, (Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Pc_ex), Pbot, Ptop, ncx, ncy, ncz)
ht3d_SetPressureBCs_v7(Pc_ex, Pbot, Ptop, ncx, ncy, ncz) =
    GC.@preserve Pc_ex ccall((:ht3d_SetPressureBCs_v7, libht3d), Cint, (Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Pc_ex), Pbot, Ptop, ncx, ncy, ncz)
ht3d_ComputeFlux(qx, qy, qz, kx, ky, kz, A, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve qx qy qz kx ky kz A ccall((:ht3d_ComputeFlux, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, A), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_UpdateT(F, T, R, qx, qy, qz, _dt, dtau, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F T R qx qy qz ccall((:ht3d_UpdateT, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F), devptr(Cdouble, T), devptr(Cdouble, R), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), _dt, dtau, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusion(F, R, qx, qy, qz, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F R qx qy qz ccall((:ht3d_ResidualDiffusion, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F), devptr(Cdouble, R), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusionTransient(F, T, R, qx, qy, qz, _dt, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F T R qx qy qz ccall((:ht3d_ResidualDiffusionTransient, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F), devptr(Cdouble, T), devptr(Cdouble, R), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), _dt, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusionPC(F, R, qx, qy, qz, PC, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F R qx qy qz PC ccall((:ht3d_ResidualDiffusionPC, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F), devptr(Cdouble, R), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), devptr(Cdouble, PC), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusionTransientPC(F, T, R, qx, qy, qz, PC, _dt, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F T R qx qy qz PC ccall((:ht3d_ResidualDiffusionTransientPC, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F), devptr(Cdouble, T), devptr(Cdouble, R), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), devptr(Cdouble, PC), _dt, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_UpdateP(F0, P, F, dampx, dtau, ncx, ncy, ncz) =
    GC.@preserve F0 P F ccall((:ht3d_UpdateP, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F0), devptr(Cdouble, P), devptr(Cdouble, F), dampx, dtau, ncx, ncy, ncz)
ht3d_DampedUpdate(F0, X, F, dampx, dtau, ncx, ncy, ncz) =
    GC.@preserve F0 X F ccall((:ht3d_DampedUpdate, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F0), devptr(Cdouble, X), devptr(Cdouble, F), dampx, dtau, ncx, ncy, ncz)
ht3d_DYRELUpdate(F0, X, F, h1, h2, dtau, ncx, ncy, ncz) =
    GC.@preserve F0 X F ccall((:ht3d_DYRELUpdate, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F0), devptr(Cdouble, X), devptr(Cdouble, F), h1, h2, dtau, ncx, ncy, ncz)
ht3d_SwapDYREL(Xc, Yc, X0c_ex, Y0c, ncx, ncy, ncz) =
    GC.@preserve Xc Yc X0c_ex Y0c ccall((:ht3d_SwapDYREL, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, Xc), devptr(Cdouble, Yc), devptr(Cdouble, X0c_ex), devptr(Cdouble, Y0c), ncx, ncy, ncz)
ht3d_GershgorinPoisson(G, D, rhoC, kx, ky, kz, transient, dt, dx, dy, dz, ncx, ncy, ncz) =
    GC.@preserve G D rhoC kx ky kz ccall((:ht3d_GershgorinPoisson, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, G), devptr(Cdouble, D), devptr(Cdouble, rhoC), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), transient, dt, dx, dy, dz, ncx, ncy, ncz)
ht3d_GershgorinHydro(G, phidrhodP, kx, ky, kz, transient, dt, dx, dy, dz, ncx, ncy, ncz) =
    GC.@preserve G phidrhodP kx ky kz ccall((:ht3d_GershgorinHydro, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, G), devptr(Cdouble, phidrhodP), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), transient, dt, dx, dy, dz, ncx, ncy, ncz)
ht3d_RayleighQuotientNumerator(G, Fc0, Fcit, Fc, dtau, ncx, ncy, ncz) =
    GC.@preserve G Fc0 Fcit Fc ccall((:ht3d_RayleighQuotientNumerator, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, G), devptr(Cdouble, Fc0), devptr(Cdouble, Fcit), devptr(Cdouble, Fc), dtau, ncx, ncy, ncz)
ht3d_RayleighQuotientDenominator(G, Fc0, dtau, ncx, ncy, ncz) =
    GC.@preserve G Fc0 ccall((:ht3d_RayleighQuotientDenominator, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, G), devptr(Cdouble, Fc0), dtau, ncx, ncy, ncz)
ht3d_ResidualTemperatureLinearised(F, Tc_ex, Tc_old, rhoC_eff, phc_ex, PC, qx, qy, qz, Qt, transient, _dt, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F Tc_ex Tc_old rhoC_eff phc_ex PC qx qy qz ccall((:ht3d_ResidualTemperatureLinearised, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F), devptr(Cdouble, Tc_ex), devptr(Cdouble, Tc_old), devptr(Cdouble, rhoC_eff), devptr(Cdouble, phc_ex), devptr(Cdouble, PC), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), Qt, transient, _dt, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualFluidPressure(F, phc_ex, phirhof, phirhof0, PC, qx, qy, qz, transient, _dt, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F phc_ex phirhof phirhof0 PC qx qy qz ccall((:ht3d_ResidualFluidPressure, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, F), devptr(Cdouble, phc_ex), devptr(Cdouble, phirhof), devptr(Cdouble, phirhof0), devptr(Cdouble, PC), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), transient, _dt, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ComputeDarcyFlux(qx, qy, qz, rhofv, kx, ky, kz, Pf, g, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve qx qy qz rhofv kx ky kz Pf ccall((:ht3d_ComputeDarcyFlux, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), devptr(Cdouble, rhofv), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, Pf), g, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Init_vel(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve Vx Vy Vz kx ky kz Pc_ex Ty ccall((:ht3d_Init_vel, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Vx), devptr(Cdouble, Vy), devptr(Cdouble, Vz), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, Pc_ex), devptr(Cdouble, Ty), Ra, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ComputeThermalConductivity(ktv, Tce, phv, phi0, scale_kt, scale_T, ncx, ncy, ncz) =
    GC.@preserve ktv Tce phv ccall((:ht3d_ComputeThermalConductivity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, ktv), devptr(Cdouble, Tce), devptr(Cdouble, phv), phi0, scale_kt, scale_T, ncx, ncy, ncz)
ht3d_ComputeHydroConductivity(kfv, Tce, Pce, phv, ymin, dy, k_fact, delta, scale_sigma, scale_T, scale_L, scale_t, scale_eta, mode, ncx, ncy, ncz) =
    GC.@preserve kfv Tce Pce phv ccall((:ht3d_ComputeHydroConductivity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint, Cint), devptr(Cdouble, kfv), devptr(Cdouble, Tce), devptr(Cdouble, Pce), devptr(Cdouble, phv), ymin, dy, k_fact, delta, scale_sigma, scale_T, scale_L, scale_t, scale_eta, mode, ncx, ncy, ncz)
ht3d_ComputeFluidDensity(rhofv, Tce, Pce, phv, scale_sigma, scale_T, scale_rho, ncx, ncy, ncz) =
    GC.@preserve rhofv Tce Pce phv ccall((:ht3d_ComputeFluidDensity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, rhofv), devptr(Cdouble, Tce), devptr(Cdouble, Pce), devptr(Cdouble, phv), scale_sigma, scale_T, scale_rho, ncx, ncy, ncz)
ht3d_SmoothConductivityV2C(ktc, ktv, ncx, ncy, ncz) =
    GC.@preserve ktc ktv ccall((:ht3d_SmoothConductivityV2C, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, ktc), devptr(Cdouble, ktv), ncx, ncy, ncz)
ht3d_SmoothConductivityC2V(ktv, ktc, ncx, ncy, ncz) =
    GC.@preserve ktv ktc ccall((:ht3d_SmoothConductivityC2V, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, ktv), devptr(Cdouble, ktc), ncx, ncy, ncz)
ht3d_ComputeRhoCeffective(rhoC_eff, Tc_ex, Pc_ex, phc_ex, phi_i, sca_T, sca_sigma, sca_C, sca_rho, ncx, ncy, ncz) =
    GC.@preserve rhoC_eff Tc_ex Pc_ex phc_ex ccall((:ht3d_ComputeRhoCeffective, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, rhoC_eff), devptr(Cdouble, Tc_ex), devptr(Cdouble, Pc_ex), devptr(Cdouble, phc_ex), phi_i, sca_T, sca_sigma, sca_C, sca_rho, ncx, ncy, ncz)
ht3d_PhiRhof(phirhofc, Tc_ex, Pc_ex, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, ncx, ncy, ncz) =
    GC.@preserve phirhofc Tc_ex Pc_ex phc_ex ccall((:ht3d_PhiRhof, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, phirhofc), devptr(Cdouble, Tc_ex), devptr(Cdouble, Pc_ex), devptr(Cdouble, phc_ex), phi_i, scale_rho, scale_T, scale_sigma, ncx, ncy, ncz)
ht3d_PhiRhof_old(phirhofc, Tc0, Pc0, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, ncx, ncy, ncz) =
    GC.@preserve phirhofc Tc0 Pc0 phc_ex ccall((:ht3d_PhiRhof_old, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, phirhofc), devptr(Cdouble, Tc0), devptr(Cdouble, Pc0), devptr(Cdouble, phc_ex), phi_i, scale_rho, scale_T, scale_sigma, ncx, ncy, ncz)
ht3d_PhiDrhoDP(phidrhodP, Tc_ex, Pc_ex, phc_ex, phi_i, scale_rho, scale_T, scale_sigma, ncx, ncy, ncz) =
    GC.@preserve phidrhodP Tc_ex Pc_ex phc_ex ccall((:ht3d_PhiDrhoDP, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, phidrhodP), devptr(Cdouble, Tc_ex), devptr(Cdouble, Pc_ex), devptr(Cdouble, phc_ex), phi_i, scale_rho, scale_T, scale_sigma, ncx, ncy, ncz)
ht3d_dim_Init_vel(Vx, Vy, Vz, qx, qy, qz, rhofv, ncx, ncy, ncz) =
    GC.@preserve Vx Vy Vz qx qy qz rhofv ccall((:ht3d_dim_Init_vel, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, Vx), devptr(Cdouble, Vy), devptr(Cdouble, Vz), devptr(Cdouble, qx), devptr(Cdouble, qy), devptr(Cdouble, qz), devptr(Cdouble, rhofv), ncx, ncy, ncz)
ht3d_dim_SetTemperatureBCs(Tc_ex, ky, wN, phc_ex, qyS, _dy, kS, TS, TN, y_plateau, a2, b2, dTdy, dx, dy, sticky_air, sca_T, sca_kT, phi_i, ncx, ncy, ncz) =
    GC.@preserve Tc_ex ky wN phc_ex ccall((:ht3d_dim_SetTemperatureBCs, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Tc_ex), devptr(Cdouble, ky), devptr(Cdouble, wN), devptr(Cdouble, phc_ex), qyS, _dy, kS, TS, TN, y_plateau, a2, b2, dTdy, dx, dy, sticky_air, sca_T, sca_kT, phi_i, ncx, ncy, ncz)
ht3d_dim_SetPressureBCs(Pc_ex, wN, phc_ex, PS, PN, y_plateau, a2, b2, rhof, g, dx, dy, sticky_air, ncx, ncy, ncz) =
    GC.@preserve Pc_ex wN phc_ex ccall((:ht3d_dim_SetPressureBCs, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Pc_ex), devptr(Cdouble, wN), devptr(Cdouble, phc_ex), PS, PN, y_plateau, a2, b2, rhof, g, dx, dy, sticky_air, ncx, ncy, ncz)
ht3d_Phase(X, phc_ex, ncx, ncy, ncz) =
    GC.@preserve X phc_ex ccall((:ht3d_Phase, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, phc_ex), ncx, ncy, ncz)
ht3d_Pressure(X, Pc_ex, phc_ex, scale_sigma, ncx, ncy, ncz) =
    GC.@preserve X Pc_ex phc_ex ccall((:ht3d_Pressure, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, Pc_ex), devptr(Cdouble, phc_ex), scale_sigma, ncx, ncy, ncz)
ht3d_Temperature(X, Tc_ex, phc_ex, scale_T, ncx, ncy, ncz) =
    GC.@preserve X Tc_ex phc_ex ccall((:ht3d_Temperature, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, Tc_ex), devptr(Cdouble, phc_ex), scale_T, ncx, ncy, ncz)
ht3d_Velocity(X, Y, Z, Vx, Vy, Vz, phc_ex, scale_V, ncx, ncy, ncz) =
    GC.@preserve X Y Z Vx Vy Vz phc_ex ccall((:ht3d_Velocity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, Y), devptr(Cdouble, Z), devptr(Cdouble, Vx), devptr(Cdouble, Vy), devptr(Cdouble, Vz), devptr(Cdouble, phc_ex), scale_V, ncx, ncy, ncz)
ht3d_EffectiveThermalConductivity(X, Tc_ex, phi0, phc_ex, scale_T, ncx, ncy, ncz) =
    GC.@preserve X Tc_ex phc_ex ccall((:ht3d_EffectiveThermalConductivity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, Tc_ex), phi0, devptr(Cdouble, phc_ex), scale_T, ncx, ncy, ncz)
ht3d_EffectiveHeatCapacity(Ceff, Tc_ex, Pc_ex, phc_ex, phi_i, sca_T, sca_sigma, sca_C, sca_rho, ncx, ncy, ncz) =
    GC.@preserve Ceff Tc_ex Pc_ex phc_ex ccall((:ht3d_EffectiveHeatCapacity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Ceff), devptr(Cdouble, Tc_ex), devptr(Cdouble, Pc_ex), devptr(Cdouble, phc_ex), phi_i, sca_T, sca_sigma, sca_C, sca_rho, ncx, ncy, ncz)
ht3d_Permeability(X, k_fact, ymin, dy, phc_ex, delta, scale_L, ncx, ncy, ncz) =
    GC.@preserve X phc_ex ccall((:ht3d_Permeability, libht3d), Cint, (Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), k_fact, ymin, dy, devptr(Cdouble, phc_ex), delta, scale_L, ncx, ncy, ncz)
ht3d_FluidDensity(X, Tc_ex, Pc_ex, phc_ex, scale_T, scale_sigma, ncx, ncy, ncz) =
    GC.@preserve X Tc_ex Pc_ex phc_ex ccall((:ht3d_FluidDensity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, Tc_ex), devptr(Cdouble, Pc_ex), devptr(Cdouble, phc_ex), scale_T, scale_sigma, ncx, ncy, ncz)
ht3d_Viscosity(X, Tc_ex, phc_ex, scale_T, ncx, ncy, ncz) =
    GC.@preserve X Tc_ex phc_ex ccall((:ht3d_Viscosity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, Tc_ex), devptr(Cdouble, phc_ex), scale_T, ncx, ncy, ncz)
ht3d_Peclet(X, Tc_ex, Pc_ex, Vx, Vy, Vz, phc_ex, phi0, scale_T, scale_sigma, scale_V, scale_L, ncx, ncy, ncz) =
    GC.@preserve X Tc_ex Pc_ex Vx Vy Vz phc_ex ccall((:ht3d_Peclet, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cdouble, Tc_ex), devptr(Cdouble, Pc_ex), devptr(Cdouble, Vx), devptr(Cdouble, Vy), devptr(Cdouble, Vz), devptr(Cdouble, phc_ex), phi0, scale_T, scale_sigma, scale_V, scale_L, ncx, ncy, ncz)
ht3d_ApplyViscosity(kx, ky, kz, Tc_ex, a, ncx, ncy, ncz) =
    GC.@preserve kx ky kz Tc_ex ccall((:ht3d_ApplyViscosity, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, Tc_ex), a, ncx, ncy, ncz)
ht3d_VxPlusMinus(Vxm, Vxp, Vx, ncx, ncy, ncz) =
    GC.@preserve Vxm Vxp Vx ccall((:ht3d_VxPlusMinus, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, Vxm), devptr(Cdouble, Vxp), devptr(Cdouble, Vx), ncx, ncy, ncz)
ht3d_VyPlusMinus(Vym, Vyp, Vy, ncx, ncy, ncz) =
    GC.@preserve Vym Vyp Vy ccall((:ht3d_VyPlusMinus, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, Vym), devptr(Cdouble, Vyp), devptr(Cdouble, Vy), ncx, ncy, ncz)
ht3d_VzPlusMinus(Vzm, Vzp, Vz, ncx, ncy, ncz) =
    GC.@preserve Vzm Vzp Vz ccall((:ht3d_VzPlusMinus, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, Vzm), devptr(Cdouble, Vzp), devptr(Cdouble, Vz), ncx, ncy, ncz)
ht3d_dFdx_Weno5(dFdxi, V1, V2, V3, V4, V5, ncx, ncy, ncz) =
    GC.@preserve dFdxi V1 V2 V3 V4 V5 ccall((:ht3d_dFdx_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, dFdxi), devptr(Cdouble, V1), devptr(Cdouble, V2), devptr(Cdouble, V3), devptr(Cdouble, V4), devptr(Cdouble, V5), ncx, ncy, ncz)
ht3d_Gradients_Weno5(v1, v2, v3, v4, v5, Fc_exxx, axis, plus, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve v1 v2 v3 v4 v5 Fc_exxx ccall((:ht3d_Gradients_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, v1), devptr(Cdouble, v2), devptr(Cdouble, v3), devptr(Cdouble, v4), devptr(Cdouble, v5), devptr(Cdouble, Fc_exxx), axis, plus, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Boundaries_Weno5(Fc_exxx, Fc, axis, src_ext, type_lo, val_lo, type_hi, val_hi, ncx, ncy, ncz) =
    GC.@preserve Fc_exxx Fc ccall((:ht3d_Boundaries_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint, Cdouble, Cint, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc_exxx), devptr(Cdouble, Fc), axis, src_ext, type_lo, val_lo, type_hi, val_hi, ncx, ncy, ncz)
ht3d_Gradients_minus_x_Weno5(v1, v2, v3, v4, v5, Fc_exxx, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve v1 v2 v3 v4 v5 Fc_exxx ccall((:ht3d_Gradients_minus_x_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, v1), devptr(Cdouble, v2), devptr(Cdouble, v3), devptr(Cdouble, v4), devptr(Cdouble, v5), devptr(Cdouble, Fc_exxx), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Gradients_plus_x_Weno5(v1, v2, v3, v4, v5, Fc_exxx, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve v1 v2 v3 v4 v5 Fc_exxx ccall((:ht3d_Gradients_plus_x_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, v1), devptr(Cdouble, v2), devptr(Cdouble, v3), devptr(Cdouble, v4), devptr(Cdouble, v5), devptr(Cdouble, Fc_exxx), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Gradients_minus_y_Weno5(v1, v2, v3, v4, v5, Fc_exxx, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve v1 v2 v3 v4 v5 Fc_exxx ccall((:ht3d_Gradients_minus_y_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, v1), devptr(Cdouble, v2), devptr(Cdouble, v3), devptr(Cdouble, v4), devptr(Cdouble, v5), devptr(Cdouble, Fc_exxx), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Gradients_plus_y_Weno5(v1, v2, v3, v4, v5, Fc_exxx, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve v1 v2 v3 v4 v5 Fc_exxx ccall((:ht3d_Gradients_plus_y_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, v1), devptr(Cdouble, v2), devptr(Cdouble, v3), devptr(Cdouble, v4), devptr(Cdouble, v5), devptr(Cdouble, Fc_exxx), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Gradients_minus_z_Weno5(v1, v2, v3, v4, v5, Fc_exxx, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve v1 v2 v3 v4 v5 Fc_exxx ccall((:ht3d_Gradients_minus_z_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, v1), devptr(Cdouble, v2), devptr(Cdouble, v3), devptr(Cdouble, v4), devptr(Cdouble, v5), devptr(Cdouble, Fc_exxx), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Gradients_plus_z_Weno5(v1, v2, v3, v4, v5, Fc_exxx, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve v1 v2 v3 v4 v5 Fc_exxx ccall((:ht3d_Gradients_plus_z_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, v1), devptr(Cdouble, v2), devptr(Cdouble, v3), devptr(Cdouble, v4), devptr(Cdouble, v5), devptr(Cdouble, Fc_exxx), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_Boundaries_x_Weno5(Fc_exxx, Fc, type_W, val_W, type_E, val_E, ncx, ncy, ncz) =
    GC.@preserve Fc_exxx Fc ccall((:ht3d_Boundaries_x_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cint, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc_exxx), devptr(Cdouble, Fc), type_W, val_W, type_E, val_E, ncx, ncy, ncz)
ht3d_Boundaries_x_Weno5_v2(Fc_exxx, Fc, type_W, val_W, type_E, val_E, ncx, ncy, ncz) =
    GC.@preserve Fc_exxx Fc ccall((:ht3d_Boundaries_x_Weno5_v2, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cint, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc_exxx), devptr(Cdouble, Fc), type_W, val_W, type_E, val_E, ncx, ncy, ncz)
ht3d_Boundaries_y_Weno5(Fc_exxx, Fc, type_S, val_S, type_N, val_N, ncx, ncy, ncz) =
    GC.@preserve Fc_exxx Fc ccall((:ht3d_Boundaries_y_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cint, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc_exxx), devptr(Cdouble, Fc), type_S, val_S, type_N, val_N, ncx, ncy, ncz)
ht3d_Boundaries_y_Weno5_v2(Fc_exxx, Fc, type_S, val_S, type_N, val_N, ncx, ncy, ncz) =
    GC.@preserve Fc_exxx Fc ccall((:ht3d_Boundaries_y_Weno5_v2, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cint, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc_exxx), devptr(Cdouble, Fc), type_S, val_S, type_N, val_N, ncx, ncy, ncz)
ht3d_Boundaries_z_Weno5(Fc_exxx, Fc, type_B, val_B, type_F, val_F, ncx, ncy, ncz) =
    GC.@preserve Fc_exxx Fc ccall((:ht3d_Boundaries_z_Weno5, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cint, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc_exxx), devptr(Cdouble, Fc), type_B, val_B, type_F, val_F, ncx, ncy, ncz)
ht3d_Boundaries_z_Weno5_v2(Fc_exxx, Fc, type_B, val_B, type_F, val_F, ncx, ncy, ncz) =
    GC.@preserve Fc_exxx Fc ccall((:ht3d_Boundaries_z_Weno5_v2, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cint, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc_exxx), devptr(Cdouble, Fc), type_B, val_B, type_F, val_F, ncx, ncy, ncz)
ht3d_Advect(Fc, Vxp, dTdxm, Vxm, dTdxp, dt, ncx, ncy, ncz) =
    GC.@preserve Fc Vxp dTdxm Vxm dTdxp ccall((:ht3d_Advect, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc), devptr(Cdouble, Vxp), devptr(Cdouble, dTdxm), devptr(Cdouble, Vxm), devptr(Cdouble, dTdxp), dt, ncx, ncy, ncz)
ht3d_TimeAveraging(Fc, Fold, order, ncx, ncy, ncz) =
    GC.@preserve Fc Fold ccall((:ht3d_TimeAveraging, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc), devptr(Cdouble, Fold), order, ncx, ncy, ncz)
ht3d_Advect_v2(Fc, X, V, axis, dt, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve Fc X V ccall((:ht3d_Advect_v2, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc), devptr(Cdouble, X), devptr(Cdouble, V), axis, dt, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_TimeAveraging_v2(Fc, Fold, order, ncx, ncy, ncz) =
    GC.@preserve Fc Fold ccall((:ht3d_TimeAveraging_v2, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Cint, Cint), devptr(Cdouble, Fc), devptr(Cdouble, Fold), order, ncx, ncy, ncz)
ht3d_dyrel_solve(f, p, rep) =
    GC.@preserve f p rep ccall((:ht3d_dyrel_solve, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), f, p, rep)
ht3d_dyrel_solve_noalloc(f, p, rep) =
    GC.@preserve f p rep ccall((:ht3d_dyrel_solve_noalloc, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), f, p, rep)
ht3d_v18_thermal_solve(f, p, rep) =
    GC.@preserve f p rep ccall((:ht3d_v18_thermal_solve, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), f, p, rep)
ht3d_v18_darcy_solve(f, p, rep) =
    GC.@preserve f p rep ccall((:ht3d_v18_darcy_solve, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), f, p, rep)
ht3d_v18_tune(force_tile, last_tile) =
    GC.@preserve last_tile ccall((:ht3d_v18_tune, libht3d), Cint, (Cint, Ptr{Cint}), force_tile, last_tile)
ht3d_v18_tile_count() =
    ccall((:ht3d_v18_tile_count, libht3d), Cint, ())
ht3d_SetTemperatureBCs_f32(Tc_ex, Tbot, Ttop, ncx, ncy, ncz) =
    GC.@preserve Tc_ex ccall((:ht3d_SetTemperatureBCs_f32, libht3d), Cint, (Ptr{Cfloat}, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, Tc_ex), Tbot, Ttop, ncx, ncy, ncz)
ht3d_SetPressureBCs_f32(Pc_ex, Pbot, Ptop, ncx, ncy, ncz) =
    GC.@preserve Pc_ex ccall((:ht3d_SetPressureBCs_f32, libht3d), Cint, (Ptr{Cfloat}, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, Pc_ex), Pbot, Ptop, ncx, ncy, ncz)
ht3d_ComputeFlux_f32(qx, qy, qz, kx, ky, kz, A, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve qx qy qz kx ky kz A ccall((:ht3d_ComputeFlux_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cfloat, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, qx), devptr(Cfloat, qy), devptr(Cfloat, qz), devptr(Cfloat, kx), devptr(Cfloat, ky), devptr(Cfloat, kz), devptr(Cfloat, A), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusion_f32(F, R, qx, qy, qz, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F R qx qy qz ccall((:ht3d_ResidualDiffusion_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cfloat, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, F), devptr(Cfloat, R), devptr(Cfloat, qx), devptr(Cfloat, qy), devptr(Cfloat, qz), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusionTransient_f32(F, T, R, qx, qy, qz, _dt, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F T R qx qy qz ccall((:ht3d_ResidualDiffusionTransient_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cfloat, Cfloat, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, F), devptr(Cfloat, T), devptr(Cfloat, R), devptr(Cfloat, qx), devptr(Cfloat, qy), devptr(Cfloat, qz), _dt, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusionPC_f32(F, R, qx, qy, qz, PC, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F R qx qy qz PC ccall((:ht3d_ResidualDiffusionPC_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cfloat, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, F), devptr(Cfloat, R), devptr(Cfloat, qx), devptr(Cfloat, qy), devptr(Cfloat, qz), devptr(Cfloat, PC), _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_ResidualDiffusionTransientPC_f32(F, T, R, qx, qy, qz, PC, _dt, _dx, _dy, _dz, ncx, ncy, ncz) =
    GC.@preserve F T R qx qy qz PC ccall((:ht3d_ResidualDiffusionTransientPC_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cfloat, Cfloat, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, F), devptr(Cfloat, T), devptr(Cfloat, R), devptr(Cfloat, qx), devptr(Cfloat, qy), devptr(Cfloat, qz), devptr(Cfloat, PC), _dt, _dx, _dy, _dz, ncx, ncy, ncz)
ht3d_DYRELUpdate_f32(F0, X, F, h1, h2, dtau, ncx, ncy, ncz) =
    GC.@preserve F0 X F ccall((:ht3d_DYRELUpdate_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cfloat, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, F0), devptr(Cfloat, X), devptr(Cfloat, F), h1, h2, dtau, ncx, ncy, ncz)
ht3d_GershgorinPoisson_f32(G, D, rhoC, kx, ky, kz, transient, dt, dx, dy, dz, ncx, ncy, ncz) =
    GC.@preserve G D rhoC kx ky kz ccall((:ht3d_GershgorinPoisson_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cfloat, Cfloat, Cfloat, Cfloat, Cfloat, Cint, Cint, Cint), devptr(Cfloat, G), devptr(Cfloat, D), devptr(Cfloat, rhoC), devptr(Cfloat, kx), devptr(Cfloat, ky), devptr(Cfloat, kz), transient, dt, dx, dy, dz, ncx, ncy, ncz)
ht3d_to_f32(out, in, n) =
    GC.@preserve out in ccall((:ht3d_to_f32, libht3d), Cint, (Ptr{Cfloat}, Ptr{Cdouble}, Csize_t), devptr(Cfloat, out), devptr(Cdouble, in), n)
ht3d_add_f32(X, e, ncx, ncy, ncz) =
    GC.@preserve X e ccall((:ht3d_add_f32, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cfloat}, Cint, Cint, Cint), devptr(Cdouble, X), devptr(Cfloat, e), ncx, ncy, ncz)
ht3d_dyrel_solve_dc(f, p, rep, outer_iters, inner_iters) =
    GC.@preserve f p rep outer_iters inner_iters ccall((:ht3d_dyrel_solve_dc, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cint}, Ptr{Cint}), f, p, rep, outer_iters, inner_iters)
ht3d_weno5_advect(Tc_ex, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot, version, ncx, ncy, ncz) =
    GC.@preserve Tc_ex Vx Vy Vz ccall((:ht3d_weno5_advect, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint, Cint), devptr(Cdouble, Tc_ex), devptr(Cdouble, Vx), devptr(Cdouble, Vy), devptr(Cdouble, Vz), dt, _dx, _dy, _dz, Ttop, Tbot, version, ncx, ncy, ncz)
ht3d_velocity_dt_nu(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, dx, dy, dz, cfl, Tc_ex, kty, Lx, Lz, out, ncx, ncy, ncz) =
    GC.@preserve Vx Vy Vz kx ky kz Pc_ex Ty Tc_ex kty out ccall((:ht3d_velocity_dt_nu, libht3d), Cint, (Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Ptr{Cdouble}, Cint, Cint, Cint), devptr(Cdouble, Vx), devptr(Cdouble, Vy), devptr(Cdouble, Vz), devptr(Cdouble, kx), devptr(Cdouble, ky), devptr(Cdouble, kz), devptr(Cdouble, Pc_ex), devptr(Cdouble, Ty), Ra, dx, dy, dz, cfl, devptr(Cdouble, Tc_ex), devptr(Cdouble, kty), Lx, Lz, devptr(Cdouble, out), ncx, ncy, ncz)
ht3d_hrl_create(ncx, ncy, ncz, Lx, Ly, Lz, Ra, cfl, out) =
    GC.@preserve out ccall((:ht3d_hrl_create, libht3d), Cint, (Cint, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}), ncx, ncy, ncz, Lx, Ly, Lz, Ra, cfl, out)
ht3d_hrl_destroy(m) =
    GC.@preserve m ccall((:ht3d_hrl_destroy, libht3d), Cint, (Ptr{Cvoid},), m)
ht3d_hrl_field(m, name, dptr, count) =
    GC.@preserve m dptr count ccall((:ht3d_hrl_field, libht3d), Cint, (Ptr{Cvoid}, Cstring, Ptr{Cvoid}, Ptr{Cvoid}), m, name, dptr, count)
ht3d_hrl_set_params(m, pT, pP, weno_version, precision) =
    GC.@preserve m pT pP ccall((:ht3d_hrl_set_params, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint), m, pT, pP, weno_version, precision)
ht3d_hrl_set_dt(m, dt) =
    GC.@preserve m ccall((:ht3d_hrl_set_dt, libht3d), Cint, (Ptr{Cvoid}, Cdouble), m, dt)
ht3d_hrl_set_dt_rule(m, rule) =
    GC.@preserve m ccall((:ht3d_hrl_set_dt_rule, libht3d), Cint, (Ptr{Cvoid}, Cint), m, rule)
ht3d_hrl_set_viscosity(m, a) =
    GC.@preserve m ccall((:ht3d_hrl_set_viscosity, libht3d), Cint, (Ptr{Cvoid}, Cdouble), m, a)
ht3d_hrl_step(m, rep) =
    GC.@preserve m rep ccall((:ht3d_hrl_step, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), m, rep)
ht3d_batch_create(models, nprob, out) =
    GC.@preserve models out ccall((:ht3d_batch_create, libht3d), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}), models, nprob, out)
ht3d_batch_destroy(b) =
    GC.@preserve b ccall((:ht3d_batch_destroy, libht3d), Cint, (Ptr{Cvoid},), b)
ht3d_batch_set_params(b, pT, pP, precision, use_graphs) =
    GC.@preserve b pT pP ccall((:ht3d_batch_set_params, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint), b, pT, pP, precision, use_graphs)
ht3d_batch_step(b, reps) =
    GC.@preserve b reps ccall((:ht3d_batch_step, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b, reps)
ht3d_batch_stats_get(b, s) =
    GC.@preserve b s ccall((:ht3d_batch_stats_get, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b, s)
ht3d_batch_stats_reset(b) =
    GC.@preserve b ccall((:ht3d_batch_stats_reset, libht3d), Cint, (Ptr{Cvoid},), b)
ht3d_batch_tune_force(tile) =
    ccall((:ht3d_batch_tune_force, libht3d), Cint, (Cint,), tile)
ht3d_batch_tile_count() =
    ccall((:ht3d_batch_tile_count, libht3d), Cint, ())
ht3d_batch_tune_force_grid(mode) =
    ccall((:ht3d_batch_tune_force_grid, libht3d), Cint, (Cint,), mode)
ht3d_batch_tune_vertex(mode) =
    ccall((:ht3d_batch_tune_vertex, libht3d), Cint, (Cint,), mode)
ht3d_batch_tune_pairs(mode) =
    ccall((:ht3d_batch_tune_pairs, libht3d), Cint, (Cint,), mode)
ht3d_batch_tune_force_pair(tile) =
    ccall((:ht3d_batch_tune_force_pair, libht3d), Cint, (Cint,), tile)
ht3d_batch_pair_tile_count() =
    ccall((:ht3d_batch_pair_tile_count, libht3d), Cint, ())
ht3d_batch_pair_tile_info(tile, rows, planes, cells_per_lane) =
    GC.@preserve rows planes cells_per_lane ccall((:ht3d_batch_pair_tile_info, libht3d), Cint, (Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}), tile, rows, planes, cells_per_lane)
ht3d_lognormal_field(kfv, ncx, ncy, ncz, xmin, ymin, zmin, dx, dy, dz, K, phi, M, sigma, fault_mult, fault_half, fault_n, fault_x0) =
    GC.@preserve kfv K phi fault_n fault_x0 ccall((:ht3d_lognormal_field, libht3d), Cint, (Ptr{Cdouble}, Cint, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}), devptr(Cdouble, kfv), ncx, ncy, ncz, xmin, ymin, zmin, dx, dy, dz, K, phi, M, sigma, fault_mult, fault_half, fault_n, fault_x0)
