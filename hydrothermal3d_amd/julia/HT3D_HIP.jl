# HT3D_HIP.jl -- one-include swap of a HydroThermal3D driver onto libht3d.so.
#
# A driver such as Dimensionless/HydroThermal3D_Alban_v6.jl replaces its
# ParallelStencil set-up (`using ParallelStencil ... @init_parallel_stencil(...)`
# and the @sqrt/@exp macros, lines 5-16) with
#
#     include("path/to/hydrothermal3d_amd/julia/HT3D_HIP.jl")
#
# and keeps everything else, including its `include`s of tools/Macros.jl,
# tools/Weno5_Routines*.jl and the kernel files.  This file provides what the
# drivers use from ParallelStencil:
#   * `Data.Array` (a device array), `Data.Number`, `@zeros`, `@ones`;
#   * `@parallel f!(args...)` and `@parallel blocks threads f!(args...)`: a call
#     of HT3D_HIP.f! (launch configurations are dropped); kernel definitions
#     written with `@parallel function` / `@parallel_indices` are swallowed,
#     because the library provides every kernel the drivers launch;
#   * Base methods on device arrays for the host reductions of tools/Macros.jl:20-25
#     (`maximum_g(A) = maximum(A)` etc. then reduce on the device), including the
#     `abs.(A)` and `A[:]` forms the drivers pass to them;
#   * the host<->device copies the drivers write (`Array(A)`, `Data.Array(h)`,
#     `A .= Data.Array(h)`, slices such as `qy[:, end, :]`, scalar stores).
# Each kernel method calls the generated binding of its C entry (HT3D_ccalls.jl,
# one literal `ccall` per entry of include/ht3d.h) with the cell extents derived
# from its arguments' staggering.
#
# Dimensionless/HydroThermal3D_v7.jl uses the SetPressureBCs! of
# Dimensionless/kernels_HT3D.jl:144-154 (Dirichlet in y over the whole face),
# Alban_v6 its own (Neumann in y); same name and arity, so the driver chooses:
#     HT3D_HIP.SETPRESSUREBCS[] = :v7        # after the include, for v7
#
# Status: not executed by the test suite (no Julia toolchain in the build
# image).  The test suite checks that every kernel the three drivers launch has
# a method here and a C entry in include/ht3d.h, and that HT3D_ccalls.jl matches
# the header; the C ABI itself is exercised through the Python ctypes binding.

module HT3D_HIP

import Statistics

export Data, DevArray, @zeros, @ones, @parallel, @parallel_indices, @sqrt, @exp,
       # kernels_HT3D.jl, HydroThermal3D_Alban_v6.jl / v7.jl, tools/Macros.jl
       ResetA!, Cpy_inn_to_all!, Cpy_all_to_inn!, ArrayEqualArray!, Multiply, InitThermal!, InitConductivity!,
       InitDarcy!, KillNorthSouthFluxes!, Compute_Rp!, SetTemperatureBCs!, SetPressureBCs!, ComputeFlux!, UpdateT!,
       ResidualDiffusion!, ResidualDiffusionTransient!, UpdateP!, DampedUpdate!, DYRELUpdate!, SwapDYREL!,
       GershgorinPoisson!, GershgorinHydro!, RayleighQuotientNumerator!, RayleighQuotientDenominator!, Init_vel!,
       # kernels_HT3D_dimensional.jl
       ComputeThermalConductivity, ComputeHydroConductivity, ComputeFluidDensity, SmoothConductivityV2C,
       SmoothConductivityC2V, ComputeρCeffective!, ϕρf, ϕρf_old, ϕdρdP!, ResidualTemperatureLinearised!,
       ResidualFluidPressure!, ComputeDarcyFlux!,
       # kernels_HT3D_postprocessing.jl
       Phase!, Pressure!, Temperature!, Velocity!, EffectiveThermalConductivity!, EffectiveHeatCapacity!,
       Permeability!, FluidDensity!, Viscosity!, Peclet!,
       # tools/Weno5_Routines.jl, tools/Weno5_Routines_v2.jl
       VxPlusMinus!, VyPlusMinus!, VzPlusMinus!, dFdx_Weno5!, Gradients_minus_x_Weno5!, Gradients_plus_x_Weno5!,
       Gradients_minus_y_Weno5!, Gradients_plus_y_Weno5!, Gradients_minus_z_Weno5!, Gradients_plus_z_Weno5!,
       Boundaries_x_Weno5!, Boundaries_y_Weno5!, Boundaries_z_Weno5!, Advect!, TimeAveraging!

const libht3d = get(ENV, "HT3D_LIB", joinpath(@__DIR__, "..", "libht3d.so"))

# --------------------------------------------------------------- device arrays
mutable struct DevArray{N} <: AbstractArray{Float64,N}
    ptr::Ptr{Float64}
    dims::NTuple{N,Int}
    function DevArray{N}(dims::NTuple{N,Int}) where {N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ht3d_alloc(8 * prod(dims), p), "ht3d_alloc")
        A = new{N}(Ptr{Float64}(p[]), dims)
        finalizer(a -> ht3d_free(a.ptr), A)
        return A
    end
end

# device pointer of a kernel argument; a host array here would be a host
# address handed to a GPU kernel, so it is refused
devptr(::Type{T}, A::DevArray) where {T} = Ptr{T}(A.ptr)
devptr(::Type{T}, p::Ptr) where {T} = Ptr{T}(p)
devptr(::Type{T}, ::Nothing) where {T} = Ptr{T}(C_NULL)
devptr(::Type, ::AbstractArray) =
    throw(ArgumentError("a host array was passed where a device array is expected (upload it with Data.Array)"))

include("HT3D_ccalls.jl")

function check(rc::Integer, name)
    rc == 0 && return nothing
    rc == 10 && return nothing      # nitmax: reported, not an error (the reference is silent)
    error("$name failed with status $rc: $(unsafe_string(ht3d_last_error()))")
end

function __init__()
    dev = parse(Int, get(ENV, "HT3D_DEVICE", "0"))
    check(ht3d_init(dev), "ht3d_init")
end

Base.size(A::DevArray) = A.dims
Base.IndexStyle(::Type{<:DevArray}) = IndexLinear()
Base.show(io::IO, A::DevArray) = print(io, "DevArray", A.dims)
Base.show(io::IO, ::MIME"text/plain", A::DevArray) = print(io, join(A.dims, "x"), " device array of Float64")
function Base.Array(A::DevArray)
    h = Array{Float64}(undef, A.dims)
    check(ht3d_d2h(h, A.ptr, 8 * length(h)), "ht3d_d2h")
    return h
end
function upload(h::AbstractArray{<:Real})
    A = DevArray{ndims(h)}(size(h))
    hh = Array{Float64}(h)
    check(ht3d_h2d(A.ptr, hh, 8 * length(hh)), "ht3d_h2d")
    return A
end
# `Ft[:]` (the norms of HydroThermal3D_v7.jl:180) is the array itself, read only
Base.getindex(A::DevArray, ::Colon) = A
Base.getindex(::DevArray, ::Int...) =
    error("scalar read of a device array; copy it with Array(A) first")
# slices (`qy[:, end, :]`, Alban_v6:346; `Pc_ex[2:end-1, ...]`, :487) are host copies
Base.getindex(A::DevArray, I...) = Array(A)[I...]
# scalar stores (`time_vec[it+1] = time`, Alban_v6:468)
function Base.setindex!(A::DevArray, v, I::Int...)
    k = LinearIndices(A.dims)[I...]
    r = Ref{Float64}(Float64(v))
    check(ht3d_h2d(A.ptr + 8 * (k - 1), r, 8), "ht3d_h2d")
    return A
end
function Base.copyto!(dst::DevArray, src::DevArray)
    length(dst) == length(src) || throw(DimensionMismatch("device copy of $(size(src)) into $(size(dst))"))
    check(ht3d_d2d(dst.ptr, src.ptr, 8 * length(src)), "ht3d_d2d")
    return dst
end
Base.copyto!(dst::DevArray, src::AbstractArray{<:Real}) = copyto!(dst, upload(src))
# `A .= Data.Array(h)` (restart, HydroThermal3D_v18.jl:384-395)
Base.Broadcast.broadcasted(::typeof(identity), A::DevArray) = A
Base.Broadcast.materialize!(dst::DevArray, src::DevArray) = copyto!(dst, src)
Base.Broadcast.materialize!(dst::DevArray, src::AbstractArray{<:Real}) = copyto!(dst, src)

module Data
    import ..HT3D_HIP: DevArray, upload
    const Number = Float64
    const Array = DevArray
    Array(h::AbstractArray{<:Real}) = upload(h)
end

macro zeros(dims...)
    :(let A = DevArray{$(length(dims))}(($(esc.(dims)...),))
        check(ht3d_memset(A.ptr, 0, 8 * length(A)), "ht3d_memset")
        A
    end)
end
macro ones(dims...)
    :(upload(ones($(esc.(dims)...))))
end
macro sqrt(args...) esc(:(Base.sqrt($(args...)))) end
macro exp(args...) esc(:(Base.exp($(args...)))) end

# `@parallel f!(args...)` / `@parallel blocks threads f!(args...)`: a call of
# this module's f!; `@parallel function f!(...) ... end`: swallowed.
macro parallel(args...)
    ex = args[end]
    if ex isa Expr && ex.head in (:function, :(=)) && ex.args[1] isa Expr && ex.args[1].head == :call
        return nothing
    end
    (ex isa Expr && ex.head == :call && ex.args[1] isa Symbol) || error("@parallel: expected a kernel call")
    return Expr(:call, GlobalRef(HT3D_HIP, ex.args[1]), map(esc, ex.args[2:end])...)
end
macro parallel_indices(args...)
    return nothing
end

# --------------------------------------------------------------- reductions
# tools/Macros.jl:20-25 defines maximum_g(A) = maximum(A), ... in the driver;
# these methods make them device reductions (ht3d_reduce: SUM 0, SUM_ABS 1,
# MAX 2, MIN 3, MAX_ABS 4).
struct AbsOf
    A::DevArray
end
Base.Broadcast.broadcasted(::typeof(abs), A::DevArray) = AbsOf(A)
Base.Broadcast.materialize(a::AbsOf) = a
function reduce_dev(A::DevArray, op)
    r = Ref{Float64}(0.0)
    check(ht3d_reduce(r, A, length(A), op), "ht3d_reduce")
    return r[]
end
Base.maximum(A::DevArray) = reduce_dev(A, 2)
Base.minimum(A::DevArray) = reduce_dev(A, 3)
Base.sum(A::DevArray) = reduce_dev(A, 0)
Statistics.mean(A::DevArray) = reduce_dev(A, 0) / length(A)
Base.maximum(a::AbsOf) = reduce_dev(a.A, 4)
Base.sum(a::AbsOf) = reduce_dev(a.A, 1)
Statistics.mean(a::AbsOf) = reduce_dev(a.A, 1) / length(a.A)

# --------------------------------------------------------------- kernels
# cell extents from an argument of known staggering
cells_E(A) = (size(A, 1) - 2, size(A, 2) - 2, size(A, 3) - 2)
cells_X6(A) = (size(A, 1) - 6, size(A, 2) - 6, size(A, 3) - 6)
cells_V(A) = (size(A, 1) - 1, size(A, 2) - 1, size(A, 3) - 1)
cells_Fx(A) = (size(A, 1) - 1, size(A, 2), size(A, 3))
cells_Fy(A) = (size(A, 1), size(A, 2) - 1, size(A, 3))
cells_c(A) = size(A)
k!(name, rc) = check(rc, name)

# tools/Macros.jl:171-199, tools/Weno5_Routines.jl:271-276, kernels_HT3D_dimensional.jl:106-138
ResetA!(A, B) = k!("ResetA!", ht3d_ResetA(A, B, size(A)...))
Cpy_inn_to_all!(A, B) = k!("Cpy_inn_to_all!", ht3d_Cpy_inn_to_all(A, B, cells_c(A)...))
Cpy_all_to_inn!(A, B) = k!("Cpy_all_to_inn!", ht3d_Cpy_all_to_inn(A, B, cells_c(B)...))
ArrayEqualArray!(F1, F2) = k!("ArrayEqualArray!", ht3d_ArrayEqualArray(F1, F2, size(F1)...))
Multiply(a, b, c) = k!("Multiply", ht3d_Multiply(a, b, c, cells_c(a)...))
SwapDYREL!(Xc, Yc, X0c_ex, Y0c) = k!("SwapDYREL!", ht3d_SwapDYREL(Xc, Yc, X0c_ex, Y0c, cells_c(Xc)...))
InitConductivity!(kx, ky, kz, ktv) = k!("InitConductivity!", ht3d_InitConductivity(kx, ky, kz, ktv, cells_V(ktv)...))

# kernels_HT3D.jl:73-154 and the closed-box drivers (Alban_v6:84-159)
InitThermal!(Rt, kx, ky, kz, Tc_ex, ktv, _dt) =
    k!("InitThermal!", ht3d_InitThermal(Rt, kx, ky, kz, Tc_ex, ktv, _dt, cells_E(Tc_ex)...))
InitThermal!(Tc0, Tc_ex) = k!("InitThermal!", ht3d_Cpy_inn_to_all(Tc0, Tc_ex, cells_c(Tc0)...))   # dimensional :128-131
InitDarcy!(Ty, kx, ky, kz, Tc_ex, kfv, _dt) =
    k!("InitDarcy!", ht3d_InitDarcy(Ty, kx, ky, kz, Tc_ex, kfv, _dt, cells_E(Tc_ex)...))
InitDarcy!(ϕρf0, ϕρf) = k!("InitDarcy!", ht3d_ArrayEqualArray(ϕρf0, ϕρf, size(ϕρf0)...))         # dimensional :261-264
KillNorthSouthFluxes!(Ty) = k!("KillNorthSouthFluxes!", ht3d_KillNorthSouthFluxes(Ty, cells_Fy(Ty)...))
Compute_Rp!(Rp, ky, Ty, Ra, _dy) = k!("Compute_Rp!", ht3d_Compute_Rp(Rp, ky, Ty, Ra, _dy, cells_c(Rp)...))
ComputeFlux!(qx, qy, qz, kx, ky, kz, A, _dx, _dy, _dz) =
    k!("ComputeFlux!", ht3d_ComputeFlux(qx, qy, qz, kx, ky, kz, A, _dx, _dy, _dz, cells_E(A)...))
UpdateT!(F, T, R, qx, qy, qz, _dt, dtau, _dx, _dy, _dz) =
    k!("UpdateT!", ht3d_UpdateT(F, T, R, qx, qy, qz, _dt, dtau, _dx, _dy, _dz, cells_c(F)...))
ResidualDiffusion!(F, R, qx, qy, qz, _dx, _dy, _dz) =
    k!("ResidualDiffusion!", ht3d_ResidualDiffusion(F, R, qx, qy, qz, _dx, _dy, _dz, cells_c(F)...))
ResidualDiffusionTransient!(F, T, R, qx, qy, qz, _dt, _dx, _dy, _dz) =
    k!("ResidualDiffusionTransient!",
       ht3d_ResidualDiffusionTransient(F, T, R, qx, qy, qz, _dt, _dx, _dy, _dz, cells_c(F)...))
UpdateP!(F0, P, F, dampx, dtau) = k!("UpdateP!", ht3d_UpdateP(F0, P, F, dampx, dtau, cells_c(F0)...))
Init_vel!(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, _dx, _dy, _dz) =
    k!("Init_vel!", ht3d_Init_vel(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, _dx, _dy, _dz, cells_E(Pc_ex)...))
# ghost fills: Alban_v6:108-115 (temperature), :120-127 (pressure, Neumann in y);
# kernels_HT3D.jl:144-154 (v7: Dirichlet in y), chosen with SETPRESSUREBCS
const SETPRESSUREBCS = Ref(:Alban_v6)
SetTemperatureBCs!(T, Tbot, Ttop) = k!("SetTemperatureBCs!", ht3d_SetTemperatureBCs(T, Tbot, Ttop, cells_E(T)...))
SetPressureBCs!(P, Pbot, Ptop) = SETPRESSUREBCS[] === :v7 ?
    k!("SetPressureBCs!", ht3d_SetPressureBCs_v7(P, Pbot, Ptop, cells_E(P)...)) :
    k!("SetPressureBCs!", ht3d_SetPressureBCs(P, Pbot, Ptop, cells_E(P)...))

# kernels_HT3D_dimensional.jl (the v18 driver)
DampedUpdate!(F0, X, F, dampx, _dτ) = k!("DampedUpdate!", ht3d_DampedUpdate(F0, X, F, dampx, _dτ, cells_c(F0)...))
DYRELUpdate!(F0, X, F, h1, h2, Δτ) = k!("DYRELUpdate!", ht3d_DYRELUpdate(F0, X, F, h1, h2, Δτ, cells_c(F0)...))
GershgorinPoisson!(G, D, ρC_eff, kx, ky, kz, transient, dt, dx, dy, dz) =
    k!("GershgorinPoisson!", ht3d_GershgorinPoisson(G, D, ρC_eff, kx, ky, kz, transient, dt, dx, dy, dz, cells_c(G)...))
GershgorinHydro!(G, ϕdρdP, kx, ky, kz, transient, dt, dx, dy, dz) =
    k!("GershgorinHydro!", ht3d_GershgorinHydro(G, ϕdρdP, kx, ky, kz, transient, dt, dx, dy, dz, cells_c(G)...))
RayleighQuotientNumerator!(G, Fc0, Fcit, Fc, Δτ) =
    k!("RayleighQuotientNumerator!", ht3d_RayleighQuotientNumerator(G, Fc0, Fcit, Fc, Δτ, cells_c(G)...))
RayleighQuotientDenominator!(G, Fc0, Δτ) =
    k!("RayleighQuotientDenominator!", ht3d_RayleighQuotientDenominator(G, Fc0, Δτ, cells_c(G)...))
ResidualTemperatureLinearised!(F, Tc_ex, Tc_old, ρC_eff, phc_ex, PC, qx, qy, qz, Qt, transient, _dt, _dx, _dy, _dz) =
    k!("ResidualTemperatureLinearised!",
       ht3d_ResidualTemperatureLinearised(F, Tc_ex, Tc_old, ρC_eff, phc_ex, PC, qx, qy, qz, Qt, transient, _dt, _dx,
                                          _dy, _dz, cells_c(F)...))
ResidualFluidPressure!(F, phc_ex, ϕρfc, ϕρfc0, PC, qx, qy, qz, transient, _dt, _dx, _dy, _dz) =
    k!("ResidualFluidPressure!",
       ht3d_ResidualFluidPressure(F, phc_ex, ϕρfc, ϕρfc0, PC, qx, qy, qz, transient, _dt, _dx, _dy, _dz, cells_c(F)...))
ComputeDarcyFlux!(qx, qy, qz, ρfv, kx, ky, kz, Pf, g, _dx, _dy, _dz) =
    k!("ComputeDarcyFlux!", ht3d_ComputeDarcyFlux(qx, qy, qz, ρfv, kx, ky, kz, Pf, g, _dx, _dy, _dz, cells_E(Pf)...))
Init_vel!(Vx, Vy, Vz, qx, qy, qz, ρfv) =
    k!("Init_vel!", ht3d_dim_Init_vel(Vx, Vy, Vz, qx, qy, qz, ρfv, cells_V(ρfv)...))
ComputeThermalConductivity(ktv, Tce, phv, ϕ0, scale_kt, scale_T) =
    k!("ComputeThermalConductivity", ht3d_ComputeThermalConductivity(ktv, Tce, phv, ϕ0, scale_kt, scale_T, cells_E(Tce)...))
ComputeHydroConductivity(kfv, Tce, Pce, phv, ymin, Δy, k_fact, δ, scale_σ, scale_T, scale_L, scale_t, scale_η, mode) =
    k!("ComputeHydroConductivity",
       ht3d_ComputeHydroConductivity(kfv, Tce, Pce, phv, ymin, Δy, k_fact, δ, scale_σ, scale_T, scale_L, scale_t,
                                     scale_η, mode, cells_E(Tce)...))
ComputeFluidDensity(ρfv, Tce, Pce, phv, scale_σ, scale_T, scale_ρ) =
    k!("ComputeFluidDensity", ht3d_ComputeFluidDensity(ρfv, Tce, Pce, phv, scale_σ, scale_T, scale_ρ, cells_E(Tce)...))
SmoothConductivityV2C(ktc, ktv) = k!("SmoothConductivityV2C", ht3d_SmoothConductivityV2C(ktc, ktv, cells_c(ktc)...))
SmoothConductivityC2V(ktv, ktc) = k!("SmoothConductivityC2V", ht3d_SmoothConductivityC2V(ktv, ktc, cells_c(ktc)...))
ComputeρCeffective!(ρC_eff, Tc_ex, Pc_ex, phc_ex, ϕi, sca_T, sca_σ, sca_C, sca_ρ) =
    k!("ComputeρCeffective!",
       ht3d_ComputeRhoCeffective(ρC_eff, Tc_ex, Pc_ex, phc_ex, ϕi, sca_T, sca_σ, sca_C, sca_ρ, cells_c(ρC_eff)...))
ϕρf(ϕρfc, Tc_ex, Pc_ex, phc_ex, ϕi, scale_ρ, scale_T, scale_σ) =
    k!("ϕρf", ht3d_PhiRhof(ϕρfc, Tc_ex, Pc_ex, phc_ex, ϕi, scale_ρ, scale_T, scale_σ, cells_c(ϕρfc)...))
ϕρf_old(ϕρfc, Tc0, Pc0, phc_ex, ϕi, scale_ρ, scale_T, scale_σ) =
    k!("ϕρf_old", ht3d_PhiRhof_old(ϕρfc, Tc0, Pc0, phc_ex, ϕi, scale_ρ, scale_T, scale_σ, cells_c(ϕρfc)...))
ϕdρdP!(ϕdρdP, Tc_ex, Pc_ex, phc_ex, ϕi, scale_ρ, scale_T, scale_σ) =
    k!("ϕdρdP!", ht3d_PhiDrhoDP(ϕdρdP, Tc_ex, Pc_ex, phc_ex, ϕi, scale_ρ, scale_T, scale_σ, cells_c(ϕdρdP)...))
# immersed-surface ghost fills (:183-205, :291-314): 19 and 13 arguments
SetTemperatureBCs!(Tc_ex, ky, wN, phc_ex, qyS, _dy, kS, TS, TN, y_plateau, a2, b2, dTdy, Δx, Δy, sticky_air, sca_T,
                   sca_kT, ϕi) =
    k!("SetTemperatureBCs!",
       ht3d_dim_SetTemperatureBCs(Tc_ex, ky, wN, phc_ex, qyS, _dy, kS, TS, TN, y_plateau, a2, b2, dTdy, Δx, Δy,
                                  sticky_air, sca_T, sca_kT, ϕi, cells_E(Tc_ex)...))
SetPressureBCs!(Pc_ex, wN, phc_ex, PS, PN, y_plateau, a2, b2, ρf, g, Δx, Δy, sticky_air) =
    k!("SetPressureBCs!",
       ht3d_dim_SetPressureBCs(Pc_ex, wN, phc_ex, PS, PN, y_plateau, a2, b2, ρf, g, Δx, Δy, sticky_air,
                               cells_E(Pc_ex)...))

# Dimensional/kernels_HT3D_postprocessing.jl (the VTK fields of HydroThermal3D_v18.jl:663-687)
Phase!(X, phc_ex) = k!("Phase!", ht3d_Phase(X, phc_ex, cells_c(X)...))
Pressure!(X, Pc_ex, phc_ex, scale_σ) = k!("Pressure!", ht3d_Pressure(X, Pc_ex, phc_ex, scale_σ, cells_c(X)...))
Temperature!(X, Tc_ex, phc_ex, scale_T) = k!("Temperature!", ht3d_Temperature(X, Tc_ex, phc_ex, scale_T, cells_c(X)...))
Velocity!(X, Y, Z, Vx, Vy, Vz, phc_ex, scale_V) =
    k!("Velocity!", ht3d_Velocity(X, Y, Z, Vx, Vy, Vz, phc_ex, scale_V, cells_c(X)...))
EffectiveThermalConductivity!(X, Tc_ex, ϕ0, phc_ex, scale_T) =
    k!("EffectiveThermalConductivity!", ht3d_EffectiveThermalConductivity(X, Tc_ex, ϕ0, phc_ex, scale_T, cells_c(X)...))
EffectiveHeatCapacity!(Ceff, Tc_ex, Pc_ex, phc_ex, ϕi, sca_T, sca_σ, sca_C, sca_ρ) =
    k!("EffectiveHeatCapacity!",
       ht3d_EffectiveHeatCapacity(Ceff, Tc_ex, Pc_ex, phc_ex, ϕi, sca_T, sca_σ, sca_C, sca_ρ, cells_c(Ceff)...))
Permeability!(X, k_fact, ymin, Δy, phc_ex, δ, scale_L) =
    k!("Permeability!", ht3d_Permeability(X, k_fact, ymin, Δy, phc_ex, δ, scale_L, cells_c(X)...))
FluidDensity!(X, Tc_ex, Pc_ex, phc_ex, scale_T, scale_σ) =
    k!("FluidDensity!", ht3d_FluidDensity(X, Tc_ex, Pc_ex, phc_ex, scale_T, scale_σ, cells_c(X)...))
Viscosity!(X, Tc_ex, phc_ex, scale_T) = k!("Viscosity!", ht3d_Viscosity(X, Tc_ex, phc_ex, scale_T, cells_c(X)...))
Peclet!(X, Tc_ex, Pc_ex, Vx, Vy, Vz, phc_ex, ϕ0, scale_T, scale_σ, scale_V, scale_L) =
    k!("Peclet!", ht3d_Peclet(X, Tc_ex, Pc_ex, Vx, Vy, Vz, phc_ex, ϕ0, scale_T, scale_σ, scale_V, scale_L,
                              cells_c(X)...))

# tools/Weno5_Routines.jl (v1) and tools/Weno5_Routines_v2.jl (v2)
VxPlusMinus!(Vxm, Vxp, Vx) = k!("VxPlusMinus!", ht3d_VxPlusMinus(Vxm, Vxp, Vx, cells_c(Vxm)...))
VyPlusMinus!(Vym, Vyp, Vy) = k!("VyPlusMinus!", ht3d_VyPlusMinus(Vym, Vyp, Vy, cells_c(Vym)...))
VzPlusMinus!(Vzm, Vzp, Vz) = k!("VzPlusMinus!", ht3d_VzPlusMinus(Vzm, Vzp, Vz, cells_c(Vzm)...))
dFdx_Weno5!(dFdxi, V1, V2, V3, V4, V5) = k!("dFdx_Weno5!", ht3d_dFdx_Weno5(dFdxi, V1, V2, V3, V4, V5, cells_c(dFdxi)...))
Gradients_minus_x_Weno5!(v1, v2, v3, v4, v5, X, _dx, _dy, _dz) =
    k!("Gradients_minus_x_Weno5!", ht3d_Gradients_minus_x_Weno5(v1, v2, v3, v4, v5, X, _dx, _dy, _dz, cells_X6(X)...))
Gradients_plus_x_Weno5!(v1, v2, v3, v4, v5, X, _dx, _dy, _dz) =
    k!("Gradients_plus_x_Weno5!", ht3d_Gradients_plus_x_Weno5(v1, v2, v3, v4, v5, X, _dx, _dy, _dz, cells_X6(X)...))
Gradients_minus_y_Weno5!(v1, v2, v3, v4, v5, X, _dx, _dy, _dz) =
    k!("Gradients_minus_y_Weno5!", ht3d_Gradients_minus_y_Weno5(v1, v2, v3, v4, v5, X, _dx, _dy, _dz, cells_X6(X)...))
Gradients_plus_y_Weno5!(v1, v2, v3, v4, v5, X, _dx, _dy, _dz) =
    k!("Gradients_plus_y_Weno5!", ht3d_Gradients_plus_y_Weno5(v1, v2, v3, v4, v5, X, _dx, _dy, _dz, cells_X6(X)...))
Gradients_minus_z_Weno5!(v1, v2, v3, v4, v5, X, _dx, _dy, _dz) =
    k!("Gradients_minus_z_Weno5!", ht3d_Gradients_minus_z_Weno5(v1, v2, v3, v4, v5, X, _dx, _dy, _dz, cells_X6(X)...))
Gradients_plus_z_Weno5!(v1, v2, v3, v4, v5, X, _dx, _dy, _dz) =
    k!("Gradients_plus_z_Weno5!", ht3d_Gradients_plus_z_Weno5(v1, v2, v3, v4, v5, X, _dx, _dy, _dz, cells_X6(X)...))
# the v1 ghost fills take the cell array, the v2 ones the extended array
extended(Fc, Fc_exxx) = size(Fc) == cells_X6(Fc_exxx) .+ 2
Boundaries_x_Weno5!(Fc_exxx, Fc, type_W, val_W, type_E, val_E) = extended(Fc, Fc_exxx) ?
    k!("Boundaries_x_Weno5!", ht3d_Boundaries_x_Weno5_v2(Fc_exxx, Fc, type_W, val_W, type_E, val_E, cells_X6(Fc_exxx)...)) :
    k!("Boundaries_x_Weno5!", ht3d_Boundaries_x_Weno5(Fc_exxx, Fc, type_W, val_W, type_E, val_E, cells_X6(Fc_exxx)...))
Boundaries_y_Weno5!(Fc_exxx, Fc, type_S, val_S, type_N, val_N) = extended(Fc, Fc_exxx) ?
    k!("Boundaries_y_Weno5!", ht3d_Boundaries_y_Weno5_v2(Fc_exxx, Fc, type_S, val_S, type_N, val_N, cells_X6(Fc_exxx)...)) :
    k!("Boundaries_y_Weno5!", ht3d_Boundaries_y_Weno5(Fc_exxx, Fc, type_S, val_S, type_N, val_N, cells_X6(Fc_exxx)...))
Boundaries_z_Weno5!(Fc_exxx, Fc, type_B, val_B, type_F, val_F) = extended(Fc, Fc_exxx) ?
    k!("Boundaries_z_Weno5!", ht3d_Boundaries_z_Weno5_v2(Fc_exxx, Fc, type_B, val_B, type_F, val_F, cells_X6(Fc_exxx)...)) :
    k!("Boundaries_z_Weno5!", ht3d_Boundaries_z_Weno5(Fc_exxx, Fc, type_B, val_B, type_F, val_F, cells_X6(Fc_exxx)...))
Advect!(Fc, Vxp, dTdxm, Vxm, dTdxp, dt) =                                              # v1, :257-262
    k!("Advect!", ht3d_Advect(Fc, Vxp, dTdxm, Vxm, dTdxp, dt, cells_c(Fc)...))
Advect!(Fc, X, V, axis, dt, _dx, _dy, _dz) =                                           # v2, _v2.jl:55-212
    k!("Advect!", ht3d_Advect_v2(Fc, X, V, axis, dt, _dx, _dy, _dz, cells_X6(X)...))
TimeAveraging!(Fc, Fold, order) = size(Fc) == size(Fold) ?                              # v1 :264-269 / v2 :214-219
    k!("TimeAveraging!", ht3d_TimeAveraging(Fc, Fold, order, cells_c(Fc)...)) :
    k!("TimeAveraging!", ht3d_TimeAveraging_v2(Fc, Fold, order, cells_c(Fold)...))

# --------------------------------------------------------------- fused entries
# Not exported: a driver calls them by name (HT3D_HIP.AdvectWithWeno5_v2(...))
# in place of its own orchestration.  WENO5 advection in one call each: v1
# (kernels_HT3D.jl:156-233, after Init_vel!) and v2 (Weno5_Routines_v2.jl:1-50);
# the scratch arguments of the reference signatures are accepted and unused.
function AdvectWithWeno5(Tc, Tc_ex, Tc_exxx, Told, dTdxm, dTdxp, Vxm, Vxp, Vym, Vyp, Vzm, Vzp, Vx, Vy, Vz,
                         v1, v2, v3, v4, v5, kx, ky, kz, dt, _dx, _dy, _dz, Ttop, Tbot, Pc_ex, Ty, Ra)
    Init_vel!(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, _dx, _dy, _dz)
    k!("AdvectWithWeno5", ht3d_weno5_advect(Tc_ex, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot, 1, cells_E(Tc_ex)...))
end
AdvectWithWeno5_v2(Tc_ex, Tc_exxx, Told, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot) =
    k!("AdvectWithWeno5_v2", ht3d_weno5_advect(Tc_ex, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot, 2, cells_E(Tc_ex)...))

# The two DYREL loops of HydroThermal3D_v18.jl (:443-478 heat, :500-537 Darcy),
# fused: a driver replaces the `for iter = 1:nitmax ... end` block (and the
# GershgorinPoisson!/ResetA! set-up before it) by one call.  `params` is
# (nitmax, nitout, ncheck, norm, tol, CFL, cfact, lmin_div); returns the report.
struct DyrelParams
    nitmax::Cint; nitout::Cint; ncheck::Cint; norm::Cint
    tol::Float64; CFL::Float64; cfact::Float64; lmin_div::Float64
end
struct DyrelReport
    iters::Cint; status::Cint; n_checks::Cint; n_lambda::Cint
    lmax::Float64; lmin::Float64; dtau::Float64; h1::Float64; h2::Float64
    nF_abs::Float64; nF_rel::Float64; nF_ini::Float64; ncheck::Cint
end
struct V18Field
    ncx::Cint; ncy::Cint; ncz::Cint
    dx::Float64; dy::Float64; dz::Float64; dt::Float64; transient::Float64
    X::Ptr{Cvoid}; kx::Ptr{Cvoid}; ky::Ptr{Cvoid}; kz::Ptr{Cvoid}; phc_ex::Ptr{Cvoid}; wN::Ptr{Cvoid}
    rhoC::Ptr{Cvoid}; PC::Ptr{Cvoid}; G::Ptr{Cvoid}; top::Float64; norm_scale::Float64
    Tc_old::Ptr{Cvoid}; qyS::Float64; Qt::Float64
    Tc_ex::Ptr{Cvoid}; phirhof0::Ptr{Cvoid}; rhofv::Ptr{Cvoid}
    g::Float64; phi_i::Float64; scale_rho::Float64; scale_T::Float64; scale_sigma::Float64
    kv::Ptr{Cvoid}   # vertex field of kx, ky, kz (C_NULL: the solve reads the face arrays)
end
# the structs are plain bits: Ref(x) hands the C side a pointer to their fields
function v18_solve!(solve, name, f::V18Field, p::DyrelParams, keep...)
    r = Ref(DyrelReport(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0))
    rc = GC.@preserve keep solve(Ref(f), Ref(p), r)
    rc == 11 && error("Nan ...")          # v18:464
    rc == 12 && error("Diverged ...")     # v18:465
    rc in (0, 10) || check(rc, name)
    return r[]
end
vp(A::DevArray) = Ptr{Cvoid}(A.ptr)
# heat loop: Tc_ex updated; ρC_ϕρ = rho*C (ComputeρCeffective!); PC, Xc0 receive D and G
function v18_thermal_solve!(Tc_ex, Tc0, ρC_ϕρ, kx, ky, kz, phc_ex, wN, PC, Xc0; dx, dy, dz, dt, transient, qyS, Ttop,
                            Qt, norm_scale, params::DyrelParams)
    n = cells_E(Tc_ex)
    f = V18Field(n..., dx, dy, dz, dt, transient, vp(Tc_ex), vp(kx), vp(ky), vp(kz), vp(phc_ex), vp(wN), vp(ρC_ϕρ),
                 vp(PC), vp(Xc0), Ttop, norm_scale, vp(Tc0), qyS, Qt, C_NULL, C_NULL, C_NULL, 0, 0, 0, 0, 0, C_NULL)
    return v18_solve!(ht3d_v18_thermal_solve, "ht3d_v18_thermal_solve", f, params,
                      Tc_ex, Tc0, ρC_ϕρ, kx, ky, kz, phc_ex, wN, PC, Xc0)
end
# Darcy loop: Pc_ex updated; ρC_ϕρ = ϕdρdP on entry, ϕρf on exit; Xc0 = ϕρf_old; k_ρf = fluid density
function v18_darcy_solve!(Pc_ex, Tc_ex, ρC_ϕρ, Xc0, k_ρf, kx, ky, kz, phc_ex, wN, PC; dx, dy, dz, dt, transient, Ptop,
                          g, ϕi, scale_ρ, scale_T, scale_σ, norm_scale, params::DyrelParams)
    n = cells_E(Pc_ex)
    f = V18Field(n..., dx, dy, dz, dt, transient, vp(Pc_ex), vp(kx), vp(ky), vp(kz), vp(phc_ex), vp(wN), vp(ρC_ϕρ),
                 vp(PC), C_NULL, Ptop, norm_scale, C_NULL, 0, 0, vp(Tc_ex), vp(Xc0), vp(k_ρf), g, ϕi, scale_ρ,
                 scale_T, scale_σ, C_NULL)
    return v18_solve!(ht3d_v18_darcy_solve, "ht3d_v18_darcy_solve", f, params,
                      Pc_ex, Tc_ex, ρC_ϕρ, Xc0, k_ρf, kx, ky, kz, phc_ex, wN, PC)
end

end # module

using .HT3D_HIP
