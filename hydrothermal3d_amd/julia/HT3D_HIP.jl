# HT3D_HIP.jl -- one-include swap of a HydroThermal3D driver onto libht3d.so.
#
# A driver such as Dimensionless/HydroThermal3D_Alban_v6.jl replaces its
# ParallelStencil set-up (`using ParallelStencil ... @init_parallel_stencil(...)`,
# lines 5-16) with
#
#     include("path/to/hydrothermal3d_amd/julia/HT3D_HIP.jl")
#
# and keeps everything else.  This file provides what the driver uses from
# ParallelStencil: `Data.Array`, `Data.Number`, `@zeros`, `@ones`, `@parallel`
# (launch configurations are dropped; kernel *definitions* written with
# `@parallel function` / `@parallel_indices` are swallowed because the library
# provides every kernel of the hot path) and the host reductions of
# tools/Macros.jl:20-25 on device arrays, including the `abs.(A)` broadcast the
# drivers pass to them.  Calls go through `ccall` into the C ABI of
# include/ht3d.h: arrays as device pointers, cell extents appended.
#
# Status: this file is not executed by the test suite (no Julia toolchain in the
# build image); the C ABI it binds is, through the Python ctypes binding.

module HT3D_HIP

export Data, DevArray, @zeros, @ones, @parallel, @parallel_indices,
       maximum_g, minimum_g, mean_g, sum_g

const libht3d = get(ENV, "HT3D_LIB", joinpath(@__DIR__, "..", "libht3d.so"))

function check(rc::Cint, name)
    rc == 0 && return nothing
    rc == 10 && return nothing      # nitmax: reported, not an error (the reference is silent)
    msg = unsafe_string(ccall((:ht3d_last_error, libht3d), Cstring, ()))
    error("$name failed with status $rc: $msg")
end

function __init__()
    dev = parse(Int, get(ENV, "HT3D_DEVICE", "0"))
    check(ccall((:ht3d_init, libht3d), Cint, (Cint,), dev), "ht3d_init")
end

# --------------------------------------------------------------- device arrays
mutable struct DevArray{N} <: AbstractArray{Float64,N}
    ptr::Ptr{Float64}
    dims::NTuple{N,Int}
    function DevArray{N}(dims::NTuple{N,Int}) where {N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:ht3d_alloc, libht3d), Cint, (Csize_t, Ptr{Ptr{Cvoid}}), 8 * prod(dims), p), "ht3d_alloc")
        A = new{N}(Ptr{Float64}(p[]), dims)
        finalizer(a -> ccall((:ht3d_free, libht3d), Cint, (Ptr{Cvoid},), a.ptr), A)
        return A
    end
end
Base.size(A::DevArray) = A.dims
Base.getindex(::DevArray, ::Int...) = error("scalar indexing of a device array; copy it with Array(A) first")
Base.Array(A::DevArray) = (h = Array{Float64}(undef, A.dims);
    check(ccall((:ht3d_d2h, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), h, A.ptr, 8 * length(h)), "ht3d_d2h"); h)
function upload(h::AbstractArray{<:Real})
    A = DevArray{ndims(h)}(size(h))
    hh = Array{Float64}(h)
    check(ccall((:ht3d_h2d, libht3d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), A.ptr, hh, 8 * length(hh)), "ht3d_h2d")
    return A
end
Base.vec(A::DevArray) = A              # `Ft[:]` in the norms of HydroThermal3D_v7.jl:180
Base.getindex(A::DevArray, ::Colon) = A

module Data
    import ..HT3D_HIP: DevArray, upload
    const Number = Float64
    const Array = DevArray
    Array(h::AbstractArray{<:Real}) = upload(h)
end

macro zeros(dims...)
    :(let A = DevArray{$(length(dims))}(($(esc.(dims)...),));
        check(ccall((:ht3d_memset, libht3d), Cint, (Ptr{Cvoid}, Cint, Csize_t), A.ptr, 0, 8 * length(A)), "ht3d_memset"); A end)
end
macro ones(dims...)
    :(upload(ones($(esc.(dims)...))))
end

# `@parallel f!(args...)` / `@parallel blocks threads f!(args...)`: call the
# library method; `@parallel function f!(...) ... end`: swallowed.
macro parallel(args...)
    ex = args[end]
    if ex isa Expr && ex.head in (:function, :(=)) && ex.args[1] isa Expr && ex.args[1].head == :call
        return nothing
    end
    return esc(ex)
end
macro parallel_indices(args...)
    return nothing
end

# --------------------------------------------------------------- reductions
struct AbsOf
    A::DevArray
end
Base.Broadcast.broadcasted(::typeof(abs), A::DevArray) = AbsOf(A)
reduce_dev(A::DevArray, op) = (r = Ref{Float64}(0.0);
    check(ccall((:ht3d_reduce, libht3d), Cint, (Ref{Float64}, Ptr{Cvoid}, Csize_t, Cint), r, A.ptr, length(A), op), "ht3d_reduce"); r[])
maximum_g(A::DevArray) = reduce_dev(A, 2)
minimum_g(A::DevArray) = reduce_dev(A, 3)
sum_g(A::DevArray) = reduce_dev(A, 0)
mean_g(A::DevArray) = reduce_dev(A, 0) / length(A)
maximum_g(a::AbsOf) = reduce_dev(a.A, 4)
mean_g(a::AbsOf) = reduce_dev(a.A, 1) / length(a.A)
sum_g(a::AbsOf) = reduce_dev(a.A, 1)
maximum_g(A) = maximum(A)
minimum_g(A) = minimum(A)
mean_g(A) = sum(A) / length(A)
sum_g(A) = sum(A)

# --------------------------------------------------------------- kernels
# Cell extents from an array of known staggering: E arrays are (nx+2, ny+2, nz+2).
cells_E(A::DevArray) = (A.dims[1] - 2, A.dims[2] - 2, A.dims[3] - 2)
cells_c(A::DevArray) = A.dims
P(A::DevArray) = A.ptr

for (jl, c, grid) in (
        (:ComputeFlux!, :ht3d_ComputeFlux, :(cells_E(args[7]))),
        (:InitThermal!, :ht3d_InitThermal, :(cells_c(args[1]))),
        (:InitDarcy!, :ht3d_InitDarcy, :(cells_E(args[5]))),
        (:Compute_Rp!, :ht3d_Compute_Rp, :(cells_c(args[1]))),
        (:UpdateT!, :ht3d_UpdateT, :(cells_c(args[1]))),
        (:ResidualDiffusion!, :ht3d_ResidualDiffusion, :(cells_c(args[1]))),
        (:ResidualDiffusionTransient!, :ht3d_ResidualDiffusionTransient, :(cells_c(args[1]))),
        (:UpdateP!, :ht3d_UpdateP, :(cells_c(args[1]))),
        (:DampedUpdate!, :ht3d_DampedUpdate, :(cells_c(args[1]))),
        (:DYRELUpdate!, :ht3d_DYRELUpdate, :(cells_c(args[1]))),
        (:GershgorinPoisson!, :ht3d_GershgorinPoisson, :(cells_c(args[1]))),
        (:Init_vel!, :ht3d_Init_vel, :(cells_E(args[7]))),
        (:KillNorthSouthFluxes!, :ht3d_KillNorthSouthFluxes, :((args[1].dims[1], args[1].dims[2] - 1, args[1].dims[3]))),
        (:Cpy_inn_to_all!, :ht3d_Cpy_inn_to_all, :(cells_c(args[1]))),
        (:Cpy_all_to_inn!, :ht3d_Cpy_all_to_inn, :(cells_c(args[2]))),
        (:Multiply, :ht3d_Multiply, :(cells_c(args[1]))),
        (:VxPlusMinus!, :ht3d_VxPlusMinus, :(cells_c(args[1]))),
        (:VyPlusMinus!, :ht3d_VyPlusMinus, :(cells_c(args[1]))),
        (:VzPlusMinus!, :ht3d_VzPlusMinus, :(cells_c(args[1]))),
        (:dFdx_Weno5!, :ht3d_dFdx_Weno5, :(cells_c(args[1]))),
        (:TimeAveraging!, :ht3d_TimeAveraging, :(cells_c(args[2]))),
        # kernels_HT3D_dimensional.jl (the v18 driver)
        (:InitConductivity!, :ht3d_InitConductivity, :(cells_c(args[1]) .- (1, 0, 0))),
        (:SwapDYREL!, :ht3d_SwapDYREL, :(cells_c(args[2]))),
        (:GershgorinHydro!, :ht3d_GershgorinHydro, :(cells_c(args[1]))),
        (:RayleighQuotientNumerator!, :ht3d_RayleighQuotientNumerator, :(cells_c(args[1]))),
        (:RayleighQuotientDenominator!, :ht3d_RayleighQuotientDenominator, :(cells_c(args[1]))),
        (:ResidualTemperatureLinearised!, :ht3d_ResidualTemperatureLinearised, :(cells_c(args[1]))),
        (:ResidualFluidPressure!, :ht3d_ResidualFluidPressure, :(cells_c(args[1]))),
        (:ComputeDarcyFlux!, :ht3d_ComputeDarcyFlux, :(cells_E(args[8]))),
        (:ComputeThermalConductivity, :ht3d_ComputeThermalConductivity, :(cells_E(args[2]))),
        (:ComputeFluidDensity, :ht3d_ComputeFluidDensity, :(cells_E(args[2]))),
        (:SmoothConductivityV2C, :ht3d_SmoothConductivityV2C, :(cells_c(args[1]))),
        (:SmoothConductivityC2V, :ht3d_SmoothConductivityC2V, :(cells_c(args[2]))),
        (:ComputeρCeffective!, :ht3d_ComputeRhoCeffective, :(cells_c(args[1]))),
        (:ϕρf, :ht3d_PhiRhof, :(cells_c(args[1]))),
        (:ϕρf_old, :ht3d_PhiRhof_old, :(cells_c(args[1]))),
        (:ϕdρdP!, :ht3d_PhiDrhoDP, :(cells_c(args[1]))),
    )
    @eval function $jl(args...)
        g = $grid
        cargs = map(a -> a isa DevArray ? a.ptr : Float64(a), args)
        types = map(a -> a isa Ptr ? Ptr{Cvoid} : Float64, cargs)
        rc = ccall(($(QuoteNode(c)), libht3d), Cint, ($(Expr(:..., :types))..., Cint, Cint, Cint),
                   cargs..., g[1], g[2], g[3])
        check(rc, $(string(c)))
    end
end

# The two ghost fills the closed-box drivers define (Alban_v6:108-130)
SetTemperatureBCs!(T::DevArray, bot, top) = check(ccall((:ht3d_SetTemperatureBCs, libht3d), Cint,
    (Ptr{Cvoid}, Float64, Float64, Cint, Cint, Cint), T.ptr, bot, top, cells_E(T)...), "ht3d_SetTemperatureBCs")
SetPressureBCs!(Pc::DevArray, bot, top) = check(ccall((:ht3d_SetPressureBCs, libht3d), Cint,
    (Ptr{Cvoid}, Float64, Float64, Cint, Cint, Cint), Pc.ptr, bot, top, cells_E(Pc)...), "ht3d_SetPressureBCs")
ResetA!(A::DevArray, B::DevArray) = check(ccall((:ht3d_ResetA, libht3d), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint), A.ptr, B.ptr, A.dims...), "ht3d_ResetA")
ArrayEqualArray!(A::DevArray, B::DevArray) = check(ccall((:ht3d_ArrayEqualArray, libht3d), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint), A.ptr, B.ptr, A.dims...), "ht3d_ArrayEqualArray")

# Dimensional kernels whose names the dimensionless files also use: the v18
# driver's methods differ in arity, so Julia dispatch picks the right entry.
InitThermal!(Tc0::DevArray, Tc_ex::DevArray) = check(ccall((:ht3d_Cpy_inn_to_all, libht3d), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint), Tc0.ptr, Tc_ex.ptr, cells_c(Tc0)...), "InitThermal!")
InitDarcy!(A::DevArray, B::DevArray) = ArrayEqualArray!(A, B)
function Init_vel!(Vx::DevArray, Vy::DevArray, Vz::DevArray, qx::DevArray, qy::DevArray, qz::DevArray, ρfv::DevArray)
    check(ccall((:ht3d_dim_Init_vel, libht3d), Cint, (ntuple(_ -> Ptr{Cvoid}, 7)..., Cint, Cint, Cint),
                Vx.ptr, Vy.ptr, Vz.ptr, qx.ptr, qy.ptr, qz.ptr, ρfv.ptr, (Vy.dims .- (0, 1, 0))...), "Init_vel!")
end
function SetTemperatureBCs!(Tc_ex::DevArray, ky::DevArray, wN::DevArray, phc_ex::DevArray, s...)
    length(s) == 15 || error("SetTemperatureBCs!: 19 arguments expected")
    check(ccall((:ht3d_dim_SetTemperatureBCs, libht3d), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, ntuple(_ -> Float64, 15)..., Cint, Cint, Cint),
                Tc_ex.ptr, ky.ptr, wN.ptr, phc_ex.ptr, Float64.(s)..., cells_E(Tc_ex)...), "SetTemperatureBCs!")
end
function SetPressureBCs!(Pc_ex::DevArray, wN::DevArray, phc_ex::DevArray, s...)
    length(s) == 10 || error("SetPressureBCs!: 13 arguments expected")
    check(ccall((:ht3d_dim_SetPressureBCs, libht3d), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, ntuple(_ -> Float64, 10)..., Cint, Cint, Cint),
                Pc_ex.ptr, wN.ptr, phc_ex.ptr, Float64.(s)..., cells_E(Pc_ex)...), "SetPressureBCs!")
end
function ComputeHydroConductivity(kfv::DevArray, Tce::DevArray, Pce::DevArray, phv::DevArray, ymin, Δy, k_fact, δ,
                                  scale_σ, scale_T, scale_L, scale_t, scale_η, mode::Integer)
    check(ccall((:ht3d_ComputeHydroConductivity, libht3d), Cint,
                (ntuple(_ -> Ptr{Cvoid}, 4)..., ntuple(_ -> Float64, 9)..., Cint, Cint, Cint, Cint),
                kfv.ptr, Tce.ptr, Pce.ptr, phv.ptr, ymin, Δy, k_fact, δ, scale_σ, scale_T, scale_L, scale_t, scale_η,
                mode, cells_E(Tce)...), "ComputeHydroConductivity")
end

# The two DYREL loops of HydroThermal3D_v18.jl (:443-478 heat, :500-537 Darcy),
# fused: a driver replaces the `for iter = 1:nitmax ... end` block (and the
# GershgorinPoisson!/ResetA! set-up before it) by one call.  `field` is the
# library's ht3d_v18_field layout, `params` (nitmax, nitout, ncheck, norm, tol,
# CFL, cfact, lmin_div); returns (iters, status, lmin, nF_abs, nF_rel).
struct DyrelParams
    nitmax::Cint; nitout::Cint; ncheck::Cint; norm::Cint
    tol::Float64; CFL::Float64; cfact::Float64; lmin_div::Float64
end
mutable struct DyrelReport
    iters::Cint; status::Cint; n_checks::Cint; n_lambda::Cint
    lmax::Float64; lmin::Float64; dtau::Float64; h1::Float64; h2::Float64
    nF_abs::Float64; nF_rel::Float64; nF_ini::Float64; ncheck::Cint
    DyrelReport() = new(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
end
struct V18Field
    ncx::Cint; ncy::Cint; ncz::Cint
    dx::Float64; dy::Float64; dz::Float64; dt::Float64; transient::Float64
    X::Ptr{Cvoid}; kx::Ptr{Cvoid}; ky::Ptr{Cvoid}; kz::Ptr{Cvoid}; phc_ex::Ptr{Cvoid}; wN::Ptr{Cvoid}
    rhoC::Ptr{Cvoid}; PC::Ptr{Cvoid}; G::Ptr{Cvoid}; top::Float64; norm_scale::Float64
    Tc_old::Ptr{Cvoid}; qyS::Float64; Qt::Float64
    Tc_ex::Ptr{Cvoid}; phirhof0::Ptr{Cvoid}; rhofv::Ptr{Cvoid}
    g::Float64; phi_i::Float64; scale_rho::Float64; scale_T::Float64; scale_sigma::Float64
end
function v18_solve!(name::Symbol, f::V18Field, p::DyrelParams)
    r = DyrelReport()
    rc = ccall((name, libht3d), Cint, (Ref{V18Field}, Ref{DyrelParams}, Ref{DyrelReport}), f, p, r)
    rc == 11 && error("Nan ...")          # v18:464
    rc == 12 && error("Diverged ...")     # v18:465
    rc in (0, 10) || check(rc, string(name))
    return r
end
# heat loop: Tc_ex updated; ρC_ϕρ = rho*C (ComputeρCeffective!); PC, Xc0 receive D and G
function v18_thermal_solve!(Tc_ex, Tc0, ρC_ϕρ, kx, ky, kz, phc_ex, wN, PC, Xc0; dx, dy, dz, dt, transient, qyS, Ttop,
                            Qt, norm_scale, params::DyrelParams)
    n = cells_E(Tc_ex)
    f = V18Field(n..., dx, dy, dz, dt, transient, Tc_ex.ptr, kx.ptr, ky.ptr, kz.ptr, phc_ex.ptr, wN.ptr, ρC_ϕρ.ptr,
                 PC.ptr, Xc0.ptr, Ttop, norm_scale, Tc0.ptr, qyS, Qt, C_NULL, C_NULL, C_NULL, 0, 0, 0, 0, 0)
    return v18_solve!(:ht3d_v18_thermal_solve, f, params)
end
# Darcy loop: Pc_ex updated; ρC_ϕρ = ϕdρdP on entry, ϕρf on exit; Xc0 = ϕρf_old; k_ρf = fluid density
function v18_darcy_solve!(Pc_ex, Tc_ex, ρC_ϕρ, Xc0, k_ρf, kx, ky, kz, phc_ex, wN, PC; dx, dy, dz, dt, transient, Ptop,
                          g, ϕi, scale_ρ, scale_T, scale_σ, norm_scale, params::DyrelParams)
    n = cells_E(Pc_ex)
    f = V18Field(n..., dx, dy, dz, dt, transient, Pc_ex.ptr, kx.ptr, ky.ptr, kz.ptr, phc_ex.ptr, wN.ptr, ρC_ϕρ.ptr,
                 PC.ptr, C_NULL, Ptop, norm_scale, C_NULL, 0, 0, Tc_ex.ptr, Xc0.ptr, k_ρf.ptr, g, ϕi, scale_ρ,
                 scale_T, scale_σ)
    return v18_solve!(:ht3d_v18_darcy_solve, f, params)
end

# WENO5: the v1 orchestration (kernels_HT3D.jl:156-233) and v2 (Weno5_Routines_v2.jl:1-50)
# as one fused call each; the scratch arguments of the reference signatures are accepted.
function AdvectWithWeno5(Tc, Tc_ex, Tc_exxx, Told, dTdxm, dTdxp, Vxm, Vxp, Vym, Vyp, Vzm, Vzp, Vx, Vy, Vz,
                         v1, v2, v3, v4, v5, kx, ky, kz, dt, _dx, _dy, _dz, Ttop, Tbot, Pc_ex, Ty, Ra)
    Init_vel!(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, _dx, _dy, _dz)
    check(ccall((:ht3d_weno5_advect, libht3d), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Float64,
                 Cint, Cint, Cint, Cint),
                Tc_ex.ptr, Vx.ptr, Vy.ptr, Vz.ptr, dt, _dx, _dy, _dz, Ttop, Tbot, 1, cells_E(Tc_ex)...), "AdvectWithWeno5")
end
function AdvectWithWeno5_v2(Tc_ex, Tc_exxx, Told, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot)
    check(ccall((:ht3d_weno5_advect, libht3d), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Float64,
                 Cint, Cint, Cint, Cint),
                Tc_ex.ptr, Vx.ptr, Vy.ptr, Vz.ptr, dt, _dx, _dy, _dz, Ttop, Tbot, 2, cells_E(Tc_ex)...), "AdvectWithWeno5_v2")
end

end # module

using .HT3D_HIP
