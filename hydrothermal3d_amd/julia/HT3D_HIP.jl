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
