# HmatrixAMD.jl -- Julia binding of libhmat_amd.so for Hmatrix.jl users.
#
# Keeps the reference's surface: build with Hmatrix.jl's `Hparams` / `NewHmat`
# (src/cluster.jl:36-43), then `G = GpuHmat(H)` and `G * y`, `G * X`,
# `y' * G` run on the MI355X.  Vectors may be host arrays (copied in and out,
# the drop-in for `H*y`, src/harithm.jl:317-329) or AMDGPU.jl `ROCArray`s
# (device-resident, no copies).  A ChainRulesCore `rrule` differentiates
# `G * x` with respect to x (G' * ȳ) and to the factors (a Tangent whose
# `data` is the gradient over `flatten`'s arena: C -> ȳ x', A -> ȳ (B'x)',
# B -> x (A'ȳ)', from `hm_vjp`), and `adjoint_gradients` returns dL/dx and
# the leaf/factor gradients of L = 0.5*||G x - b||^2.
#
# This file is shipped as source: Julia is not part of the build image, so
# the tested counterpart of every call below is the Python ctypes harness
# (hmatrix.jl_amd/_native.py, device.py) over the same C ABI.
module HmatrixAMD

using Hmatrix
using AMDGPU
using ChainRulesCore

export GpuHmat, flatten, adjoint_gradients, autotune

const libhmat = normpath(joinpath(@__DIR__, "..", "libhmat_amd.so"))

const STORAGE = Dict(:f64 => Cint(0), :f32 => Cint(1), :bf16 => Cint(2), :bf16x2 => Cint(3))

function check(rc::Cint, fn::String)
    if rc != 0
        msg = unsafe_string(ccall((:hm_last_error, libhmat), Cstring, ()))
        error("$fn -> status $rc: $msg")
    end
end

"""
    flatten(H::Hmat) -> (blocks::Matrix{Int64}, data::Vector{Float64})

Block table (8 x nleaves, i.e. row-major nleaves x 8 for the C ABI) and FP64
arena of the leaves of `H`, depth-first in the order of `hmat_matvec!`
(children [1,1], [1,2], [2,1], [2,2]; src/harithm.jl:346-352).  Offsets are
0-based, factors column-major as stored in `Hmat.C`, `Hmat.A`, `Hmat.B`.
"""
function flatten(H::Hmat)
    rows = Vector{NTuple{8,Int64}}()
    data = Float64[]
    function walk(h::Hmat, r0::Int, c0::Int)
        if h.is_fullmatrix
            push!(rows, (0, r0, c0, h.m, h.n, 0, length(data), 0))
            append!(data, vec(h.C))
        elseif h.is_rkmatrix
            k = size(h.A, 2)
            oa = length(data); append!(data, vec(h.A))
            ob = length(data); append!(data, vec(h.B))
            push!(rows, (1, r0, c0, h.m, h.n, k, oa, ob))
        else
            m, n = h.children[1, 1].m, h.children[1, 1].n
            walk(h.children[1, 1], r0, c0)
            walk(h.children[1, 2], r0, c0 + n)
            walk(h.children[2, 1], r0 + m, c0)
            walk(h.children[2, 2], r0 + m, c0 + n)
        end
    end
    walk(H, 0, 0)
    blocks = Matrix{Int64}(undef, 8, length(rows))
    for (j, r) in enumerate(rows)
        blocks[:, j] .= r
    end
    return blocks, data
end

mutable struct GpuHmat
    handle::Ptr{Cvoid}
    m::Int
    n::Int
    storage::Symbol
    ndata::Int
    data::Vector{Float64}     # the exported arena (the primal of the factor tangent)
end

"""
    GpuHmat(H::Hmat; storage=:f64)

Upload `H` (as exported by `flatten`) in the given storage (:f64, :f32, :bf16,
:bf16x2).  The plan owns its device copy; it is released by the finalizer.
"""
function GpuHmat(H::Hmat; storage::Symbol = :f64)
    blocks, data = flatten(H)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:hm_plan_create, libhmat), Cint,
               (Int64, Int64, Int64, Ptr{Int64}, Ptr{Float64}, Int64, Cint, Ref{Ptr{Cvoid}}),
               H.m, H.n, size(blocks, 2), blocks, data, length(data), STORAGE[storage], h)
    check(rc, "hm_plan_create")
    G = GpuHmat(h[], H.m, H.n, storage, length(data), data)
    finalizer(G) do g
        ccall((:hm_plan_destroy, libhmat), Cint, (Ptr{Cvoid},), g.handle)
    end
    return G
end

vectype(G::GpuHmat) = G.storage == :f64 ? Float64 : Float32
stream_ptr() = AMDGPU.stream().stream

Base.size(G::GpuHmat) = (G.m, G.n)

# device-resident products: y = beta*y + alpha*op(G)*x
function matvec!(y::ROCVector, G::GpuHmat, x::ROCVector; alpha = 1.0, beta = 0.0, transpose = false)
    fn = transpose ? :hm_matvec_t : :hm_matvec
    rc = ccall((fn, libhmat), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Cvoid}),
               G.handle, pointer(x), pointer(y), alpha, beta, stream_ptr())
    check(rc, string(fn))
    return y
end

function matmat!(Y::ROCMatrix, G::GpuHmat, X::ROCMatrix; alpha = 1.0, beta = 0.0, transpose = false)
    fn = transpose ? :hm_matmat_t : :hm_matmat
    rc = ccall((fn, libhmat), Cint,
               (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Cdouble, Cdouble, Ptr{Cvoid}),
               G.handle, size(X, 2), pointer(X), stride(X, 2), pointer(Y), stride(Y, 2), alpha, beta,
               stream_ptr())
    check(rc, string(fn))
    return Y
end

# `hmat_matvec!(r, H, v, s)` (src/harithm.jl:340): r += s*H*v
hmat_matvec!(r::ROCVector, G::GpuHmat, v::ROCVector, s::Float64) = matvec!(r, G, v; alpha = s, beta = 1.0)

Base.:*(G::GpuHmat, x::ROCVector) = matvec!(similar(x, vectype(G), G.m), G, x)
Base.:*(G::GpuHmat, X::ROCMatrix) = matmat!(similar(X, vectype(G), G.m, size(X, 2)), G, X)

# host arrays: the drop-in for the reference's `H*y` / `H*X` / `v*H`
Base.:*(G::GpuHmat, x::Vector{Float64}) = Array(G * ROCArray(vectype(G).(x)))
Base.:*(G::GpuHmat, X::Matrix{Float64}) = Array(G * ROCArray(vectype(G).(X)))
function Base.:*(v::AbstractMatrix{Float64}, G::GpuHmat)        # p x m  ->  p x n
    Vt = ROCArray(vectype(G).(permutedims(v)))
    Y = matmat!(similar(Vt, vectype(G), G.n, size(v, 1)), G, Vt; transpose = true)
    return permutedims(Array(Y))
end

# differentiation with respect to x and to the factors: the pullback of ȳ is
# x̄ = G' ȳ and the gradient of <ȳ, G x> over every arena entry (hm_vjp)
function ChainRulesCore.rrule(::typeof(*), G::GpuHmat, x::ROCVector)
    T = vectype(G)
    xs = T.(x)
    y = G * xs
    function pullback(ȳ)
        g = ROCArray(T.(unthunk(ȳ)))
        x̄ = similar(xs, T, G.n)
        grad = AMDGPU.zeros(T, G.ndata)
        rc = ccall((:hm_vjp, libhmat), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                   G.handle, pointer(xs), pointer(g), pointer(x̄), pointer(grad), stream_ptr())
        check(rc, "hm_vjp")
        return NoTangent(), Tangent{GpuHmat}(; data = Array(grad)), x̄
    end
    return y, pullback
end

"""
    adjoint_gradients(G, x, b) -> (L, res, dx, grad)

L = 0.5*||G x - b||^2 with res = G x - b, dx = G' res and `grad` the
gradient with respect to every leaf entry in the layout of `flatten`'s arena
(dense C -> res x', low-rank A -> res (B'x)', B -> x (A'res)').
"""
function adjoint_gradients(G::GpuHmat, x::ROCVector, b::ROCVector)
    T = vectype(G)
    res = similar(b, T, G.m)
    dx = similar(x, T, G.n)
    grad = AMDGPU.zeros(T, G.ndata)
    L = Ref{Cdouble}(0)
    rc = ccall((:hm_adjoint, libhmat), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}, Ptr{Cvoid}),
               G.handle, pointer(x), pointer(b), pointer(res), pointer(dx), pointer(grad), L, stream_ptr())
    check(rc, "hm_adjoint")
    return L[], res, dx, grad
end

# struct hm_config (include/hmat_amd.h)
struct HmConfig
    proj_chunk::Int32
    row_waves::Int32
    flags::Int32
    strip_shape::Int32
end

"""
    autotune(H::Hmat; storage=:f64, nrhs=1, reps=20) -> (config::HmConfig, ms)

The library's sweep of plan shapes (`hm_autotune`): for `nrhs == 1`
`proj_chunk` x `row_waves`, for `nrhs` in 2:32 (storage :f32, :bf16 or
:bf16x2) the multi-right-hand-side `strip_shape`.
"""
function autotune(H::Hmat; storage::Symbol = :f64, nrhs::Integer = 1, reps::Integer = 20)
    blocks, data = flatten(H)
    cfg = Ref(HmConfig(0, 0, 0, 0))
    ms = Ref{Cdouble}(0)
    rc = ccall((:hm_autotune, libhmat), Cint,
               (Int64, Int64, Int64, Ptr{Int64}, Ptr{Float64}, Int64, Cint, Int64, Cint, Ref{HmConfig},
                Ref{Cdouble}),
               H.m, H.n, size(blocks, 2), blocks, data, length(data), STORAGE[storage], Int64(nrhs),
               Cint(reps), cfg, ms)
    check(rc, "hm_autotune")
    return cfg[], ms[]
end

end # module
