# Julia binding of libtnhip.so for TensorNetworks.jl (ccall over the C-ABI in
# include/tnhip.h). It adds a GPU backend for the package's MPS contraction hot path:
# two-site DMRG, the truncated SVD and two-site split, centre moves, TEBD gates and the
# TEBD driver, quantum-jump trajectories, and the gradient of <psi|H(c)|psi>, keeping the
# package's own types (GMPS, OpList, GateList, Sitetypes) and keyword arguments.
#
# Routing: the reference's functions get methods whose first argument is the backend
# marker `GPU()`, e.g. `dmrg(GPU(), psi, H; maxdim=64)` or
# `tebd(GPU(), st, psi, H, dt, tmax, save, observers; cutoff=1e-10)`; the other arguments
# and keyword arguments are the reference's. `GPUProjMPS` plugs into the reference's own
# AbstractProjMPS seam instead.
#
# Not executed where this file was written (no Julia toolchain there); the Python module
# tensornetworks.jl_amd/tnhip exercises the same C-ABI calls, and a CPU test checks every
# ccall below against the declarations of include/tnhip.h.
module TensorNetworksAMD

using LinearAlgebra
using Printf
using HDF5
using TensorNetworks
import TensorNetworks: AbstractProjMPS, GMPS, GateList, OpList, Sitetypes, product,
                       buildleft!, buildright!, calculate, movecenter!, dmrg, tebd, svd,
                       replacesites!, applygate!, applygates!, measure!, checkdone!,
                       trotterize, maxbonddim, op, MPO, applyMPO, inner

const libtnhip = get(ENV, "TNHIP_LIB", joinpath(@__DIR__, "..", "..", "..", "tnhip", "libtnhip.so"))

const TNH_F64, TNH_F32, TNH_C128, TNH_C64 = Cint(0), Cint(1), Cint(2), Cint(3)
const PRECISION = Dict(:f64 => Cint(0), :f32 => Cint(1), :bf16x3 => Cint(2))

# --------------------------------------------------------------------- context
mutable struct Context
    h::Ptr{Cvoid}
end

function check(rc::Cint, ctx::Ptr{Cvoid}, fn::AbstractString)
    rc == 0 && return nothing
    msg = ctx == C_NULL ? "" :
          unsafe_string(ccall((:tnh_last_error, libtnhip), Cstring, (Ptr{Cvoid},), ctx))
    error("$fn: status $rc: $msg")  # the reference reports failures with error("...")
end

function Context(device::Integer = 0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:tnh_ctx_create, libtnhip), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r),
          C_NULL, "tnh_ctx_create")
    c = Context(r[])
    finalizer(x -> ccall((:tnh_ctx_destroy, libtnhip), Cint, (Ptr{Cvoid},), x.h), c)
    return c
end

const DEFAULT = Ref{Union{Nothing,Context}}(nothing)
default_context() = something(DEFAULT[], (DEFAULT[] = Context(0)))

synchronize(ctx::Context) =
    check(ccall((:tnh_ctx_synchronize, libtnhip), Cint, (Ptr{Cvoid},), ctx.h), ctx.h,
          "tnh_ctx_synchronize")

"""
    GPU(ctx = default_context())

Backend marker: the first argument of the methods of this module that route the
reference's `dmrg`, `tebd`, `applygates!`, `applygate!`, `replacesites!`, `svd` and
`movecenter!` onto libtnhip.
"""
struct GPU
    ctx::Context
end
GPU() = GPU(default_context())

# GEMM tile selection of a context: autotuned per size class (default) or the size rule.
gemm_autotune!(ctx::Context, on::Bool) =
    check(ccall((:tnh_gemm_autotune, libtnhip), Cint, (Ptr{Cvoid}, Cint), ctx.h, on ? 1 : 0),
          ctx.h, "tnh_gemm_autotune")

function gemm_tuning_report(ctx::Context)
    need = Ref{Int64}(0)
    check(ccall((:tnh_gemm_tuning_report, libtnhip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64, Ptr{Int64}),
                ctx.h, C_NULL, 0, need), ctx.h, "tnh_gemm_tuning_report")
    buf = Vector{UInt8}(undef, need[])
    GC.@preserve buf check(ccall((:tnh_gemm_tuning_report, libtnhip), Cint,
                                 (Ptr{Cvoid}, Ptr{UInt8}, Int64, Ptr{Int64}), ctx.h, buf, need[], need),
                           ctx.h, "tnh_gemm_tuning_report")
    return unsafe_string(pointer(buf))
end

# --------------------------------------------------------------------- buffers
mutable struct DeviceTensor
    ctx::Context
    h::Ptr{Cvoid}
end

function free!(t::DeviceTensor)
    t.h == C_NULL && return
    ccall((:tnh_buf_free, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), t.ctx.h, t.h)
    t.h = C_NULL
end

# a handle the library returned to us: freed with the Julia object
function owned(ctx::Context, h::Ptr{Cvoid})
    t = DeviceTensor(ctx, h)
    finalizer(free!, t)
    return t
end

# Julia arrays are column-major, exactly the layout the library expects.
function upload(ctx::Context, A::AbstractArray{<:Number}; real_if_possible::Bool = true)
    B = Array{ComplexF64}(A)
    dev = (real_if_possible && all(iszero ∘ imag, B)) ? TNH_F64 : TNH_C128
    dims = Int64[size(B)...]
    r = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve B dims begin
        check(ccall((:tnh_buf_upload, libtnhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Int64}, Ptr{Ptr{Cvoid}}),
                    ctx.h, pointer(B), TNH_C128, dev, ndims(B), pointer(dims), r),
              ctx.h, "tnh_buf_upload")
    end
    return owned(ctx, r[])
end

function shape(h::Ptr{Cvoid})
    dt = Ref{Cint}(0); nd = Ref{Cint}(0); dims = zeros(Int64, 8)
    ccall((:tnh_buf_shape, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cint}, Ptr{Cint}, Ptr{Int64}),
          h, dt, nd, dims)
    return Tuple(dims[1:nd[]])
end
shape(t::DeviceTensor) = shape(t.h)

function to_host(ctx::Context, h::Ptr{Cvoid})
    A = Array{ComplexF64}(undef, shape(h)...)
    GC.@preserve A check(ccall((:tnh_buf_download, libtnhip), Cint,
                               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                               ctx.h, h, pointer(A), TNH_C128), ctx.h, "tnh_buf_download")
    return A
end
to_host(t::DeviceTensor) = to_host(t.ctx, t.h)

# ------------------------------------------- AbstractProjMPS plug-in (the seam)
# GPUProjMPS slots into ProjMPSSum next to the package's ProjMPS: dmrg talks to it only
# through movecenter!, product and calculate (src/algorithms/mps/dmrg.jl:22,41,50,62).
mutable struct GPUProjMPS <: AbstractProjMPS
    objects::Vector{GMPS}                 # [psi, H, psi], held by reference
    blocks::Vector{Union{Nothing,DeviceTensor}}
    W::Vector{DeviceTensor}
    squared::Bool
    rank::Int
    center::Int
    coeff::Float64
    ctx::Context
end

function GPUProjMPS(psi::GMPS, H::GMPS; coeff::Real = 1.0, ctx::Context = default_context())
    W = [upload(ctx, H[i]) for i = 1:length(H)]
    p = GPUProjMPS([psi, H, psi], Union{Nothing,DeviceTensor}[nothing for _ = 1:length(psi)],
                   W, false, 2, 0, Float64(coeff), ctx)
    movecenter!(p, 1)
    return p
end

blockptr(p::GPUProjMPS, idx::Int) =
    (idx < 1 || idx > length(p.blocks) || p.blocks[idx] === nothing) ? C_NULL : p.blocks[idx].h

function _env!(p::GPUProjMPS, idx::Int, left::Bool)
    A = upload(p.ctx, p.objects[1][idx])
    r = Ref{Ptr{Cvoid}}(p.blocks[idx] === nothing ? C_NULL : p.blocks[idx].h)
    prev = blockptr(p, left ? idx - 1 : idx + 1)
    if left
        check(ccall((:tnh_env_left, libtnhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint),
                    p.ctx.h, prev, A.h, p.W[idx].h, r, PRECISION[:f64]), p.ctx.h, "tnh_env_left")
    else
        check(ccall((:tnh_env_right, libtnhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint),
                    p.ctx.h, prev, A.h, p.W[idx].h, r, PRECISION[:f64]), p.ctx.h, "tnh_env_right")
    end
    if p.blocks[idx] === nothing || p.blocks[idx].h != r[]
        p.blocks[idx] === nothing || (p.blocks[idx].h = C_NULL)  # replaced by the library
        p.blocks[idx] = owned(p.ctx, r[])
    end
end

buildleft!(p::GPUProjMPS, idx::Int) = _env!(p, idx, true)     # projmps.jl:50-66
buildright!(p::GPUProjMPS, idx::Int) = _env!(p, idx, false)   # projmps.jl:74-95

function product(p::GPUProjMPS, A, direction::Bool = false, nsites::Int = 2)  # projmps.jl:103-144
    nsites == 2 || error("GPUProjMPS supports nsites = 2")
    site = direction ? p.center - nsites + 1 : p.center
    x = upload(p.ctx, A; real_if_possible = false)
    y = upload(p.ctx, zeros(ComplexF64, size(A)); real_if_possible = false)
    check(ccall((:tnh_heff_apply, libtnhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Cdouble, Cdouble, Cint),
                p.ctx.h, blockptr(p, site - 1), p.W[site].h, p.W[site+1].h,
                blockptr(p, site + nsites), x.h, y.h, p.coeff, 0.0, PRECISION[:f64]),
          p.ctx.h, "tnh_heff_apply")
    return to_host(y)
end

function calculate(p::GPUProjMPS)                                  # projmps.jl:192-216
    re = Ref{Cdouble}(0); im = Ref{Cdouble}(0)
    A = upload(p.ctx, p.objects[1][p.center])
    check(ccall((:tnh_env_calculate, libtnhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble,
                 Ptr{Cdouble}, Ptr{Cdouble}),
                p.ctx.h, blockptr(p, p.center - 1), A.h, p.W[p.center].h,
                blockptr(p, p.center + 1), p.coeff, re, im), p.ctx.h, "tnh_env_calculate")
    return complex(re[], im[])
end

# ------------------------------------------------- truncated SVD and the split

"""
    svd(GPU(), x, idx; cutoff=0, maxdim=0, mindim=1) -> (U, S, V)

svd(x, idx; ...), src/tensors.jl:168-227, on the GPU (tnh_svd_trunc: FP64 Gram
eigensolve, one-sided Jacobi fallback): rows are the other indices in order, columns
`idx`; U carries the new bond at position idx, S = diagm(s), V is (k, size(x, idx)).
"""
function svd(g::GPU, x::AbstractArray, idx::Int; cutoff::Real = 0.0, maxdim::Int = 0, mindim::Int = 1)
    sz = ndims(x)
    idx = idx == -1 ? sz : idx
    perm = [[i for i = 1:sz if i != idx]..., idx]
    y = permutedims(Array{ComplexF64}(x), perm)
    M = reshape(y, :, size(x, idx))
    dM = upload(g.ctx, M)
    U = Ref{Ptr{Cvoid}}(C_NULL); S = Ref{Ptr{Cvoid}}(C_NULL); V = Ref{Ptr{Cvoid}}(C_NULL)
    k = Ref{Cint}(0); disc = Ref{Cdouble}(0); fb = Ref{Cint}(0)
    check(ccall((:tnh_svd_trunc, libtnhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Cint, Cint, Ptr{Ptr{Cvoid}},
                 Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cint}, Ptr{Cdouble}, Ptr{Cint}),
                g.ctx.h, dM.h, cutoff, maxdim, mindim, 0, U, S, V, k, disc, fb),
          g.ctx.h, "tnh_svd_trunc")
    Ud, Sd, Vd = owned(g.ctx, U[]), owned(g.ctx, S[]), owned(g.ctx, V[])
    Um = reshape(to_host(Ud), size(y)[1:end-1]..., k[])
    return permutedims(Um, invperm(perm)), diagm(real.(to_host(Sd))), to_host(Vd)
end

# --------------------------------------------------- device-resident MPS

"""
    DeviceMPS(psi::GMPS; ctx) / DeviceMPS(ctx, psi)

The site tensors of a rank-1 GMPS held on the GPU (one library buffer per site) and its
orthogonality centre. `GMPS(d)` downloads it back.
"""
mutable struct DeviceMPS
    ctx::Context
    dim::Int
    sites::Vector{Ptr{Cvoid}}
    center::Int
end

function DeviceMPS(psi::GMPS; ctx::Context = default_context())
    TensorNetworks.rank(psi) == 1 || error("DeviceMPS holds rank-1 GMPS (states).")
    sites = Ptr{Cvoid}[]
    for i = 1:length(psi)
        t = upload(ctx, psi[i])
        push!(sites, t.h)
        t.h = C_NULL  # ownership moves to the DeviceMPS
    end
    d = DeviceMPS(ctx, TensorNetworks.dim(psi), sites, psi.center)
    finalizer(free!, d)
    return d
end

function free!(d::DeviceMPS)
    for h in d.sites
        h == C_NULL || ccall((:tnh_buf_free, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), d.ctx.h, h)
    end
    empty!(d.sites)
end

Base.length(d::DeviceMPS) = length(d.sites)

function TensorNetworks.GMPS(d::DeviceMPS)
    ts = Array{ComplexF64}[to_host(d.ctx, h) for h in d.sites]
    return GMPS(1, d.dim, ts, d.center)
end

# copy a device state back into an existing GMPS (same length)
function store!(psi::GMPS, d::DeviceMPS)
    for i = 1:length(d)
        psi[i] = to_host(d.ctx, d.sites[i])
    end
    psi.center = d.center
    return psi
end

"""
    movecenter!(d::DeviceMPS, idx; cutoff=0, maxdim=0, mindim=1)

movecenter!(psi, idx; kwargs...), src/structures/mps/gmps.jl:90-112, on the device
(tnh_mps_movecenter); bond dimensions may change with truncation.
"""
function movecenter!(d::DeviceMPS, idx::Int; cutoff::Real = 0.0, maxdim::Int = 0, mindim::Int = 1)
    GC.@preserve d check(ccall((:tnh_mps_movecenter, libtnhip), Cint,
                               (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Cint, Cint, Cdouble, Cint, Cint),
                               d.ctx.h, length(d), d.sites, d.center, idx, cutoff, maxdim, mindim),
                         d.ctx.h, "tnh_mps_movecenter")
    d.center = idx
    return d
end

function movecenter!(g::GPU, psi::GMPS, idx::Int; kwargs...)
    d = DeviceMPS(psi; ctx = g.ctx)
    movecenter!(d, idx; kwargs...)
    return store!(psi, d)
end

function device_norm(d::DeviceMPS)                                         # gmps.jl:29-37
    d.center == 0 && movecenter!(d, 1)
    v = Ref{Cdouble}(0)
    check(ccall((:tnh_buf_norm, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}),
                d.ctx.h, d.sites[d.center], v), d.ctx.h, "tnh_buf_norm")
    return v[]
end

function normalize_device!(d::DeviceMPS)                                   # gmps.jl:46-51
    n = device_norm(d)
    check(ccall((:tnh_buf_scale, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble),
                d.ctx.h, d.sites[d.center], 1.0 / n, 0.0), d.ctx.h, "tnh_buf_scale")
    return d
end

"""
    replacesites!(GPU(), psi, A, site, direction=false, normalize=false; cutoff, maxdim, mindim)

replacesites!, src/structures/mps/gmps.jl:199-267, for a rank-1 GMPS: the splits of two
sites (tnh_replacesites2) and of three to six sites (tnh_replacesites_sites) run on the GPU.
"""
function replacesites!(g::GPU, psi::GMPS, A, site::Int, direction::Bool = false,
                       normalize::Bool = false; cutoff::Real = 0.0, maxdim::Int = 0, mindim::Int = 1)
    nsites = ndims(A) - 2
    if nsites == 1                                                  # gmps.jl:204-213
        psi[site] = Array{ComplexF64}(A)
        nxt = site + 1 - 2 * direction
        0 < nxt <= length(psi) && movecenter!(g, psi, nxt)
        normalize && TensorNetworks.normalize!(psi)
        return nothing
    end
    if nsites > 2                                                   # gmps.jl:215-257
        th = upload(g.ctx, A)
        hs = fill(C_NULL, nsites); ks = zeros(Cint, nsites - 1); disc = Ref{Cdouble}(0)
        try
            GC.@preserve hs ks check(ccall((:tnh_replacesites_sites, libtnhip), Cint,
                        (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cint, Cint, Cdouble, Cint, Cint,
                         Ptr{Ptr{Cvoid}}, Ptr{Cint}, Ptr{Cdouble}),
                        g.ctx.h, nsites, th.h, direction, normalize, cutoff, maxdim, mindim, hs,
                        ks, disc), g.ctx.h, "tnh_replacesites_sites")
        finally
            outs = [owned(g.ctx, h) for h in hs if h != C_NULL]  # library-owned results
        end
        for j = 1:nsites
            psi[site+j-1] = to_host(outs[j])
        end
        psi.center = direction ? site : site + nsites - 1
        return true
    end
    th = upload(g.ctx, A)
    a1 = Ref{Ptr{Cvoid}}(C_NULL); a2 = Ref{Ptr{Cvoid}}(C_NULL)
    k = Ref{Cint}(0); disc = Ref{Cdouble}(0)
    check(ccall((:tnh_replacesites2, libtnhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cdouble, Cint, Cint, Ptr{Ptr{Cvoid}},
                 Ptr{Ptr{Cvoid}}, Ptr{Cint}, Ptr{Cdouble}, Ptr{Cdouble}, Cint),
                g.ctx.h, th.h, direction, normalize, cutoff, maxdim, mindim, a1, a2, k, disc,
                C_NULL, 0), g.ctx.h, "tnh_replacesites2")
    A1, A2 = owned(g.ctx, a1[]), owned(g.ctx, a2[])
    psi[site] = to_host(A1)
    psi[site+1] = to_host(A2)
    psi.center = direction ? site : site + 1
    return true
end

# --------------------------------------------------------------- TEBD gates

"""
    DeviceGates(ctx, gl::GateList)

The gates of a GateList (trotterize, src/structures/mps/gatelist.jl:75-121) uploaded once:
per row, the first sites, the gate sizes and one device buffer per gate.
"""
struct DeviceGates
    ctx::Context
    sites::Vector{Vector{Int}}
    sizes::Vector{Vector{Int}}
    bufs::Vector{Vector{DeviceTensor}}
end

gatesize(gate) = div(ndims(gate), 2)

DeviceGates(ctx::Context, gl::GateList) =
    DeviceGates(ctx, [copy(s) for s in gl.sites], [[gatesize(g) for g in row] for row in gl.gates],
                [[upload(ctx, g) for g in row] for row in gl.gates])

"""
    applygate!(d::DeviceMPS, site, gate::DeviceTensor, rng, direction; cutoff, maxdim, mindim)

applygate!(psi, site, gate, direction; kwargs...), gatelist.jl:137-171 with error=false:
a two-site gate and the truncated split without normalisation (tnh_applygate2), a wider
gate and its rng - 1 truncated splits (tnh_applygate_sites), or a one-site gate followed by
replacesites!'s untruncated centre move.
"""
function applygate!(d::DeviceMPS, site::Int, gate::DeviceTensor, rng::Int, direction::Bool;
                    cutoff::Real = 0.0, maxdim::Int = 0, mindim::Int = 1)
    if rng == 2
        k = Ref{Cint}(0); disc = Ref{Cdouble}(0)
        a1 = Ref{Ptr{Cvoid}}(d.sites[site]); a2 = Ref{Ptr{Cvoid}}(d.sites[site+1])
        check(ccall((:tnh_applygate2, libtnhip), Cint,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Cint, Cdouble, Cint,
                     Cint, Ptr{Cint}, Ptr{Cdouble}),
                    d.ctx.h, a1, a2, gate.h, direction, cutoff, maxdim, mindim, k, disc),
              d.ctx.h, "tnh_applygate2")
        d.sites[site], d.sites[site+1] = a1[], a2[]
        d.center = direction ? site : site + 1
    elseif rng == 1
        a = Ref{Ptr{Cvoid}}(d.sites[site])
        check(ccall((:tnh_apply_site_op, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}),
                    d.ctx.h, a, gate.h), d.ctx.h, "tnh_apply_site_op")
        d.sites[site] = a[]
        d.center = site
        nxt = site + 1 - 2 * direction
        0 < nxt <= length(d) && movecenter!(d, nxt)
    else
        # wider gates: the r-site branch of replacesites! (gmps.jl:215-257) on the device
        hs = Ptr{Cvoid}[d.sites[site+j] for j = 0:rng-1]
        ks = zeros(Cint, rng - 1); disc = Ref{Cdouble}(0)
        try
            GC.@preserve hs ks check(ccall((:tnh_applygate_sites, libtnhip), Cint,
                        (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Cint, Cdouble, Cint, Cint,
                         Ptr{Cint}, Ptr{Cdouble}),
                        d.ctx.h, rng, hs, gate.h, direction, cutoff, maxdim, mindim, ks, disc),
                  d.ctx.h, "tnh_applygate_sites")
        finally
            for j = 0:rng-1  # the library replaces the handles in place, also before a failure
                d.sites[site+j] = hs[j+1]
            end
        end
        d.center = direction ? site : site + rng - 1
    end
    return d
end

"""
    applygates!(d::DeviceMPS, gates::DeviceGates; cutoff, maxdim, mindim)

applygates!(psi, gates; kwargs...), gatelist.jl:191-223: per row, the direction that
starts at the end nearer the centre, a truncated centre move before every gate.
"""
function applygates!(d::DeviceMPS, gates::DeviceGates; cutoff::Real = 0.0, maxdim::Int = 0,
                     mindim::Int = 1)
    for row = 1:length(gates.sites)
        sites, sizes = gates.sites[row], gates.sizes[row]
        isempty(sites) && continue
        firstsite = sites[1]
        lastsite = sites[end] + sizes[end] - 1
        direction = abs(d.center - firstsite) < abs(d.center - lastsite) ? false : true
        n = length(sites)
        for i = 1:n
            gi = direction ? n + 1 - i : i
            ctr = direction ? sites[gi] + sizes[gi] - 1 : sites[gi]
            movecenter!(d, ctr; cutoff = cutoff, maxdim = maxdim, mindim = mindim)
            applygate!(d, sites[gi], gates.bufs[row][gi], sizes[gi], direction;
                       cutoff = cutoff, maxdim = maxdim, mindim = mindim)
        end
    end
    return d
end

function applygates!(g::GPU, psi::GMPS, gates::GateList; cutoff::Real = 0.0, maxdim::Int = 0,
                     mindim::Int = 1)
    d = DeviceMPS(psi; ctx = g.ctx)
    applygates!(d, DeviceGates(g.ctx, gates); cutoff = cutoff, maxdim = maxdim, mindim = mindim)
    return store!(psi, d)
end

function applygate!(g::GPU, psi::GMPS, site::Int, gate, direction::Bool = false; cutoff::Real = 0.0,
                    maxdim::Int = 0, mindim::Int = 1)
    d = DeviceMPS(psi; ctx = g.ctx)
    applygate!(d, site, upload(g.ctx, gate), gatesize(gate), direction; cutoff = cutoff,
               maxdim = maxdim, mindim = mindim)
    return store!(psi, d)
end

# first sites, ranges and the operator matrices of every term (identity in a term's gaps)
function _oplist_payload(st::Sitetypes, ol::OpList)
    first = Cint[s[1] for s in ol.sites]
    rng = Cint[s[end] - s[1] + 1 for s in ol.sites]
    mats = ComplexF64[]
    for t = 1:length(ol.sites)
        for s = ol.sites[t][1]:ol.sites[t][end]
            k = findfirst(==(s), ol.sites[t])
            m = k === nothing ? op(st, "id") : op(st, ol.ops[t][k])
            append!(mats, vec(Matrix{ComplexF64}(m)))
        end
    end
    return first, rng, mats
end

"""
    inner_oplist(d::DeviceMPS, st::Sitetypes, ops::OpList) -> Vector{ComplexF64}
    inner_oplist(d::DeviceMPS, st::Sitetypes, ops::OpList, phi::DeviceMPS)

inner(st, psi, oplist, phi), src/structures/mps/mps.jl:87-134: coeff_t <psi|O_t|phi> for
every term (phi = psi in the first form), from environments on the device
(tnh_inner_oplist, tnh_inner_oplist_pair).
"""
function inner_oplist(d::DeviceMPS, st::Sitetypes, ol::OpList)
    nt = length(ol.sites)
    first, rng, mats = _oplist_payload(st, ol)
    out = zeros(Float64, 2 * nt)
    GC.@preserve d first rng mats out check(ccall((:tnh_inner_oplist, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}, Ptr{Float64}),
            d.ctx.h, length(d), d.sites, nt, first, rng, Ptr{Float64}(pointer(mats)), out),
          d.ctx.h, "tnh_inner_oplist")
    return [ol.coeffs[t] * complex(out[2t-1], out[2t]) for t = 1:nt]
end

function inner_oplist(d::DeviceMPS, st::Sitetypes, ol::OpList, phi::DeviceMPS)
    length(phi) == length(d) || error("GMPS must share the same length.")
    nt = length(ol.sites)
    first, rng, mats = _oplist_payload(st, ol)
    out = zeros(Float64, 2 * nt)
    GC.@preserve d phi first rng mats out check(ccall((:tnh_inner_oplist_pair, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cint}, Ptr{Cint},
             Ptr{Float64}, Ptr{Float64}),
            d.ctx.h, length(d), d.sites, phi.sites, nt, first, rng, Ptr{Float64}(pointer(mats)), out),
          d.ctx.h, "tnh_inner_oplist_pair")
    return [ol.coeffs[t] * complex(out[2t-1], out[2t]) for t = 1:nt]
end

"""
    tebd(GPU(), st, psi, H, dt, tmax, save, observers=[]; cutoff=1e-12, maxdim=0, mindim=1,
         evol="imag", order=2, norm=0) -> (psi, energy)

tebd(st, psi, H, dt, tmax, save, observers; kwargs...), src/algorithms/mps/tebd.jl:3-100,
with psi resident on the GPU: the gates are built on the host by the reference's
trotterize and uploaded once; applygates!, the log-norm accumulation and renormalisation
run on the device; energies come from tnh_inner_oplist; the observers receive the state
downloaded at save points. Projectors are not supported on this path.
"""
function tebd(g::GPU, st::Sitetypes, psi::GMPS, H::OpList, dt::Number, tmax::Number, save::Number,
              observers::Vector = [], projectors = []; kwargs...)
    st.dim != TensorNetworks.dim(psi) && error("Sitetypes and MPS physical dimensions do not match.")
    length(psi) != length(H) && error("OpList and MPS have different lengths.")
    isempty(projectors) || error("projectors are not supported by the GPU TEBD path")
    cutoff::Float64 = get(kwargs, :cutoff, 1e-12)
    maxdim::Int = get(kwargs, :maxdim, 0)
    mindim::Int = get(kwargs, :mindim, 1)
    evol::String = get(kwargs, :evol, "imag")
    !(evol == "imag" || evol == "real") && error("evol must be real of imag.")
    order::Int = get(kwargs, :order, 2)
    !(order == 1 || order == 2) && error("Only first or second order trotter supported.")
    gates = DeviceGates(g.ctx, trotterize(st, H, dt; evol = evol, order = order))
    nsteps = Int(round(tmax / dt))
    save = save < dt ? dt : save
    nsave = Int(round(save / dt))
    normal::Float64 = get(kwargs, :norm, 0)
    d = DeviceMPS(psi; ctx = g.ctx)
    energy = real(sum(inner_oplist(d, st, H)))
    for observer in observers
        measure!(observer, 0.0, GMPS(d), normal, energy)
    end
    step = 0
    converged = false
    while !converged                                                # tebd.jl:62-97
        applygates!(d, gates; mindim = mindim, maxdim = maxdim, cutoff = cutoff)
        psinorm = log(device_norm(d))
        normal += psinorm
        normalize_device!(d)
        step += 1
        step >= nsteps && (converged = true)
        if step % nsave == 0
            @printf("time=%.4f, energy=%.12f, maxbonddim=%d \n", step * dt, energy,
                    maximum(shape(h)[3] for h in d.sites))
            energy = real(sum(inner_oplist(d, st, H)))
            if !isempty(observers)
                snap = GMPS(d)
                for observer in observers
                    measure!(observer, step * dt, snap, normal, energy)
                    converged = converged || checkdone!(observer)
                end
            end
        end
    end
    store!(psi, d)
    return psi, energy
end

# ------------------------------------------ device-resident DMRG (whole bond step)
struct DmrgParams
    krylovdim::Cint
    maxiter::Cint
    lanczos_tol::Cdouble
    cutoff::Cdouble
    maxdim::Cint
    mindim::Cint
    precision::Cint
    nsites::Cint
    nonhermitian::Cint
end

"""
    dmrg(GPU(), psi::GMPS, Hs::GMPS...; precision=:f64, refine_sweeps=2, checkpoint="",
         kwargs...) -> (psi, cost)

dmrg(psi, H...) with the whole bond step on the GPU (tnh_dmrg_sweep): keyword arguments
and convergence rules of src/algorithms/mps/dmrg.jl:1-99. `Hs` mixes MPOs and rank-1 GMPS
(each a term coeff |phi><phi|, as in dmrg.jl:139-149), weighted by `coeffs`. `precision` (:f64, :f32,
:bf16x3) sets the arithmetic of the matvecs and environment updates; with a lower
precision the last `refine_sweeps` passes run in FP64. A non-empty `checkpoint` writes psi
after every pass in the reference's HDF5 layout (group "psi", gmps.jl:359-370), from which
a run resumes by reading it back (`read_checkpoint`) and calling dmrg again.
"""
function dmrg(g::GPU, psi::GMPS, Hs::GMPS...; nsites::Int = 2, krylovdim::Int = 3,
              kryloviter::Int = 2, ishermitian::Bool = true, minsweeps::Int = 1,
              maxsweeps::Int = 1000,
              tol::Float64 = 1e-10, tolgrad::Float64 = 1e-5, numconverges::Real = 4,
              verbose::Bool = true, cutoff::Float64 = 1e-12, maxdim::Int = 1000,
              mindim::Int = 1, coeffs = ones(length(Hs)), precision::Symbol = :f64,
              refine_sweeps::Int = 2, checkpoint::AbstractString = "")
    (nsites == 1 || nsites == 2) || error("the GPU path implements the one- and two-site updates")
    length(Hs) == 0 && error("You must provide atleast one MPS/MPO for the Hamiltonian.")
    length(coeffs) == length(Hs) || error("one coefficient per Hamiltonian term")
    ctx = g.ctx
    N = length(psi)
    # MPOs give ProjMPS(psi, H, psi), MPSs the rank-1 term ProjMPS(phi, psi; squared=true)
    # (dmrg.jl:139-149); the rank-1 terms join the sum after the MPO terms
    mpos = [j for j = 1:length(Hs) if TensorNetworks.rank(Hs[j]) == 2]
    phis = [j for j = 1:length(Hs) if TensorNetworks.rank(Hs[j]) == 1]
    length(mpos) + length(phis) == length(Hs) ||
        error("Hamiltonian must be composed of MPOs (rank 2) or MPSs (rank 1).")
    # a complex phi, or the Arnoldi of ishermitian = false, needs a complex psi
    cplx = !ishermitian || any(!all(iszero ∘ imag, Hs[j][i]) for j in phis for i = 1:N)
    dpsi = [upload(ctx, psi[i]; real_if_possible = !cplx) for i = 1:N]
    dW = [upload(ctx, Hs[j][i]) for j in mpos for i = 1:N]
    st = Ref{Ptr{Cvoid}}(C_NULL)
    c = Cdouble[coeffs[j] for j in mpos]
    hp = [t.h for t in dpsi]; hw = [t.h for t in dW]
    GC.@preserve hp hw c check(ccall((:tnh_dmrg_create, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Cdouble}, Ptr{Ptr{Cvoid}}),
            ctx.h, N, hp, length(mpos), hw, c, st), ctx.h, "tnh_dmrg_create")
    cost = 0.0
    try
        for j in phis
            dphi = [upload(ctx, Hs[j][i]; real_if_possible = !cplx) for i = 1:N]
            hf = [t.h for t in dphi]
            GC.@preserve dphi hf check(ccall((:tnh_dmrg_add_projector, libtnhip), Cint,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cdouble), st[], hf, Cdouble(coeffs[j])),
                    ctx.h, "tnh_dmrg_add_projector")
        end
        e = Ref{Cdouble}(0); D = Ref{Cint}(0); disc = Ref{Cdouble}(0)
        check(ccall((:tnh_dmrg_energy, libtnhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}),
                    st[], 0, e), ctx.h, "tnh_dmrg_energy")
        cost = lastcost = e[]; lastD = maxbonddim(psi); grad = 0.0
        direction = false; sweeps = 0; csweeps = cgrad = 0; converged = false
        prec = precision
        diff(x, y) = abs(x) < 1e-10 ? abs(x - y) : abs((x - y) / x)
        while !converged                                           # dmrg.jl:34-96
            # the last refine_sweeps passes before maxsweeps run in FP64
            if prec != :f64 && maxsweeps != 0 && sweeps >= maxsweeps - refine_sweeps
                prec = :f64
            end
            p = Ref(DmrgParams(krylovdim, kryloviter, 1e-14, cutoff, maxdim, mindim, PRECISION[prec],
                               nsites, !ishermitian))
            check(ccall((:tnh_dmrg_sweep, libtnhip), Cint,
                        (Ptr{Cvoid}, Cint, Ptr{DmrgParams}, Ptr{Cdouble}, Ptr{Cint}, Ptr{Cdouble}),
                        st[], direction, p, e, D, disc), ctx.h, "tnh_dmrg_sweep")
            cost = e[]
            if !ishermitian  # the Ritz value of smallest real part, possibly complex
                re = Ref{Cdouble}(0); im = Ref{Cdouble}(0)
                check(ccall((:tnh_dmrg_last_ritz, libtnhip), Cint,
                            (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), st[], re, im),
                      ctx.h, "tnh_dmrg_last_ritz")
                cost = complex(re[], im[])
            end
            direction = !direction; sweeps += 1
            if sweeps >= minsweeps
                dc = diff(cost, lastcost)
                csweeps = (dc < tol && lastD == D[]) ? csweeps + 1 : 0
                cgrad = ((dc + grad) != 0 && abs((dc - grad) / (dc + grad)) < tolgrad &&
                         lastD == D[]) ? cgrad + 1 : 0
                converged = max(csweeps, cgrad) >= numconverges || (sweeps >= maxsweeps && maxsweeps != 0)
            end
            grad = abs(diff(cost, lastcost)); lastcost = cost; lastD = D[]
            verbose && println("Sweep=$sweeps, energy=$(round(real(cost), digits=12)), maxbonddim=$(D[])")
            isempty(checkpoint) || (_copy_back!(psi, st[], ctx); write_checkpoint(checkpoint, psi))
        end
        _copy_back!(psi, st[], ctx)
    finally
        ccall((:tnh_dmrg_destroy, libtnhip), Cint, (Ptr{Cvoid},), st[])
    end
    return psi, cost
end

function _copy_back!(psi::GMPS, st::Ptr{Cvoid}, ctx::Context)
    for i = 1:length(psi)
        b = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:tnh_dmrg_site, libtnhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}), st, i, b),
              ctx.h, "tnh_dmrg_site")
        psi[i] = to_host(ctx, b[])  # borrowed: owned by the DMRG state
    end
    ctr = Ref{Cint}(0)
    ccall((:tnh_dmrg_center, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cint}), st, ctr)
    psi.center = ctr[]
    return psi
end

# ------------------------------------------------------------ checkpoints
"""
    write_checkpoint(path, psi::GMPS; name="psi")
    read_checkpoint(path; name="psi") -> GMPS

The MPS in the reference's HDF5 layout (HDF5.write / HDF5.read of GMPS,
src/structures/mps/gmps.jl:359-383: group attributes type = "MPS", version = 1, datasets
"length", "center", "MPS[i]"), through the reference's own methods, so files interchange
with TensorNetworks.jl. The Python side (tnhip.checkpoint) writes the same fields.
"""
function write_checkpoint(path::AbstractString, psi::GMPS; name::AbstractString = "psi")
    h5open(path, "w") do f
        write(f, name, psi)
    end
end

read_checkpoint(path::AbstractString; name::AbstractString = "psi") =
    h5open(f -> read(f, name, GMPS), path, "r")

# ------------------------------------------ <psi|H|psi> and dE/dW (tnh_mpo_expect_grad)
"""
    mpo_energy_grad(psi::GMPS, W::Vector{<:Array}; block=8) -> (E, dW::Vector{Array{ComplexF64,4}})

inner(psi, H, psi) (src/structures/mps/mpo.jl:181-217) for the MPO tensors W and its
gradient with respect to every MPO tensor: E = sum(W[i] .* dW[i]) at each site.
Environments are checkpointed every `block` sites on the device.
"""
function mpo_energy_grad(psi::GMPS, W::AbstractVector; block::Int = 8, ctx::Context = default_context())
    N = length(psi)
    dpsi = [upload(ctx, psi[i]) for i = 1:N]
    dW = [upload(ctx, W[i]) for i = 1:N]
    hp = [t.h for t in dpsi]; hw = [t.h for t in dW]
    out = fill(Ptr{Cvoid}(C_NULL), N)
    er = Ref{Cdouble}(0); ei = Ref{Cdouble}(0)
    GC.@preserve hp hw out check(ccall((:tnh_mpo_expect_grad, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cdouble}, Ptr{Cdouble},
             Ptr{Ptr{Cvoid}}),
            ctx.h, N, hp, hw, block, er, ei, out), ctx.h, "tnh_mpo_expect_grad")
    grads = [to_host(owned(ctx, out[i])) for i = 1:N]
    return complex(er[], ei[]), grads
end
mpo_energy_grad(psi::GMPS, H::GMPS; kwargs...) =
    mpo_energy_grad(psi, [H[i] for i = 1:length(H)]; kwargs...)

"""
    automaton_mpo(st::Sitetypes, ol::OpList) -> (W, slots)

The finite-state-automaton MPO of MPO(st, H), src/structures/mps/mpo.jl:323-440, before its
SVD compression (:442-457): channel 1 = nothing placed, the last = term complete, one fresh
channel per bond a multi-site term crosses. `slots[t] = (site, row, col, O_t)` says that
term t's coefficient enters only as W[site][row, :, :, col] += c_t O_t, which makes
dE/dc_t = sum(dW[site][row, :, :, col] .* O_t) exact.
"""
function automaton_mpo(st::Sitetypes, ol::OpList)
    N, d = ol.length, st.dim
    nchan = zeros(Int, N)            # channels crossing bond b (sites b, b+1)
    chans = Vector{Dict{Int,Int}}(undef, length(ol.sites))
    for (t, sites) in enumerate(ol.sites)
        chans[t] = Dict{Int,Int}()
        length(sites) == 1 && continue
        for b = sites[1]:sites[end]-1
            nchan[b] += 1
            chans[t][b] = nchan[b]
        end
    end
    dims = [1; [2 + nchan[b] for b = 1:N-1]; 1]
    Id = Matrix{ComplexF64}(op(st, "id"))
    W = [zeros(ComplexF64, dims[i], d, d, dims[i+1]) for i = 1:N]
    for i = 1:N
        i > 1 && (W[i][end, :, :, end] = Id)
        i < N && (W[i][1, :, :, 1] = Id)
    end
    slots = Vector{Tuple{Int,Int,Int,Matrix{ComplexF64}}}(undef, length(ol.sites))
    for (t, sites) in enumerate(ol.sites)
        if length(sites) == 1
            i = sites[1]
            O = Matrix{ComplexF64}(op(st, ol.ops[t][1]))
            W[i][1, :, :, end] .+= ol.coeffs[t] .* O
            slots[t] = (i, 1, size(W[i], 4), O)
            continue
        end
        for i = sites[1]:sites[end]
            k = findfirst(==(i), sites)
            O = k === nothing ? Id : Matrix{ComplexF64}(op(st, ol.ops[t][k]))
            row = i == sites[1] ? 1 : 1 + chans[t][i-1]
            col = i == sites[end] ? size(W[i], 4) : 1 + chans[t][i]
            W[i][row, :, :, col] .+= (i == sites[1] ? ol.coeffs[t] : 1.0) .* O
            i == sites[1] && (slots[t] = (i, row, col, O))
        end
    end
    return W, slots
end

"""
    coupling_energy(psi::GMPS, st::Sitetypes, ol::OpList, c::AbstractVector; block=8) -> E

E(c) = <psi|H(c)|psi> for H(c) = sum_t c_t O_t (the terms of `ol` with coefficients c),
evaluated on the GPU. Its ChainRulesCore rrule (package extension
TensorNetworksAMDChainRulesCoreExt) returns dE/dc from the device backward pass
(tnh_mpo_expect_grad) through the automaton slots.
"""
function coupling_energy(psi::GMPS, st::Sitetypes, ol::OpList, c::AbstractVector; block::Int = 8,
                         ctx::Context = default_context())
    E, _ = coupling_energy_grad(psi, st, ol, c; block = block, ctx = ctx)
    return E
end

function coupling_energy_grad(psi::GMPS, st::Sitetypes, ol::OpList, c::AbstractVector; block::Int = 8,
                              ctx::Context = default_context())
    length(c) == length(ol.coeffs) || error("one coefficient per term")
    olc = OpList(ol.length, copy(ol.ops), copy(ol.sites), ComplexF64.(c))
    W, slots = automaton_mpo(st, olc)
    E, dW = mpo_energy_grad(psi, W; block = block, ctx = ctx)
    g = [sum(dW[s][r, :, :, k] .* O) for (s, r, k, O) in slots]
    return real(E), real.(g)
end

# --------------------------------------------------------------- MPOs and products
"""
    MPO(GPU(), st::Sitetypes, H::OpList; cutoff=1e-15, maxdim=0, mindim=1) -> GMPS

MPO(st, H; kwargs), src/structures/mps/mpo.jl:323-459: the finite-state automaton built by
the library on the host, then its SVD compression sweeps on the device (tnh_mpo_from_oplist).
"""
function MPO(g::GPU, st::Sitetypes, H::OpList; cutoff::Real = 1e-15, maxdim::Int = 0,
             mindim::Int = 1)
    N, d, nt = H.length, st.dim, length(H.sites)
    first, rng, mats = _oplist_payload(st, H)
    coeffs = Float64[]
    for c in H.coeffs
        push!(coeffs, real(c), imag(c))
    end
    out = fill(C_NULL, N)
    GC.@preserve first rng mats coeffs out check(ccall((:tnh_mpo_from_oplist, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}, Ptr{Float64}, Cdouble,
             Cint, Cint, Ptr{Ptr{Cvoid}}),
            g.ctx.h, N, d, nt, first, rng, Ptr{Float64}(pointer(mats)), coeffs, cutoff, maxdim,
            mindim, out), g.ctx.h, "tnh_mpo_from_oplist")
    ts = Array{ComplexF64}[to_host(owned(g.ctx, h)) for h in out]
    return GMPS(2, d, ts, 1)
end

"""
    applyMPO(GPU(), O, psi; cutoff=0, maxdim=0, mindim=1), applyMPO(GPU(), psi, O; …),
    applyMPO(GPU(), O1, O2; …) -> GMPS

applyMPO, src/structures/mps/mpo.jl:105-169: O |psi> (tnh_mpo_apply), <psi| O as
adjoint(O) |psi> (:112), or O1 O2 (tnh_mpo_mpo_product), each with the truncating centre
move to site 1.
"""
function applyMPO(g::GPU, arg1::GMPS, arg2::GMPS; cutoff::Real = 0.0, maxdim::Int = 0,
                  mindim::Int = 1)
    r1, r2 = TensorNetworks.rank(arg1), TensorNetworks.rank(arg2)
    length(arg1) == length(arg2) || error("GMPS must share the same length.")
    TensorNetworks.dim(arg1) == TensorNetworks.dim(arg2) || error("GMPS must share the same physical dims.")
    r1 == 1 && r2 == 2 && return applyMPO(g, TensorNetworks.adjoint(arg2), arg1; cutoff = cutoff, maxdim = maxdim,
                                          mindim = mindim)
    r1 == 2 || error("applyMPO: an MPO (rank 2) must come first")
    N = length(arg1)
    a = [upload(g.ctx, arg1[i]) for i = 1:N]
    b = [upload(g.ctx, arg2[i]) for i = 1:N]
    ha = [t.h for t in a]; hb = [t.h for t in b]
    out = fill(C_NULL, N)
    fn = r2 == 1 ? :tnh_mpo_apply : :tnh_mpo_mpo_product
    if r2 == 1
        GC.@preserve a b ha hb out check(ccall((:tnh_mpo_apply, libtnhip), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cdouble, Cint, Cint, Ptr{Ptr{Cvoid}}),
                g.ctx.h, N, ha, hb, cutoff, maxdim, mindim, out), g.ctx.h, String(fn))
    else
        GC.@preserve a b ha hb out check(ccall((:tnh_mpo_mpo_product, libtnhip), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cdouble, Cint, Cint, Ptr{Ptr{Cvoid}}),
                g.ctx.h, N, ha, hb, cutoff, maxdim, mindim, out), g.ctx.h, String(fn))
    end
    ts = Array{ComplexF64}[to_host(owned(g.ctx, h)) for h in out]
    return GMPS(r2, TensorNetworks.dim(arg1), ts, 1)
end

"""
    inner(GPU(), args::GMPS...) -> ComplexF64

inner(args::GMPS...), src/structures/mps/mpo.jl:181-217: <bra| O_1 ... O_k |ket>, the bra
moved to the front past any operators before it (each becoming its adjoint, :195-204), the
operators folded into one by exact products (cutoff 0) and the chain contracted on the
device (tnh_inner_mpo). inner(GPU(), phi, psi) is the overlap.
"""
function inner(g::GPU, args::GMPS...)
    length(args) < 2 && error("There must be atleast 2 MPS arguments (rank 1).")
    all(TensorNetworks.dim(a) == TensorNetworks.dim(args[1]) for a in args) ||
        error("GMPS must share the same physical dim.")
    all(length(a) == length(args[1]) for a in args) || error("GMPS must share the same length.")
    ranks = [TensorNetworks.rank(a) for a in args]
    count(==(1), ranks) != 2 && error("The inner product must have a braket structure.")
    ranks[end] != 1 && error("The final GMPS must be rank 1 (MPS).")
    objs = collect(GMPS, args)
    idx = findfirst(==(1), ranks)
    for j = idx-1:-1:1
        psi = objs[j+1]
        objs[j+1] = TensorNetworks.adjoint(objs[j])
        objs[j] = psi
    end
    bra, ops, ket = objs[1], objs[2:end-1], objs[end]
    W = isempty(ops) ? nothing : ops[1]
    for O in ops[2:end]
        W = applyMPO(g, W, O)  # exact: cutoff 0
    end
    N = length(bra)
    db = [upload(g.ctx, bra[i]) for i = 1:N]; dk = [upload(g.ctx, ket[i]) for i = 1:N]
    dw = W === nothing ? DeviceTensor[] : [upload(g.ctx, W[i]) for i = 1:N]
    hb = [t.h for t in db]; hk = [t.h for t in dk]; hw = [t.h for t in dw]
    re = Ref{Cdouble}(0); im = Ref{Cdouble}(0)
    GC.@preserve db dk dw hb hk hw check(ccall((:tnh_inner_mpo, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cdouble}, Ptr{Cdouble}),
            g.ctx.h, N, hb, W === nothing ? C_NULL : pointer(hw), hk, re, im), g.ctx.h, "tnh_inner_mpo")
    return complex(re[], im[])
end

# ------------------------------- many QJMC trajectories (tnh_qjmc_run(_batched))
# The batched call marks a trajectory that outgrew its slots with maxbond = -1;
# qjmc_trajectories! reruns those on the per-trajectory runner.

# Layout of tnh_qjmc_problem (include/tnhip.h); the pointers refer to arrays the caller
# keeps alive (GC.@preserve) for the duration of the call.
struct QjmcProblem
    N::Cint; d::Cint
    nlayers::Cint
    layer_count::Ptr{Cint}; gate_site::Ptr{Cint}; gate_range::Ptr{Cint}
    gate_offset::Ptr{Int64}; gate_data::Ptr{Float64}
    njump::Cint
    jump_site::Ptr{Cint}; jump_coeff::Ptr{Float64}; jump_op::Ptr{Float64}
    nobs::Cint
    obs_site::Ptr{Cint}; obs_op::Ptr{Float64}
    init_state::Ptr{Float64}
    dt::Cdouble
    nsteps::Cint; savesteps::Cint
    cutoff::Cdouble
    maxdim::Cint; mindim::Cint
end

"""
    qjmc_trajectories!(prob, U, meas, njumps, jumps, jsteps, maxb; batched=true)

Many trajectories of qjmc_simulation (src/algorithms/mps/qjmc.jl:28-167, classical branch):
U[:, t] holds trajectory t's uniforms in the reference's rand() order. `batched` advances
all trajectories together (narrow slots, then wide slots of up to 512 / d for those that
outgrow them) and reruns on the per-trajectory runner those whose bond outgrew the wide
slots too.
"""
function qjmc_trajectories!(prob::Ref{QjmcProblem}, U::Matrix{Float64}, meas::Array{Float64},
                            njumps::Vector{Cint}, jumps::Matrix{Cint}, jsteps::Matrix{Cint},
                            maxb::Vector{Cint}; batched::Bool = true, nworkers::Int = 4,
                            ctx::Context = default_context())
    nu, ntraj = size(U)
    maxjumps = size(jumps, 1)
    if batched
        check(ccall((:tnh_qjmc_run_batched, libtnhip), Cint,
                    (Ptr{Cvoid}, Ptr{QjmcProblem}, Cint, Ptr{Float64}, Int64, Ptr{Float64},
                     Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Cint, Ptr{Cint}),
                    ctx.h, prob, ntraj, U, nu, meas, njumps, jumps, jsteps, maxjumps, maxb),
              ctx.h, "tnh_qjmc_run_batched")
        over = findall(<(0), maxb)  # trajectories that outgrew the batched slots
        isempty(over) && return nothing
        m2 = zeros(Float64, size(meas)[1:end-1]..., length(over))
        nj2 = zeros(Cint, length(over))
        j2 = zeros(Cint, maxjumps, length(over))
        s2 = zeros(Cint, maxjumps, length(over))
        b2 = zeros(Cint, length(over))
        _qjmc_workers!(prob, U[:, over], m2, nj2, j2, s2, b2, nworkers, ctx)
        selectdim(meas, ndims(meas), over) .= m2
        njumps[over] .= nj2
        jumps[:, over] .= j2
        jsteps[:, over] .= s2
        maxb[over] .= b2
        return nothing
    end
    _qjmc_workers!(prob, U, meas, njumps, jumps, jsteps, maxb, nworkers, ctx)
    return nothing
end

function _qjmc_workers!(prob, U, meas, njumps, jumps, jsteps, maxb, nworkers, ctx)
    nu, ntraj = size(U)
    check(ccall((:tnh_qjmc_run, libtnhip), Cint,
                (Ptr{Cvoid}, Ptr{QjmcProblem}, Cint, Ptr{Float64}, Int64, Cint, Ptr{Float64},
                 Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Cint, Ptr{Cint}),
                ctx.h, prob, ntraj, U, nu, nworkers, meas, njumps, jumps, jsteps, size(jumps, 1), maxb),
          ctx.h, "tnh_qjmc_run")
end

export Context, GPU, GPUProjMPS, DeviceMPS, DeviceGates, inner_oplist, mpo_energy_grad, automaton_mpo,
       coupling_energy, coupling_energy_grad, write_checkpoint, read_checkpoint, QjmcProblem,
       qjmc_trajectories!

end # module
