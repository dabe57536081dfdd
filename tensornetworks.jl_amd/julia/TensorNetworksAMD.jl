# Julia binding of libtnhip.so for TensorNetworks.jl (ccall over the C-ABI in
# include/tnhip.h). It adds a GPU backend for the two-site DMRG hot path and the TEBD
# gate, keeping the package's own types (GMPS, OpList, MPO) and keyword arguments.
#
# Not executed in this repository's CI: no Julia toolchain is available where it was
# written. The Python module tensornetworks.jl_amd/tnhip exercises the same C-ABI calls.
module TensorNetworksAMD

using TensorNetworks
import TensorNetworks: AbstractProjMPS, GMPS, product, buildleft!, buildright!, calculate,
                       movecenter!, dmrg

const libtnhip = get(ENV, "TNHIP_LIB", joinpath(@__DIR__, "..", "tnhip", "libtnhip.so"))

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

# Julia arrays are column-major, exactly the layout the library expects.
function upload(ctx::Context, A::Array{ComplexF64}; real_if_possible::Bool = true)
    dev = (real_if_possible && all(iszero ∘ imag, A)) ? TNH_F64 : TNH_C128
    dims = Int64[size(A)...]
    r = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve A dims begin
        check(ccall((:tnh_buf_upload, libtnhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Int64}, Ptr{Ptr{Cvoid}}),
                    ctx.h, pointer(A), TNH_C128, dev, ndims(A), pointer(dims), r),
              ctx.h, "tnh_buf_upload")
    end
    t = DeviceTensor(ctx, r[])
    finalizer(free!, t)
    return t
end

function shape(t::DeviceTensor)
    dt = Ref{Cint}(0); nd = Ref{Cint}(0); dims = zeros(Int64, 8)
    ccall((:tnh_buf_shape, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cint}, Ptr{Cint}, Ptr{Int64}),
          t.h, dt, nd, dims)
    return Tuple(dims[1:nd[]])
end

function download(t::DeviceTensor)
    A = Array{ComplexF64}(undef, shape(t)...)
    GC.@preserve A check(ccall((:tnh_buf_download, libtnhip), Cint,
                               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                               t.ctx.h, t.h, pointer(A), TNH_C128), t.ctx.h, "tnh_buf_download")
    return A
end

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
    fn = left ? :tnh_env_left : :tnh_env_right
    check(ccall((fn, libtnhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint),
                p.ctx.h, prev, A.h, p.W[idx].h, r, PRECISION[:f64]), p.ctx.h, String(fn))
    if p.blocks[idx] === nothing || p.blocks[idx].h != r[]
        p.blocks[idx] = DeviceTensor(p.ctx, r[])
        finalizer(free!, p.blocks[idx])
    end
end

buildleft!(p::GPUProjMPS, idx::Int) = _env!(p, idx, true)     # projmps.jl:50-66
buildright!(p::GPUProjMPS, idx::Int) = _env!(p, idx, false)   # projmps.jl:74-95

function product(p::GPUProjMPS, A, direction::Bool = false, nsites::Int = 2)  # projmps.jl:103-144
    nsites == 2 || error("GPUProjMPS supports nsites = 2")
    site = direction ? p.center - nsites + 1 : p.center
    x = upload(p.ctx, Array{ComplexF64}(A); real_if_possible = false)
    y = upload(p.ctx, zeros(ComplexF64, size(A)); real_if_possible = false)
    check(ccall((:tnh_heff_apply, libtnhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Cdouble, Cdouble, Cint),
                p.ctx.h, blockptr(p, site - 1), p.W[site].h, p.W[site+1].h,
                blockptr(p, site + nsites), x.h, y.h, p.coeff, 0.0, PRECISION[:f64]),
          p.ctx.h, "tnh_heff_apply")
    return download(y)
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

# ------------------------------------------ device-resident DMRG (whole bond step)
struct DmrgParams
    krylovdim::Cint
    maxiter::Cint
    lanczos_tol::Cdouble
    cutoff::Cdouble
    maxdim::Cint
    mindim::Cint
    precision::Cint
end

"""
    dmrg_gpu(psi::GMPS, Hs::GMPS...; precision=:f64, kwargs...)

dmrg(psi, H...) with the bond step on the GPU (tnh_dmrg_sweep): keyword arguments and
convergence rules of src/algorithms/mps/dmrg.jl:1-99; returns (psi, cost).
"""
function dmrg_gpu(psi::GMPS, Hs::GMPS...; nsites::Int = 2, krylovdim::Int = 3,
                  kryloviter::Int = 2, minsweeps::Int = 1, maxsweeps::Int = 1000,
                  tol::Float64 = 1e-10, tolgrad::Float64 = 1e-5, numconverges::Real = 4,
                  verbose::Bool = true, cutoff::Float64 = 1e-12, maxdim::Int = 1000,
                  mindim::Int = 1, coeffs = ones(length(Hs)), precision::Symbol = :f64,
                  ctx::Context = default_context())
    nsites == 2 || error("the GPU path implements the two-site update")
    N = length(psi)
    dpsi = [upload(ctx, psi[i]) for i = 1:N]
    dW = [upload(ctx, H[i]) for H in Hs for i = 1:N]
    st = Ref{Ptr{Cvoid}}(C_NULL)
    c = Cdouble[coeffs...]
    hp = [t.h for t in dpsi]; hw = [t.h for t in dW]
    GC.@preserve hp hw c check(ccall((:tnh_dmrg_create, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Cdouble}, Ptr{Ptr{Cvoid}}),
            ctx.h, N, hp, length(Hs), hw, c, st), ctx.h, "tnh_dmrg_create")
    e = Ref{Cdouble}(0); D = Ref{Cint}(0); disc = Ref{Cdouble}(0)
    check(ccall((:tnh_dmrg_energy, libtnhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}),
                st[], 0, e), ctx.h, "tnh_dmrg_energy")
    cost = lastcost = e[]; lastD = D0 = maxbonddim(psi); grad = 0.0
    direction = false; sweeps = 0; csweeps = cgrad = 0; converged = false
    p = Ref(DmrgParams(krylovdim, kryloviter, 1e-14, cutoff, maxdim, mindim, PRECISION[precision]))
    diff(x, y) = abs(x) < 1e-10 ? abs(x - y) : abs((x - y) / x)
    while !converged                                               # dmrg.jl:34-96
        check(ccall((:tnh_dmrg_sweep, libtnhip), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{DmrgParams}, Ptr{Cdouble}, Ptr{Cint}, Ptr{Cdouble}),
                    st[], direction, p, e, D, disc), ctx.h, "tnh_dmrg_sweep")
        cost = e[]; direction = !direction; sweeps += 1
        if sweeps >= minsweeps
            csweeps = (diff(cost, lastcost) < tol && lastD == D[]) ? csweeps + 1 : 0
            cgrad = (abs((diff(cost, lastcost) - grad) / (diff(cost, lastcost) + grad)) < tolgrad &&
                     lastD == D[]) ? cgrad + 1 : 0
            converged = max(csweeps, cgrad) >= numconverges || (sweeps >= maxsweeps && maxsweeps != 0)
        end
        grad = abs(diff(cost, lastcost)); lastcost = cost; lastD = D[]
        verbose && println("Sweep=$sweeps, energy=$(round(real(cost), digits=12)), maxbonddim=$(D[])")
    end
    for i = 1:N                                                    # copy psi back
        b = Ref{Ptr{Cvoid}}(C_NULL)
        ccall((:tnh_dmrg_site, libtnhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}), st[], i, b)
        psi[i] = download(DeviceTensor(ctx, b[]))  # borrowed: not freed here
    end
    ctr = Ref{Cint}(0)
    ccall((:tnh_dmrg_center, libtnhip), Cint, (Ptr{Cvoid}, Ptr{Cint}), st[], ctr)
    psi.center = ctr[]
    ccall((:tnh_dmrg_destroy, libtnhip), Cint, (Ptr{Cvoid},), st[])
    return psi, cost
end

# ------------------------------------------ <psi|H|psi> and dE/dW (tnh_mpo_expect_grad)
"""
    mpo_energy_grad(psi::GMPS, H::GMPS; block=8) -> (E, dW::Vector{Array{ComplexF64,4}})

inner(psi, H, psi) (src/structures/mps/mpo.jl:181-217) and its gradient with respect to
every MPO tensor: E = sum(H[i] .* dW[i]) at each site. Environments are checkpointed every
`block` sites on the device.
"""
function mpo_energy_grad(psi::GMPS, H::GMPS; block::Int = 8, ctx::Context = default_context())
    N = length(psi)
    dpsi = [upload(ctx, psi[i]) for i = 1:N]
    dW = [upload(ctx, H[i]) for i = 1:N]
    hp = [t.h for t in dpsi]; hw = [t.h for t in dW]
    out = fill(Ptr{Cvoid}(C_NULL), N)
    er = Ref{Cdouble}(0); ei = Ref{Cdouble}(0)
    GC.@preserve hp hw out check(ccall((:tnh_mpo_expect_grad, libtnhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cdouble}, Ptr{Cdouble},
             Ptr{Ptr{Cvoid}}),
            ctx.h, N, hp, hw, block, er, ei, out), ctx.h, "tnh_mpo_expect_grad")
    grads = [download(DeviceTensor(ctx, out[i])) for i = 1:N]
    return complex(er[], ei[]), grads
end

# A ChainRulesCore rule for a coupling-parametrised energy E(c) = inner(psi, H(c), psi), where
# H(c) = MPO(st, oplist(c)) places each c_t at one MPO entry (site_t, row_t, col_t) times the
# operator op_t, reads (with ChainRulesCore loaded by the caller):
#
#   function ChainRulesCore.rrule(::typeof(energy), c, psi, slots)
#       E, dW = mpo_energy_grad(psi, build_mpo(c))
#       g = [sum(dW[s][r, :, :, k] .* op) for (s, r, k, op) in slots]
#       return real(E), Ē -> (NoTangent(), Ē .* real.(g), NoTangent(), NoTangent())
#   end

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
U[:, t] holds trajectory t's uniforms in the reference's rand() order. `batched` advances all
trajectories together (bond dimension <= min(16, 32 / d)) and falls back to the
per-trajectory runner when a bond outgrows that.
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
        # rerun only those on the per-trajectory runner, each with its own uniforms
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

export Context, GPUProjMPS, dmrg_gpu, mpo_energy_grad, QjmcProblem, qjmc_trajectories!

end # module
