# TensorNetworkADHip.jl -- GPU backend for TensorNetworkAD.jl's CTMRG / TRG hot path.
#
# Binds libtnad (include/tnad.h) with `ccall` and attaches the fused forward +
# reverse calls to Zygote through ChainRulesCore rrules, so that
#   energy(h, ipeps; χ, tol, maxit)   (src/variationalipeps.jl:28-40)
#   trg(a, χ, niter; tol)              (src/trg.jl:13-30)
#   ctmrg(rt; tol, maxit)              (src/ctmrg.jl:110-117)
# run on one MI355X and remain differentiable by `Zygote.gradient`.
# Arrays cross the boundary as plain column-major Float64 buffers (Julia's own
# layout), so no transposes are needed.  Shipped as source: no Julia toolchain
# is available where the library is built.
module TensorNetworkADHip

using ChainRulesCore
import TensorNetworkAD
using TensorNetworkAD: IPEPS, SquareIPEPS, CTMRGRuntime, SquareCTMRGRuntime

const libtnad = get(ENV, "TNAD_LIB", joinpath(@__DIR__, "..", "tnad", "libtnad.so"))

# mirrors `tnad_opts` (include/tnad.h)
struct TnadOpts
    precision::Cint
    switch_tol::Cdouble
    bwd_precision::Cint
    eig_block::Cint
    eig_max_iters::Cint
    eig_warm::Cint
    eig_warm_iters::Cint
    eig_tol::Cdouble
    bwd_last::Cint
end
function default_opts()
    r = Ref{TnadOpts}()
    ccall((:tnad_opts_default, libtnad), Cvoid, (Ptr{TnadOpts},), r)
    r[]
end

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
function context()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:tnad_create, libtnad), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, r)
        st == 0 || error("tnad_create failed ($st)")
        CTX[] = r[]
    end
    CTX[]
end

function check(st)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:tnad_last_error, libtnad), Cstring, (Ptr{Cvoid},), context()))
    st == -2 ? throw(DimensionMismatch(msg)) : error("libtnad: $msg ($st)")
end

# ---------------------------------------------------------------------------
# energy: one boundary crossing per value + gradient
# ---------------------------------------------------------------------------
function energy_and_grad(h::Array{Float64,4}, bulk::Array{Float64,5}; χ::Int, tol::Real, maxit::Int,
                         opts::TnadOpts = default_opts(), want_grad::Bool = true)
    d, s = size(bulk, 1), size(bulk, 5)
    e = Ref{Cdouble}(0.0)
    it = Ref{Cint}(0)
    g = want_grad ? similar(bulk) : nothing
    o = Ref(opts)
    GC.@preserve h bulk g begin
        check(ccall((:tnad_energy, libtnad), Cint,
                    (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Cint, Cdouble, Cint, Ptr{TnadOpts},
                     Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cint}),
                    context(), h, bulk, d, s, χ, tol, maxit, o, e, want_grad ? g : C_NULL, it))
    end
    e[], g, it[]
end

gpu_energy(h, ipeps::IPEPS; χ::Int, tol::Real, maxit::Int) =
    energy_and_grad(h, ipeps.bulk; χ = χ, tol = tol, maxit = maxit, want_grad = false)[1]

function ChainRulesCore.rrule(::typeof(gpu_energy), h, ipeps::IPEPS; χ::Int, tol::Real, maxit::Int)
    e, g, _ = energy_and_grad(h, ipeps.bulk; χ = χ, tol = tol, maxit = maxit)
    pullback(ē) = (NoTangent(), NoTangent(), Tangent{typeof(ipeps)}(bulk = ē .* g))
    return e, pullback
end

# ComplexF64 bulk (tnad_energy_z): Julia's ComplexF64 memory is the interleaved (re, im) doubles the entry
# takes; the value is complex (e/n) and the gradient is that of real(energy) (src/variationalipeps.jl:71)
function energy_and_grad(h::AbstractArray{<:Number,4}, bulk::Array{ComplexF64,5}; χ::Int, tol::Real, maxit::Int,
                         want_grad::Bool = true)
    d, s = size(bulk, 1), size(bulk, 5)
    hz = Array{ComplexF64}(h)
    e = zeros(Cdouble, 2)
    it = Ref{Cint}(0)
    g = want_grad ? similar(bulk) : nothing
    GC.@preserve hz bulk g e begin
        check(ccall((:tnad_energy_z, libtnad), Cint,
                    (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Cint, Cint, Cint, Cdouble, Cint, Ptr{Cdouble},
                     Ptr{ComplexF64}, Ptr{Cint}),
                    context(), hz, bulk, d, s, χ, tol, maxit, e, want_grad ? g : C_NULL, it))
    end
    complex(e[1], e[2]), g, it[]
end

# ---------------------------------------------------------------------------
# trg
# ---------------------------------------------------------------------------
function trg_and_grad(a::Array{ComplexF64,4}, χ::Int, niter::Int; tol::Float64 = 1e-16, want_grad::Bool = true)
    lnz = zeros(Cdouble, 2)
    g = want_grad ? similar(a) : nothing
    GC.@preserve a g lnz begin
        check(ccall((:tnad_trg_z, libtnad), Cint,
                    (Ptr{Cvoid}, Ptr{ComplexF64}, Cint, Cint, Cint, Cdouble, Ptr{Cdouble}, Ptr{ComplexF64}),
                    context(), a, size(a, 1), χ, niter, tol, lnz, want_grad ? g : C_NULL))
    end
    complex(lnz[1], lnz[2]), g
end

function trg_and_grad(a::Array{Float64,4}, χ::Int, niter::Int; tol::Float64 = 1e-16, want_grad::Bool = true)
    lnz = Ref{Cdouble}(0.0)
    g = want_grad ? similar(a) : nothing
    GC.@preserve a g begin
        check(ccall((:tnad_trg, libtnad), Cint,
                    (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cint, Cint, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}),
                    context(), a, size(a, 1), χ, niter, tol, lnz, want_grad ? g : C_NULL))
    end
    lnz[], g
end

gpu_trg(a::Array{Float64,4}, χ, niter; tol::Float64 = 1e-16) = trg_and_grad(a, χ, niter; tol = tol, want_grad = false)[1]

function ChainRulesCore.rrule(::typeof(gpu_trg), a::Array{Float64,4}, χ, niter; tol::Float64 = 1e-16)
    lnz, g = trg_and_grad(a, χ, niter; tol = tol)
    pullback(z̄) = (NoTangent(), z̄ .* g, NoTangent(), NoTangent())
    return lnz, pullback
end

# trg_svd (src/trg.jl:33-44) and its pullback through svd_back (src/trg.jl:55-61,72-106)
function gpu_trg_svd(t::Array{Float64,4}, dmax::Int, tol::Float64)
    d1, d2, d3, d4 = size(t)
    kmax = min(dmax, d1 * d2, d3 * d4)
    u = zeros(d1, d2, kmax); v = zeros(kmax, d3, d4); k = Ref{Cint}(0)
    GC.@preserve t u v begin
        check(ccall((:tnad_trg_svd, libtnad), Cint,
                    (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cint, Cint, Cint, Cint, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cint}),
                    context(), t, d1, d2, d3, d4, dmax, tol, u, v, k))
    end
    u[:, :, 1:k[]], v[1:k[], :, :]
end

function ChainRulesCore.rrule(::typeof(gpu_trg_svd), t::Array{Float64,4}, dmax::Int, tol::Float64)
    u, v = gpu_trg_svd(t, dmax, tol)
    function pullback(ȳ)
        ū, v̄ = ȳ isa AbstractZero ? (ȳ, ȳ) : (ȳ[1], ȳ[2])
        k = size(u, 3)
        ūd = ū isa AbstractZero ? zero(u) : Array{Float64}(unthunk(ū))
        v̄d = v̄ isa AbstractZero ? zero(v) : Array{Float64}(unthunk(v̄))
        t̄ = similar(t)
        GC.@preserve t ūd v̄d t̄ begin
            check(ccall((:tnad_trg_svd_back, libtnad), Cint,
                        (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cint, Cint, Cint, Cint, Cdouble, Ptr{Cdouble}, Ptr{Cdouble},
                         Cint, Ptr{Cdouble}),
                        context(), t, size(t)..., dmax, tol, ūd, v̄d, k, t̄))
        end
        (NoTangent(), t̄, NoTangent(), NoTangent())
    end
    return (u, v), pullback
end

# ---------------------------------------------------------------------------
# ctmrg with a device tape: the pullback runs the reverse sweep of the same run
# ---------------------------------------------------------------------------
mutable struct Tape
    ptr::Ptr{Cvoid}
    function Tape(p)
        t = new(p)
        finalizer(t) do x
            x.ptr != C_NULL && ccall((:tnad_tape_free, libtnad), Cint, (Ptr{Cvoid},), x.ptr)
            x.ptr = C_NULL
        end
    end
end

function gpu_ctmrg_with_tape(rt::SquareCTMRGRuntime; tol::Real, maxit::Int, keep::Bool, opts::TnadOpts = default_opts())
    a, C0, T0 = rt.bulk, rt.corner, rt.edge
    D, χ = size(a, 1), size(C0, 1)
    C = similar(C0); T = similar(T0); vals = zeros(χ * D)
    it = Ref{Cint}(0)
    tp = Ref{Ptr{Cvoid}}(C_NULL)
    o = Ref(opts)
    GC.@preserve a C0 T0 C T vals begin
        check(ccall((:tnad_ctmrg, libtnad), Cint,
                    (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint, Ptr{TnadOpts},
                     Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cint}, Ptr{Ptr{Cvoid}}),
                    context(), a, D, χ, C0, T0, tol, maxit, o, C, T, vals, it, keep ? tp : C_NULL))
    end
    SquareCTMRGRuntime(a, C, T), (keep ? Tape(tp[]) : nothing)
end

gpu_ctmrg(rt::SquareCTMRGRuntime; tol::Real, maxit::Int) = gpu_ctmrg_with_tape(rt; tol = tol, maxit = maxit, keep = false)[1]

function ChainRulesCore.rrule(::typeof(gpu_ctmrg), rt::SquareCTMRGRuntime; tol::Real, maxit::Int)
    out, tape = gpu_ctmrg_with_tape(rt; tol = tol, maxit = maxit, keep = true)
    function pullback(ȳ)
        C̄ = ȳ.corner isa AbstractZero ? zero(out.corner) : collect(Float64, ȳ.corner)
        T̄ = ȳ.edge isa AbstractZero ? zero(out.edge) : collect(Float64, ȳ.edge)
        ā = similar(out.bulk)
        GC.@preserve C̄ T̄ ā begin
            check(ccall((:tnad_ctmrg_back, libtnad), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                        context(), tape.ptr, C̄, T̄, ā))
        end
        # the initial environment gets no gradient (@nograd _initializect_square, src/autodiff.jl:5)
        return NoTangent(), Tangent{typeof(rt)}(bulk = ā, corner = ZeroTangent(), edge = ZeroTangent())
    end
    return out, pullback
end

end # module
