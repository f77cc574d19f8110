# Julia package extension binding libpimi (include/pimi.h) as the :mi355x
# accelerator backend of PermutationalInvariantDynamics.jl.  Reference hooks:
# src/accelerators.jl:13-25,43,348-386; solver entry points src/krylov.jl:797,
# src/highlevel.jl:510-525.  Not run here (no Julia toolchain in this image).
using LinearAlgebra
module PermutationalInvariantDynamicsMI355XExt
using PermutationalInvariantDynamics
const PID = PermutationalInvariantDynamics
const libpimi = get(ENV, "PIMI_LIB", "libpimi.so")

# --- status handling ------------------------------------------------------
function _check(status::Int32)
    status == 0 && return nothing
    msg = unsafe_string(ccall((:pimi_last_error, libpimi), Cstring, ()))
    status == 1 && throw(ArgumentError(msg))          # PIMI_EINVAL
    status == 2 && throw(DimensionMismatch(msg))      # PIMI_EDIM
    status == 6 && throw(ArgumentError(msg))          # PIMI_ENOTCONVERGED (reference raises ArgumentError, src/krylov.jl:873)
    error("libpimi status $status: $msg")
end

# --- C structs (must match include/pimi.h) ---------------------------------
struct PimiTerm;  kind::Int32; reserved::Int32; op::Ptr{ComplexF64}; op2::Ptr{ComplexF64}; rate::Float64; hbar::Float64; end
struct PimiModel; N::Int32; d::Int32; nterms::Int32; reserved::Int32; terms::Ptr{PimiTerm}; end
struct PimiGmresOpts
    krylovdim::Int32; preconditioner::Int32; maxiter::Int64; atol::Float64; rtol::Float64
    operator_scale::Float64; inner_rtol::Float64; max_refinements::Int64; criterion::Int32; inner_steps::Int32
    weighting::Int32; reserved::Int32
end
struct PimiExpvOpts; krylovdim::Int32; reserved::Int32; atol::Float64; rtol::Float64; safety::Float64; max_steps::Int64; end
mutable struct PimiInfo
    converged::Int32; reserved::Int32; iterations::Int64; restarts::Int64; applications::Int64
    accepted::Int64; rejected::Int64; factorizations::Int64; refinements::Int64
    residual::Float64; normalized_residual::Float64; trace_error::Float64; operator_scale::Float64
    linear_residual::Float64; estimated_error::Float64; seconds::Float64; relative_residual::Float64
    PimiInfo() = new()
end

const _ctx = Ref{Ptr{Cvoid}}(C_NULL)
function _context()
    if _ctx[] == C_NULL
        out = Ref{Ptr{Cvoid}}()
        _check(ccall((:pimi_context_create, libpimi), Int32, (Int32, Ref{Ptr{Cvoid}}), 0, out))
        _ctx[] = out[]
    end
    _ctx[]
end

# --- device operator (source protocol, src/threaded_operator.jl:16-156 is the template) ---
mutable struct MI355XOperator{T}
    plan::Ptr{Cvoid}
    basis::PID.PIBasis
    workspaces::Dict{Tuple{Int32,Int32},Ptr{Cvoid}}
end
Base.size(A::MI355XOperator) = (length(A.basis), length(A.basis))
Base.eltype(::MI355XOperator{T}) where T = T
PID.isautonomous(::MI355XOperator) = true
PID._operator_basis(A::MI355XOperator) = A.basis
PID._operator_requires_matrixfree(::MI355XOperator) = true

function _workspace(A::MI355XOperator, precision::Int32, krylovdim::Integer)
    get!(A.workspaces, (precision, Int32(krylovdim))) do
        out = Ref{Ptr{Cvoid}}()
        _check(ccall((:pimi_workspace_create, libpimi), Int32, (Ptr{Cvoid}, Int32, Int32, Ref{Ptr{Cvoid}}),
                     A.plan, precision, krylovdim, out))
        out[]
    end
end

# compatibility route: one host->device->host apply per mul!
function LinearAlgebra.mul!(y::AbstractVector{ComplexF64}, A::MI355XOperator, x::AbstractVector{ComplexF64})
    ws = _workspace(A, Int32(0), 1)
    _check(ccall((:pimi_apply, libpimi), Int32, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Int32),
                 ws, x, y, 1, 0))
    y
end

# --- compile(model; backend=:mi355x) -------------------------------------
_kind(t::PID.LocalHamiltonian) = Int32(0)
_kind(t::PID.CollectiveHamiltonian) = Int32(1)
_kind(t::PID.LocalJump) = Int32(2)
_kind(t::PID.CollectiveJump) = Int32(3)

function mi355x_compile(model::PID.PIModel)
    ops = [Matrix{ComplexF64}(PID.term_operator(t)) for t in model.terms]
    terms = [PimiTerm(_kind(t), 0, pointer(ops[k]), C_NULL, Float64(PID.term_rate(t)),
                      hasproperty(t, :hbar) ? Float64(t.hbar) : 1.0) for (k, t) in enumerate(model.terms)]
    desc = PimiModel(model.basis.N, model.basis.d, length(terms), 0, pointer(terms))
    out = Ref{Ptr{Cvoid}}()
    GC.@preserve ops terms begin
        _check(ccall((:pimi_compile, libpimi), Int32, (Ptr{Cvoid}, Ref{PimiModel}, Ref{Ptr{Cvoid}}),
                     _context(), desc, out))
    end
    MI355XOperator{ComplexF64}(out[], model.basis, Dict{Tuple{Int32,Int32},Ptr{Cvoid}}())
end

# --- accelerator hooks (src/accelerators.jl:43, 348-350) ------------------
PID._accelerator_extension_capability(::Val{:mi355x}) = PID.AcceleratorCapability(
    backend=:mi355x, functional=true, supports_matrixfree_pi=true, transfer_policy=:explicit_once,
    element_types=(ComplexF32, ComplexF64), index_type=Int32)
PID._accelerate(::Val{:mi355x}, source::PID.PIModel, report; kwargs...) = mi355x_compile(source)

# --- solver interception ----------------------------------------------------
function PID.steady_state(A::MI355XOperator; krylovdim=30, maxiter=500, atol=1e-10, rtol=1e-8,
                          operator_scale=nothing, preconditioner=nothing, return_info=false, kwargs...)
    ws = _workspace(A, Int32(0), krylovdim)
    opts = PimiGmresOpts(krylovdim, preconditioner === :sector ? 1 : 0, maxiter, atol, rtol,
                         something(operator_scale, 0.0), 1e-5, 50, 0, 0, 0, 0)
    rho = zeros(ComplexF64, length(A.basis)); info = PimiInfo()
    _check(ccall((:pimi_steady_gmres, libpimi), Int32, (Ptr{Cvoid}, Ref{PimiGmresOpts}, Ptr{ComplexF64}, Ref{PimiInfo}),
                 ws, opts, rho, info))
    return_info ? (state=rho, residual=info.residual, trace_error=info.trace_error, iterations=info.iterations,
                   restarts=info.restarts, operator_scale=info.operator_scale, converged=true, method=:krylov) : rho
end

function mi355x_expv_series(A::MI355XOperator, rho0::Vector{ComplexF64}, ts::Vector{Float64},
                            obs::Matrix{ComplexF64}; krylovdim=30, atol=1e-10, rtol=1e-8, safety=0.9, max_steps=10_000)
    ws = _workspace(A, Int32(0), krylovdim)
    opts = PimiExpvOpts(krylovdim, 0, atol, rtol, safety, max_steps)
    out = zeros(ComplexF64, size(obs, 2), length(ts)); final = similar(rho0); info = PimiInfo()
    _check(ccall((:pimi_expv_series, libpimi), Int32,
                 (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Float64}, Int64, Ref{PimiExpvOpts}, Int32, Ptr{ComplexF64},
                  Ptr{ComplexF64}, Ptr{ComplexF64}, Ref{PimiInfo}),
                 ws, rho0, ts, length(ts), opts, size(obs, 2), obs, out, final, info))
    out, final, info
end

end # module
