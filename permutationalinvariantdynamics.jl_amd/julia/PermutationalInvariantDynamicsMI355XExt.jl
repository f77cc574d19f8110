# Julia package extension binding libpimi (include/pimi.h) as the :mi355x
# accelerator backend of PermutationalInvariantDynamics.jl.
#
# Reference hooks this file plugs into (file:line in the reference):
#   capability      _accelerator_extension_capability(::Val)   src/accelerators.jl:13-25,43,65-69
#   upload          _accelerate(::Val{backend}, source, report) src/accelerators.jl:211-212,348-350,366-386
#   operator traits source protocol                             src/source_protocol.jl:191-338,
#                   (template: ThreadedMatrixFreeLiouvillian)   src/threaded_operator.jl:16-156
#   stationary      steady_state(L; method=:krylov, ...)        src/liouvillian.jl:4399-4410,4553-4659,
#                                                               src/krylov.jl:797-891
#   dynamics        _highlevel_expv_interval!                   src/highlevel.jl:467-525
#   trajectories    batched_conditional_action!/rk4!, batched_channel_intensities!,
#                   batched_apply_jumps!                        src/batched_trajectories.jl:266-598
#
# Usage (no edit of the reference needed):
#   compiled = compile(model; backend=:matrixfree)
#   L = mi355x_accelerate(compiled)            # accelerate(compiled; backend=:mi355x, ...)
#   stationary_state(L; algorithm=GMRESAlgorithm(krylovdim=200, maxiter=2000, preconditioner=:sector))
#   solve_dynamics(L, rho0, (0.0, 10.0); algorithm=ExpvAlgorithm(krylovdim=40), saveat=0.1,
#                  observables=(n=spin.jp*spin.jm,))
# Julia is not installed where this file was written: it follows the cited
# reference sources and the C header, and a CPU test checks every struct and
# ccall here against include/pimi.h.
module PermutationalInvariantDynamicsMI355XExt

using LinearAlgebra
using SparseArrays
using PermutationalInvariantDynamics
const PID = PermutationalInvariantDynamics

const libpimi = get(ENV, "PIMI_LIB", "libpimi.so")

# --- status codes (include/pimi.h) -------------------------------------------
const PIMI_OK = Int32(0)
const PIMI_EINVAL = Int32(1)
const PIMI_EDIM = Int32(2)
const PIMI_ENOMEM = Int32(3)
const PIMI_EDEVICE = Int32(4)
const PIMI_EBUSY = Int32(5)
const PIMI_ENOTCONVERGED = Int32(6)
const PIMI_ERANGE = Int32(7)
const PIMI_C64 = Int32(0)
const PIMI_C32 = Int32(1)

# The reference's exception classes: ArgumentError for invalid input,
# busy workspaces (src/threaded_apply.jl:575-576) and non-convergence
# (src/krylov.jl:873) and budgets checked before allocation,
# DimensionMismatch for shapes.
function _check(status::Int32)
    status == PIMI_OK && return nothing
    msg = unsafe_string(ccall((:pimi_last_error, libpimi), Cstring, ()))
    status == PIMI_EDIM && throw(DimensionMismatch(msg))
    status == PIMI_ENOMEM && throw(ArgumentError("device memory: " * msg))
    status in (PIMI_EINVAL, PIMI_EBUSY, PIMI_ENOTCONVERGED, PIMI_ERANGE) && throw(ArgumentError(msg))
    error("libpimi status $status: $msg")
end

# --- C structs: field for field as in include/pimi.h -------------------------
struct PimiTerm
    kind::Int32
    reserved::Int32
    op::Ptr{ComplexF64}
    op2::Ptr{ComplexF64}
    rate::Float64
    hbar::Float64
end

struct PimiModelDesc
    N::Int32
    d::Int32
    nterms::Int32
    reserved::Int32
    terms::Ptr{PimiTerm}
end

struct PimiCscSet
    cp_base::Ptr{Int64}
    colptr::Ptr{Int64}
    rows::Ptr{Int32}
    vals::Ptr{ComplexF64}
end

struct PimiFusedDesc
    ns::Int32
    has_h::Int32
    has_q::Int32
    ncg::Int32
    nob::Int32
    reserved::Int32
    dims::Ptr{Int64}
    H::PimiCscSet
    Q::PimiCscSet
    K::PimiCscSet
    cg_scale::Ptr{ComplexF64}
    ob_scale::Ptr{ComplexF64}
    ob_br_base::Ptr{Int64}
    br_out::Ptr{Int32}
    br_in::Ptr{Int32}
    br_r::Ptr{Float64}
    br_use_support::Ptr{Int32}
    br_sup_base::Ptr{Int64}
    sup_rows::Ptr{Int32}
    sup_cols::Ptr{Int32}
    sup_vals::Ptr{ComplexF64}
    br_dense_base::Ptr{Int64}
    dense::Ptr{ComplexF64}
end

mutable struct PimiPlanInfo
    n::Int64
    nsectors::Int32
    max_block::Int32
    nterms::Int64
    stencil_entries::Int64
    table_bytes_c64::Int64
    table_bytes_c32::Int64
    apply_form::Int32
    banded_groups::Int32
    PimiPlanInfo() = new()
end

struct PimiGmresOpts
    krylovdim::Int32
    preconditioner::Int32
    maxiter::Int64
    atol::Float64
    rtol::Float64
    operator_scale::Float64
    inner_rtol::Float64
    max_refinements::Int64
    criterion::Int32
    inner_steps::Int32
    weighting::Int32
    reserved::Int32
end

struct PimiExpvOpts
    krylovdim::Int32
    reserved::Int32
    atol::Float64
    rtol::Float64
    safety::Float64
    max_steps::Int64
end

mutable struct PimiSolveInfo
    converged::Int32
    reserved::Int32
    iterations::Int64
    restarts::Int64
    applications::Int64
    accepted::Int64
    rejected::Int64
    factorizations::Int64
    refinements::Int64
    residual::Float64
    normalized_residual::Float64
    trace_error::Float64
    operator_scale::Float64
    linear_residual::Float64
    estimated_error::Float64
    seconds::Float64
    relative_residual::Float64
    weighted_relative_residual::Float64
    PimiSolveInfo() = new()
end

# --- context (one per device, created on the first upload) ------------------
const _contexts = Dict{Int,Ptr{Cvoid}}()
const _context_lock = ReentrantLock()
function _context(device::Integer)
    lock(_context_lock) do
        get!(_contexts, Int(device)) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            _check(ccall((:pimi_context_create, libpimi), Int32, (Int32, Ref{Ptr{Cvoid}}), Int32(device), out))
            out[]
        end
    end
end

# --- capability (src/accelerators.jl:13-25, positional, 11 fields) -----------
# The query initializes no device: it only checks that libpimi loads.  The
# device itself is opened by the upload, which raises if it is unusable.
function _library_loads()
    try
        v = ccall((:pimi_version, libpimi), Cstring, ())
        return true, "libpimi " * unsafe_string(v) * " (MI355X, gfx950)"
    catch err
        return false, "libpimi could not be loaded from $(libpimi): $(sprint(showerror, err))"
    end
end

function PID._accelerator_extension_capability(::Val{:mi355x})
    ok, msg = _library_loads()
    PID.AcceleratorCapability(
        :mi355x,                       # backend
        true,                          # extension_loaded
        ok,                            # functional
        true,                          # supports_sparse_pi: the fused Schur blocks upload as CSC
        true,                          # supports_matrixfree_pi: the device applies the Schur-block stencil
        (:vector, :matrix),            # rhs_kinds
        (ComplexF32, ComplexF64),      # scalar_types
        Int32,                         # index_type (device tables index with 32 bits)
        :explicit_once,                # transfer_policy
        ok ? :available : :library_unavailable,
        msg)
end

# --- the device operator (source protocol; template threaded_operator.jl:16-156)
mutable struct MI355XOperator{T}
    plan::Ptr{Cvoid}
    basis::PID.PIBasis
    tracevec::Any
    source::Any
    device::Int
    workspaces::Dict{Tuple{Int32,Int32},Ptr{Cvoid}}
    lock::ReentrantLock
end

function _release!(A::MI355XOperator)
    for ws in values(A.workspaces)
        ccall((:pimi_workspace_destroy, libpimi), Cvoid, (Ptr{Cvoid},), ws)
    end
    empty!(A.workspaces)
    A.plan == C_NULL || ccall((:pimi_plan_destroy, libpimi), Cvoid, (Ptr{Cvoid},), A.plan)
    A.plan = C_NULL
    nothing
end

function _operator(plan::Ptr{Cvoid}, basis, tracevec, source, device)
    A = MI355XOperator{ComplexF64}(plan, basis, tracevec, source, Int(device),
                                   Dict{Tuple{Int32,Int32},Ptr{Cvoid}}(), ReentrantLock())
    finalizer(_release!, A)
    A
end

Base.size(A::MI355XOperator) = (length(A.basis), length(A.basis))
Base.size(A::MI355XOperator, i::Integer) = i in (1, 2) ? length(A.basis) : 1
Base.eltype(::MI355XOperator{T}) where T = T
PID.isautonomous(::MI355XOperator) = true

function Base.show(io::IO, A::MI355XOperator)
    print(io, "MI355XOperator(dimension=$(size(A, 1)), device=$(A.device))")
end

# traits of src/source_protocol.jl:191-338, as ThreadedMatrixFreeLiouvillian
# defines them (src/threaded_operator.jl:139-156): the device operator owns its
# workspaces, so consumers allocate and charge no host LiouvillianWorkspace
PID._operator_basis(A::MI355XOperator) = A.basis
PID._operator_trace_functional(A::MI355XOperator) = A.tracevec
PID._operator_has_adjoint(::MI355XOperator) = true
PID._operator_requires_matrixfree(::MI355XOperator) = true
PID._linear_operator_workspace(::MI355XOperator) = nothing
PID._fixed_liouvillian_scalar_type(A::MI355XOperator) = eltype(A)
PID._performance_linear_operator_workspace_bytes(::MI355XOperator; batch_columns::Integer=0) = begin
    batch_columns >= 0 || throw(ArgumentError("batch_columns must be nonnegative"))
    big(0)
end
PID._performance_source_action_bytes(::MI355XOperator, ::Type{T}) where T = big(0)
PID._resource_source_prepared(::MI355XOperator) = true

# one workspace per (precision, Krylov dimension), owned by the operator
function _workspace(A::MI355XOperator, precision::Int32, krylovdim::Integer)
    lock(A.lock) do
        get!(A.workspaces, (precision, Int32(krylovdim))) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            _check(ccall((:pimi_workspace_create, libpimi), Int32, (Ptr{Cvoid}, Int32, Int32, Ref{Ptr{Cvoid}}),
                         A.plan, precision, Int32(krylovdim), out))
            out[]
        end
    end
end

# --- uploads ------------------------------------------------------------------
_csc(block::SparseMatrixCSC) = block
_csc(block::AbstractMatrix) = sparse(block)

# Flat CSC arrays of a block list, 0-based (pimi_csc_set)
function _flatten_csc(blocks)
    cp_base = Int64[]
    colptr = Int64[]
    rows = Int32[]
    vals = ComplexF64[]
    for B0 in blocks
        B = _csc(B0)
        push!(cp_base, length(colptr))
        base = length(rows)
        for p in B.colptr
            push!(colptr, base + p - 1)
        end
        append!(rows, Int32.(B.rowval .- 1))
        append!(vals, ComplexF64.(B.nzval))
    end
    isempty(cp_base) && push!(cp_base, 0)
    isempty(colptr) && push!(colptr, 0)
    isempty(rows) && push!(rows, Int32(0))
    isempty(vals) && push!(vals, zero(ComplexF64))
    (cp_base, colptr, rows, vals)
end

# Upload of the plan's own FusedStaticPIKernel (src/liouvillian.jl:539-545):
# the device tables are the reference's fused blocks, bit for bit.
function _upload_fused(plan::PID.LiouvillianPlan, device::Integer, source)
    kernels = plan.kernels
    kernels === nothing && throw(ArgumentError("the mi355x backend needs a plan with prepared fixed kernels"))
    (length(kernels) == 1 && first(kernels) isa PID.FusedStaticPIKernel) ||
        throw(ArgumentError("the mi355x fused upload needs one FusedStaticPIKernel; " *
                            "use mi355x_compile(model) for this model"))
    plan.autonomous || throw(ArgumentError("the mi355x backend runs autonomous generators"))
    k = first(kernels)
    isempty(k.pbody_gains) || throw(ArgumentError("p-body gains are not supported by the mi355x backend"))
    b = plan.basis
    ns = length(b.sectors)
    dims = Int64[length(b.patterns[s]) for s in 1:ns]
    empty_blocks = [spzeros(ComplexF64, m, m) for m in dims]
    H = _flatten_csc(k.hamiltonian_blocks === nothing ? empty_blocks : k.hamiltonian_blocks)
    Q = _flatten_csc(k.loss_blocks === nothing ? empty_blocks : k.loss_blocks)
    cgs = collect(k.collective_gains)
    K = _flatten_csc(isempty(cgs) ? empty_blocks : [blk for g in cgs for blk in g.blocks])
    cg_scale = isempty(cgs) ? [zero(ComplexF64)] : ComplexF64[g.scale for g in cgs]
    obs = collect(k.onebody_gains)
    ob_scale = isempty(obs) ? [zero(ComplexF64)] : ComplexF64[g.scale for g in obs]
    ob_br_base = Int64[0]
    br_out = Int32[]; br_in = Int32[]; br_r = Float64[]; use = Int32[]
    sup_base = Int64[0]; sup_rows = Int32[]; sup_cols = Int32[]; sup_vals = ComplexF64[]
    dense_base = Int64[]; dense = ComplexF64[]
    for g in obs
        for (entry, C) in zip(g.branches.entries, g.contractions)
            push!(br_out, Int32(entry.output_sector - 1))
            push!(br_in, Int32(entry.input_sector - 1))
            push!(br_r, Float64(real(entry.scale)))
            push!(use, C.use_support ? Int32(1) : Int32(0))
            if C.use_support
                append!(sup_rows, Int32.(C.output_rows .- 1))
                append!(sup_cols, Int32.(C.input_columns .- 1))
                append!(sup_vals, ComplexF64.(C.values))
            end
            push!(sup_base, length(sup_rows))
            push!(dense_base, length(dense))
            C.use_support || append!(dense, vec(ComplexF64.(C.matrix)))
        end
        push!(ob_br_base, length(br_out))
    end
    for v in (br_out, br_in, use, sup_rows, sup_cols)
        isempty(v) && push!(v, Int32(0))
    end
    isempty(br_r) && push!(br_r, 0.0)
    isempty(sup_vals) && push!(sup_vals, zero(ComplexF64))
    isempty(dense_base) && push!(dense_base, Int64(0))
    isempty(dense) && push!(dense, zero(ComplexF64))
    keep = (dims, H, Q, K, cg_scale, ob_scale, ob_br_base, br_out, br_in, br_r, use, sup_base, sup_rows, sup_cols,
            sup_vals, dense_base, dense)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep begin
        cset(t) = PimiCscSet(pointer(t[1]), pointer(t[2]), pointer(t[3]), pointer(t[4]))
        desc = PimiFusedDesc(Int32(ns), Int32(k.hamiltonian_blocks === nothing ? 0 : 1),
                             Int32(k.loss_blocks === nothing ? 0 : 1), Int32(length(cgs)), Int32(length(obs)),
                             Int32(0), pointer(dims), cset(H), cset(Q), cset(K), pointer(cg_scale),
                             pointer(ob_scale), pointer(ob_br_base), pointer(br_out), pointer(br_in), pointer(br_r),
                             pointer(use), pointer(sup_base), pointer(sup_rows), pointer(sup_cols),
                             pointer(sup_vals), pointer(dense_base), pointer(dense))
        _check(ccall((:pimi_plan_create_fused, libpimi), Int32, (Ptr{Cvoid}, Ref{PimiFusedDesc}, Ref{Ptr{Cvoid}}),
                     _context(device), desc, out))
    end
    _operator(out[], b, plan.tracevec, source, device)
end

# Native lowering from the model's terms (pimi_compile, exact rationals on
# the host).  Models with terms the native lowering does not take (a
# DirectPIHamiltonian holds a ready PI operator, not local factors) upload
# the reference's own fused plan instead.
_term_kind(::PID.LocalHamiltonian) = Int32(0)
_term_kind(::PID.CollectiveHamiltonian) = Int32(1)
_term_kind(::PID.LocalJump) = Int32(2)
_term_kind(::PID.CollectiveJump) = Int32(3)
_term_kind(t) = nothing

function mi355x_compile(model::PID.PIModel; device::Integer=0)
    kinds = [_term_kind(t) for t in model.terms]
    if any(k -> k === nothing, kinds)
        plan = PID.LiouvillianPlan(model)
        return _upload_fused(plan, device, model)
    end
    ops = [Matrix{ComplexF64}(PID.term_operator(t)) for t in model.terms]
    terms = PimiTerm[]
    GC.@preserve ops terms begin
        for (k, t) in enumerate(model.terms)
            hbar = hasproperty(t, :hbar) ? Float64(t.hbar) : 1.0
            push!(terms, PimiTerm(kinds[k], Int32(0), pointer(ops[k]), Ptr{ComplexF64}(C_NULL),
                                  Float64(PID.term_rate(t)), hbar))
        end
        desc = PimiModelDesc(Int32(model.basis.N), Int32(model.basis.d), Int32(length(terms)), Int32(0),
                             pointer(terms))
        out = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:pimi_compile, libpimi), Int32, (Ptr{Cvoid}, Ref{PimiModelDesc}, Ref{Ptr{Cvoid}}),
                     _context(device), desc, out))
        tracevec = PID._trace_functional(model.basis, ComplexF64)
        return _operator(out[], model.basis, tracevec, model, device)
    end
end

# --- upload hook (src/accelerators.jl:348-350; reached after accelerator_preflight) ---
function PID._accelerate(::Val{:mi355x},
                         source::Union{PID.CompiledPIModel,PID.LiouvillianPlan,PID.SpecializedPIModel},
                         report; device::Integer=0, kwargs...)
    plan = PID._accelerator_plan(source)
    kernels = plan.kernels
    if kernels !== nothing && length(kernels) == 1 && first(kernels) isa PID.FusedStaticPIKernel
        return _upload_fused(plan, device, source)
    end
    source isa PID.LiouvillianPlan &&
        throw(ArgumentError("this plan has no single fused kernel; accelerate its CompiledPIModel instead"))
    mi355x_compile(source.model; device)
end

# accelerate() with the host-side budget lifted: the preflight's host figure
# is the peak of a sparse materialization (src/accelerators.jl:255-272) that
# this matrix-free backend never performs.  The device budget stays the caller's.
mi355x_accelerate(source; device::Integer=0, device_memory_budget=Inf, kwargs...) =
    PID.accelerate(source; backend=:mi355x, memory_budget=Inf, device_memory_budget, device, kwargs...)

# --- applications ---------------------------------------------------------------
function _apply!(y::AbstractVecOrMat, A::MI355XOperator, x::AbstractVecOrMat, adjoint::Bool)
    size(x) == size(y) || throw(DimensionMismatch("destination and source shapes differ"))
    size(x, 1) == size(A, 1) || throw(DimensionMismatch("Liouvillian vector has the wrong length"))
    Base.mightalias(y, x) && throw(ArgumentError("source and destination must not share storage"))
    xs = x isa StridedVecOrMat{ComplexF64} && stride(x, 1) == 1 ? x : Matrix{ComplexF64}(reshape(x, size(x, 1), :))
    yb = y isa StridedVecOrMat{ComplexF64} && stride(y, 1) == 1 ? y : similar(xs, ComplexF64)
    ws = _workspace(A, PIMI_C64, 1)
    ncols = Int64(size(x, 2))
    GC.@preserve xs yb begin
        if adjoint
            _check(ccall((:pimi_apply_adjoint, libpimi), Int32,
                         (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Int32), ws, xs, yb, ncols, Int32(0)))
        else
            _check(ccall((:pimi_apply, libpimi), Int32,
                         (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Int32), ws, xs, yb, ncols, Int32(0)))
        end
    end
    yb === y || copyto!(y, yb)
    y
end

# The per-multiplication host<->device route exists for compatibility with
# generic consumers; solvers below never take it.
LinearAlgebra.mul!(y::AbstractVecOrMat, A::MI355XOperator, x::AbstractVecOrMat) = _apply!(y, A, x, false)
PID.apply!(y, A::MI355XOperator, x) = _apply!(y, A, x, false)
PID.apply!(y, A::MI355XOperator, x, t, p) = _apply!(y, A, x, false)
PID.apply_adjoint!(y, A::MI355XOperator, x) = _apply!(y, A, x, true)
PID.apply_adjoint!(y, A::MI355XOperator, x, t, p) = _apply!(y, A, x, true)
Base.:*(A::MI355XOperator, x::AbstractVector) = mul!(similar(x, ComplexF64, size(A, 1)), A, x)

# --- stationary solve (steady_state(L; method=:krylov), src/krylov.jl:797-891) ---
_preconditioner_code(::Nothing) = Int32(0)
_preconditioner_code(p::Symbol) =
    p === :sector ? Int32(1) : p === :sector_inner ? Int32(2) :
    throw(ArgumentError("the mi355x backend offers preconditioner=:sector or :sector_inner"))
_preconditioner_code(p) =
    throw(ArgumentError("the mi355x backend runs its own device preconditioners (:sector, :sector_inner)"))

function PID.steady_state(A::MI355XOperator; method=:krylov, basis=nothing, trace_vector=nothing,
                          memory_budget=nothing, initial_state=nothing, krylovdim::Integer=30,
                          maxiter::Integer=500, atol::Real=1e-10, rtol::Real=1e-8, preconditioner=nothing,
                          operator_scale=nothing, return_info::Bool=false, criterion::Symbol=:reference,
                          weighting=nothing, inner_steps::Integer=8, precision=ComplexF64,
                          workspace=nothing, recycle_dim::Integer=0)
    method in (:auto, :krylov, :gmres) ||
        throw(ArgumentError("the mi355x backend solves stationary states with method=:krylov"))
    recycle_dim == 0 || throw(ArgumentError("recycled GMRES is not offered by the mi355x backend"))
    workspace === nothing || throw(ArgumentError("the mi355x operator owns its device workspace"))
    trace_vector === nothing || trace_vector == A.tracevec ||
        throw(ArgumentError("the mi355x backend uses the plan's own trace functional"))
    n = size(A, 1)
    prec = precision === ComplexF32 ? PIMI_C32 : PIMI_C64
    ws = _workspace(A, prec, min(krylovdim, n))
    opts = PimiGmresOpts(Int32(krylovdim), _preconditioner_code(preconditioner), Int64(maxiter), Float64(atol),
                         Float64(rtol), operator_scale === nothing ? 0.0 : Float64(operator_scale), 1e-5,
                         Int64(50), criterion === :weighted ? Int32(2) : criterion === :backward ? Int32(1) : Int32(0),
                         Int32(inner_steps),
                         weighting === :sector ? Int32(1) : Int32(0), Int32(0))
    x0 = initial_state === nothing ? nothing :
         Vector{ComplexF64}(initial_state isa PID.PIState ? initial_state.data : initial_state)
    x0 === nothing || length(x0) == n || throw(DimensionMismatch("initial_state has wrong length"))
    rho = zeros(ComplexF64, n)
    info = PimiSolveInfo()
    GC.@preserve x0 rho begin
        _check(ccall((:pimi_steady_gmres, libpimi), Int32,
                     (Ptr{Cvoid}, Ref{PimiGmresOpts}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ref{PimiSolveInfo}),
                     ws, opts, x0 === nothing ? Ptr{ComplexF64}(C_NULL) : pointer(x0), rho, info))
    end
    return_info || return rho
    # the fields of the reference's info tuple (src/krylov.jl:876-889)
    (state=rho, residual=info.residual, trace_error=info.trace_error, nullity=nothing, method=:krylov,
     iterations=Int(info.iterations), eigenvalue=nothing, converged=true, linear_residual=info.linear_residual,
     unpreconditioned_linear_residual=info.linear_residual, normalized_residual=info.normalized_residual,
     operator_scale=info.operator_scale, restarts=Int(info.restarts), krylov_dimension=min(krylovdim, n),
     krylov_variant=:restarted, recycle_dimension=0, recycled_initially=false, recycle_reset=false,
     operator_applications=Int(info.applications), preconditioned=preconditioner !== nothing,
     preconditioner=preconditioner === nothing ? nothing : typeof(preconditioner), preconditioner_cost=nothing,
     relative_residual=info.relative_residual, weighted_relative_residual=info.weighted_relative_residual,
     device_seconds=info.seconds)
end

PID.krylov_steady_state(A::MI355XOperator; kwargs...) = PID.steady_state(A; method=:krylov, kwargs...)

# --- dynamics: one expv interval on the device (src/highlevel.jl:510-525) --------
function PID._highlevel_expv_interval!(current::PID.PIState,
                                      operator::PID._HighLevelExpvOperator{T,<:MI355XOperator},
                                      workspace, interval, options) where T
    iszero(interval) && return current
    interval > zero(interval) || throw(ArgumentError("Krylov exponential output times must be nondecreasing"))
    for key in (:initial_step, :minimum_step, :maximum_step)
        get(options, key, nothing) === nothing ||
            throw(ArgumentError("the mi355x backend takes no $key (the step control of krylov_expv! is used)"))
    end
    A = operator.source
    n = size(A, 1)
    krylovdim = get(options, :krylovdim, 30)
    ws = _workspace(A, PIMI_C64, min(krylovdim, n))
    opts = PimiExpvOpts(Int32(krylovdim), Int32(0), Float64(get(options, :atol, 1e-10)),
                        Float64(get(options, :rtol, 1e-8)), Float64(get(options, :safety, 0.9)),
                        Int64(get(options, :max_steps, 10_000)))
    ts = Float64[0.0, Float64(interval)]
    x = Vector{ComplexF64}(current.data)
    info = PimiSolveInfo()
    GC.@preserve x ts begin
        _check(ccall((:pimi_expv_series, libpimi), Int32,
                     (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Float64}, Int64, Ref{PimiExpvOpts}, Int32, Ptr{ComplexF64},
                      Ptr{ComplexF64}, Ptr{ComplexF64}, Ref{PimiSolveInfo}),
                     ws, x, ts, Int64(2), opts, Int32(0), Ptr{ComplexF64}(C_NULL), Ptr{ComplexF64}(C_NULL), x,
                     info))
    end
    copyto!(current.data, x)
    current
end

# Whole series in one call (observables as device dots, one host copy per
# series): out[k, o] = <obs[:, o], rho(ts[k])>.
function mi355x_expv_series(A::MI355XOperator, rho0::AbstractVector, ts::AbstractVector{<:Real},
                            obs::AbstractMatrix; krylovdim::Integer=30, atol::Real=1e-10, rtol::Real=1e-8,
                            safety::Real=0.9, max_steps::Integer=10_000)
    n = size(A, 1)
    size(obs, 1) == n || throw(DimensionMismatch("observables must have the PI dimension"))
    ws = _workspace(A, PIMI_C64, min(krylovdim, n))
    opts = PimiExpvOpts(Int32(krylovdim), Int32(0), Float64(atol), Float64(rtol), Float64(safety), Int64(max_steps))
    x = Vector{ComplexF64}(rho0)
    O = Matrix{ComplexF64}(obs)
    tv = Vector{Float64}(ts)
    out = zeros(ComplexF64, size(obs, 2), length(tv))
    final = similar(x)
    info = PimiSolveInfo()
    GC.@preserve x O tv out final begin
        _check(ccall((:pimi_expv_series, libpimi), Int32,
                     (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Float64}, Int64, Ref{PimiExpvOpts}, Int32, Ptr{ComplexF64},
                      Ptr{ComplexF64}, Ptr{ComplexF64}, Ref{PimiSolveInfo}),
                     ws, x, tv, Int64(length(tv)), opts, Int32(size(O, 2)), O, out, final, info))
    end
    permutedims(out), final, info
end

# --- conditional trajectory cohorts (src/batched_trajectories.jl:266-598) --------
# BatchedConditionalPlan(L) of a device operator is a device plan; the four
# primitives run on the device; batched_trajectory_rngs and
# batched_sample_jumps! (host RNG draws) are the reference's own, unchanged.
struct MI355XConditionalPlan
    operator::MI355XOperator
end
struct MI355XConditionalWorkspace
    plan::MI355XConditionalPlan
    capacity::Int
end

PID.BatchedConditionalPlan(A::MI355XOperator) = MI355XConditionalPlan(A)
PID.BatchedConditionalPlan(A::MI355XOperator, rho::PID.PIState) = MI355XConditionalPlan(A)
Base.size(P::MI355XConditionalPlan) = size(P.operator)
Base.size(P::MI355XConditionalPlan, i::Integer) = size(P.operator, i)
Base.eltype(P::MI355XConditionalPlan) = eltype(P.operator)
PID.isautonomous(::MI355XConditionalPlan) = true

function PID.BatchedConditionalWorkspace(plan::MI355XConditionalPlan, rho::PID.PIState, capacity::Integer;
                                         memory_budget=nothing)
    capacity > 0 || throw(ArgumentError("batch capacity must be positive"))
    rho.basis === plan.operator.basis ||
        throw(ArgumentError("state and batched conditional plan use incompatible PI bases"))
    MI355XConditionalWorkspace(plan, Int(capacity))
end

function _cohort(X::AbstractMatrix, work::MI355XConditionalWorkspace)
    size(X, 1) == size(work.plan, 1) || throw(DimensionMismatch("batched trajectory state has the wrong length"))
    size(X, 2) <= work.capacity || throw(DimensionMismatch("batch is wider than the workspace capacity"))
    X isa StridedMatrix{ComplexF64} && stride(X, 1) == 1 ||
        throw(ArgumentError("batched trajectory states must be a dense ComplexF64 matrix"))
    _workspace(work.plan.operator, PIMI_C64, 1)
end

function PID.batched_conditional_action!(Y::AbstractMatrix, intensities::AbstractVector,
                                         plan::MI355XConditionalPlan, X::AbstractMatrix, t, p,
                                         work::MI355XConditionalWorkspace)
    ws = _cohort(X, work)
    size(Y) == size(X) || throw(DimensionMismatch("destination and source shapes differ"))
    length(intensities) == size(X, 2) || throw(DimensionMismatch("one intensity per column"))
    lam = zeros(Float64, size(X, 2))
    GC.@preserve X Y lam begin
        _check(ccall((:pimi_conditional_action, libpimi), Int32,
                     (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Ptr{Float64}, Int32),
                     ws, X, Y, Int64(size(X, 2)), lam, Int32(0)))
    end
    intensities .= lam
    Y
end
PID.batched_conditional_action!(Y::AbstractMatrix, intensities::AbstractVector, plan::MI355XConditionalPlan,
                                X::AbstractMatrix, work::MI355XConditionalWorkspace) =
    PID.batched_conditional_action!(Y, intensities, plan, X, 0.0, nothing, work)

function PID.batched_conditional_rk4!(X::AbstractMatrix, plan::MI355XConditionalPlan, t, h, p,
                                      work::MI355XConditionalWorkspace)
    ws = _cohort(X, work)
    h > 0 || throw(ArgumentError("batched trajectory step must be positive"))
    GC.@preserve X begin
        _check(ccall((:pimi_conditional_rk4, libpimi), Int32, (Ptr{Cvoid}, Ptr{ComplexF64}, Int64, Float64, Int32),
                     ws, X, Int64(size(X, 2)), Float64(h), Int32(0)))
    end
    X
end

function PID.batched_channel_intensities!(rates::AbstractMatrix, plan::MI355XConditionalPlan, X::AbstractMatrix,
                                          t, p, work::MI355XConditionalWorkspace)
    ws = _cohort(X, work)
    count = Ref{Int32}(0)
    _check(ccall((:pimi_trajectory_channels, libpimi), Int32, (Ptr{Cvoid}, Ref{Int32}), plan.operator.plan, count))
    size(rates) == (Int(count[]), size(X, 2)) ||
        throw(DimensionMismatch("channel intensity destination must have dimensions $((Int(count[]), size(X, 2)))"))
    buf = zeros(Float64, Int(count[]), size(X, 2))
    GC.@preserve X buf begin
        _check(ccall((:pimi_channel_intensities, libpimi), Int32, (Ptr{Cvoid}, Ptr{ComplexF64}, Int64, Ptr{Float64}, Int32),
                     ws, X, Int64(size(X, 2)), buf, Int32(0)))
    end
    rates .= buf
    rates
end
PID.batched_channel_intensities!(rates::AbstractMatrix, plan::MI355XConditionalPlan, X::AbstractMatrix,
                                 work::MI355XConditionalWorkspace) =
    PID.batched_channel_intensities!(rates, plan, X, 0.0, nothing, work)

function PID.batched_apply_jumps!(X::AbstractMatrix, channels::AbstractVector{<:Integer},
                                  plan::MI355XConditionalPlan, t, p, work::MI355XConditionalWorkspace)
    ws = _cohort(X, work)
    length(channels) == size(X, 2) || throw(DimensionMismatch("channels must contain one entry per batch column"))
    eltype(channels) <: Bool && throw(ArgumentError("selected jump channels must be integer indices, not Bool"))
    ch = Int32.(channels)
    GC.@preserve X ch begin
        _check(ccall((:pimi_apply_jumps, libpimi), Int32, (Ptr{Cvoid}, Ptr{ComplexF64}, Int64, Ptr{Int32}, Int32),
                     ws, X, Int64(size(X, 2)), ch, Int32(0)))
    end
    X
end
PID.batched_apply_jumps!(X::AbstractMatrix, channels::AbstractVector{<:Integer}, plan::MI355XConditionalPlan,
                         work::MI355XConditionalWorkspace) =
    PID.batched_apply_jumps!(X, channels, plan, 0.0, nothing, work)

export MI355XOperator, mi355x_compile, mi355x_accelerate, mi355x_expv_series

end # module
