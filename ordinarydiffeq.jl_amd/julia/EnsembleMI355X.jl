# EnsembleMI355X.jl -- Julia front-end of libodemi (include/odemi.h).
#
# A drop-in `EnsembleAlgorithm` for SciMLBase's ensemble interface:
#
#     using OrdinaryDiffEq, EnsembleMI355X
#     sim = solve(EnsembleProblem(prob; prob_func), Tsit5(), EnsembleMI355X();
#                 trajectories = 10^6, reltol = 1e-6, abstol = 1e-6, saveat = [0.5, 1.0])
#
# The RHS (and `jac` / `tgrad` when the ODEFunction has them) is traced once
# with symbolic arguments into straight-line C++ (`odemi_f`, `odemi_jac`,
# `odemi_tgrad`, templated on the arithmetic type), compiled by hipRTC inside
# libodemi, and every trajectory is integrated by one `odemi_solve_batch` call.
# Missing Jacobians and time derivatives are differentiated symbolically.
#
# Scope of this front-end: systems that fit a lane (Tsit5 n <= 16, Rodas5P
# n <= 4) and whose RHS is straight-line arithmetic on u, p, t (+ - * / ^ and
# sin cos exp log sqrt tanh abs).  Larger and uneven-size Rodas5P batches use
# the workgroup kernels through the Python front-end (odemi/api.py), which also
# emits their rerolled forms.  This image has no Julia, so this file is not
# exercised by the test suite; the Python binding (odemi/_lib.py) makes the same
# calls and is.
module EnsembleMI355X

using SciMLBase
using OrdinaryDiffEq: Tsit5, Rodas5P

export EnsembleMI355X

const libodemi = joinpath(@__DIR__, "..", "odemi", "libodemi.so")
const TSIT5, RODAS5P = Cint(0), Cint(1)
const ENGINE_PLAIN, ENGINE_PERSISTENT = Cint(0), Cint(1)
const FP64, MIXED = Cint(0), Cint(1)
const COMPILE_TSIT5, COMPILE_RODAS5P, HAS_USER_TGRAD = Cint(1), Cint(2), Cint(4)

"""
    EnsembleMI355X(; device = 0, engine = :persistent, precision = :fp64, refine_sweeps = 1)

`engine = :plain` is the FP64 one-kernel-per-step engine; `precision = :mixed`
runs Tsit5 stages (or the Rodas5P LU) in FP32 with FP64 state, error norm and
step control.
"""
struct EnsembleMI355X <: SciMLBase.EnsembleAlgorithm
    device::Int
    engine::Symbol
    precision::Symbol
    refine_sweeps::Int
end
EnsembleMI355X(; device = 0, engine = :persistent, precision = :fp64, refine_sweeps = 1) =
    EnsembleMI355X(device, engine, precision, refine_sweeps)

# ---------------------------------------------------------------------------
# C ABI (field for field odemi_solve_args)
# ---------------------------------------------------------------------------
mutable struct OdemiSolveArgs
    alg::Cint; engine::Cint; precision::Cint; device_ptrs::Cint
    N::Int64; nmax::Cint; np::Cint
    nvec::Ptr{Int32}; u0::Ptr{Float64}; p::Ptr{Float64}; t0::Ptr{Float64}; tf::Ptr{Float64}
    reltol::Float64; abstol::Float64; dtmax::Float64; dtmin::Float64; dt::Float64
    maxiters::Int64
    saveat::Ptr{Float64}; nsaveat::Cint
    tstops::Ptr{Float64}; ntstops::Cint
    save_start::Cint; save_end::Cint; save_everystep::Cint; save_end_user::Cint
    max_saves::Cint
    t_out::Ptr{Float64}; u_out::Ptr{Float64}; nsaved::Ptr{Int32}
    u_final::Ptr{Float64}; t_final::Ptr{Float64}; retcode::Ptr{Int32}; stats::Ptr{Int32}
    refine_sweeps::Cint; block::Cint; blocks_per_cu::Cint; sort_by_cost::Cint
    stream::Ptr{Cvoid}
    kernel_ms::Float64; launches::Int64
    OdemiSolveArgs() = new()
end

odemi_create(dev) = ccall((:odemi_create, libodemi), Ptr{Cvoid}, (Cint,), dev)
odemi_last_error(ctx) = unsafe_string(ccall((:odemi_last_error, libodemi), Cstring, (Ptr{Cvoid},), ctx))
odemi_compile_rhs(ctx, src, n, np, flags) =
    ccall((:odemi_compile_rhs, libodemi), Ptr{Cvoid}, (Ptr{Cvoid}, Cstring, Cint, Cint, Cint, Ptr{Float64}),
          ctx, src, n, np, flags, C_NULL)
odemi_default_args!(a::OdemiSolveArgs, alg) =
    ccall((:odemi_default_args, libodemi), Cvoid, (Ref{OdemiSolveArgs}, Cint), a, alg)
odemi_solve_batch(ctx, rhs, a::OdemiSolveArgs) =
    ccall((:odemi_solve_batch, libodemi), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{OdemiSolveArgs}), ctx, rhs, a)

const CONTEXTS = Dict{Int,Ptr{Cvoid}}()
const COMPILED = Dict{Tuple{String,Int,Int,Cint},Ptr{Cvoid}}()

function context(dev)
    get!(CONTEXTS, dev) do
        h = odemi_create(dev)
        h == C_NULL && error("odemi_create failed")
        h
    end
end

# ---------------------------------------------------------------------------
# Tracer: symbolic expression DAG, hash-consed, constants folded
# ---------------------------------------------------------------------------
struct Node
    op::Symbol            # :u :p :t :const :add :sub :mul :div :neg :pow :sin :cos :exp :log :sqrt :tanh :abs
    args::Vector{Int}
    val::Float64          # constant value, or the index of u/p
end
mutable struct Graph
    nodes::Vector{Node}
    memo::Dict{Tuple{Symbol,Vector{Int},Float64},Int}
end
Graph() = Graph(Node[], Dict{Tuple{Symbol,Vector{Int},Float64},Int}())
const G = Ref(Graph())

struct Sym <: Real
    id::Int
end
function node(op, args = Int[], val = 0.0)
    g = G[]
    key = (op, args, val)
    haskey(g.memo, key) && return Sym(g.memo[key])
    push!(g.nodes, Node(op, args, val))
    g.memo[key] = length(g.nodes)
    Sym(length(g.nodes))
end
isconst(a::Sym) = G[].nodes[a.id].op === :const
cval(a::Sym) = G[].nodes[a.id].val
Base.convert(::Type{Sym}, x::Real) = node(:const, Int[], Float64(x))
Sym(x::Real) = convert(Sym, x)
Base.promote_rule(::Type{Sym}, ::Type{<:Real}) = Sym

function binop(op, f, a::Sym, b::Sym)
    isconst(a) && isconst(b) && return Sym(f(cval(a), cval(b)))
    node(op, [a.id, b.id])
end
Base.:+(a::Sym, b::Sym) = isconst(a) && cval(a) == 0 ? b : isconst(b) && cval(b) == 0 ? a : binop(:add, +, a, b)
Base.:-(a::Sym, b::Sym) = isconst(b) && cval(b) == 0 ? a : binop(:sub, -, a, b)
function Base.:*(a::Sym, b::Sym)
    if (isconst(a) && cval(a) == 0) || (isconst(b) && cval(b) == 0)
        return Sym(0.0)
    end
    isconst(a) && cval(a) == 1 && return b
    isconst(b) && cval(b) == 1 && return a
    binop(:mul, *, a, b)
end
Base.:/(a::Sym, b::Sym) = isconst(b) && cval(b) == 1 ? a : binop(:div, /, a, b)
Base.:-(a::Sym) = isconst(a) ? Sym(-cval(a)) : node(:neg, [a.id])
Base.:^(a::Sym, k::Integer) = k == 0 ? Sym(1.0) : k == 1 ? a : k == 2 ? a * a : k < 0 ? Sym(1.0) / a^(-k) : a^(k ÷ 2) * a^(k - k ÷ 2)
Base.:^(a::Sym, b::Sym) = binop(:pow, ^, a, b)
Base.:^(a::Sym, b::Real) = isinteger(b) ? a^Int(b) : a^Sym(b)
for f in (:sin, :cos, :exp, :log, :sqrt, :tanh, :abs)
    @eval Base.$f(a::Sym) = isconst(a) ? Sym($f(cval(a))) : node($(QuoteNode(f)), [a.id])
end

# d(e)/d(var) as a new Sym
function deriv(e::Sym, var::Sym, memo = Dict{Int,Sym}())
    haskey(memo, e.id) && return memo[e.id]
    n = G[].nodes[e.id]
    A(i) = Sym(n.args[i])
    D(i) = deriv(A(i), var, memo)
    r = if e.id == var.id
        Sym(1.0)
    elseif n.op in (:u, :p, :t, :const)
        Sym(0.0)
    elseif n.op === :add; D(1) + D(2)
    elseif n.op === :sub; D(1) - D(2)
    elseif n.op === :mul; D(1) * A(2) + A(1) * D(2)
    elseif n.op === :div; (D(1) * A(2) - A(1) * D(2)) / (A(2) * A(2))
    elseif n.op === :neg; -D(1)
    elseif n.op === :pow; e * (D(2) * log(A(1)) + A(2) * D(1) / A(1))
    elseif n.op === :sin; cos(A(1)) * D(1)
    elseif n.op === :cos; -sin(A(1)) * D(1)
    elseif n.op === :exp; e * D(1)
    elseif n.op === :log; D(1) / A(1)
    elseif n.op === :sqrt; D(1) / (Sym(2.0) * e)
    elseif n.op === :tanh; (Sym(1.0) - e * e) * D(1)
    elseif n.op === :abs; D(1) * node(:div, [A(1).id, e.id])
    else error("no derivative for $(n.op)")
    end
    memo[e.id] = r
    r
end

cliteral(x::Float64) = isfinite(x) ? "T($(repr(x)))" : (isnan(x) ? "T(NAN)" : (x > 0 ? "T(INFINITY)" : "T(-INFINITY)"))
function cexpr(id)
    n = G[].nodes[id]
    a(i) = cref(n.args[i])
    n.op === :add ? "$(a(1)) + $(a(2))" : n.op === :sub ? "$(a(1)) - $(a(2))" :
    n.op === :mul ? "$(a(1)) * $(a(2))" : n.op === :div ? "$(a(1)) / $(a(2))" :
    n.op === :neg ? "-$(a(1))" : n.op === :pow ? "pow($(a(1)), $(a(2)))" :
    n.op === :abs ? "fabs($(a(1)))" : "$(n.op)($(a(1)))"
end
function cref(id)
    n = G[].nodes[id]
    n.op === :u ? "u[$(Int(n.val))]" : n.op === :p ? "p[$(Int(n.val))]" : n.op === :t ? "t" :
    n.op === :const ? cliteral(n.val) : "v$id"
end
function emit(name, outname, outs::Vector{Sym})
    needed = falses(length(G[].nodes))
    stack = [o.id for o in outs]
    while !isempty(stack)
        i = pop!(stack)
        needed[i] && continue
        needed[i] = true
        append!(stack, G[].nodes[i].args)
    end
    io = IOBuffer()
    println(io, "template <typename T>\n__device__ __forceinline__ void $name(T* __restrict__ $outname, ",
            "const T* __restrict__ u, const T* __restrict__ p, T t) {\n    (void)u; (void)p; (void)t;")
    for i in eachindex(needed)   # ids are created in dependency order
        needed[i] || continue
        G[].nodes[i].op in (:u, :p, :t, :const) && continue
        println(io, "    const T v$i = ", cexpr(i), ";")
    end
    for (k, o) in enumerate(outs)
        println(io, "    $outname[$(k - 1)] = ", cref(o.id), ";")
    end
    println(io, "}")
    String(take!(io))
end

"""Trace `f!(du, u, p, t)` (and optional jac/tgrad) for state size n, np parameters."""
function trace(fun::SciMLBase.ODEFunction, n, np)
    G[] = Graph()
    u = [node(:u, Int[], Float64(i)) for i in 0:n-1]
    p = [node(:p, Int[], Float64(k)) for k in 0:max(np, 1)-1]
    t = node(:t)
    du = Vector{Any}(fill(Sym(0.0), n))
    fun.f(du, u, np > 0 ? p : SciMLBase.NullParameters(), t)
    f = Sym[convert(Sym, x) for x in du]
    J = if fun.jac !== nothing
        Jm = Matrix{Any}(fill(Sym(0.0), n, n))
        fun.jac(Jm, u, p, t)
        Sym[convert(Sym, Jm[i, j]) for i in 1:n for j in 1:n]     # row-major
    else
        Sym[deriv(f[i], u[j]) for i in 1:n for j in 1:n]
    end
    dT = if fun.tgrad !== nothing
        g = Vector{Any}(fill(Sym(0.0), n)); fun.tgrad(g, u, p, t); Sym[convert(Sym, x) for x in g]
    else
        Sym[deriv(f[i], t) for i in 1:n]
    end
    src = emit("odemi_f", "du", f) * emit("odemi_jac", "J", J) * emit("odemi_tgrad", "dT", dT)
    src, fun.tgrad !== nothing
end

# ---------------------------------------------------------------------------
# solve
# ---------------------------------------------------------------------------
function SciMLBase.__solve(ep::SciMLBase.EnsembleProblem, alg::Union{Tsit5,Rodas5P}, ea::EnsembleMI355X;
                           trajectories, reltol = 1e-3, abstol = 1e-6, saveat = Float64[], save_start = nothing,
                           save_end = nothing, save_everystep = isempty(saveat), dt = 0.0, dtmax = 0.0,
                           dtmin = 0.0, maxiters = 1_000_000, max_saves = 4096, tstops = Float64[], kwargs...)
    # this front-end traces f / jac / tgrad only; the Python front-end also
    # traces isoutofdomain and DiscreteCallbacks
    for k in (:callback, :isoutofdomain)
        haskey(kwargs, k) && error("EnsembleMI355X (Julia front-end): `$k` is not supported")
    end
    isempty(kwargs) || error("EnsembleMI355X: unsupported keyword(s) $(collect(keys(kwargs)))")
    probs = [ep.prob_func(deepcopy(ep.prob), i, 1) for i in 1:trajectories]
    n = length(probs[1].u0)
    np = probs[1].p isa SciMLBase.NullParameters ? 0 : length(probs[1].p)
    N = trajectories
    code = alg isa Tsit5 ? TSIT5 : RODAS5P
    src, has_tgrad = trace(probs[1].f, n, np)
    flags = (code == TSIT5 ? COMPILE_TSIT5 : COMPILE_RODAS5P) | (has_tgrad ? HAS_USER_TGRAD : Cint(0))
    ctx = context(ea.device)
    rhs = get!(COMPILED, (src, n, np, flags)) do
        h = odemi_compile_rhs(ctx, src, n, max(np, 1), flags)
        h == C_NULL && error("odemi_compile_rhs: " * odemi_last_error(ctx))
        h
    end
    # SoA inputs
    u0 = Matrix{Float64}(undef, N, n); P = zeros(Float64, N, max(np, 1))
    t0 = Vector{Float64}(undef, N); tf = Vector{Float64}(undef, N)
    for (j, pr) in enumerate(probs)
        u0[j, :] .= pr.u0
        np > 0 && (P[j, :] .= pr.p)
        t0[j], tf[j] = pr.tspan
    end
    sa = collect(Float64, saveat)
    tst = sort!(collect(Float64, tstops))
    ss = save_start === nothing ? (isempty(sa) || first(sa) == first(probs[1].tspan)) : save_start
    se = save_end === nothing ? true : save_end
    ms = save_everystep ? max_saves : length(sa) + 2
    t_out = zeros(N, ms); u_out = zeros(N, n, ms); nsaved = zeros(Int32, N)
    u_final = zeros(N, n); t_final = zeros(N); retcode = zeros(Int32, N); stats = zeros(Int32, N, 7)
    a = OdemiSolveArgs()
    odemi_default_args!(a, code)
    GC.@preserve u0 P t0 tf sa tst t_out u_out nsaved u_final t_final retcode stats begin
        a.engine = ea.engine === :plain ? ENGINE_PLAIN : ENGINE_PERSISTENT
        a.precision = ea.precision === :mixed ? MIXED : FP64
        a.refine_sweeps = ea.refine_sweeps
        a.N = N; a.nmax = n; a.np = np
        a.u0 = pointer(u0); a.p = pointer(P); a.t0 = pointer(t0); a.tf = pointer(tf)
        a.reltol = reltol; a.abstol = abstol; a.dt = dt; a.dtmax = dtmax; a.dtmin = dtmin; a.maxiters = maxiters
        a.saveat = isempty(sa) ? Ptr{Float64}(C_NULL) : pointer(sa); a.nsaveat = length(sa)
        a.tstops = isempty(tst) ? Ptr{Float64}(C_NULL) : pointer(tst); a.ntstops = length(tst)
        a.save_start = ss; a.save_end = se; a.save_everystep = save_everystep
        a.save_end_user = save_end === nothing ? -1 : Int(save_end); a.max_saves = ms
        a.t_out = pointer(t_out); a.u_out = pointer(u_out); a.nsaved = pointer(nsaved)
        a.u_final = pointer(u_final); a.t_final = pointer(t_final); a.retcode = pointer(retcode)
        a.stats = pointer(stats)
        rc = odemi_solve_batch(ctx, rhs, a)
        rc == 0 || error("odemi_solve_batch ($rc): " * odemi_last_error(ctx))
    end
    # Julia's column-major (N, ...) arrays are the C side's [..][N] SoA layout.
    codes = (1 => SciMLBase.ReturnCode.Success, 2 => SciMLBase.ReturnCode.Terminated, 3 => SciMLBase.ReturnCode.DtNaN, 4 => SciMLBase.ReturnCode.MaxIters,
             6 => SciMLBase.ReturnCode.DtLessThanMin, 7 => SciMLBase.ReturnCode.Unstable)
    rcmap = Dict(codes)
    sols = map(1:N) do j
        k = min(Int(nsaved[j]), ms)
        ts = t_out[j, 1:k]
        us = [u_out[j, :, s] for s in 1:k]
        # rows: nf naccept nreject njacs nw nsolve ncondition; DEStats field order
        # (lib/DiffEqBase/src/stats.jl:30-42): nf nf2 nw nsolve njacs nnonliniter
        # nnonlinconvfail ncondition naccept nreject maxeig
        st = SciMLBase.DEStats(stats[j, 1], 0, stats[j, 5], stats[j, 6], stats[j, 4], 0, 0, stats[j, 7],
                               stats[j, 2], stats[j, 3], 0.0)
        SciMLBase.build_solution(probs[j], alg, ts, us; retcode = get(rcmap, Int(retcode[j]), SciMLBase.ReturnCode.Failure),
                                 stats = st)
    end
    outs = [first(ep.output_func(sols[j], j)) for j in 1:N]
    u = ep.reduction === SciMLBase.DEFAULT_REDUCTION ? outs : first(ep.reduction(ep.u_init, outs, 1:N))
    SciMLBase.EnsembleSolution(u, 0.0, all(s -> SciMLBase.successful_retcode(s.retcode), sols))
end

end # module
