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
# Problems: Tsit5 for n <= 16 (lane kernels); Rodas5P on the lane kernels for
# n <= 4 and on the workgroup kernels above that or when `prob_func` returns
# states of different lengths (uneven batch, n <= 128), whose forms are emitted
# per state size with same-shape rows rerolled into thread-strided loops.  The
# RHS (and `isoutofdomain`, the callbacks' conditions and affects) must be
# straight-line arithmetic on u, p, t (+ - * / ^, sin cos exp log sqrt tanh
# abs, comparisons, & | !, ifelse, `t in values`); no `if` on state values.
# Keywords: saveat, tstops, isoutofdomain, callback (DiscreteCallback,
# ContinuousCallback, VectorContinuousCallback; on the workgroup kernels traced
# per state size), dt, dtmin, dtmax, maxiters, save_*.  This
# image has no Julia, so this file is not exercised by the test suite; the
# Python front-end (odemi/api.py, odemi/tracer.py) emits the same forms and is.
module EnsembleMI355X

using SciMLBase
using SciMLBase: DiscreteCallback, ContinuousCallback, VectorContinuousCallback, CallbackSet
using OrdinaryDiffEq: Tsit5, Rodas5P

export EnsembleMI355X

const libodemi = joinpath(@__DIR__, "..", "odemi", "libodemi.so")
const TSIT5, RODAS5P = Cint(0), Cint(1)
const ENGINE_PLAIN, ENGINE_PERSISTENT = Cint(0), Cint(1)
const FP64, MIXED = Cint(0), Cint(1)
const COMPILE_TSIT5, COMPILE_RODAS5P, HAS_USER_TGRAD, COMPILE_WG = Cint(1), Cint(2), Cint(4), Cint(8)
const LANE_MAX_TSIT5, LANE_MAX_RODAS, WG_MAX = 16, 4, 128

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
    op::Symbol            # :u :p :t :ev :const :add :sub :mul :div :neg :pow :sin :cos :exp :log :sqrt :tanh :abs
                          # :sign :lt :le :eq :ne :and :or :not :select
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
# Predicates are traced values (0 or 1), never Bools: a condition such as
# `u[1] < 0.5` or `t == 2.0` becomes a node of the DAG (comparisons of mixed
# Sym / Real arguments reach these methods through Base's promotion).
Base.sign(a::Sym) = isconst(a) ? Sym(sign(cval(a))) : node(:sign, [a.id])
Base.:<(a::Sym, b::Sym) = binop(:lt, (x, y) -> Float64(x < y), a, b)
Base.:<=(a::Sym, b::Sym) = binop(:le, (x, y) -> Float64(x <= y), a, b)
Base.:(==)(a::Sym, b::Sym) = binop(:eq, (x, y) -> Float64(x == y), a, b)
Base.:!=(a::Sym, b::Sym) = binop(:ne, (x, y) -> Float64(x != y), a, b)
Base.:&(a::Sym, b::Sym) = binop(:and, (x, y) -> Float64(x != 0 && y != 0), a, b)
Base.:|(a::Sym, b::Sym) = binop(:or, (x, y) -> Float64(x != 0 || y != 0), a, b)
Base.:&(a::Sym, b::Bool) = a & Sym(Float64(b))
Base.:&(a::Bool, b::Sym) = Sym(Float64(a)) & b
Base.:|(a::Sym, b::Bool) = a | Sym(Float64(b))
Base.:|(a::Bool, b::Sym) = Sym(Float64(a)) | b
Base.:!(a::Sym) = isconst(a) ? Sym(Float64(cval(a) == 0)) : node(:not, [a.id])
Base.:~(a::Sym) = !a
Base.ifelse(c::Sym, a::Real, b::Real) = isconst(c) ? (cval(c) != 0 ? Sym(a) : Sym(b)) :
                                        node(:select, [c.id, Sym(a).id, Sym(b).id])
Base.in(x::Sym, itr::Union{Tuple,AbstractVector}) = isempty(itr) ? Sym(0.0) : reduce(|, (x == v for v in itr))
Base.hash(a::Sym, h::UInt) = hash(a.id, hash(:Sym, h))
Base.isequal(a::Sym, b::Sym) = a.id == b.id

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
    elseif n.op in (:sign, :lt, :le, :eq, :ne, :and, :or, :not); Sym(0.0)
    elseif n.op === :select; node(:select, [n.args[1], D(2).id, D(3).id])
    else error("no derivative for $(n.op)")
    end
    memo[e.id] = r
    r
end

cliteral(x::Float64) = isfinite(x) ? "T($(repr(x)))" : (isnan(x) ? "T(NAN)" : (x > 0 ? "T(INFINITY)" : "T(-INFINITY)"))
const CMP = Dict(:lt => "<", :le => "<=", :eq => "==", :ne => "!=")
# C++ for node `id` with its arguments' names from ref(argid)
function cexpr(id, ref = cref)
    n = G[].nodes[id]
    a(i) = ref(n.args[i])
    op = n.op
    op === :add ? "$(a(1)) + $(a(2))" : op === :sub ? "$(a(1)) - $(a(2))" :
    op === :mul ? "$(a(1)) * $(a(2))" : op === :div ? "$(a(1)) / $(a(2))" :
    op === :neg ? "-$(a(1))" : op === :pow ? "pow($(a(1)), $(a(2)))" :
    op === :abs ? "fabs($(a(1)))" : op === :sign ? "T(($(a(1)) > T(0)) - ($(a(1)) < T(0)))" :
    haskey(CMP, op) ? "T(($(a(1)) $(CMP[op]) $(a(2))) ? 1 : 0)" :
    op === :and ? "T((($(a(1)) != T(0)) && ($(a(2)) != T(0))) ? 1 : 0)" :
    op === :or ? "T((($(a(1)) != T(0)) || ($(a(2)) != T(0))) ? 1 : 0)" :
    op === :not ? "T(($(a(1)) == T(0)) ? 1 : 0)" :
    op === :select ? "($(a(1)) != T(0)) ? $(a(2)) : $(a(3))" : "$(op)($(a(1)))"
end
function cref(id)
    n = G[].nodes[id]
    n.op === :u ? "u[$(Int(n.val))]" : n.op === :p ? "p[$(Int(n.val))]" : n.op === :t ? "t" :
    n.op === :ev ? "ev[$(Int(n.val))]" : n.op === :const ? cliteral(n.val) : "v$id"
end

# Temporaries of the DAG that `outs` need, in creation (dependency) order, then
# the stores `lvalue = value`.
function emit_body(outs::Vector{Pair{String,Sym}}, indent = "    ")
    needed = falses(length(G[].nodes))
    stack = [o.second.id for o in outs]
    while !isempty(stack)
        i = pop!(stack)
        needed[i] && continue
        needed[i] = true
        append!(stack, G[].nodes[i].args)
    end
    io = IOBuffer()
    for i in eachindex(needed)
        needed[i] || continue
        G[].nodes[i].op in (:u, :p, :t, :ev, :const) && continue
        println(io, indent, "const T v$i = ", cexpr(i), ";")
    end
    for (lv, v) in outs
        println(io, indent, lv, " = ", cref(v.id), ";")
    end
    String(take!(io))
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

# ---------------------------------------------------------------------------
# Tracing the problem and its hooks
# ---------------------------------------------------------------------------
symbols(n, np) = ([node(:u, Int[], Float64(i)) for i in 0:n-1], [node(:p, Int[], Float64(k)) for k in 0:max(np, 1)-1],
                  node(:t))
params_arg(p, np) = np > 0 ? p : SciMLBase.NullParameters()

function trace_f(fun, u, p, t, n, np)
    du = Vector{Any}(fill(Sym(0.0), n))
    fun.f(du, u, params_arg(p, np), t)
    Sym[convert(Sym, x) for x in du]
end
function trace_jac(fun, f, u, p, t, n, np)          # row-major entries
    if fun.jac !== nothing
        Jm = Matrix{Any}(fill(Sym(0.0), n, n))
        fun.jac(Jm, u, params_arg(p, np), t)
        return Sym[convert(Sym, Jm[i, j]) for i in 1:n for j in 1:n]
    end
    Sym[deriv(f[i], u[j]) for i in 1:n for j in 1:n]
end
function trace_tgrad(fun, f, u, p, t, n, np)
    if fun.tgrad !== nothing
        g = Vector{Any}(fill(Sym(0.0), n)); fun.tgrad(g, u, params_arg(p, np), t)
        return Sym[convert(Sym, x) for x in g]
    end
    Sym[deriv(f[i], t) for i in 1:n]
end

"""The integrator a traced callback sees: `u` (a vector of traced entries it
may overwrite), `p`, `t`; `terminate!` records the call."""
mutable struct TracedIntegrator
    u::Vector{Any}
    p::Any
    t::Sym
    term::Bool
end
TracedIntegrator(u, p, t) = TracedIntegrator(Vector{Any}(u), p, t, false)
SciMLBase.terminate!(itg::TracedIntegrator, retcode = nothing) = (itg.term = true; nothing)

# affect!(integrator[, events]) traced: the stores of its modified entries,
# computed from the pre-affect state, and whether it called terminate!
function trace_affect(aff, u, p, t, np, extra...)
    itg = TracedIntegrator(copy(u), params_arg(p, np), t)
    aff(itg, extra...)
    length(itg.u) == length(u) || error("an affect! must keep the state length")
    outs = [(i - 1, convert(Sym, itg.u[i])) for i in eachindex(u) if convert(Sym, itg.u[i]).id != u[i].id]
    outs, itg.term
end
function affect_body(outs, n, indent)
    isempty(outs) && return ""
    nu = emit_body(Pair{String,Sym}["nu_[$i]" => v for (i, v) in outs], indent)
    indent * "T nu_[$n];\n" * nu * join(("$(indent)u[$i] = nu_[$i];\n" for (i, _) in outs))
end
rootfind_code(cb) = cb.rootfind == SciMLBase.LeftRootFind ? 0 : cb.rootfind == SciMLBase.RightRootFind ? 1 : 2

split_callbacks(cb::Nothing) = (Any[], Any[])
split_callbacks(cb::DiscreteCallback) = (Any[], Any[cb])
split_callbacks(cb::Union{ContinuousCallback,VectorContinuousCallback}) = (Any[cb], Any[])
split_callbacks(cb::CallbackSet) = (collect(Any, cb.continuous_callbacks), collect(Any, cb.discrete_callbacks))

"""Lane forms: `odemi_f` / `odemi_jac` / `odemi_tgrad`, and when given the
isoutofdomain predicate (`odemi_isout`), DiscreteCallbacks (`OdemiDcb<K>`)
and Continuous/VectorContinuousCallbacks (`OdemiCcb<K>`) -- the forms of
include/odemi.h, as the Python tracer emits them."""
function trace(fun::SciMLBase.ODEFunction, n, np; isoutofdomain = nothing, callback = nothing)
    G[] = Graph()
    u, p, t = symbols(n, np)
    f = trace_f(fun, u, p, t, n, np)
    J = trace_jac(fun, f, u, p, t, n, np)
    dT = trace_tgrad(fun, f, u, p, t, n, np)
    src = emit("odemi_f", "du", f) * emit("odemi_jac", "J", J) * emit("odemi_tgrad", "dT", dT)
    if isoutofdomain !== nothing
        r = convert(Sym, isoutofdomain(u, params_arg(p, np), t))
        src *= "#define ODEMI_HAS_ISOUT 1\ntemplate <typename T, typename Q>\n__device__ __forceinline__ bool " *
               "odemi_isout(const T* __restrict__ u, const Q* __restrict__ p, T t) {\n    (void)u; (void)p; (void)t;\n" *
               "    T r_[1];\n" * emit_body(["r_[0]" => r]) * "    return r_[0] != T(0);\n}\n"
    end
    ccbs, dcbs = split_callbacks(callback)
    hdr = "    template <typename T, typename Q>\n    static __device__ __forceinline__ "
    if !isempty(dcbs)
        src *= "#define ODEMI_NDCB $(length(dcbs))\ntemplate <int K> struct OdemiDcb;\n"
        for (k, cb) in enumerate(dcbs)
            c = convert(Sym, cb.condition(u, t, TracedIntegrator(u, params_arg(p, np), t)))
            outs, term = trace_affect(cb.affect!, u, p, t, np)
            s0, s1 = Int.(cb.save_positions)
            src *= "template <> struct OdemiDcb<$(k - 1)> {\n    static constexpr int save0 = $s0, save1 = $s1, " *
                   "term = $(Int(term));\n" *
                   hdr * "bool cond(const T* __restrict__ u, const Q* __restrict__ p, T t) {\n" *
                   "        (void)u; (void)p; (void)t;\n        T r_[1];\n" * emit_body(["r_[0]" => c], "        ") *
                   "        return r_[0] != T(0);\n    }\n" *
                   hdr * "void affect(T* __restrict__ u, const Q* __restrict__ p, T t) {\n" *
                   "        (void)u; (void)p; (void)t;\n" * affect_body(outs, n, "        ") * "    }\n};\n"
        end
    end
    if !isempty(ccbs)
        vmax = maximum([cb isa VectorContinuousCallback ? cb.len : 1 for cb in ccbs])
        length(ccbs) <= 255 || error("at most 255 continuous callbacks")
        src *= "#define ODEMI_NCCB $(length(ccbs))\n#define ODEMI_VCB_MAXLEN $vmax\ntemplate <int K> struct OdemiCcb;\n"
        for (k, cb) in enumerate(ccbs)
            s0, s1 = Int.(cb.save_positions)
            consts = "    static constexpr int rootfind = $(rootfind_code(cb)), interp_points = $(cb.interp_points);\n" *
                     "    static constexpr double abstol = $(cb.abstol), repeat_nudge = $(cb.repeat_nudge);\n"
            if cb isa VectorContinuousCallback
                out = Vector{Any}(fill(Sym(0.0), cb.len))
                cb.condition(out, u, t, TracedIntegrator(u, params_arg(p, np), t))
                evs = [node(:ev, Int[], Float64(i)) for i in 0:cb.len-1]
                outs, term = cb.affect! === nothing ? (Tuple{Int,Sym}[], false) : trace_affect(cb.affect!, u, p, t, np, evs)
                src *= "template <> struct OdemiCcb<$(k - 1)> {\n    static constexpr int vlen = $(cb.len), " *
                       "has_affect = $(Int(cb.affect! !== nothing)), has_affect_neg = 0, save0 = $s0, save1 = $s1, " *
                       "term = 0, term_neg = 0;\n" * consts *
                       hdr * "void vcond(T* __restrict__ out, const T* __restrict__ u, const Q* __restrict__ p, T t) {\n" *
                       "        (void)u; (void)p; (void)t;\n" *
                       emit_body(Pair{String,Sym}["out[$(i - 1)]" => convert(Sym, out[i]) for i in 1:cb.len], "        ") *
                       "    }\n" * hdr * "bool vaffect(T* __restrict__ u, const Q* __restrict__ p, T t, " *
                       "const double* __restrict__ ev) {\n        (void)u; (void)p; (void)t; (void)ev;\n" *
                       affect_body(outs, n, "        ") * "        return $(term ? "true" : "false");\n    }\n};\n"
            else
                c = convert(Sym, cb.condition(u, t, TracedIntegrator(u, params_arg(p, np), t)))
                ao, at = cb.affect! === nothing ? (Tuple{Int,Sym}[], false) : trace_affect(cb.affect!, u, p, t, np)
                no, nt = cb.affect_neg! === nothing ? (Tuple{Int,Sym}[], false) : trace_affect(cb.affect_neg!, u, p, t, np)
                src *= "template <> struct OdemiCcb<$(k - 1)> {\n    static constexpr int vlen = 0, " *
                       "has_affect = $(Int(cb.affect! !== nothing)), has_affect_neg = $(Int(cb.affect_neg! !== nothing)), " *
                       "save0 = $s0, save1 = $s1, term = $(Int(at)), term_neg = $(Int(nt));\n" * consts *
                       hdr * "T cond(const T* __restrict__ u, const Q* __restrict__ p, T t) {\n" *
                       "        (void)u; (void)p; (void)t;\n        T r_[1];\n" * emit_body(["r_[0]" => c], "        ") *
                       "        return r_[0];\n    }\n" *
                       hdr * "void affect(T* __restrict__ u, const Q* __restrict__ p, T t) {\n" *
                       "        (void)u; (void)p; (void)t;\n" * affect_body(ao, n, "        ") * "    }\n" *
                       hdr * "void affect_neg(T* __restrict__ u, const Q* __restrict__ p, T t) {\n" *
                       "        (void)u; (void)p; (void)t;\n" * affect_body(no, n, "        ") * "    }\n};\n"
            end
        end
    end
    src, fun.tgrad !== nothing
end

# ---------------------------------------------------------------------------
# Workgroup forms for uneven batches (one case per state size n): outputs
# whose expressions have the same shape relative to their row index are
# rerolled into thread-strided loops (o = row), each output keeping its own
# operation order, so values equal the straight-line form's.
# ---------------------------------------------------------------------------
function rel_sig(id, row, memo)
    haskey(memo, id) && return memo[id]
    n = G[].nodes[id]
    r = n.op in (:u, :v) ? "$(n.op)[$(Int(n.val) - row)]" : n.op === :p ? "p[$(Int(n.val))]" : n.op === :t ? "t" :
        n.op === :const ? repr(n.val) : string(n.op) * "(" * join((rel_sig(a, row, memo) for a in n.args), ",") * ")"
    memo[id] = r
end
# entries (row, col or nothing, expr) -> runs (row0, dcol, count, expr of the first row)
function group_runs(entries)
    groups = Dict{Tuple{Any,String},Vector{Tuple{Int,Any,Sym}}}()
    for (row, col, e) in entries
        col !== nothing && isconst(e) && cval(e) == 0 && continue   # Jacobian zeros are pre-filled
        key = (col === nothing ? nothing : col - row, rel_sig(e.id, row, Dict{Int,String}()))
        push!(get!(groups, key, Tuple{Int,Any,Sym}[]), (row, col, e))
    end
    runs = Tuple{Int,Any,Int,Sym}[]
    for items in values(groups)
        sort!(items; by = first)
        start = 1
        for i in 2:length(items)+1
            if i > length(items) || items[i][1] != items[i-1][1] + 1
                r0, c0, e0 = items[start]
                push!(runs, (r0, c0 === nothing ? nothing : c0 - r0, i - start, e0))
                start = i
            end
        end
    end
    sort!(runs; by = r -> (-r[3], r[1]))
end
function emit_runs(runs, store)
    io = IOBuffer()
    for (r0, dcol, cnt, e) in runs
        println(io, "        for (int q = tid; q < $cnt; q += nt) {\n            const int o = $r0 + q;")
        needed = falses(length(G[].nodes))
        stack = [e.id]
        while !isempty(stack)
            i = pop!(stack)
            needed[i] && continue
            needed[i] = true
            append!(stack, G[].nodes[i].args)
        end
        function ref(id)
            nd = G[].nodes[id]
            if nd.op in (:u, :v)
                off = Int(nd.val) - r0
                return off == 0 ? "$(nd.op)[o]" : "$(nd.op)[o + ($off)]"
            end
            nd.op === :const ? cliteral(nd.val) : nd.op === :p ? "p[$(Int(nd.val))]" : nd.op === :t ? "t" : "v$id"
        end
        for i in eachindex(needed)
            needed[i] || continue
            G[].nodes[i].op in (:u, :v, :p, :t, :const) && continue
            println(io, "            const T v$i = ", cexpr(i, ref), ";")
        end
        println(io, "            ", store(dcol), " = ", ref(e.id), ";\n        }")
    end
    String(take!(io))
end

"""Workgroup forms for the state sizes `ns`: `odemi_f_wg`, `odemi_jac_wg`,
`odemi_tgrad_wg`, `odemi_jvp_wg` (J v over each row's nonzeros in column
order), `odemi_isout_wg` and the callbacks' `OdemiDcbWg<K>` / `OdemiCcbWg<K>`
(scalar and vector ContinuousCallbacks), each a switch on n; plus empty lane
forms and the Gauss-Jordan tile mask."""
function trace_wg(fun::SciMLBase.ODEFunction, ns, np; isoutofdomain = nothing, callback = nothing)
    ns = sort(unique(ns))
    ccbs, dcbs = split_callbacks(callback)
    fc, jc, tc, vc, ic = String[], String[], String[], String[], String[]
    dcase = [String[] for _ in dcbs]; dterm = [Bool[] for _ in dcbs]
    ccase = [String[] for _ in ccbs]; cterm = [Tuple{Bool,Bool}[] for _ in ccbs]
    for n in ns
        G[] = Graph()
        u, p, t = symbols(n, np)
        v = [node(:v, Int[], Float64(i)) for i in 0:n-1]
        f = trace_f(fun, u, p, t, n, np)
        Jf = trace_jac(fun, f, u, p, t, n, np)
        dT = trace_tgrad(fun, f, u, p, t, n, np)
        jent = [(i - 1, j - 1, Jf[(i - 1) * n + j]) for i in 1:n for j in 1:n if !(isconst(Jf[(i - 1) * n + j]) && cval(Jf[(i - 1) * n + j]) == 0)]
        jv = Sym[]
        for i in 0:n-1
            acc = nothing
            for (r, c, e) in jent
                r == i || continue
                term = e * v[c + 1]
                acc = acc === nothing ? term : acc + term
            end
            push!(jv, acc === nothing ? Sym(0.0) : acc)
        end
        push!(fc, "    case $n: {\n" * emit_runs(group_runs([(i - 1, nothing, f[i]) for i in 1:n]), d -> "du[o]") * "        break;\n    }")
        push!(jc, "    case $n: {\n" * emit_runs(group_runs(jent), d -> "J[o * ld + o + ($d)]") * "        break;\n    }")
        push!(tc, "    case $n: {\n" * emit_runs(group_runs([(i - 1, nothing, dT[i]) for i in 1:n if !(isconst(dT[i]) && cval(dT[i]) == 0)]),
                                                d -> "dT[o]") * "        break;\n    }")
        push!(vc, "    case $n: {\n" * emit_runs(group_runs([(i - 1, nothing, jv[i]) for i in 1:n]), d -> "Jv[o]") * "        break;\n    }")
        if isoutofdomain !== nothing
            r = convert(Sym, isoutofdomain(u, params_arg(p, np), t))
            push!(ic, "    case $n: {\n        T r_[1];\n" * emit_body(["r_[0]" => r], "        ") * "        return r_[0] != T(0);\n    }")
        end
        for (k, cb) in enumerate(dcbs)
            c = convert(Sym, cb.condition(u, t, TracedIntegrator(u, params_arg(p, np), t)))
            outs, term = trace_affect(cb.affect!, u, p, t, np)
            push!(dcase[k], "        case $n: {\n        T r_[1];\n" * emit_body(["r_[0]" => c], "        ") *
                            "        return r_[0] != T(0);\n        }\n" * "\x00" *
                            "        case $n: {\n" * affect_body(outs, n, "        ") * "        break;\n        }")
            push!(dterm[k], term)
        end
        for (k, cb) in enumerate(ccbs)
            if cb isa VectorContinuousCallback
                # vcond writes out[0:len]; vaffect sees the event directions ev[i] and returns terminate!
                out = Vector{Any}(fill(Sym(0.0), cb.len))
                cb.condition(out, u, t, TracedIntegrator(u, params_arg(p, np), t))
                evs = [node(:ev, Int[], Float64(i)) for i in 0:cb.len-1]
                vo, vt = cb.affect! === nothing ? (Tuple{Int,Sym}[], false) : trace_affect(cb.affect!, u, p, t, np, evs)
                push!(ccase[k], "        case $n: {\n" *
                                emit_body(Pair{String,Sym}["out[$(i - 1)]" => convert(Sym, out[i]) for i in 1:cb.len], "        ") *
                                "        break;\n        }" * "\x00" *
                                "        case $n: {\n" * affect_body(vo, n, "        ") *
                                "        return $(vt ? "true" : "false");\n        }")
                push!(cterm[k], (false, false))
                continue
            end
            c = convert(Sym, cb.condition(u, t, TracedIntegrator(u, params_arg(p, np), t)))
            ao, at = cb.affect! === nothing ? (Tuple{Int,Sym}[], false) : trace_affect(cb.affect!, u, p, t, np)
            no, nt = cb.affect_neg! === nothing ? (Tuple{Int,Sym}[], false) : trace_affect(cb.affect_neg!, u, p, t, np)
            push!(ccase[k], "        case $n: {\n        T r_[1];\n" * emit_body(["r_[0]" => c], "        ") *
                            "        return r_[0];\n        }\n" * "\x00" *
                            "        case $n: {\n" * affect_body(ao, n, "        ") * "        break;\n        }" * "\x00" *
                            "        case $n: {\n" * affect_body(no, n, "        ") * "        break;\n        }")
            push!(cterm[k], (at, nt))
        end
    end
    rmask = 0
    for n in ns
        n <= WG_MAX && (rmask |= 1 << ((n + 15) ÷ 16 - 1))
    end
    hdr = "template <typename T>\n__device__ __noinline__ void "
    args = "const T* __restrict__ u, const T* __restrict__ p, T t, int n, int tid, int nt"
    sw(cases) = "    switch (n) {\n" * join(cases, "\n") * "\n    default: break;\n    }\n}\n"
    src = "template <typename T> __device__ __forceinline__ void odemi_f(T* du, const T* u, const T* p, T t) {}\n" *
          "template <typename T> __device__ __forceinline__ void odemi_jac(T* J, const T* u, const T* p, T t) {}\n" *
          "template <typename T> __device__ __forceinline__ void odemi_tgrad(T* dT, const T* u, const T* p, T t) {}\n" *
          "#define ODEMI_GJ_RMASK 0x$(string(rmask; base = 16))\n" *
          hdr * "odemi_f_wg(T* __restrict__ du, $args) {\n    (void)u; (void)p; (void)t;\n" * sw(fc) *
          hdr * "odemi_jac_wg(T* __restrict__ J, int ld, $args) {\n    (void)u; (void)p; (void)t;\n" * sw(jc) *
          hdr * "odemi_tgrad_wg(T* __restrict__ dT, $args) {\n    (void)u; (void)p; (void)t;\n" * sw(tc) *
          hdr * "odemi_jvp_wg(T* __restrict__ Jv, const T* __restrict__ v, $args) {\n    (void)u; (void)v; (void)p; (void)t;\n" * sw(vc)
    if isoutofdomain !== nothing
        src *= "#define ODEMI_HAS_ISOUT_WG 1\ntemplate <typename T>\n__device__ __noinline__ bool odemi_isout_wg(" *
               "const T* __restrict__ u, const T* __restrict__ p, T t, int n) {\n    (void)u; (void)p; (void)t;\n" *
               "    switch (n) {\n" * join(ic, "\n") * "\n    default: return false;\n    }\n}\n"
    end
    mhdr = "    template <typename T>\n    static __device__ __noinline__ "
    part(cases, i) = join((split(c, '\x00')[i] for c in cases), "\n")
    mfn(kind, name, a, cases, i, default) = mhdr * "$kind $name($a, int n) {\n        (void)u; (void)p; (void)t;\n" *
                                             "        switch (n) {\n" * part(cases, i) * "\n        default: $default\n        }\n    }\n"
    ua = "const T* __restrict__ u, const T* __restrict__ p, T t"
    wa = "T* __restrict__ u, const T* __restrict__ p, T t"
    if !isempty(dcbs)
        src *= "#define ODEMI_NDCB_WG $(length(dcbs))\ntemplate <int K> struct OdemiDcbWg;\n"
        for (k, cb) in enumerate(dcbs)
            all(==(dterm[k][1]), dterm[k]) || error("an affect! must call terminate! for every state size or for none")
            s0, s1 = Int.(cb.save_positions)
            src *= "template <> struct OdemiDcbWg<$(k - 1)> {\n    static constexpr int save0 = $s0, save1 = $s1, " *
                   "term = $(Int(dterm[k][1]));\n" * mfn("bool", "cond", ua, dcase[k], 1, "return false;") *
                   mfn("void", "affect", wa, dcase[k], 2, "break;") * "};\n"
        end
    end
    if !isempty(ccbs)
        length(ccbs) <= 255 || error("at most 255 continuous callbacks")
        src *= "#define ODEMI_NCCB_WG $(length(ccbs))\ntemplate <int K> struct OdemiCcbWg;\n"
        vmax = maximum([cb isa VectorContinuousCallback ? cb.len : 0 for cb in ccbs])
        vmax > 0 && (src *= "#define ODEMI_VCB_MAXLEN_WG $vmax\n")
        for (k, cb) in enumerate(ccbs)
            s0, s1 = Int.(cb.save_positions)
            if cb isa VectorContinuousCallback
                src *= "template <> struct OdemiCcbWg<$(k - 1)> {\n    static constexpr int vlen = $(cb.len), " *
                       "has_affect = $(Int(cb.affect! !== nothing)), has_affect_neg = 0, save0 = $s0, save1 = $s1;\n" *
                       "    static constexpr int rootfind = $(rootfind_code(cb)), interp_points = $(cb.interp_points), " *
                       "term = 0, term_neg = 0;\n" *
                       "    static constexpr double abstol = $(cb.abstol), repeat_nudge = $(cb.repeat_nudge);\n" *
                       mfn("void", "vcond", "T* __restrict__ out, $ua", ccase[k], 1, "break;") *
                       mfn("bool", "vaffect", "$wa, const double* __restrict__ ev", ccase[k], 2, "return false;") * "};\n"
                continue
            end
            all(==(cterm[k][1]), cterm[k]) || error("an affect! must call terminate! for every state size or for none")
            at, nt = cterm[k][1]
            src *= "template <> struct OdemiCcbWg<$(k - 1)> {\n    static constexpr int vlen = 0, has_affect = $(Int(cb.affect! !== nothing)), " *
                   "has_affect_neg = $(Int(cb.affect_neg! !== nothing)), save0 = $s0, save1 = $s1;\n" *
                   "    static constexpr int rootfind = $(rootfind_code(cb)), interp_points = $(cb.interp_points), " *
                   "term = $(Int(at)), term_neg = $(Int(nt));\n" *
                   "    static constexpr double abstol = $(cb.abstol), repeat_nudge = $(cb.repeat_nudge);\n" *
                   mfn("T", "cond", ua, ccase[k], 1, "return T(0);") * mfn("void", "affect", wa, ccase[k], 2, "break;") *
                   mfn("void", "affect_neg", wa, ccase[k], 3, "break;") * "};\n"
        end
    end
    src, fun.tgrad !== nothing
end

# ---------------------------------------------------------------------------
# solve
# ---------------------------------------------------------------------------
function SciMLBase.__solve(ep::SciMLBase.EnsembleProblem, alg::Union{Tsit5,Rodas5P}, ea::EnsembleMI355X;
                           trajectories, reltol = 1e-3, abstol = 1e-6, saveat = Float64[], save_start = nothing,
                           save_end = nothing, save_everystep = isempty(saveat), dt = 0.0, dtmax = 0.0,
                           dtmin = 0.0, maxiters = 1_000_000, max_saves = 4096, tstops = Float64[],
                           isoutofdomain = nothing, callback = nothing, kwargs...)
    isempty(kwargs) || error("EnsembleMI355X: unsupported keyword(s) $(collect(keys(kwargs)))")
    probs = [ep.prob_func(deepcopy(ep.prob), i, 1) for i in 1:trajectories]
    nvec = Int32[length(pr.u0) for pr in probs]
    n = Int(maximum(nvec))
    uneven = any(!=(n), nvec)
    np = probs[1].p isa SciMLBase.NullParameters ? 0 : length(probs[1].p)
    N = trajectories
    code = alg isa Tsit5 ? TSIT5 : RODAS5P
    wg = code == RODAS5P && (uneven || n > LANE_MAX_RODAS)
    code == TSIT5 && (uneven || n > LANE_MAX_TSIT5) && error("Tsit5 runs states of one size n <= $LANE_MAX_TSIT5")
    wg && n > WG_MAX && error("Rodas5P runs n <= $WG_MAX")
    src, has_tgrad = wg ? trace_wg(probs[1].f, unique(nvec), np; isoutofdomain, callback) :
                          trace(probs[1].f, n, np; isoutofdomain, callback)
    flags = (wg ? COMPILE_WG : code == TSIT5 ? COMPILE_TSIT5 : COMPILE_RODAS5P) | (has_tgrad ? HAS_USER_TGRAD : Cint(0))
    ctx = context(ea.device)
    rhs = get!(COMPILED, (src, n, np, flags)) do
        h = odemi_compile_rhs(ctx, src, n, max(np, 1), flags)
        h == C_NULL && error("odemi_compile_rhs: " * odemi_last_error(ctx))
        h
    end
    # SoA inputs (states of an uneven batch zero-padded to n)
    u0 = zeros(Float64, N, n); P = zeros(Float64, N, max(np, 1))
    t0 = Vector{Float64}(undef, N); tf = Vector{Float64}(undef, N)
    for (j, pr) in enumerate(probs)
        u0[j, 1:length(pr.u0)] .= pr.u0
        np > 0 && (P[j, :] .= pr.p)
        t0[j], tf[j] = pr.tspan
    end
    sa = sort!(collect(Float64, saveat))
    tst = sort!(collect(Float64, tstops))
    ss = save_start === nothing ? (isempty(sa) || first(probs[1].tspan) in sa) : save_start
    se = save_end === nothing ? true : save_end
    ncb = callback === nothing ? 0 : sum(length, split_callbacks(callback))
    ms = (save_everystep || ncb > 0) ? max_saves : length(sa) + 2
    t_out = zeros(N, ms); u_out = zeros(N, n, ms); nsaved = zeros(Int32, N)
    u_final = zeros(N, n); t_final = zeros(N); retcode = zeros(Int32, N); stats = zeros(Int32, N, 7)
    a = OdemiSolveArgs()
    odemi_default_args!(a, code)
    GC.@preserve u0 P t0 tf nvec sa tst t_out u_out nsaved u_final t_final retcode stats begin
        a.engine = ea.engine === :plain ? ENGINE_PLAIN : ENGINE_PERSISTENT
        a.precision = ea.precision === :mixed ? MIXED : FP64
        a.refine_sweeps = ea.refine_sweeps
        a.N = N; a.nmax = n; a.np = np
        a.nvec = wg ? pointer(nvec) : Ptr{Int32}(C_NULL)
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
        nj = Int(nvec[j])
        ts = t_out[j, 1:k]
        us = [u_out[j, 1:nj, s] for s in 1:k]
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
