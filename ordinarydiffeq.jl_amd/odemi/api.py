"""Host-side mirror of the reference's ensemble surface.

    solve(EnsembleProblem(prob; prob_func, output_func, reduction),
          Tsit5() | Rodas5P(), EnsembleMI355X(); trajectories, reltol, abstol, saveat, ...)

returns an EnsembleSolution, like SciMLBase's ``EnsembleThreads`` driver does
(call sites: lib/DiffEqBase/test/downstream/ensemble_thread_safety.jl:1-31;
``prob_func(prob, ctx)`` / ``output_func(sol, ctx)`` signatures: NEWS.md:209-225).
Keyword defaults and save rules follow lib/OrdinaryDiffEqCore/src/solve.jl:128-203,
381-393 and 596-599; saveat filtering follows solve.jl:1103-1123.  The numerics
run in libodemi.so on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import builtins
import enum
import math
import time
import warnings
from fractions import Fraction

import numpy as np

from . import _lib
from .tracer import trace_problem, trace_problem_wg

__all__ = ["ReturnCode", "ODEFunction", "ODEProblem", "remake", "Tsit5", "Rodas5P", "EnsembleMI355X",
           "EnsembleContext", "EnsembleProblem", "DEStats", "ODESolution", "EnsembleSolution", "solve",
           "solve_arrays", "get_context", "compile_problem"]


class ReturnCode(enum.IntEnum):
    """SciMLBase.ReturnCode (values in enum order)."""
    Default = 0
    Success = 1
    Terminated = 2
    DtNaN = 3
    MaxIters = 4
    MaxNumSub = 5
    DtLessThanMin = 6
    Unstable = 7


class ODEFunction:
    """In-place RHS f(du, u, p, t) with optional analytic jac(J, u, p, t) and
    tgrad(dT, u, p, t) (the reference's ODEFunction fields f, jac, tgrad)."""

    def __init__(self, f, jac=None, tgrad=None):
        if isinstance(f, ODEFunction):
            f, jac, tgrad = f.f, jac or f.jac, tgrad or f.tgrad
        self.f, self.jac, self.tgrad = f, jac, tgrad


class ODEProblem:
    def __init__(self, f, u0, tspan, p=()):
        self.f = f if isinstance(f, ODEFunction) else ODEFunction(f)
        self.u0 = np.array(u0, dtype=np.float64).reshape(-1)
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.p = np.array(p if p is not None else (), dtype=np.float64).reshape(-1)


def remake(prob, u0=None, tspan=None, p=None, f=None):
    return ODEProblem(prob.f if f is None else f, prob.u0 if u0 is None else u0,
                      prob.tspan if tspan is None else tspan, prob.p if p is None else p)


class Tsit5:
    code = _lib.TSIT5
    order = 5

    def __repr__(self):
        return "Tsit5()"


class Rodas5P:
    code = _lib.RODAS5P
    order = 5

    def __init__(self, autodiff=None, linsolve=None):
        self.autodiff, self.linsolve = autodiff, linsolve

    def __repr__(self):
        return "Rodas5P()"


class EnsembleMI355X:
    """Ensemble algorithm running every trajectory on one MI355X.

    engine:    "persistent" (one launch, work queue, state in registers) or
               "plain" (FP64, one launch per step attempt, state in HBM).
    precision: "fp64", or "mixed" (Tsit5: FP32 stages, FP64 state/t/error/controller;
               Rodas5P: FP32 LU with FP64 iterative refinement).
    allocate:  False runs odemi_solve_batch_noalloc: no device allocation by the
               library (scratch from its fixed 64 MiB per-GPU array), device-resident
               (torch) arrays and the persistent engine only; the solve is enqueued
               on torch's current stream.
    """

    def __init__(self, device=0, engine="persistent", precision="fp64", refine_sweeps=1, sort_by_cost=True,
                 block=0, blocks_per_cu=0, autotune=False, allocate=True):
        if engine not in ("persistent", "plain"):
            raise ValueError("engine must be 'persistent' or 'plain'")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if engine == "plain" and precision != "fp64":
            raise ValueError("the plain engine is FP64 only")
        if not allocate and (engine != "persistent" or autotune):
            raise ValueError("allocate=False runs the persistent engine without the autotuner")
        self.device, self.engine, self.precision = device, engine, precision
        self.refine_sweeps, self.sort_by_cost, self.block = refine_sweeps, sort_by_cost, block
        self.blocks_per_cu, self.autotune = int(blocks_per_cu), bool(autotune)
        self.allocate = bool(allocate)


class MixedPrecisionWarning(UserWarning):
    """A mixed-precision solve that is likely slower than the FP64 engine."""


def _mixed_precision_advice(ea, algname, callback):
    """The reason a mixed-precision solve is likely slower than FP64, or None.

    Tsit5 with continuous callbacks: near an event the FP32 stage rounding puts
    a floor under the error estimate, so the controller keeps rejecting and
    shrinking steps (1,048,576 bouncing balls: 4.1x the step attempts and 3.3x
    the time of the FP64 engine, whose results are also bit-identical to the
    reference restatement)."""
    if ea.precision != "mixed" or algname != "tsit5":
        return None
    if any(isinstance(c, (ContinuousCallback, VectorContinuousCallback)) for c in _callback_list(callback)):
        return ("precision='mixed' with continuous callbacks under Tsit5: FP32 stage rounding floors the error "
                "estimate near events, costing several times the step attempts of precision='fp64' "
                "(4.1x on 1M bouncing balls); precision='fp64' is faster here and bit-identical to the reference")
    return None


# Autotuner choices, per context: (problem handle, alg, precision, N, nmax, save
# mode) -> (blocks_per_cu, {candidate: kernel ms}).
_TUNED = {}


def _autotune(ctx, h, a, key, wg):
    """Pick the resident workgroups per CU of the persistent kernel (the number
    of lanes, or of concurrent trajectories for the workgroup kernels) by timing
    each candidate on the caller's own batch; results do not depend on it.  Runs
    once per key; the outputs of the last run are overwritten by the caller's."""
    hit = _TUNED.get(key)
    if hit is not None:
        return hit[0]
    cands = (1, 2, 3, 4) if wg else (2, 3, 4, 5, 6, 8)
    times = {}
    for b in cands:
        a.blocks_per_cu = b
        ctx.solve(h, a)           # first run of a candidate also warms the module
        ctx.solve(h, a)
        times[b] = a.kernel_ms
    best = min(times, key=times.get)
    _TUNED[key] = (best, times)
    return best


class EnsembleContext:
    """Per-trajectory context handed to prob_func/output_func (v7 API)."""

    def __init__(self, sim_id, repeat, rng):
        self.sim_id, self.repeat, self.rng = sim_id, repeat, rng

    @property
    def i(self):
        return self.sim_id


class EnsembleProblem:
    def __init__(self, prob, prob_func=None, output_func=None, reduction=None, u_init=None, safetycopy=True):
        self.prob = prob
        self.prob_func = prob_func or (lambda prob, ctx: prob)
        self.output_func = output_func or (lambda sol, ctx: (sol, False))
        self.reduction = reduction
        self.u_init = u_init
        self.safetycopy = safetycopy


class DEStats:
    __slots__ = ("nf", "naccept", "nreject", "njacs", "nw", "nsolve", "ncondition")

    def __init__(self, nf, naccept, nreject, njacs, nw, nsolve, ncondition=0):
        self.nf, self.naccept, self.nreject = int(nf), int(naccept), int(nreject)
        self.njacs, self.nw, self.nsolve = int(njacs), int(nw), int(nsolve)
        self.ncondition = int(ncondition)

    def __repr__(self):
        return (f"DEStats(nf={self.nf}, naccept={self.naccept}, nreject={self.nreject}, "
                f"njacs={self.njacs}, nw={self.nw}, nsolve={self.nsolve}, ncondition={self.ncondition})")


class DiscreteCallback:
    """DiscreteCallback(condition, affect!; save_positions) (lib/DiffEqBase/src/
    callbacks.jl:645-683).  After every accepted step, when condition(u, t,
    integrator) holds, affect(integrator) runs; it may rewrite integrator.u (by
    index, slice or whole) and call odemi.terminate(integrator).  Both are
    traced like f: conditions are predicates (==, <, &, |, ~, odemi.sym_in /
    sym_any / sym_all), the affect is branch-free.  save_positions = (save
    before, save after) the affect."""

    def __init__(self, condition, affect, save_positions=(True, True)):
        if len(save_positions) != 2:
            raise ValueError("save_positions is a pair of Bools")
        self.condition, self.affect = condition, affect
        self.save_positions = (bool(save_positions[0]), bool(save_positions[1]))


_SAME = object()
builtins_len = builtins.len


class ContinuousCallback:
    """ContinuousCallback(condition, affect!, affect_neg! = affect!; rootfind,
    save_positions, interp_points, abstol, repeat_nudge) (SciMLBase; handled by
    lib/DiffEqBase/src/callbacks.jl:239-640), lane kernels.  condition(u, t,
    integrator) is a traced value whose sign change within an accepted step is
    an event: the step is cut back to the root (found to adjacent floats; the
    float before the root for rootfind="left", after it for "right", the first
    detected sample point for "none") and affect (upcrossing, - to +) or
    affect_neg (downcrossing) runs there; None makes that direction no event.
    Sign changes are looked for at the step's ends and, when those show none,
    at interp_points equally spaced times."""

    def __init__(self, condition, affect, affect_neg=_SAME, *, rootfind="left", save_positions=(True, True),
                 interp_points=10, abstol=10 * 2.220446049250313e-16, repeat_nudge=1 / 100):
        if rootfind not in ("left", "right", "none"):
            raise ValueError('rootfind is "left", "right" or "none"')
        if int(interp_points) == 1 or int(interp_points) < 0:
            raise ValueError("interp_points is 0 or at least 2")
        if len(save_positions) != 2:
            raise ValueError("save_positions is a pair of Bools")
        self.condition, self.affect = condition, affect
        self.affect_neg = affect if affect_neg is _SAME else affect_neg
        self.rootfind, self.interp_points = rootfind, int(interp_points)
        self.save_positions = (bool(save_positions[0]), bool(save_positions[1]))
        self.abstol, self.repeat_nudge = float(abstol), float(repeat_nudge)


class VectorContinuousCallback:
    """VectorContinuousCallback(condition, affect!, len; rootfind,
    save_positions, interp_points, abstol, repeat_nudge) (SciMLBase; handled by
    lib/DiffEqBase/src/callbacks.jl:239-362,499-640), lane and workgroup
    kernels (traced per state size in the latter).
    condition(out, u, t, integrator) fills out[0:len] with traced values; a
    sign change of any of them within an accepted step is an event: the step
    is cut back to the earliest root and affect(integrator, events) runs once,
    events[i] the direction of component i (-1 down, +1 up, 0 none; several
    when simultaneous).  The affect is traced like the others, choosing by
    event with odemi.sym_ifelse, and may call odemi.terminate(integrator,
    when=predicate).  len <= 16."""

    def __init__(self, condition, affect, len, *, rootfind="left", save_positions=(True, True),
                 interp_points=10, abstol=10 * 2.220446049250313e-16, repeat_nudge=1 / 100):
        if not 1 <= int(len) <= 16:
            raise ValueError("len is 1..16")
        if rootfind not in ("left", "right", "none"):
            raise ValueError('rootfind is "left", "right" or "none"')
        if int(interp_points) == 1 or int(interp_points) < 0:
            raise ValueError("interp_points is 0 or at least 2")
        if builtins_len(save_positions) != 2:
            raise ValueError("save_positions is a pair of Bools")
        self.condition, self.affect, self.affect_neg, self.len = condition, affect, None, int(len)
        self.rootfind, self.interp_points = rootfind, int(interp_points)
        self.save_positions = (bool(save_positions[0]), bool(save_positions[1]))
        self.abstol, self.repeat_nudge = float(abstol), float(repeat_nudge)


class CallbackSet:
    """CallbackSet(cb...): continuous callbacks are checked first (the earliest
    event of the step wins, in the order given on ties), then the discrete ones
    in order."""

    def __init__(self, *callbacks):
        cbs = _callback_list(list(callbacks))
        self.continuous_callbacks = [c for c in cbs if isinstance(c, (ContinuousCallback, VectorContinuousCallback))]
        self.discrete_callbacks = [c for c in cbs if isinstance(c, DiscreteCallback)]


def _callback_list(cb):
    if cb is None:
        return []
    if isinstance(cb, (DiscreteCallback, ContinuousCallback, VectorContinuousCallback)):
        return [cb]
    if isinstance(cb, CallbackSet):
        return list(cb.continuous_callbacks) + list(cb.discrete_callbacks)
    if isinstance(cb, (list, tuple)):
        out = []
        for c in cb:
            out += _callback_list(c)
        return out
    raise TypeError(f"unsupported callback {type(cb).__name__} (DiscreteCallback / ContinuousCallback / "
                    "CallbackSet)")


def terminate(integrator, when=None):
    """terminate!(integrator) inside a traced affect: the trajectory stops after
    this step with retcode Terminated (integrator_interface.jl:443-446).
    `when`, a traced predicate, makes it conditional (VectorContinuousCallback
    affects)."""
    integrator.terminate(when)


class ODESolution:
    def __init__(self, t, u, retcode, stats, prob=None, alg=None):
        self.t, self.u, self.retcode, self.stats, self.prob, self.alg = t, u, retcode, stats, prob, alg


class EnsembleSolution:
    """Ensemble result.  ``u`` is a list of per-trajectory outputs (ODESolution by
    default); the packed arrays (``u_final``, ``t_out``, ``u_out``, ``retcodes``,
    ``stats``) are kept for bulk access."""

    def __init__(self, u, elapsed, arrays, converged=True):
        self.u, self.elapsedTime, self.arrays, self.converged = u, elapsed, arrays, converged

    def __len__(self):
        return len(self.u)


_CTX = {}


def get_context(device=0):
    c = _CTX.get(device)
    if c is None:
        c = _lib.Context(device)
        _CTX[device] = c
    return c


_TRACES = {}


LANE_MAX_RODAS = 4      # Rodas5P: lane-per-trajectory kernel up to this n, workgroup kernel above
LANE_MAX_TSIT5 = 16

_LANE_STUBS = """template <typename T> __device__ __forceinline__ void odemi_f(T* du, const T* u, const T* p, T t) {}
template <typename T> __device__ __forceinline__ void odemi_jac(T* J, const T* u, const T* p, T t) {}
template <typename T> __device__ __forceinline__ void odemi_tgrad(T* dT, const T* u, const T* p, T t) {}
"""


def _callback_key(c):
    """Trace-cache key of a callback.  It holds the traced callables themselves
    (hashed and compared by identity), never their id(): the cache keeps them
    alive, so a later callable cannot reuse a freed one's address and be given
    its trace."""
    if isinstance(c, DiscreteCallback):
        return ("d", c.condition, c.affect, c.save_positions)
    if isinstance(c, VectorContinuousCallback):
        return ("v", c.condition, c.affect, c.len, c.rootfind, c.interp_points, c.save_positions,
                c.abstol, c.repeat_nudge)
    return ("c", c.condition, c.affect, c.affect_neg, c.rootfind, c.interp_points,
            c.save_positions, c.abstol, c.repeat_nudge)


def compile_problem(fun, n, np_, algs=("tsit5", "rodas5p"), ctx=None, ns=None, isoutofdomain=None,
                    callback=None):
    """Trace fun (ODEFunction) for state size n (or the sizes `ns` of an uneven
    batch, n = max) and hipRTC-compile it.  isoutofdomain(u, p, t): a traceable
    predicate (comparisons, &, |, ~, odemi.sym_any / sym_all) for the lane
    kernels; callback: DiscreteCallback / ContinuousCallback / CallbackSet,
    traced likewise.
    Returns (rhs, traced, seconds)."""
    tr, wg = trace_for(fun, n, np_, algs=algs, ns=ns, isoutofdomain=isoutofdomain, callback=callback)
    flags = 0
    if wg:
        flags |= _lib.COMPILE_WG
    else:
        if "tsit5" in algs:
            flags |= _lib.COMPILE_TSIT5
        if "rodas5p" in algs:
            flags |= _lib.COMPILE_RODAS5P
    if tr.has_user_tgrad:
        flags |= _lib.HAS_USER_TGRAD
    ctx = ctx or get_context()
    h, secs = ctx.compile(tr.src, n, max(np_, 1), flags)
    return h, tr, secs


def trace_for(fun, n, np_, algs=("tsit5", "rodas5p"), ns=None, isoutofdomain=None, callback=None):
    """The traced HIP source of a problem (cached), and whether it targets the
    workgroup kernels.  No device is touched."""
    wg = "rodas5p" in algs and (ns is not None or n > LANE_MAX_RODAS)
    cbs = _callback_list(callback)
    if not wg and "tsit5" in algs and n > LANE_MAX_TSIT5:
        raise ValueError(f"Tsit5 supports state dimension <= {LANE_MAX_TSIT5}")
    if ns is not None and "tsit5" in algs:
        raise ValueError("uneven-size batches (nvec) are supported with Rodas5P")
    # the key holds the callables (identity hash), so they stay alive while cached
    key = (fun.f, fun.jac, fun.tgrad, n, np_, tuple(ns) if ns is not None else None, wg,
           isoutofdomain, tuple(_callback_key(c) for c in cbs))
    tr = _TRACES.get(key)
    if tr is None:
        if wg:
            tr = trace_problem_wg(fun.f, list(ns) if ns is not None else [n], np_, jac=fun.jac, tgrad=fun.tgrad,
                                  isout=isoutofdomain, callbacks=cbs)
            tr.src = _LANE_STUBS + tr.src
            tr.has_user_tgrad = fun.tgrad is not None
        else:
            tr = trace_problem(fun.f, n, np_, jac=fun.jac, tgrad=fun.tgrad, isout=isoutofdomain, callbacks=cbs)
        _TRACES[key] = tr
    return tr, wg


def _decimal(x):
    return Fraction(repr(float(x)))


def saveat_times(saveat, t0, tf):
    """Saveat list, sorted ascending, for the span from t0 to tf: a scalar
    interval d gives (t0 + tdir d):(tdir d):tf (decimal-exact like Julia's float
    ranges; initialize_saveat, solve.jl:1103-1123, tdir = sign(tf - t0)); a
    collection is sorted (each trajectory keeps the points between its own t0,
    exclusive, and tf, inclusive)."""
    if saveat is None:
        return np.zeros(0)
    if np.isscalar(saveat):
        d = abs(float(saveat))
        if d == 0:
            raise ValueError("saveat interval must be nonzero")
        tdir = -1 if tf < t0 else 1
        step = tdir * _decimal(d)
        start = _decimal(t0) + step
        end = _decimal(tf)
        slack = abs(end) * Fraction(1, 10 ** 12)
        out = []
        k = 0
        while True:
            v = start + k * step
            if tdir * (v - end) > slack:
                break
            fv = float(v)
            if tdir * fv <= tdir * tf:
                out.append(fv)
            k += 1
        return np.sort(np.array(out, dtype=np.float64))
    arr = np.sort(np.asarray(list(saveat), dtype=np.float64))
    return arr


def _tstops_array(tstops):
    """User tstops (solve.jl:440-446) as one sorted FP64 list shared by every
    trajectory; a scalar is one time.  Callables (late-binding tstops of a single
    problem) are not supported for an ensemble."""
    if tstops is None:
        return np.zeros(0)
    if callable(tstops):
        raise ValueError("tstops must be times shared by all trajectories (callables are not supported)")
    arr = np.atleast_1d(np.asarray(tstops, dtype=np.float64))
    if arr.ndim != 1:
        raise ValueError("tstops must be a 1-D list of times")
    return np.sort(arr)


def _resolve_save_flags(saveat, save_start, save_end, save_everystep, t0, tf):
    is_num = saveat is not None and np.isscalar(saveat)
    empty = saveat is None or (not is_num and len(saveat) == 0)
    sv = list(saveat) if (not empty and not is_num) else []
    if save_everystep is None:
        save_everystep = empty
    if save_start is None:
        save_start = save_everystep or empty or is_num or (t0 in sv)
    save_end_user = -1 if save_end is None else int(bool(save_end))
    if save_end is None:
        save_end = save_everystep or empty or is_num or (tf in sv)
    return bool(save_start), bool(save_end), bool(save_everystep), save_end_user


def _is_torch(x):
    return type(x).__module__.startswith("torch")


_PIN_MIN_BYTES = 1 << 20


def _host_empty(shape, dtype):
    """np.empty for an output of a host-array solve; from 1 MiB up a view of a
    pinned (page-locked) tensor from torch's caching host allocator, so the
    device-to-host copy runs at DMA speed into pages that are already mapped.
    Fresh pageable pages take a page fault each: 4-17 GB/s for the copy on
    the MI355X box against 57 GB/s pinned.  The array keeps its tensor (and
    the block) alive; freed blocks are reused by later solves."""
    nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
    if nbytes >= _PIN_MIN_BYTES:
        import torch
        if torch.cuda.is_available():
            tdt = {np.float64: torch.float64, np.int32: torch.int32}[dtype]
            return torch.empty(shape, dtype=tdt, pin_memory=True).numpy()
    return np.empty(shape, dtype=dtype)


def solve_arrays(fun, alg, u0, p, t0, tf, ensemblealg=None, *, reltol=None, abstol=None, saveat=None,
                 save_start=False, save_end=True, save_everystep=False, save_end_user=-1, max_saves=None,
                 dt=None, dtmax=None, dtmin=0.0, maxiters=1_000_000, ctx=None, outputs=None, nvec=None,
                 tstops=None, isoutofdomain=None, callback=None):
    """Bulk entry point on packed SoA arrays: u0 (nmax, N), p (np, N), t0/tf (N,);
    nvec (N,) gives per-trajectory dimensions for uneven batches (Rodas5P);
    tstops: stopping times shared by all trajectories (each keeps those strictly
    inside its own (t0, tf)); isoutofdomain(u, p, t): traced predicate that
    rejects a step whose end state it flags (the next dt is dt * qmin);
    callback: DiscreteCallback / CallbackSet (forced saves of save_positions need
    max_saves room; nsaved beyond it is reported, the extra points dropped).

    Arrays may be numpy (host) or torch CUDA tensors (device-resident, zero
    copies).  Returns a dict of output arrays of the same kind; host outputs of
    1 MiB and more are numpy views of pinned (page-locked) torch tensors, which
    go back to torch's pinned-block cache when the arrays are freed."""
    ea = ensemblealg or EnsembleMI355X()
    ctx = ctx or get_context(ea.device)
    fun = fun if isinstance(fun, ODEFunction) else ODEFunction(fun)
    dev = _is_torch(u0)
    n, N = int(u0.shape[0]), int(u0.shape[1])
    np_ = int(p.shape[0]) if p is not None else 0
    algname = "tsit5" if alg.code == _lib.TSIT5 else "rodas5p"
    ns = None
    if nvec is not None:
        nv_host = nvec.cpu().numpy() if _is_torch(nvec) else np.asarray(nvec)
        ns = sorted(set(int(x) for x in nv_host))
        if ns[-1] > n or ns[0] < 1:
            raise ValueError("nvec entries must lie in [1, u0.shape[0]]")
    h, tr, csecs = compile_problem(fun, n, np_, algs=(algname,), ctx=ctx, ns=ns, isoutofdomain=isoutofdomain,
                                   callback=callback)
    advice = _mixed_precision_advice(ea, algname, callback)
    if advice:
        warnings.warn(advice, MixedPrecisionWarning, stacklevel=2)
    reltol = 1e-3 if reltol is None else float(reltol)
    abstol = 1e-6 if abstol is None else float(abstol)
    sa = np.zeros(0) if saveat is None else np.asarray(saveat, dtype=np.float64)
    tsv = _tstops_array(tstops)
    if max_saves is None:
        max_saves = (1 if save_end else 0) + (1 if save_start else 0) + len(sa)
        if save_everystep:
            max_saves = 1024
    a = ctx.default_args(alg.code)
    a.engine = _lib.ENGINE_PLAIN if ea.engine == "plain" else _lib.ENGINE_PERSISTENT
    a.precision = _lib.MIXED if ea.precision == "mixed" else _lib.FP64
    a.refine_sweeps = ea.refine_sweeps
    a.sort_by_cost = int(ea.sort_by_cost)
    a.block = ea.block
    a.blocks_per_cu = ea.blocks_per_cu
    a.N, a.nmax, a.np = N, n, np_
    a.reltol, a.abstol = reltol, abstol
    a.dtmax = 0.0 if dtmax is None else float(dtmax)
    a.dtmin = float(dtmin)
    a.dt = 0.0 if dt is None else float(dt)
    a.maxiters = int(maxiters)
    a.save_start, a.save_end, a.save_everystep = int(save_start), int(save_end), int(save_everystep)
    a.save_end_user = int(save_end_user)
    a.max_saves = int(max_saves)
    keep = []
    if dev:
        import torch
        devc = u0.device
        dd = dict(dtype=torch.float64, device=devc)
        u0c = u0.contiguous()
        pc = (p.contiguous() if p is not None else torch.zeros((1, N), **dd))
        t0c, tfc = t0.contiguous(), tf.contiguous()
        sat = torch.as_tensor(sa, **dd) if len(sa) else torch.zeros(1, **dd)
        tst = torch.as_tensor(tsv, **dd) if len(tsv) else torch.zeros(1, **dd)
        out = outputs or {}
        o = dict(
            t_out=out.get("t_out", torch.empty((max(max_saves, 1), N), **dd)),
            u_out=out.get("u_out", torch.empty((max(max_saves, 1), n, N), **dd)),
            nsaved=out.get("nsaved", torch.empty(N, dtype=torch.int32, device=devc)),
            u_final=out.get("u_final", torch.empty((n, N), **dd)),
            t_final=out.get("t_final", torch.empty(N, **dd)),
            retcode=out.get("retcode", torch.empty(N, dtype=torch.int32, device=devc)),
            stats=out.get("stats", torch.empty((_lib.NSTATS, N), dtype=torch.int32, device=devc)),
        )
        keep += [u0c, pc, t0c, tfc, sat, tst]
        if nvec is not None:
            nvc = torch.as_tensor(nvec, dtype=torch.int32, device=devc).contiguous()
            keep.append(nvc)
            a.nvec = nvc.data_ptr()
        ptr = lambda x: x.data_ptr()
        a.device_ptrs = 1
        a.stream = torch.cuda.current_stream(devc).cuda_stream
    else:
        u0c = np.ascontiguousarray(u0, dtype=np.float64)
        pc = np.ascontiguousarray(p if p is not None else np.zeros((1, N)), dtype=np.float64)
        t0c = np.ascontiguousarray(np.broadcast_to(t0, (N,)), dtype=np.float64)
        tfc = np.ascontiguousarray(np.broadcast_to(tf, (N,)), dtype=np.float64)
        sat = np.ascontiguousarray(sa if len(sa) else np.zeros(1))
        tst = np.ascontiguousarray(tsv if len(tsv) else np.zeros(1))
        f8, i4 = np.float64, np.int32
        o = dict(t_out=_host_empty((max(max_saves, 1), N), f8), u_out=_host_empty((max(max_saves, 1), n, N), f8),
                 nsaved=_host_empty((N,), i4), u_final=_host_empty((n, N), f8), t_final=_host_empty((N,), f8),
                 retcode=_host_empty((N,), i4), stats=_host_empty((_lib.NSTATS, N), i4))
        keep += [u0c, pc, t0c, tfc, sat, tst]
        if nvec is not None:
            nvc = np.ascontiguousarray(nvec, dtype=np.int32)
            keep.append(nvc)
            a.nvec = nvc.ctypes.data
        ptr = lambda x: x.ctypes.data
        a.device_ptrs = 0
    a.u0, a.p, a.t0, a.tf = ptr(u0c), ptr(pc), ptr(t0c), ptr(tfc)
    a.saveat, a.nsaveat = ptr(sat), len(sa)
    a.tstops, a.ntstops = ptr(tst), len(tsv)
    a.t_out, a.u_out, a.nsaved = ptr(o["t_out"]), ptr(o["u_out"]), ptr(o["nsaved"])
    a.u_final, a.t_final, a.retcode, a.stats = ptr(o["u_final"]), ptr(o["t_final"]), ptr(o["retcode"]), ptr(o["stats"])
    if ea.autotune and ea.engine != "plain" and not ea.blocks_per_cu:
        key = (id(ctx), int(h), alg.code, ea.precision, N, n, int(save_everystep), len(sa), len(tsv),
               ns is not None)
        a.blocks_per_cu = _autotune(ctx, h, a, key, wg=(alg.code == _lib.RODAS5P and (ns is not None or n > LANE_MAX_RODAS)))
        o["tuning"] = dict(zip(("blocks_per_cu", "kernel_ms"), _TUNED[key]))
    t_s = time.perf_counter()
    if not ea.allocate:
        if not dev:
            raise ValueError("allocate=False needs device-resident (torch CUDA) arrays")
        ctx.solve(h, a, noalloc=True)
    else:
        ctx.solve(h, a)
    o["wall_s"] = time.perf_counter() - t_s
    o["kernel_ms"] = a.kernel_ms
    o["launches"] = a.launches
    o["compile_s"] = csecs
    o["traced"] = tr
    o["max_saves"] = max_saves
    return o


def _concat_outputs(parts):
    """Per-batch output dicts joined along the trajectory axis (the last one);
    save buffers are padded with NaN to the largest batch capacity."""
    if len(parts) == 1:
        return parts[0]
    out = dict(parts[-1])
    ms = max(q["t_out"].shape[0] for q in parts)

    def pad(a):
        if a.shape[0] == ms:
            return a
        fill = np.full((ms - a.shape[0],) + a.shape[1:], np.nan)
        return np.concatenate([a, fill], axis=0)
    out["t_out"] = np.concatenate([pad(q["t_out"]) for q in parts], axis=-1)
    out["u_out"] = np.concatenate([pad(q["u_out"]) for q in parts], axis=-1)
    out["max_saves"] = max(q["max_saves"] for q in parts)
    for k in ("nsaved", "u_final", "t_final", "retcode", "stats"):
        out[k] = np.concatenate([q[k] for q in parts], axis=-1)
    out["wall_s"] = sum(q["wall_s"] for q in parts)
    out["kernel_ms"] = sum(q["kernel_ms"] for q in parts)
    out["launches"] = sum(q["launches"] for q in parts)
    return out


def solve(prob, alg, ensemblealg=None, *, trajectories=1, reltol=None, abstol=None, saveat=None,
          save_start=None, save_end=None, save_everystep=None, dt=None, dtmax=None, dtmin=0.0,
          maxiters=1_000_000, seed=0, batch_size=None, max_saves=None, tstops=None, isoutofdomain=None,
          callback=None, **kw):
    """solve(EnsembleProblem, alg, EnsembleMI355X(); trajectories, ...) -> EnsembleSolution.

    SciMLBase's ensemble semantics (NEWS.md:209-225 for the prob_func /
    output_func contexts): trajectories run in batches of `batch_size` (one
    launch each, default all); after a batch, output_func(sol, ctx) runs per
    trajectory and a trajectory whose output_func asks for a rerun is redrawn
    with prob_func at ctx.repeat + 1 and solved again; then reduction(u,
    batch_outputs, I) -> (u, converged) folds the batch and a converged
    reduction ends the solve.  ctx.rng is seeded from (seed, sim_id, repeat).
    EnsembleSolution.arrays holds the packed outputs of the batches' first
    launches.  A plain ODEProblem is solved as a one-trajectory ensemble and
    returns its ODESolution."""
    if kw:
        raise TypeError(f"unsupported keyword(s): {sorted(kw)}")
    single = isinstance(prob, ODEProblem)
    ens = EnsembleProblem(prob) if single else prob
    if single:
        trajectories = 1
    ea = ensemblealg or EnsembleMI355X()
    t_start = time.perf_counter()
    base = ens.prob
    N = int(trajectories)
    bs = N if not batch_size else max(1, int(batch_size))
    n = len(base.u0)
    np_ = len(base.p)
    # save rules are resolved against the template tspan (shared by the common case)
    ss, se, sev, seu = _resolve_save_flags(saveat, save_start, save_end, save_everystep,
                                           base.tspan[0], base.tspan[1])
    cbs = _callback_list(callback)
    forced = any(any(c.save_positions) for c in cbs)
    if max_saves is None and (sev or forced):
        max_saves = 4096

    def draw(i, repeat):
        rng = np.random.default_rng([int(seed), i + 1] if repeat == 1 else [int(seed), i + 1, repeat])
        pi = ens.prob_func(remake(base) if ens.safetycopy else base, EnsembleContext(i + 1, repeat, rng))
        if len(pi.u0) != n or len(pi.p) != np_:
            raise ValueError("all trajectories must share the state and parameter dimensions")
        if pi.f.f is not base.f.f:
            raise ValueError("prob_func must not change the right-hand side function")
        return pi

    def run(probs):
        Nb = len(probs)
        u0 = np.stack([pi.u0 for pi in probs], axis=1) if Nb else np.zeros((n, 0))
        pp = np.stack([pi.p for pi in probs], axis=1) if np_ else None
        t0 = np.array([pi.tspan[0] for pi in probs])
        tf = np.array([pi.tspan[1] for pi in probs])
        if saveat is not None and np.isscalar(saveat):
            back = tf < t0
            if back.any() and not back.all():
                raise ValueError("a saveat interval needs every trajectory to integrate in the same direction")
            if back.any():
                sa = saveat_times(saveat, float(t0.max()), float(tf.min()))
            else:
                sa = saveat_times(saveat, float(t0.min(initial=base.tspan[0])), float(tf.max(initial=base.tspan[1])))
        elif saveat is not None and len(saveat):
            sa = saveat_times(saveat, 0, 0)
        else:
            sa = np.zeros(0)
        o = solve_arrays(base.f, alg, u0, pp, t0, tf, ea, reltol=reltol, abstol=abstol, saveat=sa,
                         save_start=ss, save_end=se, save_everystep=sev, save_end_user=seu, max_saves=max_saves,
                         dt=dt, dtmax=dtmax, dtmin=dtmin, maxiters=maxiters, tstops=tstops,
                         isoutofdomain=isoutofdomain, callback=cbs)
        if (sev or forced) and int(o["nsaved"].max(initial=0)) > o["max_saves"]:
            raise RuntimeError(f"the saves (save_everystep / callback save_positions) exceeded max_saves="
                               f"{o['max_saves']} points; pass a larger max_saves")
        sols = []
        for j in range(Nb):
            ns = min(int(o["nsaved"][j]), o["max_saves"])
            st = o["stats"][:, j]
            sols.append(ODESolution(o["t_out"][:ns, j].copy(), [o["u_out"][k, :, j].copy() for k in range(ns)],
                                    ReturnCode(int(o["retcode"][j])), DEStats(*st), probs[j], alg))
        return sols, o

    u_acc = ens.u_init
    converged = False
    collected, parts = [], []
    for b0 in range(0, N, bs):
        idx = list(range(b0, min(N, b0 + bs)))
        sols, o = run([draw(i, 1) for i in idx])
        parts.append(o)
        outs = [None] * len(idx)
        pending = list(range(len(idx)))
        repeat = 1
        while pending:
            redo = []
            for q in pending:
                res, rerun = ens.output_func(sols[q], EnsembleContext(idx[q] + 1, repeat, None))
                outs[q] = res
                if rerun:
                    redo.append(q)
            if not redo:
                break
            repeat += 1
            resols, _ = run([draw(idx[q], repeat) for q in redo])
            for q, sq in zip(redo, resols):
                sols[q] = sq
            pending = redo
        if ens.reduction is not None:
            u_acc, converged = ens.reduction(u_acc, outs, [i + 1 for i in idx])
            if converged:
                break
        else:
            collected += outs
    o = _concat_outputs(parts) if parts else run([])[1]
    result = u_acc if ens.reduction is not None else collected
    elapsed = time.perf_counter() - t_start
    if single:
        return result[0]
    return EnsembleSolution(result, elapsed, o, converged=bool(converged) if ens.reduction is not None else True)
