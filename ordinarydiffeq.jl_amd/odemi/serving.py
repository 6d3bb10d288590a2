"""Continuous-batching ensemble server (odemi_server_* in include/odemi.h).

Many independent callers submit small ensembles (a request = N trajectories of
one compiled problem); the native worker thread admits them into a
device-resident slot pool as slots free up, advances every active trajectory
together in ticks of a few step attempts, and hands each request its final
states as soon as its last trajectory finishes.  The pool size is the
admission bound: it is set from an HBM budget (bytes) or a slot count.

The reference has no serving layer; the per-trajectory semantics are the
reference's `solve` with save_everystep=false, identical to odemi.solve_arrays
on the same inputs: the final state, and (lane servers created with max_saves
/ max_tstops) the request's own save times and stopping times.
"""
import ctypes

import numpy as np

from . import _lib
from .api import ODEFunction, compile_problem, get_context


class Request:
    """Handle of one submitted ensemble; outputs are filled at completion."""

    def __init__(self, server, rid, N, u_final, retcode, stats, keep, t_out=None, u_out=None, nsaved=None):
        self.server, self.id, self.N = server, rid, N
        self.u_final, self.retcode, self.stats = u_final, retcode, stats
        self.t_out, self.u_out, self.nsaved = t_out, u_out, nsaved   # saved points (requests with saves)
        self._keep = keep

    def done(self):
        rc = self.server._L.odemi_poll(self.server._handle(), self.id)
        if rc < 0:
            raise RuntimeError(f"request {self.id} failed: {self.server.error()}")
        if rc == 1:
            self.server._live.pop(self.id, None)
        return rc == 1

    def wait(self, timeout=None):
        rc = self.server._L.odemi_wait(self.server._handle(), self.id, 0.0 if timeout is None else float(timeout))
        if rc < 0:
            raise RuntimeError(f"request {self.id} failed: {self.server.error()}")
        if rc == 1:
            self.server._live.pop(self.id, None)
        return rc == 1

    def times(self):
        """(t_submit, t_admit, t_done) on the steady clock, seconds."""
        a, b, c = _lib.c_d(), _lib.c_d(), _lib.c_d()
        if self.server._L.odemi_request_times(self.server._handle(), self.id, ctypes.byref(a), ctypes.byref(b),
                                              ctypes.byref(c)) != 0:
            raise RuntimeError(f"unknown request {self.id}")
        return a.value, b.value, c.value

    def latency(self):
        s, _, d = self.times()
        return d - s

    def release(self):
        """Drop the server's record of this finished request (times() then fails)."""
        if self.server._h and self.server._L.odemi_request_release(self.server._h, self.id) == 0:
            self.server._live.pop(self.id, None)


class EnsembleServer:
    """Serve one problem type (ODEFunction + state size + algorithm).

    `ns` (Rodas5P): the state sizes requests may use (uneven n, workgroup
    kernels, n = max(ns)); each request then passes its per-trajectory `nvec`.
    `max_saves` (lane servers) / `max_tstops`: the saved points per trajectory
    and the stopping times per request a request may ask for (the pool keeps
    max_saves points per slot, counted in the HBM budget).
    `callback` / `isoutofdomain`: traced into the served problem, as for
    solve_arrays; every request of the server runs them."""

    def __init__(self, fun, n, np_, alg, *, precision="fp64", reltol=1e-3, abstol=1e-6, dtmax=None, dtmin=0.0,
                 maxiters=1_000_000, max_slots=1 << 20, hbm_budget_bytes=0, tick_attempts=64, refine_sweeps=1,
                 ns=None, ctx=None, max_saves=0, max_tstops=0, callback=None, isoutofdomain=None):
        fun = fun if isinstance(fun, ODEFunction) else ODEFunction(fun)
        self.ctx = ctx or get_context()
        algname = "tsit5" if alg.code == _lib.TSIT5 else "rodas5p"
        self.ns = sorted(set(int(x) for x in ns)) if ns is not None else None
        if self.ns is not None and (self.ns[0] < 1 or self.ns[-1] > n):
            raise ValueError("ns entries must lie in [1, n]")
        self._rhs, _, _ = compile_problem(fun, n, np_, algs=(algname,), ctx=self.ctx, ns=self.ns,
                                          callback=callback, isoutofdomain=isoutofdomain)
        self.n, self.np = n, np_
        self._L = self.ctx.L
        o = _lib.ServerOpts()
        self._L.odemi_server_default_opts(ctypes.byref(o), alg.code)
        o.precision = _lib.MIXED if precision == "mixed" else _lib.FP64
        o.reltol, o.abstol = float(reltol), float(abstol)
        o.dtmax = 0.0 if dtmax is None else float(dtmax)
        o.dtmin = float(dtmin)
        o.maxiters = int(maxiters)
        o.max_slots = int(max_slots)
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        o.tick_attempts = int(tick_attempts)
        o.refine_sweeps = int(refine_sweeps)
        o.max_saves = int(max_saves)
        o.max_tstops = int(max_tstops)
        self.max_saves = int(max_saves)
        self._h = self._L.odemi_server_create(self.ctx.h, self._rhs, ctypes.byref(o))
        if not self._h:
            raise RuntimeError("odemi_server_create failed: " + self.ctx.error())
        self._live = {}   # request id -> output buffers the worker writes at completion

    def _handle(self):
        if not self._h:
            raise RuntimeError("the server is closed")
        return self._h

    def error(self):
        return self._L.odemi_server_last_error(self._handle()).decode(errors="replace")

    def submit(self, u0, p, t0, tf, stats=True, nvec=None, reltol=None, abstol=None, saveat=None, tstops=None,
               save_start=False, save_end=True, save_end_user=-1):
        """u0 (n, N), p (np, N) or None, t0/tf (N,) or scalars, nvec (N,) per-
        trajectory sizes (servers created with `ns`); reltol / abstol: this
        request's tolerances (None: the server's); saveat / tstops: its save
        times and stopping times (lane servers with max_saves / max_tstops),
        save_start / save_end / save_end_user as in solve_arrays: with saveat
        or save_start the request's Request carries t_out (max_saves, N),
        u_out (max_saves, n, N) and nsaved (N,).  Inputs are copied before
        return; returns a Request."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        if u0.ndim == 1:
            u0 = u0[:, None]
        n, N = u0.shape
        if n != self.n:
            raise ValueError(f"state size {n} != served size {self.n}")
        t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (N,)))
        tf = np.ascontiguousarray(np.broadcast_to(np.asarray(tf, dtype=np.float64), (N,)))
        pp = None
        if self.np > 0:
            pp = np.ascontiguousarray(p, dtype=np.float64).reshape(self.np, N)
        nv = None
        if nvec is not None:
            nv = np.ascontiguousarray(nvec, dtype=np.int32).reshape(N)
            if self.ns is None and not (nv == self.n).all():
                raise ValueError("uneven nvec needs a server created with ns")
            if self.ns is not None and not np.isin(nv, self.ns).all():
                raise ValueError("nvec entries must be among the server's ns")
        elif self.ns is not None and self.n not in self.ns:
            raise ValueError("pass nvec: n itself is not among the server's ns")
        uf = np.empty((n, N), dtype=np.float64)
        rc = np.zeros(N, dtype=np.int32)
        st = np.zeros((_lib.NSTATS, N), dtype=np.int32) if stats else None
        ptr = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
        ro = _lib.RequestOpts(0.0 if reltol is None else float(reltol), 0.0 if abstol is None else float(abstol))
        if (reltol is not None and not ro.reltol > 0) or (abstol is not None and not ro.abstol > 0):
            raise ValueError("request tolerances must be positive")
        sa = np.ascontiguousarray(np.sort(np.atleast_1d(np.asarray(saveat, dtype=np.float64)))) \
            if saveat is not None else np.zeros(0)
        ts = np.ascontiguousarray(np.sort(np.atleast_1d(np.asarray(tstops, dtype=np.float64)))) \
            if tstops is not None else np.zeros(0)
        tout = uout = nsv = None
        if len(sa) or save_start:
            if self.max_saves <= 0:
                raise ValueError("this server keeps no saved points (create it with max_saves)")
            tout = np.full((self.max_saves, N), np.nan)
            uout = np.full((self.max_saves, n, N), np.nan)
            nsv = np.zeros(N, dtype=np.int32)
            ro.saveat, ro.nsaveat = ptr(sa) if len(sa) else None, len(sa)
            ro.save_start, ro.save_end, ro.save_end_user = int(bool(save_start)), int(bool(save_end)), int(save_end_user)
            ro.t_out, ro.u_out, ro.nsaved = ptr(tout), ptr(uout), ptr(nsv)
        else:
            ro.save_end_user = -1
        if len(ts):
            ro.tstops, ro.ntstops = ptr(ts), len(ts)
        rid = self._L.odemi_submit_opts(self._handle(), N, ptr(nv), ptr(u0), ptr(pp), ptr(t0), ptr(tf),
                                        ctypes.byref(ro), ptr(uf), ptr(rc), ptr(st))
        if rid <= 0:
            raise RuntimeError(f"odemi_submit failed ({rid}): {self.error()}")
        keep = (uf, rc, st, tout, uout, nsv)
        self._live[rid] = keep
        return Request(self, rid, N, uf, rc, st, keep, tout, uout, nsv)

    def info(self):
        v = [_lib.c_i64() for _ in range(5)]
        self._L.odemi_server_info(self._handle(), *[ctypes.byref(x) for x in v])
        return dict(zip(("slots", "ticks", "pending", "active", "finished"), (x.value for x in v)))

    def close(self):
        if self._h:
            self._L.odemi_server_destroy(self._h)   # joins the worker: no writes after this
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
