"""ctypes front-end to the CPU restatement in odeoracle.c.

TEST INFRASTRUCTURE ONLY: the parity checker used by tests/, by the smoke
check and by bench.py's CPU-baseline leg.  The product never imports this.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libodeoracle.so")

TSIT5, RODAS5P = 0, 1
PROB_USER, PROB_LORENZ, PROB_ROBERTSON, PROB_CHAIN, PROB_LINEAR, PROB_EXPBLOWUP, PROB_BALL = range(7)
RC_NAMES = {0: "Default", 1: "Success", 2: "Terminated", 3: "DtNaN", 4: "MaxIters",
            5: "MaxNumSub", 6: "DtLessThanMin", 7: "Unstable"}

CCB_COND_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.c_double, ctypes.c_int, ctypes.c_void_p)
CCB_AFFECT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)
VCB_COND_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_int, ctypes.c_void_p)
VCB_AFFECT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                 ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
DCB_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                          ctypes.c_double, ctypes.c_int, ctypes.c_void_p)
ISOUT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                            ctypes.c_double, ctypes.c_void_p)
RHS_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_void_p)


class Opts(ctypes.Structure):
    _fields_ = [("alg", ctypes.c_int), ("reltol", ctypes.c_double), ("abstol", ctypes.c_double),
                ("dtmax", ctypes.c_double), ("dtmin", ctypes.c_double), ("dt", ctypes.c_double),
                ("maxiters", ctypes.c_int64), ("save_start", ctypes.c_int), ("save_end", ctypes.c_int),
                ("save_everystep", ctypes.c_int), ("save_end_user", ctypes.c_int),
                ("saveat", ctypes.POINTER(ctypes.c_double)), ("nsaveat", ctypes.c_int),
                ("has_tgrad", ctypes.c_int), ("adaptive", ctypes.c_int),
                ("tstops", ctypes.POINTER(ctypes.c_double)), ("ntstops", ctypes.c_int),
                ("isoutofdomain", ctypes.c_void_p), ("isout_user", ctypes.c_void_p),
                ("ndcb", ctypes.c_int), ("dcb_cond", ctypes.c_void_p), ("dcb_affect", ctypes.c_void_p),
                ("dcb_save", ctypes.POINTER(ctypes.c_int)), ("dcb_user", ctypes.c_void_p),
                ("nccb", ctypes.c_int), ("ccb_cond", ctypes.c_void_p), ("ccb_affect", ctypes.c_void_p),
                ("ccb_flags", ctypes.POINTER(ctypes.c_int)), ("ccb_tol", ctypes.POINTER(ctypes.c_double)),
                ("ccb_user", ctypes.c_void_p), ("vcb_cond", ctypes.c_void_p), ("vcb_affect", ctypes.c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("retcode", ctypes.c_int), ("nf", ctypes.c_int64), ("naccept", ctypes.c_int64),
                ("nreject", ctypes.c_int64), ("njacs", ctypes.c_int64), ("nw", ctypes.c_int64),
                ("nsolve", ctypes.c_int64), ("nsaved", ctypes.c_int), ("t_final", ctypes.c_double),
                ("dt_last", ctypes.c_double), ("dt_init", ctypes.c_double), ("ncondition", ctypes.c_int64)]


def build():
    """Compile libodeoracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.POINTER
        d, i64, i32 = ctypes.c_double, ctypes.c_int64, ctypes.c_int32
        L.oo_default_opts.argtypes = [P(Opts), ctypes.c_int]
        L.oo_solve.argtypes = [ctypes.c_int, RHS_FN, RHS_FN, ctypes.c_void_p, ctypes.c_void_p,
                               P(d), P(d), d, d, P(Opts), P(d), P(d), ctypes.c_int, P(d), P(Result)]
        L.oo_solve_builtin.argtypes = [ctypes.c_int, ctypes.c_int, P(d), P(d), d, d, P(Opts),
                                       P(d), P(d), ctypes.c_int, P(d), P(Result)]
        L.oo_ensemble_builtin.argtypes = [ctypes.c_int, i64, ctypes.c_int, P(ctypes.c_int), P(d),
                                          ctypes.c_int, P(d), P(d), P(d), P(Opts), ctypes.c_int, i64,
                                          P(d), P(i32), P(i64), P(i64), P(i64)]
        L.oo_ensemble_builtin.restype = d
        L.oo_fastpower.argtypes = [d, d]
        L.oo_fastpower.restype = d
        L.oo_fastlog2.argtypes = [ctypes.c_float]
        L.oo_fastlog2.restype = ctypes.c_float
        L.oo_jl_eps.argtypes = [d]
        L.oo_jl_eps.restype = d
        L.oo_det_log.argtypes = [d]
        L.oo_det_log.restype = d
        L.oo_det_exp.argtypes = [d]
        L.oo_det_exp.restype = d
        L.oo_lu_solve.argtypes = [ctypes.c_int, P(d), P(ctypes.c_int), P(d)]
        L.oo_tsit5_tableau.argtypes = [P(d)] * 4
        L.oo_rodas5p_tableau.argtypes = [P(d)] * 8
        L.oo_tsit5_interp_coeffs.argtypes = [P(d)]
        L.oo_vern9_builtin.argtypes = [ctypes.c_int, ctypes.c_int, P(d), P(d), d, d, d, d, i64, P(d), P(i64)]
        L.oo_vern9_builtin.restype = ctypes.c_int
        L.oo_vern9_ensemble.argtypes = [ctypes.c_int, i64, ctypes.c_int, P(ctypes.c_int), P(d), ctypes.c_int, P(d),
                                        P(d), P(d), d, d, i64, ctypes.c_int, P(d), P(ctypes.c_int32), P(i64)]
        L.oo_vern9_ensemble.restype = d
        L.oo_vern9_tableau.argtypes = [P(d)] * 4
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _CallbackIntegrator:
    """The integrator a callback sees in the oracle: u (numpy, writable or
    reassignable), p, t, terminate()."""

    def __init__(self, u, p, t):
        self.u, self.p, self.t, self._term = u, p, t, False

    def terminate(self, when=None):
        if when is None or bool(when):
            self._term = True


def make_opts(alg=TSIT5, reltol=1e-3, abstol=1e-6, dtmax=None, dtmin=0.0, dt=0.0, maxiters=1000000,
              save_start=True, save_end=True, save_everystep=False, save_end_user=None, saveat=None,
              has_tgrad=False, adaptive=True, tstops=None, isoutofdomain=None, isout_shape=(0, 0),
              callbacks=None, ccallbacks=None):
    """callbacks: objects with condition(u, t, integrator), affect(integrator)
    and save_positions (DiscreteCallback's fields), run as plain Python on numpy
    copies of the state.  ccallbacks: objects with ContinuousCallback's fields
    condition (a value), affect / affect_neg (callable or None), rootfind
    ("left" / "right" / "none"), interp_points, save_positions, abstol,
    repeat_nudge; a VectorContinuousCallback also has len, its condition
    fills out[0:len] and its affect takes (integrator, events)."""
    o = Opts()
    lib().oo_default_opts(ctypes.byref(o), alg)
    o.reltol, o.abstol = reltol, abstol
    o.dtmax = float("nan") if dtmax is None else dtmax
    o.dtmin, o.dt, o.maxiters = dtmin, dt, maxiters
    o.save_start, o.save_end, o.save_everystep = int(save_start), int(save_end), int(save_everystep)
    o.save_end_user = -1 if save_end_user is None else int(save_end_user)
    keep = None
    if saveat is not None and len(saveat) > 0:
        keep = np.ascontiguousarray(saveat, dtype=np.float64)
        o.saveat, o.nsaveat = _dp(keep), len(keep)
    o.has_tgrad = int(has_tgrad)
    o.adaptive = int(adaptive)
    if tstops is not None and len(tstops) > 0:
        ts = np.ascontiguousarray(tstops, dtype=np.float64)
        o.tstops, o.ntstops = _dp(ts), len(ts)
        keep = (keep, ts)
    if isoutofdomain is not None:
        # isoutofdomain(u, p, t) -> bool on numpy copies of the state and parameters
        n, np_ = isout_shape

        def _io(u, p, t, _user, _f=isoutofdomain):
            uu = np.array(np.ctypeslib.as_array(u, shape=(max(n, 1),))[:n])
            pp = np.array(np.ctypeslib.as_array(p, shape=(max(np_, 1),))[:np_])
            return 1 if _f(uu, pp, t) else 0
        cb = ISOUT_FN(_io)
        o.isoutofdomain = ctypes.cast(cb, ctypes.c_void_p)
        keep = (keep, cb)
    if callbacks:
        n, np_ = isout_shape
        cbl = list(callbacks)

        def _cond(u, p, t, k, _user):
            uu = np.array(np.ctypeslib.as_array(u, shape=(max(n, 1),))[:n])
            pp = np.array(np.ctypeslib.as_array(p, shape=(max(np_, 1),))[:np_])
            return 1 if cbl[k].condition(uu, t, _CallbackIntegrator(uu.copy(), pp, t)) else 0

        def _affect(u, p, t, k, _user):
            ua = np.ctypeslib.as_array(u, shape=(max(n, 1),))
            pp = np.array(np.ctypeslib.as_array(p, shape=(max(np_, 1),))[:np_])
            itg = _CallbackIntegrator(np.array(ua[:n]), pp, t)
            cbl[k].affect(itg)
            new = np.asarray(itg.u, dtype=np.float64).reshape(-1)
            if len(new) != n:
                raise ValueError("affect! must keep the state length")
            ua[:n] = new
            return 1 if itg._term else 0
        cc, ca = DCB_FN(_cond), DCB_FN(_affect)
        sv = np.ascontiguousarray([int(bool(x)) for c in cbl for x in c.save_positions], dtype=np.int32)
        o.ndcb = len(cbl)
        o.dcb_cond, o.dcb_affect = ctypes.cast(cc, ctypes.c_void_p), ctypes.cast(ca, ctypes.c_void_p)
        o.dcb_save = sv.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        keep = (keep, cc, ca, sv)
    if ccallbacks:
        n, np_ = isout_shape
        ccl = list(ccallbacks)

        def _ccond(u, p, t, k, _user):
            uu = np.array(np.ctypeslib.as_array(u, shape=(max(n, 1),))[:n])
            pp = np.array(np.ctypeslib.as_array(p, shape=(max(np_, 1),))[:np_])
            return float(ccl[k].condition(uu, t, _CallbackIntegrator(uu.copy(), pp, t)))

        def _caffect(u, p, t, k, neg, _user):
            ua = np.ctypeslib.as_array(u, shape=(max(n, 1),))
            pp = np.array(np.ctypeslib.as_array(p, shape=(max(np_, 1),))[:np_])
            itg = _CallbackIntegrator(np.array(ua[:n]), pp, t)
            (ccl[k].affect_neg if neg else ccl[k].affect)(itg)
            ua[:n] = np.asarray(itg.u, dtype=np.float64).reshape(-1)
            return 1 if itg._term else 0

        def _vcond(out, u, p, t, k, _user):
            uu = np.array(np.ctypeslib.as_array(u, shape=(max(n, 1),))[:n])
            pp = np.array(np.ctypeslib.as_array(p, shape=(max(np_, 1),))[:np_])
            o_ = np.zeros(ccl[k].len)
            ccl[k].condition(o_, uu, t, _CallbackIntegrator(uu.copy(), pp, t))
            for i in range(ccl[k].len):
                out[i] = float(o_[i])

        def _vaffect(u, p, t, k, ev, _user):
            ua = np.ctypeslib.as_array(u, shape=(max(n, 1),))
            pp = np.array(np.ctypeslib.as_array(p, shape=(max(np_, 1),))[:np_])
            evs = np.array(np.ctypeslib.as_array(ev, shape=(ccl[k].len,)))
            itg = _CallbackIntegrator(np.array(ua[:n]), pp, t)
            ccl[k].affect(itg, evs)
            ua[:n] = np.asarray(itg.u, dtype=np.float64).reshape(-1)
            return 1 if itg._term else 0
        rf = {"left": 0, "right": 1, "none": 2}
        fl = np.ascontiguousarray([[int(c.affect is not None), int(getattr(c, "affect_neg", None) is not None),
                                    int(bool(c.save_positions[0])), int(bool(c.save_positions[1])),
                                    rf[c.rootfind], int(c.interp_points), int(getattr(c, "len", 0))]
                                   for c in ccl], dtype=np.int32)
        tl = np.ascontiguousarray([[float(c.abstol), float(c.repeat_nudge)] for c in ccl], dtype=np.float64)
        c1, c2 = CCB_COND_FN(_ccond), CCB_AFFECT_FN(_caffect)
        c3, c4 = VCB_COND_FN(_vcond), VCB_AFFECT_FN(_vaffect)
        o.nccb = len(ccl)
        o.ccb_cond, o.ccb_affect = ctypes.cast(c1, ctypes.c_void_p), ctypes.cast(c2, ctypes.c_void_p)
        o.vcb_cond, o.vcb_affect = ctypes.cast(c3, ctypes.c_void_p), ctypes.cast(c4, ctypes.c_void_p)
        o.ccb_flags = fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        o.ccb_tol = tl.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        keep = (keep, c1, c2, c3, c4, fl, tl)
    return o, keep


def solve(f, u0, p, tspan, jac=None, tgrad=None, max_out=4096, **kw):
    """Solve one IVP with Python callables f(du,u,p,t) / jac(J,u,p,t) (J is n x n).

    Returns dict(t, u, u_final, retcode, nf, naccept, nreject, njacs, nw, nsolve, t_final)."""
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    n = len(u0)
    pp = np.ascontiguousarray(p if p is not None and len(p) else [0.0], dtype=np.float64)

    def _f(du, u, pr, t, _user):
        uu = np.ctypeslib.as_array(u, shape=(max(n, 1),))[:n]
        out = np.zeros(n)
        f(out, uu.copy(), pp, t)
        for i in range(n):
            du[i] = out[i]

    def _j(J, u, pr, t, _user):
        uu = np.ctypeslib.as_array(u, shape=(max(n, 1),))[:n]
        out = np.zeros((n, n))
        jac(out, uu.copy(), pp, t)
        flat = out.reshape(-1)
        for i in range(n * n):
            J[i] = flat[i]

    def _tg(dT, u, pr, t, _user):
        uu = np.ctypeslib.as_array(u, shape=(max(n, 1),))[:n]
        out = np.zeros(n)
        tgrad(out, uu.copy(), pp, t)
        for i in range(n):
            dT[i] = out[i]

    cf = RHS_FN(_f)
    cj = RHS_FN(_j) if jac is not None else RHS_FN(lambda *a: None)
    ct = ctypes.cast(RHS_FN(_tg), ctypes.c_void_p) if tgrad is not None else None
    keep_tg = ct
    opts, keep = make_opts(has_tgrad=tgrad is not None, isout_shape=(n, len(pp)), **kw)
    t_out = np.zeros(max_out)
    u_out = np.zeros((max_out, max(n, 1)))
    uf = np.zeros(max(n, 1))
    res = Result()
    lib().oo_solve(n, cf, cj, keep_tg, None, _dp(u0), _dp(pp), float(tspan[0]), float(tspan[1]),
                   ctypes.byref(opts), _dp(t_out), _dp(u_out), max_out, _dp(uf), ctypes.byref(res))
    ns = min(res.nsaved, max_out)
    return dict(t=t_out[:ns].copy(), u=u_out[:ns, :n].copy(), u_final=uf[:n].copy(), retcode=res.retcode,
                nf=res.nf, naccept=res.naccept, nreject=res.nreject, njacs=res.njacs, nw=res.nw,
                nsolve=res.nsolve, t_final=res.t_final, nsaved=res.nsaved, dt_init=res.dt_init,
                ncondition=res.ncondition)


def solve_builtin(prob, u0, p, tspan, max_out=4096, **kw):
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    n = len(u0)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    opts, keep = make_opts(isout_shape=(n, len(pp)), **kw)
    t_out = np.zeros(max_out)
    u_out = np.zeros((max_out, max(n, 1)))
    uf = np.zeros(max(n, 1))
    res = Result()
    lib().oo_solve_builtin(prob, n, _dp(u0), _dp(pp), float(tspan[0]), float(tspan[1]), ctypes.byref(opts),
                           _dp(t_out), _dp(u_out), max_out, _dp(uf), ctypes.byref(res))
    ns = min(res.nsaved, max_out)
    return dict(t=t_out[:ns].copy(), u=u_out[:ns, :n].copy(), u_final=uf[:n].copy(), retcode=res.retcode,
                nf=res.nf, naccept=res.naccept, nreject=res.nreject, njacs=res.njacs, nw=res.nw,
                nsolve=res.nsolve, t_final=res.t_final, nsaved=res.nsaved, dt_init=res.dt_init,
                ncondition=res.ncondition)


def ensemble_builtin(prob, u0_soa, p_soa, t0, tf, nvec=None, threads=None, chunk=64, **kw):
    """Threaded ensemble of a built-in problem.  u0_soa: (nmax, N), p_soa: (np, N).

    Returns dict(u_final (nmax,N), retcode, naccept, nreject, nf, seconds, threads)."""
    u0_soa = np.ascontiguousarray(u0_soa, dtype=np.float64)
    p_soa = np.ascontiguousarray(p_soa, dtype=np.float64)
    nmax, N = u0_soa.shape
    npar = p_soa.shape[0]
    t0 = np.ascontiguousarray(np.broadcast_to(t0, (N,)), dtype=np.float64)
    tf = np.ascontiguousarray(np.broadcast_to(tf, (N,)), dtype=np.float64)
    if threads is None:
        threads = os.cpu_count() or 1
    opts, keep = make_opts(**kw)
    uf = np.zeros((nmax, N))
    rc = np.zeros(N, dtype=np.int32)
    na = np.zeros(N, dtype=np.int64)
    nr = np.zeros(N, dtype=np.int64)
    nf = np.zeros(N, dtype=np.int64)
    nv = None
    if nvec is not None:
        nv_arr = np.ascontiguousarray(nvec, dtype=np.int32)
        nv = nv_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    P64 = ctypes.POINTER(ctypes.c_int64)
    secs = lib().oo_ensemble_builtin(prob, N, nmax, nv, _dp(u0_soa), npar, _dp(p_soa), _dp(t0), _dp(tf),
                                     ctypes.byref(opts), threads, chunk, _dp(uf),
                                     rc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     na.ctypes.data_as(P64), nr.ctypes.data_as(P64), nf.ctypes.data_as(P64))
    return dict(u_final=uf, retcode=rc, naccept=na, nreject=nr, nf=nf, seconds=secs, threads=threads)


def fastpower(x, y):
    return lib().oo_fastpower(x, y)


def fastlog2(x):
    return lib().oo_fastlog2(x)


def jl_eps(x):
    return lib().oo_jl_eps(x)


def lu_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    n = len(b)
    piv = (ctypes.c_int * max(n, 1))()
    lib().oo_lu_solve(n, _dp(A), piv, _dp(b))
    return b


def tsit5_tableau():
    c, a, b, bt = np.zeros(7), np.zeros(49), np.zeros(7), np.zeros(7)
    lib().oo_tsit5_tableau(_dp(c), _dp(a), _dp(b), _dp(bt))
    return c, a.reshape(7, 7), b, bt


def rodas5p_tableau():
    A, C, c, d, H, b, bt = (np.zeros(64), np.zeros(64), np.zeros(8), np.zeros(8), np.zeros(24),
                            np.zeros(8), np.zeros(8))
    g = ctypes.c_double()
    lib().oo_rodas5p_tableau(_dp(A), _dp(C), _dp(c), _dp(d), _dp(H), _dp(b), _dp(bt), ctypes.byref(g))
    return dict(A=A.reshape(8, 8), C=C.reshape(8, 8)[:, :7], c=c, d=d, H=H.reshape(3, 8), b=b,
                btilde=bt, gamma=g.value)


def tsit5_interp_coeffs():
    r = np.zeros(28)
    lib().oo_tsit5_interp_coeffs(_dp(r))
    return r.reshape(7, 4)


def det_log(x):
    return lib().oo_det_log(x)


def det_exp(x):
    return lib().oo_det_exp(x)


def vern9_tableau():
    """Vern9 (c[16], A[16][16], b[16], btilde[16]) as held by the C restatement."""
    c, a, b, bt = np.zeros(16), np.zeros((16, 16)), np.zeros(16), np.zeros(16)
    lib().oo_vern9_tableau(_dp(c), _dp(a), _dp(b), _dp(bt))
    return c, a, b, bt


def vern9_builtin(prob, u0, p, tspan, reltol=1e-12, abstol=1e-12, maxsteps=10_000_000):
    """High-accuracy Vern9 reference solve of a built-in problem; returns (u_final, status, nsteps)."""
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    p = np.ascontiguousarray(p if p is not None and len(p) else [0.0], dtype=np.float64)
    uf = np.zeros(len(u0))
    ns = ctypes.c_int64()
    st = lib().oo_vern9_builtin(prob, len(u0), _dp(u0), _dp(p), float(tspan[0]), float(tspan[1]), reltol, abstol,
                                maxsteps, _dp(uf), ctypes.byref(ns))
    return uf, st, ns.value


def vern9_ensemble(prob, u0_soa, p_soa, t0, tf, nvec=None, reltol=1e-12, abstol=1e-12, maxsteps=10_000_000,
                   threads=None):
    """Vern9 reference for every trajectory of an SoA batch (threaded)."""
    u0_soa = np.ascontiguousarray(u0_soa, dtype=np.float64)
    p_soa = np.ascontiguousarray(p_soa, dtype=np.float64)
    nmax, N = u0_soa.shape
    t0 = np.ascontiguousarray(np.broadcast_to(t0, (N,)), dtype=np.float64)
    tf = np.ascontiguousarray(np.broadcast_to(tf, (N,)), dtype=np.float64)
    nv = None if nvec is None else np.ascontiguousarray(nvec, dtype=np.int32)
    uf = np.zeros((nmax, N))
    st = np.zeros(N, dtype=np.int32)
    ns = np.zeros(N, dtype=np.int64)
    threads = threads or min(len(os.sched_getaffinity(0)), 16)
    secs = lib().oo_vern9_ensemble(prob, N, nmax, nv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if nv is not None
                                   else None, _dp(u0_soa), p_soa.shape[0], _dp(p_soa), _dp(t0), _dp(tf), reltol,
                                   abstol, maxsteps, threads, _dp(uf),
                                   st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return dict(u_final=uf, status=st, nsteps=ns, seconds=secs)
