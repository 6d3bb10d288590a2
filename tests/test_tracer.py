"""Host-side tracing (no GPU): the HIP source the tracer emits for f, the
isoutofdomain predicate and DiscreteCallbacks, checked by structure and by
hipRTC-compiling it for gfx950 (odemi_rtc_codeobject needs no device)."""
import ctypes
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "ordinarydiffeq.jl_amd"))
import odemi  # noqa: E402
from odemi import _lib  # noqa: E402
from odemi.tracer import trace_problem  # noqa: E402
from odemi.api import trace_for  # noqa: E402


def _decay(du, u, p, t):
    du[0] = -0.5 * u[0] + 10
    du[1] = -p[0] * u[1]


def _dose_cond(u, t, itg):
    return odemi.sym_in(t, [5.0, 8.0])


def _dose(itg):
    itg.u[1] = itg.u[1] + itg.p[1]


def _reset_cond(u, t, itg):
    return (u[0] >= 19.0) & ~(t < 1.0)


def _reset_stop(itg):
    itg.u[:] = 1.0
    odemi.terminate(itg)


def test_equality_builds_predicates():
    tr = trace_problem(_decay, 2, 2, isout=lambda u, p, t: (t == 3.0) | (u[0] != p[0]))
    assert "#define ODEMI_HAS_ISOUT 1" in tr.src
    assert "(t == T(3.0))" in tr.src and "(u[0] != T(p[0]))" in tr.src
    assert odemi.sym_in(2.0, [1.0, 2.0]) is True and odemi.sym_in(2.5, [1.0, 2.0]) is False


def test_discrete_callbacks_emit_condition_affect_and_flags():
    cbs = odemi.CallbackSet(odemi.DiscreteCallback(_dose_cond, _dose),
                            odemi.DiscreteCallback(_reset_cond, _reset_stop, save_positions=(False, True)))
    tr = trace_problem(_decay, 2, 2, callbacks=cbs.discrete_callbacks)
    src = tr.src
    assert "#define ODEMI_NDCB 2" in src
    assert "static constexpr int save0 = 1, save1 = 1, term = 0;" in src
    assert "static constexpr int save0 = 0, save1 = 1, term = 1;" in src
    first = src[src.index("struct OdemiDcb<0>"):src.index("struct OdemiDcb<1>")]
    # the dose writes u[1] only, from the pre-affect state
    assert "u[1] = nu_[1];" in first and "u[0] = nu_[0];" not in first
    second = src[src.index("struct OdemiDcb<1>"):]
    assert "u[0] = nu_[0];" in second and "u[1] = nu_[1];" in second


def test_callback_module_compiles_for_gfx950():
    cbs = [odemi.DiscreteCallback(_dose_cond, _dose),
           odemi.DiscreteCallback(_reset_cond, _reset_stop, save_positions=(False, True))]
    tr = trace_problem(_decay, 2, 2, callbacks=cbs, isout=lambda u, p, t: u[1] < 0.0)
    L = _lib.load()
    for flags in (_lib.COMPILE_TSIT5, _lib.COMPILE_RODAS5P):
        path = ctypes.create_string_buffer(512)
        log = ctypes.create_string_buffer(1 << 16)
        secs = ctypes.c_double()
        rc = L.odemi_rtc_codeobject(tr.src.encode(), 2, 2, flags, path, 512, ctypes.byref(secs), log, 1 << 16)
        assert rc == 0, log.value.decode()[:2000]


def test_vector_callbacks_in_workgroup_kernels_emit_and_compile_and_unknown_types():
    """A VectorContinuousCallback traced for the workgroup kernels (n = 8 and
    10): its condition per state size writes out[0:len], its affect reads the
    event directions ev[i] and returns its conditional terminate!; the source
    compiles for gfx950 (hipRTC, no device).  Unknown callback types and
    Python branches on traced values are refused."""
    def vc(out, u, t, itg):
        out[0] = u[0]
        out[1] = u[1] - 0.5

    def va(itg, ev):
        m = len(itg.u) // 2
        itg.u[m] = odemi.sym_ifelse(ev[0] == -1, -itg.u[m], itg.u[m])
        odemi.terminate(itg, when=itg.t > 2.0)
    cb = odemi.VectorContinuousCallback(vc, va, 2)

    def big(du, u, p, t):
        for i in range(len(u)):
            du[i] = -p[0] * u[i]
    tr, wg = trace_for(odemi.ODEFunction(big), 10, 1, algs=("rodas5p",), ns=[8, 10], callback=cb)
    assert wg and "#define ODEMI_VCB_MAXLEN_WG 2" in tr.src and "vlen = 2" in tr.src
    vc_src = tr.src[tr.src.index("void vcond("):tr.src.index("bool vaffect(")]
    assert "case 8:" in vc_src and "case 10:" in vc_src and "out[1] =" in vc_src
    va_src = tr.src[tr.src.index("bool vaffect("):]
    assert "ev[0]" in va_src and "u[4] = " in va_src and "u[5] = " in va_src and "return nu_[8] != T(0);" in va_src
    path, _ = _lib.precompile(tr.src, 10, 1, _lib.COMPILE_WG)
    assert os.path.getsize(path) > 0
    with pytest.raises(TypeError):
        odemi.compile_problem(odemi.ODEFunction(_decay), 2, 2, algs=("tsit5",), callback=object())
    with pytest.raises(TypeError):
        trace_problem(_decay, 2, 2, callbacks=[odemi.DiscreteCallback(lambda u, t, i: u[0] if u[0] > 1 else 0,
                                                                     _dose)])


def test_continuous_callbacks_emit_and_compile():
    def fall(du, u, p, t):
        du[0] = u[1]
        du[1] = -p[0]

    def bounce(itg):
        itg.u[1] = -p_e(itg) * itg.u[1]

    def p_e(itg):
        return itg.p[1]
    cbs = odemi.CallbackSet(odemi.ContinuousCallback(lambda u, t, i: u[0], None, bounce, interp_points=0),
                            odemi.ContinuousCallback(lambda u, t, i: t - 7.5, lambda i: odemi.terminate(i),
                                                     rootfind="right", save_positions=(False, True)),
                            odemi.DiscreteCallback(_dose_cond, _dose))
    assert len(cbs.continuous_callbacks) == 2 and len(cbs.discrete_callbacks) == 1
    tr = trace_problem(fall, 2, 2, callbacks=cbs.continuous_callbacks + cbs.discrete_callbacks)
    src = tr.src
    assert "#define ODEMI_NCCB 2" in src and "#define ODEMI_NDCB 1" in src
    c0 = src[src.index("struct OdemiCcb<0>"):src.index("struct OdemiCcb<1>")]
    assert "has_affect = 0, has_affect_neg = 1" in c0 and "interp_points = 0" in c0 and "affect_neg(" in c0
    c1 = src[src.index("struct OdemiCcb<1>"):]
    assert "rootfind = 1" in c1 and "term = 1, term_neg = 1" in c1 and "save0 = 0, save1 = 1" in c1
    L = _lib.load()
    path = ctypes.create_string_buffer(512)
    log = ctypes.create_string_buffer(1 << 16)
    secs = ctypes.c_double()
    rc = L.odemi_rtc_codeobject(tr.src.encode(), 2, 2, _lib.COMPILE_TSIT5, path, 512, ctypes.byref(secs), log,
                                1 << 16)
    assert rc == 0, log.value.decode()[:2000]
    rc = L.odemi_rtc_codeobject(tr.src.encode(), 2, 2, _lib.COMPILE_RODAS5P, path, 512, ctypes.byref(secs), log,
                                1 << 16)
    assert rc == 0, log.value.decode()[:2000]
    with pytest.raises(ValueError):
        odemi.ContinuousCallback(lambda u, t, i: u[0], bounce, interp_points=1)


def test_vector_continuous_callback_emits_and_compiles():
    def box(du, u, p, t):
        du[0] = u[1]
        du[1] = -9.81
        du[2] = u[3]
        du[3] = 0.0

    def cond(out, u, t, itg):
        out[0] = u[0]
        out[1] = (u[2] - 10.0) * u[2]

    def aff(itg, ev):
        itg.u[1] = odemi.sym_ifelse(ev[0] == -1, -itg.u[1], itg.u[1])
        odemi.terminate(itg, when=ev[1] != 0)
    tr = trace_problem(box, 4, 1, callbacks=[odemi.VectorContinuousCallback(cond, aff, 2)])
    src = tr.src
    assert "#define ODEMI_VCB_MAXLEN 2" in src and "vlen = 2" in src
    assert "ev[0] == T(-1.0)" in src and "return nu_[4] != T(0);" in src and "u[1] = nu_[1];" in src
    L = _lib.load()
    path = ctypes.create_string_buffer(512)
    log = ctypes.create_string_buffer(1 << 16)
    secs = ctypes.c_double()
    rc = L.odemi_rtc_codeobject(src.encode(), 4, 1, _lib.COMPILE_TSIT5, path, 512, ctypes.byref(secs), log, 1 << 16)
    assert rc == 0, log.value.decode()[:2000]
    with pytest.raises(TypeError, match="conditional terminate"):
        trace_problem(box, 4, 1, callbacks=[odemi.DiscreteCallback(
            lambda u, t, i: t == 1.0, lambda i: odemi.terminate(i, when=i.u[0] > 1.0))])
    with pytest.raises(ValueError):
        odemi.VectorContinuousCallback(cond, aff, 17)


def _window(hi):
    return lambda u, p, t: (t > 0.3) & (t < hi)


def test_trace_cache_never_reuses_a_freed_callables_trace():
    """Predicates made one after another (each dropped by the caller before the
    next exists, so CPython may hand the next one the same address) each get
    their own trace: the cache keys on the callable, not on its id()."""
    from odemi import problems as P
    from odemi.api import trace_for
    his = [0.31, 0.3001, 0.30001, 0.3000001]
    srcs = []
    for hi in his:
        tr, wg = trace_for(P.CHAIN, 8, 3, algs=("rodas5p",), ns=[8], isoutofdomain=_window(hi))
        assert wg
        srcs.append(tr.src)
    for hi, src in zip(his, srcs):
        assert f"T({float(hi).hex()})" in src, hi
    assert len(set(srcs)) == len(his)
    # the same object is still a cache hit
    f = _window(0.5)
    a, _ = trace_for(P.CHAIN, 8, 3, algs=("rodas5p",), ns=[8], isoutofdomain=f)
    b, _ = trace_for(P.CHAIN, 8, 3, algs=("rodas5p",), ns=[8], isoutofdomain=f)
    assert a is b


def test_callback_trace_cache_keys_on_the_callables():
    def make(level):
        return odemi.DiscreteCallback(lambda u, t, itg: u[0] >= level, _reset_stop)
    srcs = [trace_for_cb(make(lv)) for lv in (19.0, 18.5, 18.25)]
    assert len(set(srcs)) == 3


def trace_for_cb(cb):
    from odemi.api import trace_for
    tr, _ = trace_for(odemi.ODEFunction(_decay), 2, 2, algs=("tsit5",), callback=cb)
    return tr.src


def test_codeobjects_compile_concurrently_from_two_threads(tmp_path):
    """odemi_rtc_codeobject from two host threads at once (distinct problems):
    both succeed and produce distinct cache files (the compile log buffer is
    per call)."""
    import threading
    env_old = os.environ.get("ODEMI_CACHE_DIR")
    L = _lib.load()
    srcs = [trace_problem(_decay, 2, 2).src,
            trace_problem(_decay, 2, 2, isout=lambda u, p, t: u[0] < 0.0).src]
    res = [None, None]

    def work(i):
        path = ctypes.create_string_buffer(4096)
        log = ctypes.create_string_buffer(65536)
        secs = ctypes.c_double()
        rc = L.odemi_rtc_codeobject(srcs[i].encode(), 2, 2, _lib.COMPILE_TSIT5, path, 4096, ctypes.byref(secs),
                                    log, 65536)
        res[i] = (rc, path.value.decode(), log.value.decode(errors="replace"))

    os.environ["ODEMI_CACHE_DIR"] = str(tmp_path)
    try:
        th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    finally:
        if env_old is None:
            os.environ.pop("ODEMI_CACHE_DIR", None)
        else:
            os.environ["ODEMI_CACHE_DIR"] = env_old
    assert res[0][0] == 0 and res[1][0] == 0, res
    assert res[0][1] != res[1][1]
    assert all(os.path.getsize(r[1]) > 0 for r in res)


def test_rtc_version_reported():
    major, minor = _lib.rtc_version()
    assert major >= 1 and minor >= 0


def _kick(itg):
    m = len(itg.u) // 2   # the first mass's velocity, for every chain length
    itg.u[m] = itg.u[m] + itg.p[1]


def test_workgroup_callbacks_emit_per_size_and_compile():
    """Discrete and continuous callbacks of an uneven batch are traced once per
    state size into OdemiDcbWg / OdemiCcbWg (a switch on n), with the affect's
    index following n, and compile with the workgroup kernels."""
    from odemi import problems as P
    from odemi.api import trace_for
    cbs = odemi.CallbackSet(
        odemi.ContinuousCallback(lambda u, t, itg: u[0], None, affect_neg=lambda itg: odemi.terminate(itg)),
        odemi.DiscreteCallback(lambda u, t, itg: odemi.sym_in(t, [0.25]), _kick))
    tr, wg = trace_for(P.CHAIN, 12, 3, algs=("rodas5p",), ns=[8, 12], callback=cbs)
    assert wg
    src = tr.src
    assert "#define ODEMI_NCCB_WG 1" in src and "#define ODEMI_NDCB_WG 1" in src
    assert "term = 0, term_neg = 1" in src
    assert "u[4] = nu_[4];" in src and "u[6] = nu_[6];" in src   # m = n / 2 for n = 8 and 12
    L = _lib.load()
    path = ctypes.create_string_buffer(512)
    log = ctypes.create_string_buffer(1 << 16)
    secs = ctypes.c_double()
    rc = L.odemi_rtc_codeobject(src.encode(), 12, 3, _lib.COMPILE_WG, path, 512, ctypes.byref(secs), log, 1 << 16)
    assert rc == 0, log.value.decode()[:3000]


def test_mixed_precision_advice_for_tsit5_continuous_callbacks():
    """Host check: the mixed engine is flagged for Tsit5 with a continuous
    callback (FP32 stage rounding floors the error estimate near events), and
    not for FP64, for Rodas5P, or for discrete callbacks alone."""
    from odemi import api
    cond = lambda u, t, integ: u[0]  # noqa: E731

    def affect(integ):
        integ.u[1] = -integ.u[1]

    cc = api.ContinuousCallback(cond, affect)
    dc = api.DiscreteCallback(lambda u, t, integ: t > 1.0, affect)
    mixed, fp64 = api.EnsembleMI355X(precision="mixed"), api.EnsembleMI355X(precision="fp64")
    assert "precision='fp64'" in api._mixed_precision_advice(mixed, "tsit5", cc)
    assert api._mixed_precision_advice(mixed, "tsit5", api.CallbackSet(dc, cc)) is not None
    assert api._mixed_precision_advice(fp64, "tsit5", cc) is None
    assert api._mixed_precision_advice(mixed, "rodas5p", cc) is None
    assert api._mixed_precision_advice(mixed, "tsit5", dc) is None
    assert api._mixed_precision_advice(mixed, "tsit5", None) is None
