"""odemi_solve_batch_noalloc (the allocation-free entry point) against
odemi_solve_batch on the same device-resident inputs, through the C ABI:
bit-identical outputs on two calls in a row, the second with a smaller batch,
for the lane kernels (Tsit5 FP64 and mixed, Rodas5P) and the workgroup kernels
(uneven chains, FP64 LU and mixed Gauss-Jordan, whose Jacobian blocks live in
the scratch array); and its refusals (host pointers, the plain engine, a solve
whose scratch exceeds the fixed array)."""
import ctypes

import numpy as np
import pytest

import odemi
from odemi import _lib
from odemi import problems as P

pytestmark = pytest.mark.gpu


def _dev(d, cols):
    import torch
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")  # noqa: E731
    return T(d["u0"][:, cols]), T(d["p"][:, cols]), T(d["t0"][cols]), T(d["tf"][cols])


def _pair(fun, alg, d, cols, prec, **kw):
    import torch
    args = _dev(d, cols)
    ref = odemi.solve_arrays(fun, alg, *args, odemi.EnsembleMI355X(precision=prec), **kw)
    new = odemi.solve_arrays(fun, alg, *args, odemi.EnsembleMI355X(precision=prec, allocate=False), **kw)
    torch.cuda.synchronize()
    return ({k: v.cpu().numpy() for k, v in ref.items() if hasattr(v, "cpu")},
            {k: v.cpu().numpy() for k, v in new.items() if hasattr(v, "cpu")})


def _same(ref, new, nvec=None):
    for k in ("u_final", "t_final", "retcode", "stats", "nsaved"):
        assert np.array_equal(ref[k], new[k]), k
    m = int(ref["nsaved"].max()) if ref["nsaved"].size else 0
    assert np.array_equal(ref["t_out"][:m], new["t_out"][:m])
    if nvec is None:
        assert np.array_equal(ref["u_out"][:m], new["u_out"][:m])
    else:   # rows past a trajectory's own n are not written
        for j, n in enumerate(nvec):
            assert np.array_equal(ref["u_out"][:m, :n, j], new["u_out"][:m, :n, j]), j


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_lorenz_equals_solve_batch_twice(prec):
    d = P.lorenz_large(N=8192)
    for N in (8192, 1000):
        print(f"[case] noalloc lorenz tsit5 {prec} N={N}", flush=True)
        ref, new = _pair(P.LORENZ, odemi.Tsit5(), d, np.arange(N), prec, reltol=d["reltol"], abstol=d["abstol"])
        assert (ref["retcode"] == 1).all()
        _same(ref, new)


def test_noalloc_robertson_saveat_equals_solve_batch_twice():
    d = P.robertson_ensemble(N=2048)
    sa = np.array([1e-3, 1.0, 1e3])
    for N in (2048, 333):
        print(f"[case] noalloc robertson rodas5p fp64 saveat N={N}", flush=True)
        ref, new = _pair(P.ROBERTSON, odemi.Rodas5P(), d, np.arange(N), "fp64", reltol=d["reltol"],
                         abstol=d["abstol"], saveat=sa)
        assert (ref["retcode"] == 1).all()
        _same(ref, new)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_uneven_chains_equals_solve_batch_twice(prec):
    d = P.chain_ensemble(N=96, seed=7, m_lo=4, m_hi=64)
    d["tf"] = np.full(96, 0.5)
    for N in (96, 37):
        cols = np.arange(N)
        print(f"[case] noalloc chains rodas5p wg {prec} N={N}", flush=True)
        ref, new = _pair(P.CHAIN, odemi.Rodas5P(), d, cols, prec, reltol=d["reltol"], abstol=d["abstol"],
                         nvec=d["nvec"][cols])
        assert (ref["retcode"] == 1).all()
        _same(ref, new, nvec=d["nvec"][cols])


def test_noalloc_refusals():
    import torch
    ctx = odemi.get_context(0)
    h, _, _ = odemi.compile_problem(P.LORENZ, 3, 3, algs=("tsit5",), ctx=ctx)

    def args(N, dev=True, engine=_lib.ENGINE_PERSISTENT):
        a = ctx.default_args(_lib.TSIT5)
        a.engine, a.N, a.nmax, a.np = engine, N, 3, 3
        a.reltol = a.abstol = 1e-6
        a.save_end, a.max_saves = 1, 1
        bufs = [torch.zeros((k, N), dtype=torch.float64, device="cuda:0") for k in (3, 3, 1, 1, 1, 3, 1)]
        bufs[3] += 1.0
        rc = torch.zeros(N, dtype=torch.int32, device="cuda:0")
        ns = torch.zeros(N, dtype=torch.int32, device="cuda:0")
        st = torch.zeros((_lib.NSTATS, N), dtype=torch.int32, device="cuda:0")
        uo = torch.zeros((1, 3, N), dtype=torch.float64, device="cuda:0")
        a.u0, a.p, a.t0, a.tf, a.t_out, a.u_final, a.t_final = (b.data_ptr() for b in bufs)
        a.u_out, a.retcode, a.nsaved, a.stats = uo.data_ptr(), rc.data_ptr(), ns.data_ptr(), st.data_ptr()
        a.device_ptrs = int(dev)
        a.sort_by_cost = 1
        return a, (bufs, rc, ns, st, uo)

    L = ctx.L
    print("[case] noalloc refusals", flush=True)
    a, keep = args(1 << 22)   # order 16 MiB + start-up buffers 128 MiB: above the 64 MiB array
    assert L.odemi_solve_batch_noalloc(ctx.h, ctypes.c_void_p(h), ctypes.byref(a)) == -2
    assert "bytes of scratch" in ctx.error()
    a, keep = args(64, dev=False)
    assert L.odemi_solve_batch_noalloc(ctx.h, ctypes.c_void_p(h), ctypes.byref(a)) == -2
    assert "device pointers" in ctx.error()
    a, keep = args(64, engine=_lib.ENGINE_PLAIN)
    assert L.odemi_solve_batch_noalloc(ctx.h, ctypes.c_void_p(h), ctypes.byref(a)) == -2
    assert "persistent" in ctx.error()
    torch.cuda.synchronize()
