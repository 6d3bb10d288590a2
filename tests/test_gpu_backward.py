"""Backward integration (tf < t0) on the GPU engines against the CPU
restatement, which is pinned to the forward path by time reversal
(tests/test_oracle.py): tdir = sign(tf - t0) in initdt, the dt bounds, the
tstop and saveat cursors and the loop condition (solve.jl:268-273,401,904-946,
968-985,1019-1123; integrator_utils.jl:268-324,340-414,1199-1256,1290-1314).
Forward and backward trajectories share one launch."""
import numpy as np
import pytest

import oracle as O
import odemi
from odemi import problems as P

pytestmark = pytest.mark.gpu


def _lorenz_mixed_directions(N, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    u0 = rng.uniform(-1.0, 1.0, (3, N))
    p = np.stack([rng.uniform(8.0, 12.0, N), rng.uniform(0.0, 28.0, N), np.full(N, 8.0 / 3.0)])
    t0 = rng.uniform(-0.2, 0.2, N)
    span = rng.uniform(0.3, 0.6, N)
    back = np.arange(N) % 2 == 1
    tf = np.where(back, t0 - span, t0 + span)
    return u0, p, t0, tf, back


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_backward_and_forward_in_one_batch_bit_identical(alg, engine):
    """Lorenz, half the trajectories integrated backward, with a shared saveat
    list and shared tstops on both sides of t0: saves, final states, retcodes
    and step counts bit-identical to the restatement."""
    N = 192
    u0, p, t0, tf, back = _lorenz_mixed_directions(N, 51)
    sa = [-0.5, -0.31, -0.1, 0.05, 0.25, 0.4]
    ts = [-0.35, -0.02, 0.12, 0.33]
    kw = dict(reltol=1e-7, abstol=1e-7, saveat=sa, tstops=ts, save_start=True, save_end=True)
    print(f"[case] backward+forward lorenz {alg.code} {engine}", flush=True)
    out = odemi.solve_arrays(P.LORENZ, alg, u0, p, t0, tf, odemi.EnsembleMI355X(engine=engine), max_saves=16,
                             **kw)
    for j in range(N):
        r = O.solve_builtin(O.PROB_LORENZ, u0[:, j], p[:, j], (t0[j], tf[j]), alg=alg.code, **kw)
        ns = int(out["nsaved"][j])
        assert int(out["retcode"][j]) == r["retcode"] == 1 and ns == r["nsaved"], j
        assert np.array_equal(out["t_out"][:ns, j], r["t"]) and np.array_equal(out["u_out"][:ns, :, j], r["u"]), j
        assert np.array_equal(out["u_final"][:, j], r["u_final"]) and out["t_final"][j] == tf[j], j
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2]) == (r["nf"], r["naccept"], r["nreject"]), j
        if back[j]:
            assert np.all(np.diff(r["t"]) < 0)   # saves in the integration direction


def _rotation(du, u, p, t):
    du[0] = p[0] * u[1]
    du[1] = -p[0] * u[0]


def test_backward_mixed_engines_within_tolerance():
    """The mixed engines (FP32 Tsit5 stages; FP32 LU + FP64 refinement) on a
    rotation u' = w (u1, -u0), as well conditioned backward as forward (the
    dissipative Lorenz flow amplifies every difference by e^(13.7 |t|)
    backward): final states within a few tolerance units of the FP64 engine's,
    half the batch backward."""
    N = 512
    rng = np.random.Generator(np.random.PCG64(52))
    u0 = rng.uniform(-1.0, 1.0, (2, N))
    p = rng.uniform(0.5, 3.0, (1, N))
    t0 = np.zeros(N)
    tf = np.where(np.arange(N) % 2 == 1, -1.0, 1.0) * rng.uniform(2.0, 6.0, N)
    fun = odemi.ODEFunction(_rotation)
    for alg in (odemi.Tsit5(), odemi.Rodas5P()):
        print(f"[case] backward mixed rotation {alg.code}", flush=True)
        kw = dict(reltol=1e-6, abstol=1e-6)
        o64 = odemi.solve_arrays(fun, alg, u0, p, t0, tf, **kw)
        omx = odemi.solve_arrays(fun, alg, u0, p, t0, tf, odemi.EnsembleMI355X(precision="mixed"), **kw)
        assert (omx["retcode"] == 1).all() and np.array_equal(omx["t_final"], tf)
        err = np.abs(omx["u_final"] - o64["u_final"]) / (1e-6 + 1e-6 * np.abs(o64["u_final"]))
        e = err.max(axis=0)
        print(f"scaled err p99 forward {np.quantile(e[::2], 0.99):.3e} backward {np.quantile(e[1::2], 0.99):.3e}")
        assert np.quantile(e, 0.99) <= 10.0


def test_backward_uneven_chains_workgroup_bit_identical():
    """Uneven chains (n = 8..24) integrated from 0.6 back to 0 on the FP64
    workgroup engine, with saveat: bit-identical to the restatement."""
    d = P.chain_ensemble(N=64, seed=53, m_lo=4, m_hi=12)
    N = 64
    t0, tf = np.full(N, 0.6), np.zeros(N)
    sa = [0.1, 0.45]
    print("[case] backward chains wg fp64", flush=True)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], t0, tf, reltol=1e-6, abstol=1e-6,
                             nvec=d["nvec"], saveat=sa, save_start=False, save_end=True)
    for j in range(N):
        n = int(d["nvec"][j])
        r = O.solve_builtin(O.PROB_CHAIN, d["u0"][:n, j], d["p"][:, j], (0.6, 0.0), alg=O.RODAS5P, reltol=1e-6,
                            abstol=1e-6, saveat=sa, save_start=False, save_end=True)
        ns = int(out["nsaved"][j])
        assert int(out["retcode"][j]) == r["retcode"] == 1 and ns == r["nsaved"] == 3, j
        assert np.array_equal(out["t_out"][:ns, j], r["t"]) and np.array_equal(out["u_out"][:ns, :n, j], r["u"]), j
        assert np.array_equal(out["u_final"][:n, j], r["u_final"]), j
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2]) == (r["nf"], r["naccept"], r["nreject"]), j
