"""GPU tests of the solve surface (through libodemi.so's C ABI) against the CPU
restatement: saveat / save_start / save_end / save_everystep, dense output of
both methods, return codes, stats, user dt, and the EnsembleProblem hooks."""
import math

import numpy as np
import pytest

import oracle as O
import odemi
from odemi import problems as P

pytestmark = pytest.mark.gpu


def _f2(du, u, p, t):
    du[0] = -odemi.sym_cos(u[0]) * u[0]


def _f2_py(du, u, p, t):
    du[0] = -math.cos(u[0]) * u[0]


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()])
def test_saveat_rules_match_reference_lists(alg):
    """test/InterfaceI/ode_saveat_tests.jl:246-267 on the GPU engine."""
    print(f"[case] saveat rules {alg}", flush=True)
    prob = odemi.ODEProblem(_f2, [10.0], (0.0, 0.4))
    sa = [0.0, 0.1, 0.2, 0.3, 0.4]
    assert list(odemi.solve(prob, alg, saveat=sa).t) == sa
    assert list(odemi.solve(prob, alg, saveat=sa, save_start=True, save_end=True).t) == sa
    assert list(odemi.solve(prob, alg, saveat=sa, save_start=False, save_end=False).t) == [0.1, 0.2, 0.3]
    assert list(odemi.solve(prob, alg, saveat=[0.2]).t) == [0.2]
    assert list(odemi.solve(prob, alg, saveat=[0.2], save_start=True, save_end=True).t) == [0.0, 0.2, 0.4]
    ts = odemi.solve(prob, alg).t
    assert 0.0 in ts and 0.4 in ts
    ts = odemi.solve(prob, alg, save_start=False, save_end=False).t
    assert 0.0 not in ts and 0.4 not in ts
    # values agree with the oracle (cos differs by at most an ulp between libms)
    sol = odemi.solve(prob, alg, saveat=sa)
    ref = O.solve(_f2_py, [10.0], [0.0], (0.0, 0.4), alg=alg.code, saveat=sa, save_start=True, save_end=True,
                  jac=lambda J, u, p, t: J.__setitem__((0, 0), math.sin(u[0]) * u[0] - math.cos(u[0])))
    assert np.allclose(np.array(sol.u)[:, 0], ref["u"][:, 0], rtol=1e-9, atol=0)


@pytest.mark.parametrize("alg,prob_id,fun,dgen", [
    (odemi.Tsit5(), O.PROB_LORENZ, P.LORENZ, lambda: P.lorenz_large(N=512)),
    (odemi.Rodas5P(), O.PROB_ROBERTSON, P.ROBERTSON, lambda: P.robertson_ensemble(N=256)),
])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_ensemble_saveat_bit_identical(alg, prob_id, fun, dgen, engine):
    d = dgen()
    N = d["u0"].shape[1]
    sa = np.linspace(0.05, 0.95, 7) * (1.0 if alg.code == O.TSIT5 else 1e4)
    print(f"[case] ensemble saveat {alg} {engine} N={N}", flush=True)
    out = odemi.solve_arrays(fun, alg, d["u0"], d["p"], d["t0"], d["tf"], odemi.EnsembleMI355X(engine=engine),
                             reltol=d["reltol"], abstol=d["abstol"], saveat=sa, save_start=True, save_end=True)
    for j in range(0, N, 37):
        r = O.solve_builtin(prob_id, d["u0"][:, j], d["p"][:, j], (d["t0"][j], d["tf"][j]), alg=alg.code,
                            reltol=d["reltol"], abstol=d["abstol"], saveat=sa, save_start=True, save_end=True)
        ns = int(out["nsaved"][j])
        assert ns == r["nsaved"], (j, ns, r["nsaved"])
        assert np.array_equal(out["t_out"][:ns, j], r["t"])
        assert np.array_equal(out["u_out"][:ns, :, j], r["u"])
        assert int(out["retcode"][j]) == r["retcode"]


@pytest.mark.parametrize("alg,prob_id,fun,dgen", [
    (odemi.Tsit5(), O.PROB_LORENZ, P.LORENZ, lambda: P.lorenz_small(N=64)),
    (odemi.Rodas5P(), O.PROB_ROBERTSON, P.ROBERTSON, lambda: P.robertson_ensemble(N=64)),
])
def test_save_everystep_bit_identical(alg, prob_id, fun, dgen):
    d = dgen()
    print(f"[case] save_everystep {alg}", flush=True)
    out = odemi.solve_arrays(fun, alg, d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"], abstol=d["abstol"],
                             save_everystep=True, save_start=True, save_end=True, max_saves=2048)
    for j in range(0, 64, 9):
        r = O.solve_builtin(prob_id, d["u0"][:, j], d["p"][:, j], (d["t0"][j], d["tf"][j]), alg=alg.code,
                            reltol=d["reltol"], abstol=d["abstol"], save_everystep=True, max_out=2048)
        ns = int(out["nsaved"][j])
        assert ns == r["nsaved"]
        assert np.array_equal(out["t_out"][:ns, j], r["t"])
        assert np.array_equal(out["u_out"][:ns, :, j], r["u"])


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_saveat_inside_final_clipped_step(alg, engine):
    """Saveat points inside the last step, which modify_dt_for_tstops! clips to
    tf: the accept restores dt = dtpropose, and the dtpropose the reference's two
    clips (apply_step! and the loop header) leave is the clipped dt, so the
    interpolation is the step's own (R10).  Bit-identical to the CPU restatement."""
    d = P.lorenz_small(N=64)
    sa = [0.97, 0.99, 0.999, 0.9999, 0.99999]
    print(f"[case] saveat in final step {alg.code} {engine}", flush=True)
    out = odemi.solve_arrays(P.LORENZ, alg, d["u0"], d["p"], d["t0"], d["tf"], odemi.EnsembleMI355X(engine=engine),
                             reltol=d["reltol"], abstol=d["abstol"], saveat=sa, save_start=False, save_end=True)
    for j in range(0, 64, 3):
        r = O.solve_builtin(O.PROB_LORENZ, d["u0"][:, j], d["p"][:, j], (0.0, 1.0), alg=alg.code,
                            reltol=d["reltol"], abstol=d["abstol"], saveat=sa, save_start=False, save_end=True)
        ns = int(out["nsaved"][j])
        assert ns == r["nsaved"] == len(sa) + 1
        assert np.array_equal(out["t_out"][:ns, j], r["t"])
        assert np.array_equal(out["u_out"][:ns, :, j], r["u"])


@pytest.mark.parametrize("alg,prob_id,fun,dgen,ts", [
    (odemi.Tsit5(), O.PROB_LORENZ, P.LORENZ, lambda: P.lorenz_small(N=128), [0.13, 0.5, 0.5, 0.77, 0.0, 2.0]),
    (odemi.Rodas5P(), O.PROB_ROBERTSON, P.ROBERTSON, lambda: P.robertson_ensemble(N=128),
     [1e-3, 0.37, 40.0, 40.0, 5e3, 1e6]),
], ids=["tsit5-lorenz", "rodas5p-robertson"])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_tstops_bit_identical(alg, prob_id, fun, dgen, ts, engine):
    """User tstops (shared list; times outside (t0, tf) and duplicates dropped):
    every step time, state and counter equals the CPU restatement, which runs
    the reference's nested tstops loop (solve.jl:904-927), and every tstop
    inside the span is a step time."""
    d = dgen()
    N = d["u0"].shape[1]
    print(f"[case] tstops {alg.code} {engine} N={N}", flush=True)
    out = odemi.solve_arrays(fun, alg, d["u0"], d["p"], d["t0"], d["tf"], odemi.EnsembleMI355X(engine=engine),
                             reltol=d["reltol"], abstol=d["abstol"], save_everystep=True, save_start=True,
                             save_end=True, max_saves=4096, tstops=ts)
    for j in range(0, N, 11):
        r = O.solve_builtin(prob_id, d["u0"][:, j], d["p"][:, j], (d["t0"][j], d["tf"][j]), alg=alg.code,
                            reltol=d["reltol"], abstol=d["abstol"], save_everystep=True, tstops=ts)
        ns = int(out["nsaved"][j])
        assert ns == r["nsaved"], (j, ns, r["nsaved"])
        assert np.array_equal(out["t_out"][:ns, j], r["t"])
        assert np.array_equal(out["u_out"][:ns, :, j], r["u"])
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2]) == (r["nf"], r["naccept"], r["nreject"])
        for s in ts:
            if d["t0"][j] < s < d["tf"][j]:
                assert s in out["t_out"][:ns, j]


def test_tstops_workgroup_and_mixed():
    """tstops through the workgroup kernel (uneven-n chains, FP64 bit-identical
    to the CPU restatement) and the mixed lane engine (tstops hit exactly,
    final states within the tolerance scale of FP64)."""
    ts = [0.05, 0.2, 0.2, 0.61]
    d = P.chain_ensemble(N=24, seed=5, m_lo=2, m_hi=12)
    d["tf"] = np.full(24, 1.0)
    print("[case] tstops workgroup chains", flush=True)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"], save_everystep=True, save_start=True,
                             max_saves=4096, tstops=ts)
    for j in range(0, 24, 5):
        n = int(d["nvec"][j])
        r = O.solve_builtin(O.PROB_CHAIN, d["u0"][:n, j], d["p"][:, j], (0.0, 1.0), alg=O.RODAS5P,
                            reltol=d["reltol"], abstol=d["abstol"], save_everystep=True, tstops=ts)
        ns = int(out["nsaved"][j])
        assert ns == r["nsaved"]
        assert np.array_equal(out["t_out"][:ns, j], r["t"])
        assert np.array_equal(out["u_out"][:ns, :n, j], r["u"])
        assert all(s in out["t_out"][:ns, j] for s in ts)
    d = P.lorenz_large(N=2048)
    print("[case] tstops mixed lane", flush=True)
    kw = dict(reltol=1e-6, abstol=1e-6, save_everystep=True, save_start=True, max_saves=4096, tstops=ts)
    o64 = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"], **kw)
    om = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), **kw)
    assert (om["retcode"] == 1).all()
    for j in range(0, 2048, 97):
        ns = int(om["nsaved"][j])
        assert all(s in om["t_out"][:ns, j] for s in ts)
    uf64 = np.stack([o64["u_out"][int(o64["nsaved"][j]) - 1, :, j] for j in range(2048)], axis=1)
    ufm = np.stack([om["u_out"][int(om["nsaved"][j]) - 1, :, j] for j in range(2048)], axis=1)
    err = np.abs(ufm - uf64) / (1e-6 + 1e-6 * np.abs(uf64))
    print(f"tstops mixed scaled err median {np.median(err):.3e}")
    assert np.median(err.max(axis=0)) < 10.0


def _lin(du, u, p, t):
    du[0] = p[0] * u[0]


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_isoutofdomain_bit_identical(alg, engine):
    """isoutofdomain through the lane engines, bit-identical to the CPU
    restatement: (a) decay u' = lambda u with a loose abstol and u < 0 flagged
    (Rodas5P leaves the domain without it), (b) Lorenz with y > 8 flagged, where
    trajectories that would cross stall with the reference's Unstable code."""
    N = 96
    rng = np.random.default_rng(7)
    lam = rng.uniform(-50.0, -10.0, N)
    print(f"[case] isoutofdomain {alg.code} {engine}", flush=True)
    ea = odemi.EnsembleMI355X(engine=engine)
    out = odemi.solve_arrays(_lin, alg, np.ones((1, N)), lam[None, :], np.zeros(N), np.full(N, 5.0), ea,
                             reltol=1e-3, abstol=0.1, save_everystep=True, save_start=True, max_saves=4096,
                             isoutofdomain=lambda u, p, t: u[0] < 0.0)
    for j in range(0, N, 7):
        r = O.solve(lambda du, u, p, t: du.__setitem__(0, p[0] * u[0]), [1.0], [lam[j]], (0.0, 5.0),
                    jac=lambda J, u, p, t: J.__setitem__((0, 0), p[0]), alg=alg.code, reltol=1e-3,
                    abstol=0.1, save_everystep=True, isoutofdomain=lambda u, p, t: u[0] < 0.0)
        ns = int(out["nsaved"][j])
        assert ns == r["nsaved"] and int(out["retcode"][j]) == r["retcode"] == 1
        assert np.array_equal(out["t_out"][:ns, j], r["t"]) and np.array_equal(out["u_out"][:ns, 0, j], r["u"][:, 0])
        assert (out["u_out"][:ns, 0, j] >= 0).all()
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2]) == (r["nf"], r["naccept"], r["nreject"])
    d = P.lorenz_small(N=64)
    out = odemi.solve_arrays(P.LORENZ, alg, d["u0"], d["p"], d["t0"], d["tf"], ea, reltol=1e-6, abstol=1e-6,
                             isoutofdomain=lambda u, p, t: u[1] > 8.0)
    codes = set()
    for j in range(0, 64, 3):
        r = O.solve_builtin(O.PROB_LORENZ, d["u0"][:, j], d["p"][:, j], (0.0, 1.0), alg=alg.code, reltol=1e-6,
                            abstol=1e-6, save_start=False, isoutofdomain=lambda u, p, t: u[1] > 8.0)
        assert int(out["retcode"][j]) == r["retcode"]
        assert np.array_equal(out["u_final"][:, j], r["u_final"])
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2]) == (r["nf"], r["naccept"], r["nreject"])
        codes.add(r["retcode"])
    assert codes == {1, 7}   # some trajectories cross y = 8 and stall, some never reach it


def _pk(du, u, p, t):
    du[0] = -p[0] * u[0]
    du[1] = p[0] * u[0] - p[1] * u[1]


def _pk_jac(J, u, p, t):
    J[0, 0] = -p[0]
    J[0, 1] = 0.0
    J[1, 0] = p[0]
    J[1, 1] = -p[1]


_DOSES = [2.0, 4.0, 6.0]


def _dose_cond(u, t, itg):
    return odemi.sym_in(t, _DOSES)


def _dose(itg):
    itg.u[0] = itg.u[0] + itg.p[2]


def _cap_cond(u, t, itg):
    return u[1] > itg.p[3]


def _stop(itg):
    odemi.terminate(itg)


_PK_CBS = odemi.CallbackSet(odemi.DiscreteCallback(_dose_cond, _dose),
                            odemi.DiscreteCallback(_cap_cond, _stop, save_positions=(False, True)))


def _pk_batch(N, seed=11):
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(0.5, 2.0, N), rng.uniform(0.1, 0.5, N), rng.uniform(1.0, 3.0, N),
                  rng.uniform(1.5, 4.0, N)])
    u0 = np.stack([p[2].copy(), np.zeros(N)])
    return u0, p


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_discrete_callbacks_bit_identical(alg, engine):
    """A CallbackSet of two DiscreteCallbacks (lib/DiffEqBase/src/callbacks.jl:
    645-683) through the lane engines, bit-identical to the CPU restatement:
    two-compartment dosing, u[0] += dose at the tstops 2, 4, 6 (saved before
    and after), and terminate! once u[1] exceeds a per-trajectory cap (saved
    after), with save_everystep, with saveat, and final state only (forced
    saves still land in the outputs).  Retcodes (Success / Terminated), every
    saved point and nf / naccept / nreject / ncondition agree."""
    N = 96
    u0, p = _pk_batch(N)
    fun = odemi.ODEFunction(_pk, jac=_pk_jac)
    ea = odemi.EnsembleMI355X(engine=engine)
    modes = [dict(save_everystep=True, save_start=True, save_end=True),
             dict(saveat=[1.0, 2.0, 3.0, 5.0, 7.5], save_start=False, save_end=True),
             dict(save_start=False, save_end=True)]
    codes = set()
    for m in modes:
        print(f"[case] discrete callbacks {alg.code} {engine} {sorted(m)}", flush=True)
        out = odemi.solve_arrays(fun, alg, u0, p, np.zeros(N), np.full(N, 10.0), ea, reltol=1e-6, abstol=1e-8,
                                 tstops=_DOSES, callback=_PK_CBS, max_saves=2048, **m)
        for j in range(0, N, 5):
            r = O.solve(lambda du, u, pp, t: _pk(du, u, pp, t), u0[:, j], p[:, j], (0.0, 10.0),
                        jac=lambda J, u, pp, t: _pk_jac(J, u, pp, t), alg=alg.code, reltol=1e-6, abstol=1e-8,
                        tstops=_DOSES, callbacks=_PK_CBS.discrete_callbacks, **m)
            ns = int(out["nsaved"][j])
            assert int(out["retcode"][j]) == r["retcode"] and ns == r["nsaved"], (j, m)
            assert np.array_equal(out["t_out"][:ns, j], r["t"]), (j, m)
            assert np.array_equal(out["u_out"][:ns, :, j], r["u"]), (j, m)
            assert np.array_equal(out["u_final"][:, j], r["u_final"])
            st = out["stats"][:, j]
            assert (st[0], st[1], st[2], st[6]) == (r["nf"], r["naccept"], r["nreject"], r["ncondition"])
            codes.add(r["retcode"])
            if r["retcode"] == 1:   # every dose saved twice, before and after
                assert all(list(r["t"]).count(d) == 2 for d in _DOSES)
    assert codes == {1, 2}


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
def test_discrete_callbacks_mixed_and_surface(alg):
    """The same callbacks on the mixed engines (within the FP32-stage error of
    the FP64 engine, same save times) and through solve(EnsembleProblem):
    ODESolution.retcode Terminated / Success and stats.ncondition."""
    N = 256
    u0, p = _pk_batch(N, seed=12)
    fun = odemi.ODEFunction(_pk, jac=_pk_jac)
    kw = dict(reltol=1e-6, abstol=1e-8, tstops=_DOSES, callback=_PK_CBS, save_start=False, save_end=True,
              max_saves=64)
    print(f"[case] discrete callbacks mixed {alg.code}", flush=True)
    o64 = odemi.solve_arrays(fun, alg, u0, p, np.zeros(N), np.full(N, 10.0), **kw)
    om = odemi.solve_arrays(fun, alg, u0, p, np.zeros(N), np.full(N, 10.0),
                            odemi.EnsembleMI355X(precision="mixed"), **kw)
    same = (om["retcode"] == o64["retcode"])
    assert same.mean() > 0.95 and set(np.unique(o64["retcode"])) == {1, 2}
    for j in np.nonzero(same & (o64["retcode"] == 1))[0][:40]:
        ns = int(om["nsaved"][j])
        assert ns == int(o64["nsaved"][j]) and np.array_equal(om["t_out"][:ns, j], o64["t_out"][:ns, j])
    ok = same & (o64["retcode"] == 1)
    err = np.abs(om["u_final"][:, ok] - o64["u_final"][:, ok]) / (1e-8 + 1e-6 * np.abs(o64["u_final"][:, ok]))
    print(f"callbacks mixed scaled err median {np.median(err):.3e} max {err.max():.3e}")
    assert np.median(err) < 10.0
    # the EnsembleProblem surface
    base = odemi.ODEProblem(fun, list(u0[:, 0]), (0.0, 10.0), list(p[:, 0]))
    ens = odemi.EnsembleProblem(base, prob_func=lambda pr, ctx: odemi.remake(
        pr, u0=list(u0[:, ctx.sim_id - 1]), p=list(p[:, ctx.sim_id - 1])))
    sim = odemi.solve(ens, alg, odemi.EnsembleMI355X(), trajectories=32, reltol=1e-6, abstol=1e-8,
                      tstops=_DOSES, callback=_PK_CBS, save_everystep=False)
    for j, sol in enumerate(sim.u):
        assert int(sol.retcode) == int(o64["retcode"][j])
        assert sol.stats.ncondition == 2 * sol.stats.naccept
        assert np.array_equal(sol.u[-1], o64["u_final"][:, j])


def _ball(du, u, p, t):
    du[0] = u[1]
    du[1] = -9.81


def _ball_jac(J, u, p, t):
    J[0, 0] = 0.0
    J[0, 1] = 1.0
    J[1, 0] = 0.0
    J[1, 1] = 0.0


def _bounce(itg):
    itg.u[1] = -itg.p[0] * itg.u[1]


_BALL_CBS = odemi.CallbackSet(
    odemi.ContinuousCallback(lambda u, t, itg: u[0], None, _bounce),
    odemi.ContinuousCallback(lambda u, t, itg: t - itg.p[1], lambda itg: odemi.terminate(itg), rootfind="right"))


def _ball_batch(N, seed=21):
    rng = np.random.default_rng(seed)
    u0 = np.stack([rng.uniform(10.0, 50.0, N), rng.uniform(-5.0, 5.0, N)])
    p = np.stack([rng.uniform(0.8, 0.95, N), rng.uniform(4.0, 14.0, N)])
    return u0, p


class _PyCCB:
    """The same callbacks for the CPU restatement (plain Python on numpy)."""

    def __init__(self, cb):
        self.__dict__.update(cb.__dict__)


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_continuous_callbacks_bit_identical(alg, engine):
    """Two ContinuousCallbacks through the lane engines, bit-identical to the
    CPU restatement (event search at tprev / t / 9 interpolated points, root to
    adjacent floats, change_t_via_interpolation! with the method's stage
    re-step, affect_neg! only, then terminate! at a per-trajectory time with
    RightRootFind): bouncing balls with restitution p[0], stopped at p[1] when
    that comes before tf = 10; save_everystep, saveat and final state only.
    Save times and states, retcodes, nf / naccept / nreject / ncondition."""
    N = 96
    u0, p = _ball_batch(N)
    ea = odemi.EnsembleMI355X(engine=engine)
    fun = odemi.ODEFunction(_ball, jac=_ball_jac)
    modes = [dict(save_everystep=True, save_start=True, save_end=True),
             dict(saveat=list(np.linspace(0.5, 10.0, 20)), save_start=False, save_end=True),
             dict(save_start=False, save_end=True)]
    codes = set()
    ccbs = [_PyCCB(c) for c in _BALL_CBS.continuous_callbacks]
    for m in modes:
        print(f"[case] continuous callbacks {alg.code} {engine} {sorted(m)}", flush=True)
        out = odemi.solve_arrays(fun, alg, u0, p, np.zeros(N), np.full(N, 10.0), ea, reltol=1e-6,
                                 abstol=1e-6, callback=_BALL_CBS, max_saves=2048, **m)
        for j in range(0, N, 5):
            r = O.solve(_ball, u0[:, j], p[:, j], (0.0, 10.0), jac=_ball_jac, alg=alg.code, reltol=1e-6,
                        abstol=1e-6, ccallbacks=ccbs, **m)
            ns = int(out["nsaved"][j])
            assert int(out["retcode"][j]) == r["retcode"] and ns == r["nsaved"], (j, m)
            assert np.array_equal(out["t_out"][:ns, j], r["t"]), (j, m)
            assert np.array_equal(out["u_out"][:ns, :, j], r["u"]), (j, m)
            assert np.array_equal(out["u_final"][:, j], r["u_final"]) and out["t_final"][j] == r["t_final"]
            st = out["stats"][:, j]
            assert (st[0], st[1], st[2], st[6]) == (r["nf"], r["naccept"], r["nreject"], r["ncondition"])
            codes.add(r["retcode"])
            if r["retcode"] == 2:
                assert r["t_final"] == np.nextafter(p[1, j], np.inf) or r["t_final"] == p[1, j]
    assert codes == {1, 2}


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
def test_continuous_callbacks_mixed_and_surface(alg):
    """The same events on the mixed engines (FP32 stages / FP32 LU): bounce
    times within the integration error of the FP64 engine's, same retcodes; and
    through solve(EnsembleProblem): every bounce saved twice with the
    velocity's sign flipped."""
    N = 256
    u0, p = _ball_batch(N, seed=22)
    fun = odemi.ODEFunction(_ball, jac=_ball_jac)
    kw = dict(reltol=1e-6, abstol=1e-6, callback=_BALL_CBS, save_start=False, save_end=True, max_saves=256)
    print(f"[case] continuous callbacks mixed {alg.code}", flush=True)
    o64 = odemi.solve_arrays(fun, alg, u0, p, np.zeros(N), np.full(N, 10.0), **kw)
    om = odemi.solve_arrays(fun, alg, u0, p, np.zeros(N), np.full(N, 10.0),
                            odemi.EnsembleMI355X(precision="mixed"), **kw)
    assert np.array_equal(om["retcode"], o64["retcode"]) and np.array_equal(om["nsaved"], o64["nsaved"])
    for j in range(N):
        ns = int(o64["nsaved"][j])
        assert np.allclose(om["t_out"][:ns, j], o64["t_out"][:ns, j], rtol=1e-5, atol=1e-5)
    base = odemi.ODEProblem(fun, list(u0[:, 0]), (0.0, 10.0), list(p[:, 0]))
    ens = odemi.EnsembleProblem(base, prob_func=lambda pr, ctx: odemi.remake(
        pr, u0=list(u0[:, ctx.sim_id - 1]), p=list(p[:, ctx.sim_id - 1])))
    sim = odemi.solve(ens, alg, odemi.EnsembleMI355X(), trajectories=16, reltol=1e-6, abstol=1e-6,
                      callback=_BALL_CBS, save_everystep=False)
    for j, sol in enumerate(sim.u):
        assert int(sol.retcode) == int(o64["retcode"][j])
        ts = list(sol.t)
        bounces = [t for t in set(ts) if ts.count(t) == 2 and t != sol.t[-1]]
        for t in bounces:
            i = ts.index(t)
            assert sol.u[i][1] < 0 < sol.u[i + 1][1] and abs(sol.u[i][0]) < 1e-9


def _expb(du, u, p, t):
    du[0] = odemi.sym_exp(u[0])


def test_return_codes():
    print("[case] return codes", flush=True)
    for alg in (odemi.Tsit5(), odemi.Rodas5P()):
        sol = odemi.solve(odemi.ODEProblem(_expb, [0.0], (0.0, 10.0)), alg, reltol=1e-8, abstol=1e-8)
        assert sol.retcode in (odemi.ReturnCode.MaxIters, odemi.ReturnCode.Unstable)
        ref = O.solve_builtin(O.PROB_EXPBLOWUP, [0.0], [0.0], (0.0, 10.0), alg=alg.code, reltol=1e-8, abstol=1e-8)
        assert int(sol.retcode) == ref["retcode"]
    lor = odemi.ODEProblem(P.LORENZ, [1.0, 0.0, 0.0], (0.0, 10.0), [10.0, 28.0, 8 / 3])
    sol = odemi.solve(lor, odemi.Tsit5(), maxiters=5)
    assert sol.retcode == odemi.ReturnCode.MaxIters and sol.stats.naccept + sol.stats.nreject == 5
    sol = odemi.solve(odemi.remake(lor, u0=[float("nan"), 0.0, 0.0]), odemi.Tsit5())
    assert sol.retcode == odemi.ReturnCode.Unstable
    sol = odemi.solve(lor, odemi.Tsit5(), dt=float("nan"))
    assert sol.retcode == odemi.ReturnCode.DtNaN


@pytest.mark.parametrize("alg,prec", [(odemi.Tsit5(), "mixed"), (odemi.Tsit5(), "fp64"), (odemi.Rodas5P(), "fp64")],
                         ids=["tsit5-mixed", "tsit5-fp64", "rodas5p-fp64"])
def test_work_order_is_a_permutation(alg, prec):
    """The longest-first work order (a counting sort over cost bins) hands out
    every trajectory exactly once: results equal the identity order's for a
    batch that is not a multiple of the sort's 4,096-trajectory tile, with
    zero-length spans and a NaN end time (cost bin 0) mixed in."""
    N = 4096 + 777
    d = P.lorenz_large(N=N)
    d["tf"][::97] = d["t0"][::97]
    d["tf"][5] = float("nan")
    print(f"[case] work order permutation {alg.code} {prec} N={N}", flush=True)
    outs = []
    for srt in (True, False):
        ea = odemi.EnsembleMI355X(precision=prec, sort_by_cost=srt)
        outs.append(odemi.solve_arrays(P.LORENZ, alg, d["u0"], d["p"], d["t0"], d["tf"], ea, reltol=1e-6,
                                       abstol=1e-6))
    a, b = outs
    assert np.array_equal(a["u_final"], b["u_final"], equal_nan=True)
    assert np.array_equal(a["stats"], b["stats"]) and np.array_equal(a["retcode"], b["retcode"])
    assert (a["retcode"] == 1).all() and (a["stats"][1] > 0).sum() == N - len(range(0, N, 97)) - 1


def test_autotune_lanes_keeps_results():
    """The autotuner times the resident workgroups per CU on the caller's batch and
    keeps the fastest; the launch shape does not change any result."""
    d = P.lorenz_large(N=20000)
    print("[case] autotune blocks_per_cu", flush=True)
    base = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"],
                              odemi.EnsembleMI355X(precision="mixed"), reltol=1e-6, abstol=1e-6)
    tuned = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"],
                               odemi.EnsembleMI355X(precision="mixed", autotune=True), reltol=1e-6, abstol=1e-6)
    best, times = tuned["tuning"]["blocks_per_cu"], tuned["tuning"]["kernel_ms"]
    print(f"autotune: best {best} per CU, kernel ms {times}")
    assert best in times and times[best] == min(times.values())
    assert np.array_equal(base["u_final"], tuned["u_final"]) and np.array_equal(base["stats"], tuned["stats"])
    fixed = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"],
                               odemi.EnsembleMI355X(precision="mixed", blocks_per_cu=2), reltol=1e-6, abstol=1e-6)
    assert np.array_equal(base["u_final"], fixed["u_final"])


def test_stats_and_user_dt_match_oracle():
    print("[case] stats + user dt", flush=True)
    d = P.lorenz_small(N=128)
    for alg, pid, fun in ((odemi.Tsit5(), O.PROB_LORENZ, P.LORENZ), (odemi.Rodas5P(), O.PROB_LORENZ, P.LORENZ)):
        out = odemi.solve_arrays(fun, alg, d["u0"], d["p"], d["t0"], d["tf"], reltol=1e-6, abstol=1e-6, dt=1e-3)
        for j in range(0, 128, 17):
            r = O.solve_builtin(pid, d["u0"][:, j], d["p"][:, j], (0.0, 1.0), alg=alg.code, reltol=1e-6,
                                abstol=1e-6, dt=1e-3, save_start=False)
            st = out["stats"][:, j]
            assert (st[0], st[1], st[2], st[3], st[4], st[5]) == (r["nf"], r["naccept"], r["nreject"], r["njacs"],
                                                                   r["nw"], r["nsolve"])
            assert np.array_equal(out["u_final"][:, j], r["u_final"])


def test_ensemble_problem_hooks():
    """prob_func(prob, ctx) sets the initial conditions (ensemble_thread_safety.jl);
    output_func and reduction see every trajectory."""
    print("[case] EnsembleProblem hooks", flush=True)

    def f(du, u, p, t):
        du[0] = 1.01 * u[0]
        du[1] = 0.0 * u[1]

    prob = odemi.ODEProblem(f, [0.0, 0.0], (0.0, 1.0))
    ics = np.linspace(0, 1, 100)

    def prob_func(prob, ctx):
        return odemi.remake(prob, u0=[ics[ctx.sim_id - 1], 0.0])

    sim = odemi.solve(odemi.EnsembleProblem(prob, prob_func=prob_func), odemi.Tsit5(), odemi.EnsembleMI355X(),
                      trajectories=100)
    assert len(sim.u) == 100
    assert np.array_equal([s.u[0][0] for s in sim.u], ics)           # get_timepoint(sim, 0)
    assert all(s.retcode == odemi.ReturnCode.Success for s in sim.u)
    assert np.allclose([s.u[-1][0] for s in sim.u], ics * math.exp(1.01), rtol=1e-3)
    red = odemi.solve(odemi.EnsembleProblem(prob, prob_func=prob_func,
                                            output_func=lambda sol, ctx: (sol.u[-1][0], False),
                                            reduction=lambda u, data, I: ((u or 0.0) + sum(data), False)),
                      odemi.Tsit5(), odemi.EnsembleMI355X(), trajectories=100, saveat=[1.0])
    assert red.u == pytest.approx(float(np.sum(ics)) * math.exp(1.01), rel=1e-3)


def test_mixed_precision_saveat_within_tolerance():
    d = P.lorenz_small(N=256)
    sa = np.linspace(0.1, 0.9, 9)
    print("[case] mixed saveat", flush=True)
    o64 = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"], reltol=1e-6, abstol=1e-6,
                             saveat=sa)
    om = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), reltol=1e-6, abstol=1e-6, saveat=sa)
    assert np.array_equal(o64["nsaved"], om["nsaved"])
    ns = int(o64["nsaved"].min())
    err = np.abs(om["u_out"][:ns] - o64["u_out"][:ns]) / (1e-6 + 1e-6 * np.abs(o64["u_out"][:ns]))
    assert np.quantile(err, 0.99) < 20.0


def _ring16(du, u, p, t):
    for i in range(16):
        du[i] = -(0.1 * (i + 1)) * u[i] + 0.3 * u[(i + 1) % 16] + p[0] * t


def _decay_nop(du, u, p, t):
    du[0] = -2.0 * u[0]
    du[1] = u[0] - u[1]


@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_edge_cases_match_oracle(engine):
    """Edge cases of the ensemble surface against the CPU restatement, FP64
    bit for bit: the largest lane state (Tsit5, n = 16, non-autonomous);
    spans of zero length (tf == t0), negative and shifted t0 in one batch; a
    trajectory whose u0 holds a NaN among finite ones (only it fails, with the
    restatement's code); an empty batch; a problem without parameters."""
    ea = odemi.EnsembleMI355X(engine=engine)
    rng = np.random.default_rng(31)
    N = 40
    u0 = rng.uniform(-1.0, 1.0, (16, N))
    p = rng.uniform(-0.5, 0.5, (1, N))
    t0 = rng.choice([0.0, -1.5, 2.0], N)
    tf = t0 + rng.choice([0.0, 0.7, 3.0], N)
    u0[3, 7] = np.nan
    print(f"[case] edge cases n=16 {engine}", flush=True)
    out = odemi.solve_arrays(_ring16, odemi.Tsit5(), u0, p, t0, tf, ea, reltol=1e-7, abstol=1e-9,
                             save_everystep=True, save_start=True, max_saves=4096)
    seen = set()
    for j in range(N):
        r = O.solve(_ring16, u0[:, j], p[:, j], (t0[j], tf[j]), alg=O.TSIT5, reltol=1e-7, abstol=1e-9,
                    save_everystep=True)
        ns = int(out["nsaved"][j])
        assert int(out["retcode"][j]) == r["retcode"] and ns == r["nsaved"], j
        assert np.array_equal(out["t_out"][:ns, j], r["t"], equal_nan=True), j
        assert np.array_equal(out["u_out"][:ns, :, j], r["u"], equal_nan=True), j
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2]) == (r["nf"], r["naccept"], r["nreject"]), j
        seen.add((tf[j] == t0[j], r["retcode"]))
    assert (True, 1) in seen and (False, 1) in seen and int(out["retcode"][7]) != 1
    assert all(int(out["retcode"][j]) == 1 for j in range(N) if j != 7)
    # an empty batch
    e = odemi.solve_arrays(_ring16, odemi.Tsit5(), np.zeros((16, 0)), np.zeros((1, 0)), np.zeros(0), np.zeros(0),
                           ea)
    assert e["retcode"].shape == (0,) and e["u_final"].shape == (16, 0)
    # no parameters: ODEProblem without p
    for alg in (odemi.Tsit5(), odemi.Rodas5P()):
        sol = odemi.solve(odemi.ODEProblem(_decay_nop, [1.0, 0.0], (0.0, 2.0)), alg, ea, reltol=1e-6, abstol=1e-8)
        r = O.solve(_decay_nop, [1.0, 0.0], [], (0.0, 2.0), alg=alg.code, reltol=1e-6, abstol=1e-8,
                    save_everystep=True, jac=lambda J, u, pp, t: J.__setitem__(slice(None), [[-2.0, 0.0],
                                                                                          [1.0, -1.0]]))
        assert int(sol.retcode) == r["retcode"] == 1
        assert np.array_equal(np.array(sol.t), r["t"]) and np.array_equal(np.array(sol.u), r["u"])


def _box(du, u, p, t):
    du[0] = u[1]
    du[1] = -9.81
    du[2] = u[3]
    du[3] = 0.0


def _box_jac(J, u, p, t):
    for i in range(4):
        for k in range(4):
            J[i, k] = 1.0 if (i, k) in ((0, 1), (2, 3)) else 0.0


def _box_cond(out, u, t, itg):
    out[0] = u[0]
    out[1] = (u[2] - 10.0) * u[2]


def _box_affect(itg, ev):
    itg.u[1] = odemi.sym_ifelse(ev[0] == -1, -itg.p[0] * itg.u[1], itg.u[1])
    itg.u[3] = odemi.sym_ifelse(ev[1] != 0, -itg.u[3], itg.u[3])
    odemi.terminate(itg, when=itg.t > itg.p[1])


_BOX_CBS = odemi.CallbackSet(odemi.VectorContinuousCallback(_box_cond, _box_affect, 2),
                             odemi.ContinuousCallback(lambda u, t, itg: u[2] - 5.0, None, None))


@pytest.mark.parametrize("alg", [odemi.Tsit5(), odemi.Rodas5P()], ids=["tsit5", "rodas5p"])
@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_vector_continuous_callbacks_bit_identical(alg, engine):
    """A VectorContinuousCallback (floor and walls of a box, restitution p[0],
    a conditional terminate! once t > p[1]) in a CallbackSet with a scalar
    ContinuousCallback that is never an event (affect! = affect_neg! =
    nothing) through the lane engines, bit-identical to the CPU restatement:
    save_everystep and final state only; times, states, retcodes, nf / naccept
    / nreject / ncondition."""
    N = 64
    rng = np.random.default_rng(41)
    u0 = np.stack([rng.uniform(10.0, 50.0, N), rng.uniform(-3.0, 3.0, N), rng.uniform(1.0, 9.0, N),
                   rng.uniform(-6.0, 6.0, N)])
    p = np.stack([rng.uniform(0.8, 0.95, N), rng.uniform(3.0, 14.0, N)])
    fun = odemi.ODEFunction(_box, jac=_box_jac)
    ea = odemi.EnsembleMI355X(engine=engine)
    codes = set()
    for m in (dict(save_everystep=True, save_start=True, save_end=True), dict(save_start=False, save_end=True)):
        print(f"[case] vector continuous callbacks {alg.code} {engine} {sorted(m)}", flush=True)
        out = odemi.solve_arrays(fun, alg, u0, p, np.zeros(N), np.full(N, 10.0), ea, reltol=1e-6, abstol=1e-6,
                                 callback=_BOX_CBS, max_saves=4096, **m)
        for j in range(0, N, 3):
            r = O.solve(_box, u0[:, j], p[:, j], (0.0, 10.0), jac=_box_jac, alg=alg.code, reltol=1e-6,
                        abstol=1e-6, ccallbacks=_BOX_CBS.continuous_callbacks, **m)
            ns = int(out["nsaved"][j])
            assert int(out["retcode"][j]) == r["retcode"] and ns == r["nsaved"], (j, m)
            assert np.array_equal(out["t_out"][:ns, j], r["t"]), (j, m)
            assert np.array_equal(out["u_out"][:ns, :, j], r["u"]), (j, m)
            st = out["stats"][:, j]
            assert (st[0], st[1], st[2], st[6]) == (r["nf"], r["naccept"], r["nreject"], r["ncondition"])
            codes.add(r["retcode"])
            assert r["u"][:, 2].min() >= -1e-9 and r["u"][:, 2].max() <= 10.0 + 1e-9
    assert codes == {1, 2}


def test_ensemble_batches_rerun_and_converged_reduction():
    """SciMLBase's ensemble loop: batch_size batches, output_func's rerun flag
    (the trajectory is redrawn by prob_func with ctx.repeat + 1 and solved
    again), reduction(u, batch, I) -> (u, converged) stopping at convergence;
    outputs equal to one unbatched solve."""
    print("[case] ensemble batches", flush=True)
    prob = odemi.ODEProblem(_lin, [1.0], (0.0, 1.0), [-1.0])
    pf = lambda pr, ctx: odemi.remake(pr, p=[-0.1 * ctx.sim_id - (ctx.repeat - 1)])
    full = odemi.solve(odemi.EnsembleProblem(prob, prob_func=pf), odemi.Tsit5(), trajectories=10,
                       reltol=1e-8, abstol=1e-10)
    ends = [s.u[-1][0] for s in full.u]
    seen = []

    def red(u, batch, I):
        seen.append(list(I))
        u = u + [s.u[-1][0] for s in batch]
        return u, len(u) >= 6
    sim = odemi.solve(odemi.EnsembleProblem(prob, prob_func=pf, reduction=red, u_init=[]), odemi.Tsit5(),
                      trajectories=10, batch_size=3, reltol=1e-8, abstol=1e-10)
    assert seen == [[1, 2, 3], [4, 5, 6]] and sim.converged and sim.u == ends[:6]
    calls = {}

    def of(sol, ctx):   # rerun trajectory 2 twice: its third draw has p = -0.2 - 2
        calls[ctx.sim_id] = calls.get(ctx.sim_id, 0) + 1
        return (ctx.repeat, sol.u[-1][0]), ctx.sim_id == 2 and ctx.repeat < 3
    sim = odemi.solve(odemi.EnsembleProblem(prob, prob_func=pf, output_func=of), odemi.Tsit5(), trajectories=4,
                      batch_size=2, reltol=1e-8, abstol=1e-10)
    assert [r for r, _ in sim.u] == [1, 3, 1, 1] and calls == {1: 1, 2: 3, 3: 1, 4: 1}
    assert abs(sim.u[1][1] - math.exp(-2.2)) < 1e-7 and sim.u[0][1] == ends[0]


def test_events_workload_matches_builtin_restatement():
    """The events benchmark's problem (bouncing balls with the floor
    ContinuousCallback) on the FP64 persistent engine equals the restatement's
    built-in C floor event bit for bit (final state, retcode, stats)."""
    print("[case] events workload vs built-in restatement", flush=True)
    d = P.ball_ensemble(N=512)
    out = odemi.solve_arrays(P.BALL, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"],
                             odemi.EnsembleMI355X(precision="fp64"), reltol=d["reltol"], abstol=d["abstol"],
                             save_start=False, callback=P.BALL_BOUNCE)
    ref = O.ensemble_builtin(O.PROB_BALL, d["u0"], d["p"], d["t0"], d["tf"], alg=O.TSIT5, reltol=d["reltol"],
                             abstol=d["abstol"], save_start=False)
    assert np.array_equal(out["u_final"], ref["u_final"]) and np.array_equal(out["retcode"], ref["retcode"])
    assert np.array_equal(out["stats"][1], ref["naccept"]) and np.array_equal(out["stats"][0], ref["nf"])
