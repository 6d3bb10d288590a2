"""Pins the CPU restatement (oracle/) before it is trusted as the parity target.

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign or index, a transposed operand) fails it:
  * tableaus: Butcher / Rosenbrock order conditions in mpmath at 50 digits;
  * one step on u' = lambda u equals the method's stability function, which
    must agree with exp(z) to the method order and (Rodas5P) vanish at -inf;
  * fixed-step convergence order 5 +- 0.2 (the reference's own test:
    lib/OrdinaryDiffEqRosenbrock/test/ode_rosenbrock_tests.jl:195-215);
  * reference known answers: saveat/save_start/save_end time lists
    (test/InterfaceI/ode_saveat_tests.jl:246-267), qmax_first_step growth and
    u(100) = exp(-100) (test/InterfaceI/controllers.jl:29-73), the exp(u)
    blow-up return codes (test/Integrators_I/check_error.jl:1-44), nf accounting
    (test/InterfaceIII/stats_tests.jl:1-40), empty-state dt = 1e-6
    (test/Regression_I/ode_adaptive_tests.jl:111-118), DiscreteCallback save
    and terminate! rules (test/Integrators_I/ode_event_tests.jl);
  * the textbook Hairer-Norsett-Wanner initial step recomputed in mpmath;
  * invariants: Robertson mass conservation; agreement with scipy's DOP853 /
    Radau at tight tolerance; LU against numpy; Julia eps/min semantics.
"""
import math
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest

import oracle as O

mp.mp.dps = 50


# ---------------------------------------------------------------------------
# rooted trees and Butcher order conditions
# ---------------------------------------------------------------------------
def _trees(order):
    """All rooted trees with `order` nodes, as sorted tuples of children."""
    if order == 1:
        return [()]
    out = set()

    def partitions(n, maxpart):
        if n == 0:
            yield []
            return
        for k in range(min(n, maxpart), 0, -1):
            for rest in partitions(n - k, k):
                yield [k] + rest

    for parts in partitions(order - 1, order - 1):
        lists = [_trees(k) for k in parts]

        def combos(i):
            if i == len(lists):
                yield []
                return
            for t in lists[i]:
                for r in combos(i + 1):
                    yield [t] + r

        for ch in combos(0):
            out.add(tuple(sorted(ch)))
    return sorted(out)


def _gamma(t):
    n = 1 + sum(_size(c) for c in t)
    g = n
    for c in t:
        g *= _gamma(c)
    return g


def _size(t):
    return 1 + sum(_size(c) for c in t)


def _stage_weights(t, A, s):
    """Vector Phi_i(t) = prod_children (A Phi(child))_i."""
    v = [mp.mpf(1)] * s
    for c in t:
        w = _stage_weights(c, A, s)
        aw = [mp.fsum(A[i][j] * w[j] for j in range(s)) for i in range(s)]
        v = [v[i] * aw[i] for i in range(s)]
    return v


def _order_residuals(A, b, p):
    s = len(b)
    res = []
    for order in range(1, p + 1):
        for t in _trees(order):
            phi = _stage_weights(t, A, s)
            res.append(abs(mp.fsum(b[i] * phi[i] for i in range(s)) - mp.mpf(1) / _gamma(t)))
    return res


def test_tree_counts():
    assert [len(_trees(k)) for k in range(1, 6)] == [1, 1, 2, 4, 9]


def test_tsit5_tableau_order_conditions():
    c, a, b, bt = O.tsit5_tableau()
    A = [[mp.mpf(float(x)) for x in row] for row in a]
    B = [mp.mpf(float(x)) for x in b]
    # row sums: c_i = sum_j a_ij
    for i in range(7):
        assert abs(mp.fsum(A[i]) - mp.mpf(float(c[i]))) < 1e-15
    # FSAL: last row of A is b, b_7 = 0
    assert b[6] == 0.0 and np.all(a[6] == b)
    # 5th order solution (17 conditions through order 5) to double rounding
    assert max(_order_residuals(A, B, 5)) < 1e-13
    # embedded 4th order: bhat = b - btilde satisfies order 4 but not 5
    Bh = [B[i] - mp.mpf(float(bt[i])) for i in range(7)]
    assert max(_order_residuals(A, Bh, 4)) < 1e-13
    assert max(_order_residuals(A, Bh, 5)) > 1e-6


def _tsit5_stability(z):
    c, a, b, _ = O.tsit5_tableau()
    A = mp.matrix([[mp.mpf(float(x)) for x in row] for row in a])
    B = mp.matrix([[mp.mpf(float(x)) for x in b]])
    one = mp.matrix([[1]] * 7)
    M = mp.eye(7) - z * A
    return 1 + z * (B * mp.lu_solve(M, one))[0]


def test_tsit5_stability_polynomial_matches_exp_to_order5():
    # R(z) - e^z = O(z^6): check the Taylor coefficients through z^5
    for z in (mp.mpf("1e-2"), mp.mpf("-2e-2")):
        assert abs(_tsit5_stability(z) - mp.e ** z) < 5 * abs(z) ** 6


@pytest.mark.parametrize("lam,h", [(-1.0, 0.1), (-3.0, 0.05), (0.7, 0.2)])
def test_tsit5_single_step_is_stability_polynomial(lam, h):
    # one fixed step of u' = lam u
    r = O.solve_builtin(O.PROB_LINEAR, [1.0], [lam], (0.0, h), dt=h, adaptive=False, save_start=False)
    assert r["naccept"] == 1
    exact = _tsit5_stability(mp.mpf(lam) * mp.mpf(h))
    assert abs(r["u_final"][0] - float(exact)) < 4e-16


def _rodas_stability(z):
    """R(z) of the transformed Rosenbrock scheme on u' = lam u from (A, C, gamma, b)."""
    T = O.rodas5p_tableau()
    A = [[mp.mpf(float(x)) for x in row] for row in T["A"]]
    C = [[mp.mpf(float(x)) for x in row] for row in T["C"]]
    g = mp.mpf(float(T["gamma"]))
    b = [mp.mpf(float(x)) for x in T["b"]]
    # with h = 1: (1/gamma - z) k_s = z (1 + sum_j A_sj k_j) + sum_j C_sj k_j
    k = []
    for s in range(8):
        rhs = z * (1 + mp.fsum(A[s][j] * k[j] for j in range(s))) + mp.fsum(C[s][j] * k[j] for j in range(s))
        k.append(rhs / (1 / g - z))
    return 1 + mp.fsum(b[s] * k[s] for s in range(8))


def test_rodas5p_stability_function():
    # order 5: R(z) = e^z + O(z^6)
    for z in (mp.mpf("1e-2"), mp.mpf("-1e-2")):
        assert abs(_rodas_stability(z) - mp.e ** z) < 50 * abs(z) ** 6
    # L-stable: R(-inf) = 0
    assert abs(_rodas_stability(mp.mpf("-1e12"))) < 1e-9


@pytest.mark.parametrize("lam,h", [(-1.0, 0.1), (-50.0, 0.1), (1.01, 0.25)])
def test_rodas5p_single_step_is_stability_function(lam, h):
    r = O.solve_builtin(O.PROB_LINEAR, [1.0], [lam], (0.0, h), alg=O.RODAS5P, dt=h, adaptive=False,
                        save_start=False)
    exact = _rodas_stability(mp.mpf(lam) * mp.mpf(h))
    assert abs(r["u_final"][0] - float(exact)) < 1e-14 * max(1.0, abs(float(exact)))


def test_tsit5_convergence_order_5_linear():
    """test_convergence on prob_ode_linear (u' = 1.01 u, u0 = 1/2), fixed dt = 2^-k: order 5 +- 0.2."""
    dts = [0.5 ** k for k in (4, 3, 2, 1)]
    errs = []
    for dt in dts:
        r = O.solve_builtin(O.PROB_LINEAR, [0.5], [1.01], (0.0, 1.0), dt=dt, adaptive=False, save_start=False)
        errs.append(abs(r["u_final"][0] - 0.5 * math.exp(1.01)))
    slope = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    assert abs(slope - 5.0) < 0.2, slope


def _nl_f(du, u, p, t):
    du[0] = -u[0] * u[0] + math.sin(t)


def _nl_j(J, u, p, t):
    J[0, 0] = -2.0 * u[0]


def _nl_tg(dT, u, p, t):
    dT[0] = math.cos(t)


@pytest.mark.parametrize("alg", [O.TSIT5, O.RODAS5P])
def test_convergence_nonlinear_nonautonomous(alg):
    """u' = -u^2 + sin t exercises the nonlinear stage terms, J and dT/dt (Rodas):
    a wrong coefficient, sign or index drops the observed order to <= 3."""
    ref = O.solve(_nl_f, [1.0], [0.0], (0.0, 2.0), jac=_nl_j, tgrad=_nl_tg, reltol=1e-13, abstol=1e-14,
                  save_start=False)["u_final"][0]
    dts = [0.5 ** k for k in (6, 5, 4, 3)]
    errs = [abs(O.solve(_nl_f, [1.0], [0.0], (0.0, 2.0), jac=_nl_j, tgrad=_nl_tg, alg=alg, dt=dt,
                        adaptive=False, save_start=False)["u_final"][0] - ref) for dt in dts]
    slope = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    assert slope > 4.6, slope


def test_rodas5p_convergence_linear():
    dts = [0.5 ** k for k in (4, 3, 2, 1)]
    errs = [abs(O.solve_builtin(O.PROB_LINEAR, [0.5], [1.01], (0.0, 1.0), alg=O.RODAS5P, dt=dt, adaptive=False,
                                save_start=False)["u_final"][0] - 0.5 * math.exp(1.01)) for dt in dts]
    assert np.polyfit(np.log(dts), np.log(errs), 1)[0] > 4.8


def test_rodas5p_adaptive_linear_few_steps():
    # ode_rosenbrock_tests.jl: the adaptive solve of prob_ode_linear needs < 20 steps
    r = O.solve_builtin(O.PROB_LINEAR, [0.5], [1.01], (0.0, 1.0), alg=O.RODAS5P, save_everystep=True)
    assert r["retcode"] == 1 and r["nsaved"] < 20


# ---------------------------------------------------------------------------
# step control known answers
# ---------------------------------------------------------------------------
def test_qmax_first_step_then_qmax():
    """u' = 0: EEst == 0 every step, so q = 1/qmax: the first step grows dt by
    qmax_first_step = 1e4, later ones by qmax = 10 (controllers.jl:288-293, 805-821)."""
    for alg in (O.TSIT5, O.RODAS5P):
        r = O.solve_builtin(O.PROB_LINEAR, [1.0], [0.0], (0.0, 100.0), alg=alg, save_everystep=True)
        # initdt: d1 = 0 -> dt0 = 1e-6, f0 == f1 -> 100*dt0 = 1e-4
        assert r["dt_init"] == pytest.approx(1e-4, rel=1e-15)
        t = r["t"]
        assert t[0] == 0.0 and t[1] == pytest.approx(1e-4, rel=1e-15)
        assert t[2] - t[1] == pytest.approx(1.0, rel=1e-12)      # x 1e4
        assert t[3] - t[2] == pytest.approx(10.0, rel=1e-12)     # x 10
        assert t[-1] == 100.0 and len(t) == 5                    # dtmax = 100, clipped at tf


@pytest.mark.parametrize("alg", [O.TSIT5, O.RODAS5P])
def test_exponential_decay_to_100(alg):
    r = O.solve_builtin(O.PROB_LINEAR, [1.0], [-1.0], (0.0, 100.0), alg=alg)
    assert r["retcode"] == 1
    assert abs(r["u_final"][0] - math.exp(-100.0)) < 1e-6


def test_empty_state_initial_dt():
    r = O.solve(lambda du, u, p, t: None, np.zeros(0), [0.0], (0.0, 1.0))
    assert r["dt_init"] == pytest.approx(1e-6)


def _hnw_initdt(f, u0, p, t0, tf, reltol, abstol, order=5):
    """Textbook Hairer-Norsett-Wanner starting step (Solving ODEs I, II.4) in mpmath."""
    u0 = [mp.mpf(x) for x in u0]
    n = len(u0)
    sk = [abstol + abs(x) * reltol for x in u0]
    nrm = lambda v: mp.sqrt(mp.fsum(x * x for x in v) / n)
    f0 = np.zeros(n)
    f(f0, np.array([float(x) for x in u0]), p, t0)
    d0 = nrm([u0[i] / sk[i] for i in range(n)])
    d1 = nrm([mp.mpf(f0[i]) / sk[i] for i in range(n)])
    h0 = mp.mpf("1e-6") if (d0 < 1e-5 or d1 < 1e-5) else d0 / d1 / 100
    h0 = min(h0, mp.mpf(tf - t0))
    u1 = np.array([float(u0[i] + h0 * f0[i]) for i in range(n)])
    f1 = np.zeros(n)
    f(f1, u1, p, float(t0 + h0))
    d2 = nrm([(mp.mpf(f1[i]) - f0[i]) / sk[i] for i in range(n)]) / h0
    m = max(d1, d2)
    h1 = max(mp.mpf("1e-6"), h0 * mp.mpf("1e-3")) if m <= 1e-15 else mp.power(10, -(2 + mp.log10(m)) / order)
    return float(min(100 * h0, h1, tf - t0))


def _lorenz(du, u, p, t):
    du[0] = p[0] * (u[1] - u[0])
    du[1] = u[0] * (p[1] - u[2]) - u[1]
    du[2] = u[0] * u[1] - p[2] * u[2]


@pytest.mark.parametrize("u0,p,tol", [([1.0, 0.0, 0.0], [10.0, 28.0, 8 / 3], 1e-6),
                                      ([0.3, -0.2, 0.9], [9.0, 14.0, 8 / 3], 1e-8),
                                      ([1.0, 0.0, 0.0], [10.0, 0.5, 8 / 3], 1e-3)])
def test_initdt_matches_textbook(u0, p, tol):
    r = O.solve_builtin(O.PROB_LORENZ, u0, p, (0.0, 10.0), reltol=tol, abstol=tol)
    ref = _hnw_initdt(_lorenz, u0, p, 0.0, 10.0, tol, tol)
    assert r["dt_init"] == pytest.approx(ref, rel=1e-12)


def test_saveat_save_start_save_end_rules():
    """test/InterfaceI/ode_saveat_tests.jl:246-267 (resolved save flags, f = -cos(u) u)."""
    f2 = lambda du, u, p, t: du.__setitem__(0, -math.cos(u[0]) * u[0])
    kw = dict(tspan=(0.0, 0.4), p=[0.0])
    sa = [0.0, 0.1, 0.2, 0.3, 0.4]
    # defaults: t0 and tf are in saveat -> save_start = save_end = true
    r = O.solve(f2, [10.0], saveat=sa, save_start=True, save_end=True, **kw)
    assert list(r["t"]) == [0.0, 0.1, 0.2, 0.3, 0.4]
    r = O.solve(f2, [10.0], saveat=sa, save_start=False, save_end=False, **kw)
    assert list(r["t"]) == [0.1, 0.2, 0.3]
    r = O.solve(f2, [10.0], saveat=[0.2], save_start=False, save_end=False, **kw)
    assert list(r["t"]) == [0.2]
    r = O.solve(f2, [10.0], saveat=[0.2], save_start=True, save_end=True, save_end_user=True, **kw)
    assert list(r["t"]) == [0.0, 0.2, 0.4]
    r = O.solve(f2, [10.0], save_everystep=True, **kw)
    assert r["t"][0] == 0.0 and r["t"][-1] == 0.4
    r = O.solve(f2, [10.0], save_everystep=True, save_start=False, save_end=False, **kw)
    assert 0.0 not in r["t"] and 0.4 not in r["t"]


def test_tstops_reference_known_answers():
    """test/InterfaceI/ode_tstops_tests.jl:8-62 on prob_ode_linear (u' = 1.01 u,
    u0 = 1/2, tspan (0, 1)): tstops are hit exactly, a fixed step is split at a
    tstop, endpoints and duplicates are dropped, and a 0:0.1:1 grid is a subset
    of the solution times."""
    lin = lambda du, u, p, t: du.__setitem__(0, 1.01 * u[0])
    kw = dict(tspan=(0.0, 1.0), p=[0.0], save_everystep=True)
    for alg in (O.TSIT5, O.RODAS5P):
        r = O.solve(lin, [0.5], alg=alg, dt=1 / 64, tstops=[0.5], **kw)
        assert 0.5 in list(r["t"])
        # fixed steps of 1/3: [0, 1/3, 1/2, 1/3 + 1/2, 1] (the reference's RK4 case)
        r = O.solve(lin, [0.5], alg=alg, dt=1 / 3, adaptive=False, tstops=[0.5], **kw)
        assert list(r["t"]) == [0.0, 1 / 3, 0.5, 1 / 3 + 0.5, 1.0]
        ts = [0, 1 / 5, 1 / 4, 1 / 3, 1 / 2, 3 / 4, 1, 1 / 2]
        r = O.solve(lin, [0.5], alg=alg, tstops=ts, **kw)
        t = list(r["t"])
        assert all(s in t for s in ts) and len(set(t)) == len(t)
        grid = [k / 10 for k in range(11)]
        r = O.solve(lin, [0.5], alg=alg, tstops=grid, **kw)
        assert set(grid) <= set(r["t"])
        # same final state as without tstops to the tolerance, more (shorter) steps
        r0 = O.solve(lin, [0.5], alg=alg, reltol=1e-8, abstol=1e-8, **kw)
        r1 = O.solve(lin, [0.5], alg=alg, reltol=1e-8, abstol=1e-8, tstops=grid, **kw)
        exact = 0.5 * math.exp(1.01)
        assert abs(r1["u"][-1][0] - exact) < 1e-6 and abs(r0["u"][-1][0] - exact) < 1e-6
        assert r1["naccept"] >= r0["naccept"]
    # initdt is bounded by the first tstop, not by tf (integrator_interface.jl:644-662)
    r = O.solve(lin, [0.5], tspan=(0.0, 1.0), p=[0.0], tstops=[1e-4], save_everystep=True)
    assert r["t"][1] <= 1e-4


def test_isoutofdomain_rejects_and_keeps_the_domain():
    """isoutofdomain (integrator_utils.jl:612-614, 135-136): a flagged step is
    rejected and the next dt is dt * qmin.  Rodas5P on u' = -30 u with a loose
    abstol goes below zero without it; with u < 0 flagged every saved state
    stays >= 0 at the price of rejections; a predicate that is never true
    changes nothing; one that is always true past t = 0.5 stalls the
    integrator there with the reference's Unstable code (dt shrinks by qmin
    each attempt until dt <= eps(t))."""
    lin = lambda du, u, p, t: du.__setitem__(0, p[0] * u[0])
    kw = dict(tspan=(0.0, 10.0), p=[-30.0], save_everystep=True, reltol=1e-3, abstol=0.1, alg=O.RODAS5P)
    r0 = O.solve(lin, [1.0], **kw)
    assert r0["u"].min() < 0
    r1 = O.solve(lin, [1.0], isoutofdomain=lambda u, p, t: u[0] < 0, **kw)
    assert r1["retcode"] == 1 and r1["u"].min() >= 0 and r1["nreject"] > r0["nreject"]   # Success
    r2 = O.solve(lin, [1.0], isoutofdomain=lambda u, p, t: False, **kw)
    assert np.array_equal(r2["u"], r0["u"]) and r2["naccept"] == r0["naccept"] and r2["nreject"] == r0["nreject"]
    for alg in (O.TSIT5, O.RODAS5P):
        r3 = O.solve(lin, [1.0], tspan=(0.0, 1.0), p=[-1.0], alg=alg, isoutofdomain=lambda u, p, t: t > 0.5)
        assert r3["retcode"] == 7 and 0.4 < r3["t_final"] <= 0.5   # Unstable


class _CB:
    """DiscreteCallback's fields (condition, affect!, save_positions)."""

    def __init__(self, condition, affect, save_positions=(True, True)):
        self.condition, self.affect, self.save_positions = condition, affect, save_positions


def test_discrete_callback_reference_known_answers():
    """DiscreteCallback handling (integrator_utils.jl:1081-1132,
    lib/DiffEqBase/src/callbacks.jl:645-683) against the reference's tests:
    test/Integrators_I/ode_event_tests.jl:400-419 (u .= 1 at tstops 5 and 8:
    each tstop twice in sol.t, also with saveat = prevfloat.(tstop)), 235-258
    (always-true callbacks with save_positions (true, false) on free fall: one
    adds nothing to save_everystep, two give 2 len - 1 points), 463-475 (two
    no-save callbacks: no repeated time), test/InterfaceI/ode_tstops_tests.jl:
    64-81 (sol.t[end] == 1.5574) and lib/DiffEqBase/test/downstream/
    stats_tests.jl (ncondition = condition calls).  The pre-affect save is the
    integrated state (closed form u1 = 20 - 10 exp(-t/2) at t = 5), the
    post-affect one exactly the affect's value; a modified state costs one FSAL
    re-evaluation at the next accept; terminate! ends with Terminated after the
    endpoint save."""
    def f(du, u, p, t):
        du[0] = -0.5 * u[0] + 10
        du[1] = -0.5 * u[1]
    tstop = [5.0, 8.0]
    calls = [0]

    def cond(u, t, itg):
        calls[0] += 1
        return t in tstop

    def ones(itg):
        itg.u[:] = 1.0
    for alg in (O.TSIT5, O.RODAS5P):
        jac = (lambda J, u, p, t: J.__setitem__(slice(None), [[-0.5, 0.0], [0.0, -0.5]]))
        kw = dict(alg=alg, tstops=tstop, callbacks=[_CB(cond, ones)], save_start=False, save_end=False,
                  save_everystep=False)
        calls[0] = 0
        r = O.solve(f, [10.0, 10.0], [0.0], (0.0, 10.0), jac=jac, saveat=tstop, **kw)
        assert list(r["t"]) == [5.0, 5.0, 8.0, 8.0]
        assert abs(r["u"][0, 0] - (20 - 10 * math.exp(-2.5))) < 1e-2 and list(r["u"][1]) == [1.0, 1.0]
        assert r["ncondition"] == calls[0] == r["naccept"] and r["retcode"] == 1
        sa = list(np.nextafter(np.array(tstop), -np.inf))
        r = O.solve(f, [10.0, 10.0], [0.0], (0.0, 10.0), jac=jac, saveat=sa, **kw)
        assert [list(r["t"]).count(x) for x in tstop] == [2, 2] and len(r["t"]) == 6

    def fall(du, u, p, t):
        du[0] = u[1]
        du[1] = -9.81
    always, nop = (lambda u, t, itg: True), (lambda itg: None)
    r0 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), save_everystep=True)
    r1 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), save_everystep=True,
                 callbacks=[_CB(always, nop, (True, False))])
    r2 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), save_everystep=True,
                 callbacks=[_CB(always, nop, (True, False))] * 2)
    assert np.array_equal(r1["t"], r0["t"]) and len(r2["t"]) == 2 * len(r1["t"]) - 1
    assert r1["nf"] == r0["nf"] + r0["naccept"] - 1 and r1["ncondition"] == r1["naccept"]
    assert r2["ncondition"] == 2 * r2["naccept"]

    def grow(du, u, p, t):
        du[0] = 1.01 * u[0]
    r = O.solve(grow, [1.0], [0.0], (0.0, 10.0), tstops=[2.5], save_everystep=True,
                callbacks=[_CB(lambda u, t, itg: t == 2.5, nop, (False, False))] * 2)
    assert 2.5 in r["t"] and not np.any(np.diff(r["t"]) == 0)

    def de(du, u, p, t):
        du[0] = p[0] * u[0]
        du[1] = p[1] * u[1]

    def bump(itg):
        itg.u[0] += 1.0
    r = O.solve(de, [0.0, 0.0], [0.3, 0.7], (-1.0, 3.0), tstops=[0.010823], saveat=[0.0, 0.0094777, 1.5574],
                callbacks=[_CB(lambda u, t, itg: t == 0.010823, bump)], save_start=False, save_end=False)
    assert list(r["t"]) == [0.0, 0.0094777, 0.010823, 0.010823, 1.5574]
    assert list(r["u"][2]) == [0.0, 0.0] and list(r["u"][3]) == [1.0, 0.0]

    def stop(itg):
        itg.terminate()
    for alg in (O.TSIT5, O.RODAS5P):
        r = O.solve(f, [10.0, 10.0], [0.0], (0.0, 10.0), alg=alg, save_start=False,
                    jac=lambda J, u, p, t: J.__setitem__(slice(None), [[-0.5, 0.0], [0.0, -0.5]]),
                    callbacks=[_CB(lambda u, t, itg: u[0] > 15.0, stop, (False, False))])
        assert r["retcode"] == 2 and 1.0 < r["t_final"] < 10.0   # Terminated
        assert list(r["t"]) == [r["t_final"]] and r["u"][0, 0] > 15.0
        assert abs(r["u"][0, 0] - 20 + 10 * math.exp(-r["t_final"] / 2)) < 1e-2


class _CCB:
    """ContinuousCallback's fields with the reference's defaults."""

    def __init__(self, condition, affect, affect_neg="same", rootfind="left", interp_points=10,
                 save_positions=(True, True), abstol=10 * 2.220446049250313e-16, repeat_nudge=1 / 100):
        self.condition, self.affect = condition, affect
        self.affect_neg = affect if affect_neg == "same" else affect_neg
        self.rootfind, self.interp_points, self.save_positions = rootfind, interp_points, save_positions
        self.abstol, self.repeat_nudge = abstol, repeat_nudge


def test_continuous_callback_reference_known_answers():
    """ContinuousCallback handling (lib/DiffEqBase/src/callbacks.jl:157-640,
    change_t_via_interpolation!, the Tsit5 re-step) against the reference's
    bouncing-ball tests (test/Integrators_I/ode_event_tests.jl:100-200,
    300-316): u = (height, velocity), u' = (v, -9.81), u0 = (50, 0); with
    affect! = nothing, affect_neg! flipping v and interp_points = 100 the middle
    save of save_everystep = false is an event, saveat at it or at t - eps(t)
    gives that time exactly twice, v < 0 before and > 0 after; terminate! on the
    ground (tspan (0, Inf)) ends Terminated at t = sqrt(100 / 9.81) with height
    < 3e-12.  Closed forms: the bounce times are sqrt(100/9.81) (1 + 2k) to the
    root's float; RightRootFind lands just past the root (height < 0),
    LeftRootFind just before it."""
    def fall(du, u, p, t):
        du[0] = u[1]
        du[1] = -9.81

    def flip(itg):
        itg.u[1] = -itg.u[1]
    t1 = math.sqrt(100 / 9.81)
    cb = _CCB(lambda u, t, itg: u[0], None, flip, interp_points=100)
    r = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), ccallbacks=[cb], save_everystep=False)
    assert r["retcode"] == 1 and len(r["t"]) == 6
    t = r["t"][len(r["t"]) // 2 - 1]
    assert abs(t - t1) < 1e-13 and abs(r["t"][3] - 3 * t1) < 1e-12
    for sa in ([t], [t - np.spacing(t)]):
        r2 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), ccallbacks=[cb], saveat=sa, save_start=False,
                     save_end=False)
        ts = list(r2["t"])
        assert ts.count(t) == 2
        i = ts.index(t)
        assert r2["u"][i, 1] < 0 < r2["u"][i + 1, 1]

    def stop(itg):
        itg.terminate()
    for rf in ("left", "right"):
        r5 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, math.inf), ccallbacks=[_CCB(lambda u, t, itg: u[0], stop,
                                                                                   rootfind=rf)])
        assert r5["retcode"] == 2 and abs(r5["t"][-1] - t1) < 1e-13
        assert abs(r5["u"][-1][0]) < 3e-12 and (r5["u"][-1][0] > 0) == (rf == "left")
    # Rodas5P: its dense output and event re-step (stiff_addsteps.jl:140-224) find the same bounces
    jac = lambda J, u, p, t: J.__setitem__(slice(None), [[0.0, 1.0], [0.0, 0.0]])
    r8 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), jac=jac, alg=O.RODAS5P, ccallbacks=[cb],
                 save_everystep=False)
    assert r8["retcode"] == 1 and len(r8["t"]) == 6
    assert abs(r8["t"][1] - t1) < 1e-12 and abs(r8["t"][3] - 3 * t1) < 1e-12 and r8["t"][1] == r8["t"][2]
    assert r8["u"][1][1] < 0 < r8["u"][2][1]
    # no event: a condition that never changes sign changes nothing but the condition count
    r6 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 2.0), save_everystep=True)
    r7 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 2.0), save_everystep=True,
                 ccallbacks=[_CCB(lambda u, t, itg: u[0] + 100.0, flip)])
    assert np.array_equal(r6["t"], r7["t"]) and np.array_equal(r6["u"], r7["u"])
    assert r7["ncondition"] == r7["naccept"] * 11   # tprev, t and 9 interpolated points per step


class _VCB:
    """VectorContinuousCallback's fields with the reference's defaults."""

    def __init__(self, condition, affect, len_, rootfind="left", interp_points=10, save_positions=(True, True),
                 abstol=10 * 2.220446049250313e-16, repeat_nudge=1 / 100):
        self.condition, self.affect, self.affect_neg, self.len = condition, affect, None, len_
        self.rootfind, self.interp_points, self.save_positions = rootfind, interp_points, save_positions
        self.abstol, self.repeat_nudge = abstol, repeat_nudge


def test_vector_continuous_callback_reference_known_answers():
    """VectorContinuousCallback (callbacks.jl:239-362, 499-521, 557-640)
    against the reference's vector bouncing-ball tests
    (test/Integrators_I/ode_event_tests.jl:121-140, 168-190, 262-316): the same
    bounce times and saves as the scalar form, saveat at a bounce (or
    t - eps(t)) gives it twice, terminate! on the ground at sqrt(100/9.81);
    a ball in a box (floor and walls as two components) stays inside, each
    bounce's affect sees only its own component's direction; two components
    crossing at the same time are simultaneous."""
    def fall(du, u, p, t):
        du[0] = u[1]
        du[1] = -9.81

    def vcond(out, u, t, itg):
        out[0] = u[0]

    def vflip(itg, ev):
        if ev[0] == -1:
            itg.u[1] = -itg.u[1]
    t1 = math.sqrt(100 / 9.81)
    cb = _VCB(vcond, vflip, 1, interp_points=100)
    r = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), ccallbacks=[cb], save_everystep=False)
    rs = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), save_everystep=False,
                 ccallbacks=[_CCB(lambda u, t, itg: u[0], None, lambda itg: itg.u.__setitem__(1, -itg.u[1]),
                                  interp_points=100)])
    assert np.array_equal(r["t"], rs["t"]) and np.array_equal(r["u"], rs["u"]) and len(r["t"]) == 6
    t = r["t"][len(r["t"]) // 2 - 1]
    for sa in ([t], [t - np.spacing(t)]):
        r2 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, 15.0), ccallbacks=[cb], saveat=sa, save_start=False,
                     save_end=False)
        assert list(r2["t"]).count(t) == 2
    r5 = O.solve(fall, [50.0, 0.0], [0.0], (0.0, math.inf),
                 ccallbacks=[_VCB(vcond, lambda itg, ev: itg.terminate(when=ev[0] != 0), 1)])
    assert r5["retcode"] == 2 and abs(r5["t"][-1] - t1) < 1e-13 and abs(r5["u"][-1][0]) < 3e-12

    def box(du, u, p, t):
        du[0] = u[1]
        du[1] = -9.81
        du[2] = u[3]
        du[3] = 0.0
    seen = []

    def bcond(out, u, t, itg):
        out[0] = u[0]
        out[1] = (u[2] - 10.0) * u[2]

    def baff(itg, ev):
        seen.append(tuple(ev))
        if ev[0] == -1:
            itg.u[1] = -0.9 * itg.u[1]
        if ev[1] != 0:
            itg.u[3] = -itg.u[3]
    r6 = O.solve(box, [50.0, 0.0, 5.0, 4.0], [0.0], (0.0, 15.0), ccallbacks=[_VCB(bcond, baff, 2)],
                 save_everystep=True)
    assert r6["retcode"] == 1 and r6["u"][:, 2].min() >= -1e-12 and r6["u"][:, 2].max() <= 10 + 1e-12
    assert r6["u"][:, 0].min() >= -1e-12
    # the floor is a downcrossing of u[0]; (x - 10) x < 0 inside, so either wall is an upcrossing
    assert set(seen) == {(-1.0, 0.0), (0.0, 1.0)}
    # walls hit at x = 10 and 0 alternately every 2.5 s from x = 5 at 4 m/s: 1.25, 3.75, 6.25, ...
    wall_t = [tt for tt in r6["t"] if min(abs(tt - (1.25 + 2.5 * m)) for m in range(6)) < 1e-9]
    assert len(set(wall_t)) == 6
    seen.clear()
    r7 = O.solve(box, [50.0, 0.0, 5.0, 0.0], [0.0], (0.0, 4.0), save_everystep=False,
                 ccallbacks=[_VCB(lambda out, u, t, itg: (out.__setitem__(0, u[0]),
                                                          out.__setitem__(1, u[0] - 0.0 * t)), baff, 2)])
    assert (-1.0, -1.0) in seen   # two identical components: one simultaneous event


def test_saveat_interpolant_accuracy():
    # dense output (Tsit5 free 4th-order interpolant, Rodas 3rd-order H-rows) vs exp
    sa = list(np.linspace(0.05, 0.95, 19))
    for alg, tol in ((O.TSIT5, 1e-6), (O.RODAS5P, 1e-5)):
        r = O.solve_builtin(O.PROB_LINEAR, [1.0], [-2.0], (0.0, 1.0), alg=alg, saveat=sa, reltol=1e-8,
                            abstol=1e-10, save_start=False, save_end=False)
        assert np.allclose(r["t"], sa)
        assert np.max(np.abs(r["u"][:, 0] - np.exp(-2.0 * r["t"]))) < tol


def test_exp_blowup_retcode():
    """u' = exp(u), u0 = 0 on (0, 10): explodes at t = 1 -> MaxIters or Unstable."""
    r = O.solve_builtin(O.PROB_EXPBLOWUP, [0.0], [0.0], (0.0, 10.0), reltol=1e-8, abstol=1e-8)
    assert O.RC_NAMES[r["retcode"]] in ("MaxIters", "Unstable")
    r = O.solve_builtin(O.PROB_EXPBLOWUP, [0.0], [0.0], (0.0, 10.0), alg=O.RODAS5P, reltol=1e-8, abstol=1e-8)
    assert O.RC_NAMES[r["retcode"]] in ("MaxIters", "Unstable")


def test_maxiters_and_nan_retcodes():
    r = O.solve_builtin(O.PROB_LORENZ, [1.0, 0.0, 0.0], [10.0, 28.0, 8 / 3], (0.0, 10.0), maxiters=5)
    assert O.RC_NAMES[r["retcode"]] == "MaxIters" and r["naccept"] + r["nreject"] == 5
    # NaN in the state: initdt returns dtmin (initdt.jl "First function call produced NaNs"),
    # every attempt is rejected until dt <= eps(t) -> Unstable
    r = O.solve_builtin(O.PROB_LORENZ, [float("nan"), 0.0, 0.0], [10.0, 28.0, 8 / 3], (0.0, 1.0))
    assert O.RC_NAMES[r["retcode"]] == "Unstable"
    # a NaN step size is caught first by check_error -> DtNaN
    r = O.solve_builtin(O.PROB_LORENZ, [1.0, 0.0, 0.0], [10.0, 28.0, 8 / 3], (0.0, 1.0), dt=float("nan"))
    assert O.RC_NAMES[r["retcode"]] == "DtNaN"


def test_stats_nf_accounting():
    """Tsit5: nf = 1 (FSAL) + 2 (initdt) + 6 per attempt; Rodas5P with analytic J and
    AD time derivative: nf = 2 + 8 per attempt + 1 per new J (stats_tests.jl)."""
    rng = np.random.default_rng(5)
    for _ in range(5):
        u0 = rng.uniform(-1, 1, 3)
        p = [rng.uniform(8, 12), rng.uniform(0, 28), 8 / 3]
        r = O.solve_builtin(O.PROB_LORENZ, u0, p, (0.0, 3.0), reltol=1e-6, abstol=1e-6)
        assert r["nf"] == 3 + 6 * (r["naccept"] + r["nreject"])
        r = O.solve_builtin(O.PROB_LORENZ, u0, p, (0.0, 3.0), alg=O.RODAS5P, reltol=1e-6, abstol=1e-6)
        att = r["naccept"] + r["nreject"]
        assert r["nf"] == 2 + 8 * att + r["njacs"]
        assert r["nw"] == att and r["nsolve"] == 8 * att
        # every rejected attempt is followed by one attempt that reuses J
        assert r["njacs"] == att - r["nreject"]


def test_rodas_reuses_jacobian_after_reject():
    # do_newJW: J is recomputed every attempt except right after a rejection
    r = O.solve_builtin(O.PROB_ROBERTSON, [1.0, 0.0, 0.0], [0.04, 3e7, 1e4], (0.0, 1e5), alg=O.RODAS5P,
                        reltol=1e-6, abstol=1e-10)
    assert r["nreject"] > 0
    assert r["njacs"] == r["naccept"]


def test_robertson_conservation_and_radau():
    from scipy.integrate import solve_ivp
    k = [0.04, 3e7, 1e4]
    r = O.solve_builtin(O.PROB_ROBERTSON, [1.0, 0.0, 0.0], k, (0.0, 1e5), alg=O.RODAS5P, reltol=1e-6,
                        abstol=1e-10)
    assert abs(sum(r["u_final"]) - 1.0) < 1e-8
    ref = solve_ivp(lambda t, u: [-k[0] * u[0] + k[2] * u[1] * u[2],
                                  k[0] * u[0] - k[1] * u[1] ** 2 - k[2] * u[1] * u[2], k[1] * u[1] ** 2],
                    (0, 1e5), [1, 0, 0], method="Radau", rtol=1e-12, atol=1e-16)
    err = np.abs(r["u_final"] - ref.y[:, -1]) / (1e-10 + 1e-6 * np.abs(ref.y[:, -1]))
    assert err.max() < 50


def test_tsit5_tight_tolerance_vs_dop853():
    from scipy.integrate import solve_ivp
    p = [10.0, 28.0, 8 / 3]
    r = O.solve_builtin(O.PROB_LORENZ, [1.0, 0.0, 0.0], p, (0.0, 1.0), reltol=1e-11, abstol=1e-11)
    ref = solve_ivp(lambda t, u: [p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1], u[0] * u[1] - p[2] * u[2]],
                    (0, 1), [1, 0, 0], method="DOP853", rtol=1e-13, atol=1e-14)
    assert np.max(np.abs(r["u_final"] - ref.y[:, -1]) / np.abs(ref.y[:, -1])) < 1e-8


def test_vern9_tableau_order_conditions():
    """The Vern9 reference: all 486 order conditions through order 9 for b, the
    embedded pair b - btilde to order 8 only, row sums = c."""
    c, a, b, bt = O.vern9_tableau()
    A = [[mp.mpf(float(x)) for x in row] for row in a]
    B = [mp.mpf(float(x)) for x in b]
    for i in range(16):
        assert abs(mp.fsum(A[i]) - mp.mpf(float(c[i]))) < 1e-14
    assert max(_order_residuals(A, B, 9)) < 1e-12
    Bh = [B[i] - mp.mpf(float(bt[i])) for i in range(16)]
    assert max(_order_residuals(A, Bh, 8)) < 1e-12
    assert max(_order_residuals(A, Bh, 9)) > 1e-8


def test_vern9_reference_accuracy():
    """u' = 1.01 u to machine precision; Lorenz to t = 1 against scipy DOP853 at
    rtol 1e-13; an oscillator chain against the converged Rodas5P restatement."""
    from scipy.integrate import solve_ivp
    uf, st, ns = O.vern9_builtin(O.PROB_LINEAR, [0.5], [1.01], (0.0, 1.0))
    assert st == 0 and abs(uf[0] / (0.5 * math.exp(1.01)) - 1) < 1e-14
    p = [10.0, 28.0, 8 / 3]
    uf, st, _ = O.vern9_builtin(O.PROB_LORENZ, [1.0, 0.0, 0.0], p, (0.0, 1.0))
    ref = solve_ivp(lambda t, u: [p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1], u[0] * u[1] - p[2] * u[2]],
                    (0, 1), [1, 0, 0], method="DOP853", rtol=1e-13, atol=1e-14)
    assert st == 0 and np.max(np.abs(uf - ref.y[:, -1]) / np.abs(ref.y[:, -1])) < 1e-10
    # chain, m = 4 (n = 8), moderate stiffness
    u0 = [0.1, -0.05, 0.02, 0.0, 0, 0, 0, 0]
    pc = [100.0, 0.3, 0.5]
    uf, st, _ = O.vern9_builtin(O.PROB_CHAIN, u0, pc, (0.0, 2.0))
    r = O.solve_builtin(O.PROB_CHAIN, u0, pc, (0.0, 2.0), alg=O.RODAS5P, reltol=1e-10, abstol=1e-12)
    assert st == 0 and np.linalg.norm(uf - r["u_final"]) / np.linalg.norm(uf) < 1e-7


def test_lu_matches_numpy():
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 5, 8, 17):
        A = rng.normal(size=(n, n))
        b = rng.normal(size=n)
        x = O.lu_solve(A, b)
        assert np.allclose(A @ x, b, rtol=1e-10, atol=1e-10)
    # pivoting: a zero leading entry must still solve
    A = np.array([[0.0, 1.0], [1.0, 0.0]])
    assert np.allclose(O.lu_solve(A, [2.0, 3.0]), [3.0, 2.0])


def test_julia_scalar_semantics():
    for x in (1.0, 0.0, 1e300, 3.5e-310, 0.1, 123456.789, 2.0 ** -1022):
        assert O.jl_eps(x) == np.spacing(abs(x)) or (x == 0.0 and O.jl_eps(x) == 5e-324)
    assert math.isnan(O.jl_eps(float("inf")))
    # fast log2: exact at powers of two, within 2e-3 of log2 elsewhere
    for x in (1.0, 2.0, 0.5, 1024.0):
        assert O.fastlog2(x) == math.log2(x)
    xs = np.exp(np.linspace(-20, 20, 2001))
    assert max(abs(O.fastlog2(float(x)) - math.log2(x)) for x in xs) < 2e-3
    # fastpower: FastPower semantics (0 -> 0), Float32 result
    assert O.fastpower(0.0, 0.14) == 0.0
    assert O.fastpower(1.0, 0.14) == 1.0
    v = O.fastpower(0.37, 0.14)
    assert v == float(np.float32(v)) and abs(v - 0.37 ** 0.14) < 1e-3


def test_det_log_exp_accuracy():
    rng = np.random.default_rng(11)
    xs = np.exp(rng.uniform(-700, 700, 20000))
    ys = rng.uniform(-700, 700, 20000)
    el = max(abs(O.det_log(float(x)) - math.log(x)) / max(abs(math.log(x)), 1e-300) for x in xs)
    ee = max(abs(O.det_exp(float(y)) - math.exp(y)) / math.exp(y) for y in ys)
    assert el < 3e-16 and ee < 3e-16
    assert O.det_exp(0.0) == 1.0 and O.det_log(1.0) == 0.0


def test_builtin_ball_event_matches_traced_form():
    """The restatement's built-in bouncing ball (C floor event, the CPU
    baseline of the events workload) equals the same ContinuousCallback run
    through the Python callback path, bit for bit, and bounces first at the
    closed-form time (h0 + v0 t - 9.81 t^2 / 2 = 0)."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "ordinarydiffeq.jl_amd"))
    from odemi import problems as P
    d = P.ball_ensemble(N=32)
    for j in range(0, 32, 5):
        h0, v0 = d["u0"][0, j], d["u0"][1, j]
        a = O.solve_builtin(O.PROB_BALL, d["u0"][:, j], d["p"][:, j], (0.0, 10.0), alg=O.TSIT5, reltol=1e-6,
                            abstol=1e-6, save_start=False)
        b = O.solve(P.ball, d["u0"][:, j], d["p"][:, j], (0.0, 10.0), alg=O.TSIT5, reltol=1e-6, abstol=1e-6,
                    save_start=False, ccallbacks=[P.BALL_BOUNCE])
        assert np.array_equal(a["u_final"], b["u_final"]) and a["ncondition"] == b["ncondition"]
        assert a["naccept"] == b["naccept"] and a["nf"] == b["nf"]
        tb = (v0 + math.sqrt(v0 * v0 + 2 * 9.81 * h0)) / 9.81
        c = O.solve_builtin(O.PROB_BALL, d["u0"][:, j], d["p"][:, j], (0.0, tb + 0.5), alg=O.TSIT5, reltol=1e-6,
                            abstol=1e-6, save_start=False, save_everystep=True)
        vb = v0 - 9.81 * tb
        # just after the bounce the velocity is -restitution * impact velocity
        assert abs(c["u"][-1][1] - (-d["p"][0, j] * vb - 9.81 * (tb + 0.5 - tb))) < 1e-6


# ---------------------------------------------------------------------------
# Backward integration (tf < t0): tdir = sign(tf - t0) threads through
# initdt, the dt bounds, the tstop and saveat queues (stored as tdir * t) and
# the loop condition (solve.jl:268-273,401,904-946,968-985,1019-1123;
# integrator_utils.jl:268-324,340-414,1199-1256,1290-1314; initdt.jl:17-298).
# ---------------------------------------------------------------------------
def _lorenz_py(du, u, p, t):
    du[0] = p[0] * (u[1] - u[0])
    du[1] = u[0] * (p[1] - u[2]) - u[1]
    du[2] = u[0] * u[1] - p[2] * u[2]


def _lorenz_py_jac(J, u, p, t):
    J[0, 0], J[0, 1], J[0, 2] = -p[0], p[0], 0.0
    J[1, 0], J[1, 1], J[1, 2] = p[1] - u[2], -1.0, -u[0]
    J[2, 0], J[2, 1], J[2, 2] = u[1], u[0], -p[2]


def _forced(du, u, p, t):
    du[0] = -p[0] * u[0] + math.sin(3.0 * t)
    du[1] = u[0] - p[1] * u[1] * t


def _forced_jac(J, u, p, t):
    J[0, 0], J[0, 1] = -p[0], 0.0
    J[1, 0], J[1, 1] = 1.0, -p[1] * t


def _forced_tgrad(dT, u, p, t):
    dT[0] = 3.0 * math.cos(3.0 * t)
    dT[1] = -p[1] * u[1]


def _reversed(f):
    """v' = -f(v, -s): the time-reversed problem, v(s) = u(-s)."""
    def g(du, u, p, s):
        f(du, u, p, -s)
        du[:] = -du
    return g


def _reversed_mat(jf):
    def g(J, u, p, s):
        jf(J, u, p, -s)
        J[:] = -J
    return g


def _reversed_tgrad(tg):
    def g(dT, u, p, s):
        tg(dT, u, p, -s)   # d/ds (-f(v, -s)) = +f_t(v, -s)
    return g


@pytest.mark.parametrize("alg", [O.TSIT5, O.RODAS5P], ids=["tsit5", "rodas5p"])
@pytest.mark.parametrize("prob", ["lorenz", "forced"])
def test_backward_solve_is_the_forward_solve_of_the_reversed_problem(alg, prob):
    """Integrating u' = f(u, t) from T down to 0 is, step for step, integrating
    v' = -f(v, -s) from -T up to 0: IEEE arithmetic is symmetric under
    negation, so the two solves agree bit for bit (states, save times negated,
    step and function counts) — pinning every tdir factor to the forward path.
    With saveat, tstops and a user dtmax."""
    if prob == "lorenz":   # backward in time the Lorenz flow expands volumes by e^(13.7 t): a short span
        f, jac, tg, u0, p, T = _lorenz_py, _lorenz_py_jac, None, [1.0, -0.5, 20.0], [10.0, 28.0, 8 / 3], 0.5
        sa, ts, dtmax = [0.05, 0.2, 0.45, T + 1.0], [0.4, 0.15, -1.0], 0.03
    else:
        f, jac, tg, u0, p, T = _forced, _forced_jac, _forced_tgrad, [1.0, 0.3], [2.0, 0.7], 4.0
        sa, ts, dtmax = [0.3, 1.1, 2.9, T + 1.0], [2.5, 0.75, -1.0], 0.5
    kw = dict(alg=alg, reltol=1e-7, abstol=1e-9, saveat=sa, tstops=ts, dtmax=dtmax, save_start=True)
    back = O.solve(f, u0, p, (T, 0.0), jac=jac, tgrad=tg, **kw)
    kw_f = dict(kw, saveat=[-s for s in sa], tstops=[-s for s in ts])
    fwd = O.solve(_reversed(f), u0, p, (-T, 0.0), jac=_reversed_mat(jac),
                  tgrad=None if tg is None else _reversed_tgrad(tg), **kw_f)
    assert back["retcode"] == fwd["retcode"] == 1
    assert np.array_equal(back["u_final"], fwd["u_final"]) and back["t_final"] == -fwd["t_final"] == 0.0
    assert np.array_equal(back["t"], -fwd["t"]) and np.array_equal(back["u"], fwd["u"])
    assert [back[k] for k in ("nf", "naccept", "nreject", "njacs")] == [fwd[k] for k in ("nf", "naccept", "nreject",
                                                                                            "njacs")]
    # saveat inside [0, T) in descending order, tstops hit exactly, the end point last
    assert list(back["t"][:4]) == [T] + sorted(sa[:3], reverse=True) and back["t"][-1] == 0.0
    assert back["dt_init"] < 0


@pytest.mark.parametrize("alg", [O.TSIT5, O.RODAS5P], ids=["tsit5", "rodas5p"])
def test_backward_exponential_closed_form(alg):
    """u' = lam u from t = 2 back to 0: u(0) = u(2) exp(-2 lam)."""
    lam = -0.8
    r = O.solve(lambda du, u, p, t: du.__setitem__(0, p[0] * u[0]), [math.exp(2 * lam)], [lam], (2.0, 0.0),
                jac=lambda J, u, p, t: J.__setitem__((0, 0), p[0]), alg=alg, reltol=1e-10, abstol=1e-12)
    assert r["retcode"] == 1 and r["t_final"] == 0.0
    assert abs(r["u_final"][0] - 1.0) < 1e-8
