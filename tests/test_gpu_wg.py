"""GPU parity of the workgroup-per-trajectory Rodas5P kernel (uneven batches,
n up to 128) against the CPU restatement, through the C ABI."""
import numpy as np
import pytest

import oracle as O
import odemi
from odemi import problems as P

pytestmark = pytest.mark.gpu


def _chain(N, m_lo=4, m_hi=12, seed=4):
    d = P.chain_ensemble(N=N, seed=seed, m_lo=m_lo, m_hi=m_hi)
    d["tf"] = np.full(N, 1.0)
    return d


def _oracle(d, **kw):
    return O.ensemble_builtin(O.PROB_CHAIN, d["u0"], d["p"], d["t0"], d["tf"], nvec=d["nvec"], alg=O.RODAS5P,
                              reltol=d["reltol"], abstol=d["abstol"], save_start=False, **kw)


def test_chain_uneven_fp64_bit_identical():
    d = _chain(256)
    print("[case] chain uneven n 8..24 wg fp64", flush=True)
    ref = _oracle(d)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"])
    assert (out["retcode"] == 1).all()
    same = (out["stats"][1] == ref["naccept"]) & (out["stats"][2] == ref["nreject"])
    exact = np.all(out["u_final"] == ref["u_final"], axis=0)
    print(f"steps equal {same.mean():.4f} exact {exact.mean():.4f}")
    assert same.mean() >= 0.99 and exact.mean() >= 0.99
    assert (out["stats"][0] == ref["nf"]).all()


def test_chain_uneven_plain_engine_bit_identical():
    """The plain FP64 engine of uneven batches (one launch per step attempt, the
    integrator state in HBM between launches): final states, step counts and nf
    identical to the restatement and to the persistent FP64 workgroup engine,
    n = 8..64 in one batch, with save_start; a batch without nvec runs every
    trajectory at nmax."""
    d = _chain(96, m_lo=4, m_hi=32, seed=21)
    print("[case] chain uneven n 8..64 plain wg fp64", flush=True)
    ref = _oracle(d)
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], nvec=d["nvec"], save_start=True)
    pl = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(engine="plain"), **kw)
    ps = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], **kw)
    assert (pl["retcode"] == 1).all() and pl["launches"] > 100
    assert np.array_equal(pl["u_final"], ps["u_final"]) and np.array_equal(pl["stats"], ps["stats"])
    assert np.array_equal(pl["nsaved"], ps["nsaved"]) and (pl["nsaved"] == 2).all()
    assert np.array_equal(pl["u_out"][:2], ps["u_out"][:2]) and np.array_equal(pl["t_out"][:2], ps["t_out"][:2])
    same = (pl["stats"][1] == ref["naccept"]) & (pl["stats"][2] == ref["nreject"])
    exact = np.all(pl["u_final"] == ref["u_final"], axis=0)
    print(f"steps equal {same.mean():.4f} exact {exact.mean():.4f} launches {pl['launches']}")
    assert same.mean() >= 0.99 and exact.mean() >= 0.99
    assert (pl["stats"][0] == ref["nf"]).all()
    # no nvec: every trajectory at the compiled n
    e = _chain(32, m_lo=6, m_hi=6, seed=22)
    print("[case] chain n 12 plain wg fp64 without nvec", flush=True)
    a = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), e["u0"], e["p"], e["t0"], e["tf"],
                           odemi.EnsembleMI355X(engine="plain"), reltol=e["reltol"], abstol=e["abstol"])
    b = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), e["u0"], e["p"], e["t0"], e["tf"], reltol=e["reltol"],
                           abstol=e["abstol"], nvec=e["nvec"])
    assert np.array_equal(a["u_final"], b["u_final"]) and np.array_equal(a["stats"], b["stats"])


def test_chain_large_n_fp64_and_mixed():
    d = _chain(64, m_lo=48, m_hi=64, seed=9)
    print("[case] chain n 96..128 wg fp64 + mixed", flush=True)
    ref = _oracle(d)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"])
    rel = np.abs(out["u_final"] - ref["u_final"]) / np.maximum(np.abs(ref["u_final"]), 1e-300)
    same = (out["stats"][1] == ref["naccept"]) & (out["stats"][2] == ref["nreject"])
    print(f"fp64 steps equal {same.mean():.4f} exact {np.mean(np.all(out['u_final'] == ref['u_final'], axis=0)):.4f}")
    assert same.mean() >= 0.99
    assert np.quantile(rel.max(axis=0), 0.99) <= 1e-10
    mx = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), reltol=d["reltol"], abstol=d["abstol"],
                            nvec=d["nvec"])
    assert (mx["retcode"] == 1).all()
    err = np.abs(mx["u_final"] - ref["u_final"]) / (d["abstol"] + d["reltol"] * np.abs(ref["u_final"]))
    print(f"mixed scaled err p99 {np.quantile(err.max(axis=0), 0.99):.3e} max {err.max():.3e}")
    assert np.quantile(err.max(axis=0), 0.99) <= 10.0


def test_chain_uneven_mixed_gauss_jordan_within_tolerance():
    """Mixed variant (register-tiled Gauss-Jordan, FP32, FP64 refinement) on an
    uneven batch n = 2..48 against the FP64 restatement."""
    d = _chain(192, m_lo=1, m_hi=24, seed=12)
    print("[case] chain uneven n 2..48 wg mixed", flush=True)
    ref = _oracle(d)
    mx = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), reltol=d["reltol"], abstol=d["abstol"],
                            nvec=d["nvec"])
    assert (mx["retcode"] == 1).all()
    err = np.abs(mx["u_final"] - ref["u_final"]) / (d["abstol"] + d["reltol"] * np.abs(ref["u_final"]))
    same = (mx["stats"][1] == ref["naccept"]).mean()
    print(f"mixed scaled err p99 {np.quantile(err.max(axis=0), 0.99):.3e} max {err.max():.3e} same steps {same:.3f}")
    assert np.quantile(err.max(axis=0), 0.99) <= 10.0
    assert same >= 0.8   # the mixed arithmetic may move a step boundary; the error bound is the criterion


def test_chain_mid_n_mixed_gauss_jordan_tiles_4_to_6():
    """The middle Gauss-Jordan tiles (n = 50..94: R = 4, 5, 6) in one uneven
    mixed batch, against the FP64 restatement."""
    d = _chain(96, m_lo=25, m_hi=47, seed=21)
    r = (d["nvec"] + 15) // 16
    assert set(np.unique(r)) == {4, 5, 6}
    print("[case] chain uneven n 50..94 wg mixed", flush=True)
    ref = _oracle(d)
    mx = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), reltol=d["reltol"], abstol=d["abstol"],
                            nvec=d["nvec"])
    assert (mx["retcode"] == 1).all()
    err = np.abs(mx["u_final"] - ref["u_final"]) / (d["abstol"] + d["reltol"] * np.abs(ref["u_final"]))
    print(f"mixed scaled err p99 {np.quantile(err.max(axis=0), 0.99):.3e} max {err.max():.3e}")
    assert np.quantile(err.max(axis=0), 0.99) <= 10.0


@pytest.mark.parametrize("prec,width", [("fp64", 0.01), ("mixed", 1e-4)])
def test_chain_isoutofdomain_workgroup(prec, width):
    """isoutofdomain in the workgroup kernels: a step whose end time falls in
    (0.3, 0.3 + width) is rejected with dt *= qmin (handle_step_rejection!).
    The 0.01 window also traps most trajectories below 0.3 until dt underflows
    (the restatement's return codes and counts, bit for bit, FP64); the 1e-4
    window only adds rejections (mixed, within tolerance)."""
    d = _chain(128)
    hi = 0.3 + width
    isout = lambda u, p, t: (t > 0.3) & (t < hi)
    print(f"[case] chain isoutofdomain wg {prec} width {width}", flush=True)
    # the predicate reads t only: a (1, 1) view stays inside every trajectory's state
    ref = _oracle(d, isoutofdomain=lambda u, p, t: 0.3 < t < hi, isout_shape=(1, 1))
    base = _oracle(d)
    assert (ref["nreject"] > base["nreject"]).any()   # the predicate does reject steps
    ea = odemi.EnsembleMI355X(precision=prec)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], ea, reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"], isoutofdomain=isout)
    assert (out["retcode"] == ref["retcode"]).all()
    if prec == "fp64":
        assert (out["stats"][1] == ref["naccept"]).all() and (out["stats"][2] == ref["nreject"]).all()
        assert np.array_equal(out["u_final"], ref["u_final"], equal_nan=True)
    else:
        assert (ref["retcode"] == 1).all()
        err = np.abs(out["u_final"] - ref["u_final"]) / (d["abstol"] + d["reltol"] * np.abs(ref["u_final"]))
        print(f"mixed scaled err p99 {np.quantile(err.max(axis=0), 0.99):.3e}")
        assert np.quantile(err.max(axis=0), 0.99) <= 10.0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_wg_repeat_solves_bit_identical(prec):
    """The same uneven chain batch solved twice in one process (and once more
    with a predicate that never fires, which compiles a second module) gives
    bit-identical results: nothing in the workgroup kernels depends on state
    left by an earlier launch."""
    d = _chain(96, m_lo=4, m_hi=40, seed=5)
    print(f"[case] chain wg repeat {prec}", flush=True)
    ea = odemi.EnsembleMI355X(precision=prec)
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], nvec=d["nvec"])
    a = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], ea, **kw)
    b = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], ea, **kw)
    c = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], ea,
                           isoutofdomain=lambda u, p, t: t > 1e9, **kw)
    for o in (b, c):
        assert np.array_equal(a["u_final"], o["u_final"])
        assert np.array_equal(a["stats"], o["stats"]) and np.array_equal(a["retcode"], o["retcode"])


def test_wg_kernel_robertson_matches_lane_kernel():
    """n = 3 through the workgroup kernel (nvec given) equals the lane kernel bit for bit."""
    d = P.robertson_ensemble(N=128)
    print("[case] robertson via wg kernel", flush=True)
    lane = odemi.solve_arrays(P.ROBERTSON, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                              abstol=d["abstol"])
    wg = odemi.solve_arrays(P.ROBERTSON, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                            abstol=d["abstol"], nvec=np.full(128, 3, dtype=np.int32))
    assert np.array_equal(lane["u_final"], wg["u_final"])
    assert np.array_equal(lane["stats"], wg["stats"])


def test_wg_saveat_bit_identical():
    d = _chain(32)
    sa = np.linspace(0.1, 0.9, 5)
    print("[case] chain wg saveat", flush=True)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"], saveat=sa, save_start=True, save_end=True)
    for j in range(0, 32, 5):
        n = int(d["nvec"][j])
        r = O.solve_builtin(O.PROB_CHAIN, d["u0"][:n, j], d["p"][:, j], (0.0, 1.0), alg=O.RODAS5P,
                            reltol=d["reltol"], abstol=d["abstol"], saveat=sa, save_start=True, save_end=True)
        ns = int(out["nsaved"][j])
        assert ns == r["nsaved"]
        assert np.array_equal(out["t_out"][:ns, j], r["t"])
        assert np.array_equal(out["u_out"][:ns, :n, j], r["u"])


def test_mixed_saveat_within_tolerance_lane_and_workgroup():
    """Mixed-precision Rodas5P with saveat (lane kernel: Robertson; workgroup
    Gauss-Jordan kernel: uneven chains) against the FP64 engine at the same
    save points: same number of points, within a few tolerance units."""
    print("[case] mixed saveat rodas lane + wg", flush=True)
    d = P.robertson_ensemble(N=128)
    sa = np.array([1e-3, 1.0, 10.0, 1e3, 1e5])
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], saveat=sa, save_start=True, save_end=True)
    o64 = odemi.solve_arrays(P.ROBERTSON, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], **kw)
    om = odemi.solve_arrays(P.ROBERTSON, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), **kw)
    assert np.array_equal(o64["nsaved"], om["nsaved"]) and np.array_equal(o64["t_out"], om["t_out"])
    ns = int(o64["nsaved"].min())
    err = np.abs(om["u_out"][:ns] - o64["u_out"][:ns]) / (d["abstol"] + d["reltol"] * np.abs(o64["u_out"][:ns]))
    assert np.quantile(err, 0.99) < 10.0
    c = _chain(64, m_lo=4, m_hi=40, seed=17)
    sa = np.linspace(0.1, 0.9, 5)
    kw = dict(reltol=c["reltol"], abstol=c["abstol"], saveat=sa, save_start=True, save_end=True, nvec=c["nvec"])
    o64 = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), c["u0"], c["p"], c["t0"], c["tf"], **kw)
    om = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), c["u0"], c["p"], c["t0"], c["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), **kw)
    assert (om["retcode"] == 1).all()
    assert np.array_equal(o64["nsaved"], om["nsaved"]) and np.array_equal(o64["t_out"], om["t_out"])
    ns = int(o64["nsaved"].min())
    err = np.abs(om["u_out"][:ns] - o64["u_out"][:ns]) / (c["abstol"] + c["reltol"] * np.abs(o64["u_out"][:ns]))
    print(f"chain saveat mixed scaled err p99 {np.quantile(err, 0.99):.3e}")
    assert np.quantile(err, 0.99) < 10.0


@pytest.fixture(scope="module")
def chain_cfg4_sample():
    """128 trajectories of the 16,384-chain workload as stated: m ~ U{4..64}
    (n = 8..128), tspan (0, 10), reltol = abstol = 1e-6; the FP64 restatement
    and the Vern9 (reltol 1e-12) solutions of each."""
    d = P.chain_ensemble(N=128, seed=40)
    assert d["nvec"].min() <= 24 and d["nvec"].max() >= 112 and (d["tf"] == 10.0).all()
    ref = _oracle(d)
    v9 = O.vern9_ensemble(O.PROB_CHAIN, d["u0"], d["p"], d["t0"], d["tf"], nvec=d["nvec"])
    assert (v9["status"] == 0).all()
    return d, ref, v9["u_final"]


def test_chain_full_span_uneven_one_launch_fp64_bit_identical(chain_cfg4_sample):
    """FP64 workgroup engine, n = 8..128 in one launch over tspan (0, 10):
    final states and step counts bit-identical to the restatement."""
    d, ref, _ = chain_cfg4_sample
    print(f"[case] chain cfg4 sample n {d['nvec'].min()}..{d['nvec'].max()} tspan (0,10) wg fp64", flush=True)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"])
    assert out["launches"] == 1
    assert (out["retcode"] == 1).all() and (ref["retcode"] == 1).all()
    same = (out["stats"][1] == ref["naccept"]) & (out["stats"][2] == ref["nreject"])
    exact = np.all(out["u_final"] == ref["u_final"], axis=0)
    print(f"steps equal {same.mean():.4f} exact {exact.mean():.4f} mean attempts "
          f"{(ref['naccept'] + ref['nreject']).mean():.0f}")
    assert same.mean() >= 0.99 and exact.mean() >= 0.99
    assert (out["stats"][0] == ref["nf"]).all()


def test_chain_full_span_uneven_one_launch_mixed_accuracy(chain_cfg4_sample):
    """Mixed (FP32 Gauss-Jordan + FP64 refinement) engine on the same sample:
    p99 of the tolerance-scaled error against the FP64 restatement at most 10,
    and its norm-wise error against Vern9 no larger than 1.5x the FP64
    restatement's (median and p99)."""
    d, ref, v9 = chain_cfg4_sample
    print("[case] chain cfg4 sample tspan (0,10) wg mixed", flush=True)
    mx = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(precision="mixed"), reltol=d["reltol"], abstol=d["abstol"],
                            nvec=d["nvec"])
    assert mx["launches"] == 1 and (mx["retcode"] == 1).all()
    err = np.abs(mx["u_final"] - ref["u_final"]) / (d["abstol"] + d["reltol"] * np.abs(ref["u_final"]))
    nrm = np.maximum(np.linalg.norm(v9, axis=0), 1e-300)
    e64 = np.linalg.norm(ref["u_final"] - v9, axis=0) / nrm
    emx = np.linalg.norm(mx["u_final"] - v9, axis=0) / nrm
    print(f"mixed scaled err vs fp64 p99 {np.quantile(err.max(axis=0), 0.99):.3e}; vs Vern9: fp64 median "
          f"{np.median(e64):.3e} p99 {np.quantile(e64, 0.99):.3e}, mixed median {np.median(emx):.3e} "
          f"p99 {np.quantile(emx, 0.99):.3e}")
    assert np.quantile(err.max(axis=0), 0.99) <= 10.0
    assert np.median(emx) <= 1.5 * np.median(e64) and np.quantile(emx, 0.99) <= 1.5 * np.quantile(e64, 0.99)


def _wall(itg):
    m = len(itg.u) // 2   # the first mass's velocity, reversed and halved
    itg.u[m] = -0.5 * itg.u[m]


def _kick(itg):
    m = len(itg.u) // 2
    itg.u[m] = itg.u[m] + itg.p[1]


class _PyCB:
    """The same callback for the CPU restatement (plain Python on numpy)."""

    def __init__(self, cb):
        self.__dict__.update(cb.__dict__)


def _chain_events(N, seed):
    d = _chain(N, m_lo=4, m_hi=12, seed=seed)
    d["p"][0] = np.random.Generator(np.random.PCG64(seed + 100)).uniform(1.0, 100.0, N)   # a few events each
    return d


_WG_CBS = odemi.CallbackSet(
    odemi.ContinuousCallback(lambda u, t, itg: u[0], None, affect_neg=_wall),
    odemi.DiscreteCallback(lambda u, t, itg: odemi.sym_in(t, [0.25, 0.5]), _kick))


@pytest.mark.parametrize("saves", ["saveat", "final"])
def test_wg_callbacks_bit_identical(saves):
    """Callbacks in the workgroup kernels (uneven chains n = 8..24, FP64 LU):
    a ContinuousCallback on the first mass's position (downcrossings only: its
    velocity reversed and halved; event search on the Rosenbrock dense output,
    root to adjacent floats, the step's stages recomputed for the event time)
    and a DiscreteCallback kicking that velocity at tstops 0.25 and 0.5, with
    forced saves around each affect.  Bit-identical to the CPU restatement:
    save times and states, final states, retcodes, nf / naccept / nreject /
    ncondition."""
    d = _chain_events(48, 31)
    m = dict(saveat=[0.1, 0.6], save_start=False, save_end=True) if saves == "saveat" else \
        dict(save_start=False, save_end=True)
    print(f"[case] chain wg callbacks fp64 {saves}", flush=True)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"], tstops=[0.25, 0.5], callback=_WG_CBS,
                             max_saves=512, **m)
    ccb = [_PyCB(c) for c in _WG_CBS.continuous_callbacks]
    dcb = [_PyCB(c) for c in _WG_CBS.discrete_callbacks]
    events = 0
    for j in range(d["u0"].shape[1]):
        n = int(d["nvec"][j])
        r = O.solve_builtin(O.PROB_CHAIN, d["u0"][:n, j], d["p"][:, j], (0.0, float(d["tf"][j])), alg=O.RODAS5P,
                            reltol=d["reltol"], abstol=d["abstol"], tstops=[0.25, 0.5], callbacks=dcb,
                            ccallbacks=ccb, **m)
        ns = int(out["nsaved"][j])
        assert int(out["retcode"][j]) == r["retcode"] == 1 and ns == r["nsaved"], j
        assert np.array_equal(out["t_out"][:ns, j], r["t"]), j
        assert np.array_equal(out["u_out"][:ns, :n, j], r["u"]), j
        assert np.array_equal(out["u_final"][:n, j], r["u_final"]) and out["t_final"][j] == r["t_final"], j
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2], st[6]) == (r["nf"], r["naccept"], r["nreject"], r["ncondition"]), j
        events += ns
    print(f"saved points {events} over {d['u0'].shape[1]} chains")
    assert events > 5 * d["u0"].shape[1]   # two kicks (4 forced saves) plus wall events


def test_wg_callbacks_terminate_and_mixed():
    """terminate! from a ContinuousCallback's affect (the first upcrossing of
    the first mass's position past 0.02) ends each chain at the event time with
    Terminated, the end state saved; FP64 bit-identical to the restatement, the
    mixed (Gauss-Jordan) engine with the same retcodes and event times within
    1e-6 of FP64's."""
    d = _chain_events(40, 32)
    d["u0"][0] = -0.05   # every chain starts below the threshold
    cb = odemi.ContinuousCallback(lambda u, t, itg: u[0] - 0.02, lambda itg: odemi.terminate(itg), None,
                                  rootfind="right")
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], nvec=d["nvec"], callback=cb, save_start=False, save_end=True,
              max_saves=8)
    print("[case] chain wg callback terminate fp64 + mixed", flush=True)
    o64 = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], **kw)
    omx = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                             odemi.EnsembleMI355X(precision="mixed"), **kw)
    codes = set()
    for j in range(d["u0"].shape[1]):
        n = int(d["nvec"][j])
        r = O.solve_builtin(O.PROB_CHAIN, d["u0"][:n, j], d["p"][:, j], (0.0, float(d["tf"][j])), alg=O.RODAS5P,
                            reltol=d["reltol"], abstol=d["abstol"], ccallbacks=[_PyCB(cb)], save_start=False,
                            save_end=True)
        assert int(o64["retcode"][j]) == r["retcode"] and o64["t_final"][j] == r["t_final"], j
        assert np.array_equal(o64["u_final"][:n, j], r["u_final"]), j
        assert int(o64["nsaved"][j]) == r["nsaved"] and np.array_equal(o64["t_out"][:r["nsaved"], j], r["t"]), j
        assert int(omx["retcode"][j]) == r["retcode"], j
        assert abs(omx["t_final"][j] - r["t_final"]) <= 1e-6, (j, omx["t_final"][j], r["t_final"])
        codes.add(r["retcode"])
    assert 2 in codes


def test_wg_callbacks_plain_engine_and_server_equal_persistent():
    """The workgroup callbacks (ContinuousCallback wall + DiscreteCallback kicks
    at tstops) through the plain FP64 engine (one launch per step attempt, the
    event state in HBM between launches) and through an uneven-n server (the
    kicks' tstops passed per request, several requests in flight): final
    states, retcodes and all seven stats counters equal the persistent
    engine's, bit for bit."""
    d = _chain_events(40, 33)
    nv = np.asarray(d["nvec"])
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], nvec=nv, tstops=[0.25, 0.5], callback=_WG_CBS,
              save_start=False, save_end=True)
    print("[case] chain wg callbacks persistent vs plain", flush=True)
    ref = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], odemi.EnsembleMI355X(),
                             max_saves=512, **kw)
    pl = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                            odemi.EnsembleMI355X(engine="plain"), max_saves=512, **kw)
    for j in range(len(nv)):
        assert np.array_equal(pl["u_final"][:nv[j], j], ref["u_final"][:nv[j], j]), j
    assert np.array_equal(pl["stats"], ref["stats"]) and np.array_equal(pl["retcode"], ref["retcode"])
    assert (ref["stats"][6] > 0).all()
    print("[case] chain wg callbacks server", flush=True)
    n = d["u0"].shape[0]
    ns = sorted(set(int(x) for x in nv))
    parts = (np.arange(0, 13), np.arange(13, 40))
    with odemi.EnsembleServer(P.CHAIN, n, d["p"].shape[0], odemi.Rodas5P(), reltol=d["reltol"], abstol=d["abstol"],
                              tick_attempts=8, ns=ns, max_tstops=2, callback=_WG_CBS) as srv:
        reqs = [srv.submit(d["u0"][:, idx], d["p"][:, idx], d["t0"][idx], d["tf"][idx], nvec=nv[idx],
                           tstops=[0.25, 0.5]) for idx in parts]
        for r in reqs:
            assert r.wait(timeout=300)
    for idx, r in zip(parts, reqs):
        for c, j in enumerate(idx):
            assert np.array_equal(r.u_final[:nv[j], c], ref["u_final"][:nv[j], j]), j
        assert np.array_equal(r.stats, ref["stats"][:, idx]) and np.array_equal(r.retcode, ref["retcode"][idx])


def _walls_cond(out, u, t, itg):
    out[0] = u[0]                 # the first mass at its rest position
    out[1] = u[1] - 0.02          # the second mass past +0.02


def _walls_affect(itg, ev):
    m = len(itg.u) // 2
    itg.u[m] = odemi.sym_ifelse(ev[0] == -1, -0.5 * itg.u[m], itg.u[m])
    itg.u[m + 1] = odemi.sym_ifelse(ev[1] != 0, -itg.u[m + 1], itg.u[m + 1])
    odemi.terminate(itg, when=itg.t > 0.8)


_WG_VCBS = odemi.CallbackSet(odemi.VectorContinuousCallback(_walls_cond, _walls_affect, 2),
                             odemi.ContinuousCallback(lambda u, t, itg: u[0] - 1.0, None, None))


@pytest.mark.parametrize("saves", ["everystep", "final"])
def test_wg_vector_continuous_callbacks_bit_identical(saves):
    """A VectorContinuousCallback in the workgroup kernels (uneven chains n =
    8..24, FP64 LU): two components (the first mass's downcrossing of its rest
    position halves and reverses its velocity, the second mass crossing +0.02
    either way reverses its own), simultaneous events as the lane engines
    handle them, a conditional terminate! once t > 0.8, in a CallbackSet with
    a scalar ContinuousCallback that is never an event; bit-identical to the
    CPU restatement: save times and states, final states, retcodes, nf /
    naccept / nreject / ncondition."""
    d = _chain_events(40, 35)
    m = dict(save_everystep=True, save_start=True, save_end=True) if saves == "everystep" else \
        dict(save_start=False, save_end=True)
    print(f"[case] chain wg vector callbacks fp64 {saves}", flush=True)
    out = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"], nvec=d["nvec"], callback=_WG_VCBS, max_saves=4096, **m)
    ccb = [_PyCB(c) for c in _WG_VCBS.continuous_callbacks]
    codes, events = set(), 0
    for j in range(d["u0"].shape[1]):
        n = int(d["nvec"][j])
        r = O.solve_builtin(O.PROB_CHAIN, d["u0"][:n, j], d["p"][:, j], (0.0, float(d["tf"][j])), alg=O.RODAS5P,
                            reltol=d["reltol"], abstol=d["abstol"], ccallbacks=ccb, **m)
        ns = int(out["nsaved"][j])
        assert int(out["retcode"][j]) == r["retcode"] and ns == r["nsaved"], j
        assert np.array_equal(out["t_out"][:ns, j], r["t"]), j
        assert np.array_equal(out["u_out"][:ns, :n, j], r["u"]), j
        assert np.array_equal(out["u_final"][:n, j], r["u_final"]) and out["t_final"][j] == r["t_final"], j
        st = out["stats"][:, j]
        assert (st[0], st[1], st[2], st[6]) == (r["nf"], r["naccept"], r["nreject"], r["ncondition"]), j
        codes.add(r["retcode"])
    print("retcodes", sorted(codes))
    assert 2 in codes   # some chains terminate


def _vterm_cond(out, u, t, itg):
    out[0] = u[0] - 0.02          # the first mass past +0.02
    out[1] = u[1] - 10.0          # never reached: a second component that stays silent


def _vterm_affect(itg, ev):
    odemi.terminate(itg, when=ev[0] == 1)   # upcrossings only


def test_wg_vector_callbacks_mixed_and_plain_engine():
    """A VectorContinuousCallback (RightRootFind) whose affect terminates on an
    upcrossing of its first component: the mixed Gauss-Jordan engine ends the
    same chains with the same retcodes, at event times within 1e-6 of the FP64
    engine's; and the mirrored-wall vector callbacks through the plain FP64
    workgroup engine (the event state in HBM between launches) equal the
    persistent engine's final states, retcodes and stats bit for bit."""
    d = _chain_events(40, 36)
    d["u0"][0] = -0.05
    cb = odemi.VectorContinuousCallback(_vterm_cond, _vterm_affect, 2, rootfind="right")
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], nvec=d["nvec"], callback=cb, save_start=False, save_end=True,
              max_saves=8)
    print("[case] chain wg vector callback terminate fp64 + mixed", flush=True)
    o64 = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], **kw)
    omx = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                             odemi.EnsembleMI355X(precision="mixed"), **kw)
    assert 2 in set(int(x) for x in o64["retcode"])
    assert np.array_equal(omx["retcode"], o64["retcode"])
    term = o64["retcode"] == 2
    assert np.abs(omx["t_final"][term] - o64["t_final"][term]).max() <= 1e-6
    d = _chain_events(40, 37)
    nv = np.asarray(d["nvec"])
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], nvec=nv, callback=_WG_VCBS, save_start=False, save_end=True)
    print("[case] chain wg vector callbacks plain fp64 vs persistent", flush=True)
    per = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"], **kw)
    pln = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                             odemi.EnsembleMI355X(engine="plain"), **kw)
    for j in range(len(nv)):
        n = int(nv[j])
        assert np.array_equal(pln["u_final"][:n, j], per["u_final"][:n, j]), j
    assert np.array_equal(pln["retcode"], per["retcode"]) and np.array_equal(pln["t_final"], per["t_final"])
    assert np.array_equal(pln["stats"], per["stats"])
