"""GPU tests of the continuous-batching server (odemi_server_* through the C ABI):
a request's result is the same alone and in a mix of concurrent requests, and
the same as the batch engine and the CPU restatement on the same inputs;
admission under a slot budget smaller than the offered load."""
import threading

import numpy as np
import pytest

import oracle as O
import odemi
from odemi import problems as P

pytestmark = pytest.mark.gpu


def _cols(d, idx):
    return (d["u0"][:, idx], d["p"][:, idx], d["t0"][idx], d["tf"][idx])


CASES = [
    ("tsit5-fp64", odemi.Tsit5(), "fp64", P.LORENZ, O.PROB_LORENZ, lambda: P.lorenz_large(N=768)),
    ("tsit5-mixed", odemi.Tsit5(), "mixed", P.LORENZ, O.PROB_LORENZ, lambda: P.lorenz_large(N=768)),
    ("rodas-fp64", odemi.Rodas5P(), "fp64", P.ROBERTSON, O.PROB_ROBERTSON, lambda: P.robertson_ensemble(N=768)),
    ("rodas-mixed", odemi.Rodas5P(), "mixed", P.ROBERTSON, O.PROB_ROBERTSON, lambda: P.robertson_ensemble(N=768)),
]


@pytest.mark.parametrize("name,alg,prec,fun,pid,dgen", CASES, ids=[c[0] for c in CASES])
def test_alone_equals_in_mix(name, alg, prec, fun, pid, dgen):
    d = dgen()
    n, N = d["u0"].shape
    print(f"[case] serving alone vs mix {name} N={N}", flush=True)
    probe = np.arange(40, 104)                       # the request under test: 64 trajectories
    kw = dict(precision=prec, reltol=d["reltol"], abstol=d["abstol"], tick_attempts=16)
    with odemi.EnsembleServer(fun, n, d["p"].shape[0], alg, **kw) as srv:
        alone = srv.submit(*_cols(d, probe))
        assert alone.wait(timeout=120)
        # the same trajectories again, now among concurrent requests of ragged sizes
        reqs, k = [], 0
        others = np.setdiff1d(np.arange(N), probe)
        sizes = [1, 7, 64, 3, 130, 31, 256, 2]
        for i, sz in enumerate(sizes):
            idx = others[k:k + sz]
            k += sz
            reqs.append(srv.submit(*_cols(d, idx)))
            if i == 3:
                mixed = srv.submit(*_cols(d, probe))
        for r in reqs + [mixed]:
            assert r.wait(timeout=120)
        info = srv.info()
        assert info["finished"] == len(sizes) + 2
    assert np.array_equal(alone.u_final, mixed.u_final)
    assert np.array_equal(alone.retcode, mixed.retcode) and np.array_equal(alone.stats, mixed.stats)
    # identical to the batch engine on the same trajectories
    ea = odemi.EnsembleMI355X(precision=prec)
    b = odemi.solve_arrays(fun, alg, *_cols(d, probe), ea, reltol=d["reltol"], abstol=d["abstol"])
    assert np.array_equal(alone.u_final, b["u_final"])
    assert np.array_equal(alone.stats, b["stats"]) and np.array_equal(alone.retcode, b["retcode"])
    # and, for FP64, to the CPU restatement
    if prec == "fp64":
        for c, j in enumerate(probe[::9]):
            r = O.solve_builtin(pid, d["u0"][:, j], d["p"][:, j], (d["t0"][j], d["tf"][j]), alg=alg.code,
                                reltol=d["reltol"], abstol=d["abstol"], save_start=False)
            col = int(np.where(probe == j)[0][0])
            assert np.array_equal(alone.u_final[:, col], r["u_final"])
            assert int(alone.retcode[col]) == r["retcode"]
            assert alone.stats[1, col] == r["naccept"] and alone.stats[2, col] == r["nreject"]


def test_admission_under_slot_budget_and_threads():
    """More offered trajectories than slots: requests queue, are admitted as
    slots free up, and every result matches the batch engine.  Submissions come
    from several host threads at once."""
    d = P.lorenz_large(N=1200)
    n, N = d["u0"].shape
    print("[case] serving admission budget 300 slots, 4 submit threads", flush=True)
    ref = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"], reltol=d["reltol"],
                             abstol=d["abstol"])
    chunks = np.array_split(np.arange(N), 24)
    out = [None] * len(chunks)
    with odemi.EnsembleServer(P.LORENZ, n, 3, odemi.Tsit5(), reltol=d["reltol"], abstol=d["abstol"],
                              max_slots=300, tick_attempts=8) as srv:
        assert srv.info()["slots"] == 300

        def client(t):
            for c in range(t, len(chunks), 4):
                out[c] = srv.submit(*_cols(d, chunks[c]))

        th = [threading.Thread(target=client, args=(t,)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for r in out:
            assert r.wait(timeout=120)
        assert srv.info()["finished"] == len(chunks)
        for r in out:
            ts, ta, td = r.times()
            assert ts <= ta <= td
    with pytest.raises(RuntimeError):
        out[0].times()                                   # the server is closed
    for c, idx in enumerate(chunks):
        assert np.array_equal(out[c].u_final, ref["u_final"][:, idx])
        assert np.array_equal(out[c].stats, ref["stats"][:, idx])


def test_hbm_budget_sizes_pool_and_rejects_oversized_request():
    print("[case] serving hbm budget", flush=True)
    with odemi.EnsembleServer(P.LORENZ, 3, 3, odemi.Tsit5(), hbm_budget_bytes=1 << 20) as srv:
        slots = srv.info()["slots"]
        assert 0 < slots < 10000
        d = P.lorenz_large(N=slots + 1)
        with pytest.raises(RuntimeError):
            srv.submit(*_cols(d, np.arange(slots + 1)))
        r = srv.submit(*_cols(d, np.arange(slots)))
        assert r.wait(timeout=120) and (r.retcode == 1).all()
        r.release()                                        # the record is gone, the outputs stay
        with pytest.raises(RuntimeError):
            r.times()
        assert (r.retcode == 1).all()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_chain_serving_uneven_n_alone_equals_in_mix(prec):
    """Uneven-n Rodas5P requests through the workgroup tick kernel: a request's
    result is the same alone, in a mix, and in the one-launch batch engine
    (and, for FP64, the CPU restatement), across many short ticks."""
    N = 96
    d = P.chain_ensemble(N=N, seed=21, m_lo=2, m_hi=16)
    d["tf"] = np.full(N, 0.5)
    nmax = d["u0"].shape[0]
    ns = sorted(set(int(x) for x in d["nvec"]))
    print(f"[case] chain serving {prec} n {ns[0]}..{ns[-1]}", flush=True)
    cols = lambda idx: (d["u0"][:, idx], d["p"][:, idx], d["t0"][idx], d["tf"][idx])  # noqa: E731
    probe = np.arange(10, 26)
    kw = dict(precision=prec, reltol=d["reltol"], abstol=d["abstol"], tick_attempts=8, ns=ns)
    with odemi.EnsembleServer(P.CHAIN, nmax, 3, odemi.Rodas5P(), **kw) as srv:
        alone = srv.submit(*cols(probe), nvec=d["nvec"][probe])
        assert alone.wait(timeout=150)
        others = np.setdiff1d(np.arange(N), probe)
        reqs, k = [], 0
        for i, sz in enumerate([1, 5, 20, 3, 30]):
            idx = others[k:k + sz]
            k += sz
            reqs.append(srv.submit(*cols(idx), nvec=d["nvec"][idx]))
            if i == 2:
                mixed = srv.submit(*cols(probe), nvec=d["nvec"][probe])
        for r in reqs + [mixed]:
            assert r.wait(timeout=150)
    assert (alone.retcode == 1).all()
    assert np.array_equal(alone.u_final, mixed.u_final) and np.array_equal(alone.stats, mixed.stats)
    b = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), d["u0"], d["p"], d["t0"], d["tf"],
                           odemi.EnsembleMI355X(precision=prec), reltol=d["reltol"], abstol=d["abstol"], nvec=d["nvec"])
    assert np.array_equal(alone.u_final, b["u_final"][:, probe])
    assert np.array_equal(alone.stats, b["stats"][:, probe])
    if prec == "fp64":
        for col, j in enumerate(probe[::5]):
            n = int(d["nvec"][j])
            r = O.solve_builtin(O.PROB_CHAIN, d["u0"][:n, j], d["p"][:, j], (0.0, 0.5), alg=O.RODAS5P,
                                reltol=d["reltol"], abstol=d["abstol"], save_start=False)
            c = int(np.where(probe == j)[0][0])
            assert np.array_equal(alone.u_final[:n, c], r["u_final"])


def test_poisson_mixed_types_under_budget_alone_equals_batch():
    """Open-loop Poisson arrivals (seeded) of requests of mixed type — Lorenz
    Tsit5, Robertson Rodas5P (lane servers) and uneven chains n = 4..32
    (workgroup server) — at 2,000 req/s for 1 s, each server's pool smaller than
    the load it is offered.  Requests are admitted as slots free up and run
    overlapped on the device; every request's result equals the batch engine's
    (that is, its result alone) on the same trajectories, bit for bit."""
    import time
    rng = np.random.Generator(np.random.PCG64(11))
    pools = {
        "lorenz": (P.lorenz_large(N=65536), P.LORENZ, odemi.Tsit5(), 64),
        "robertson": (P.robertson_ensemble(N=65536), P.ROBERTSON, odemi.Rodas5P(), 64),
        "chain": (P.chain_ensemble(N=2048, seed=12, m_lo=2, m_hi=16), P.CHAIN, odemi.Rodas5P(), 4),
    }
    pools["chain"][0]["tf"] = np.full(2048, 0.5)
    rate, dur = 2000.0, 1.0
    t_arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * dur * 2)))
    t_arr = t_arr[t_arr < dur]
    kinds = list(pools)
    kind = rng.choice(3, len(t_arr), p=[0.45, 0.45, 0.10])
    size = np.array([int(rng.integers(1, pools[kinds[k]][3] + 1)) for k in kind])
    print(f"[case] serving poisson {len(t_arr)} requests mixed types", flush=True)
    servers, cursor, taken = {}, {k: 0 for k in kinds}, {k: [] for k in kinds}
    for k, (d, fun, alg, _) in pools.items():
        ns = sorted(set(int(x) for x in d["nvec"])) if "nvec" in d else None
        servers[k] = odemi.EnsembleServer(fun, d["u0"].shape[0], d["p"].shape[0], alg, reltol=d["reltol"],
                                          abstol=d["abstol"], ns=ns, tick_attempts=16,
                                          max_slots=96 if k == "chain" else 2048)
    reqs = []
    try:
        t0 = time.monotonic() + 0.02
        for q in range(len(t_arr)):
            while time.monotonic() < t0 + t_arr[q]:
                time.sleep(0.0002)
            k = kinds[kind[q]]
            d = pools[k][0]
            a, b = cursor[k], cursor[k] + int(size[q])
            cursor[k] = b
            idx = np.arange(a, b)
            taken[k].append(idx)
            nv = d["nvec"][idx] if "nvec" in d else None
            reqs.append((k, idx, servers[k].submit(*_cols(d, idx), nvec=nv)))
        for k, idx, r in reqs:
            assert r.wait(timeout=300)
        late = sum(1 for _, _, r in reqs if r.times()[1] > r.times()[0])
        offered = {k: cursor[k] for k in kinds}
        print(f"offered trajectories {offered}; admitted after arrival: {late} of {len(reqs)}")
        for k in kinds:
            assert servers[k].info()["finished"] == len(taken[k])
    finally:
        for s in servers.values():
            s.close()
    for k in kinds:
        d, fun, alg, _ = pools[k]
        m = cursor[k]
        if m == 0:
            continue
        cols = np.arange(m)
        kw = dict(reltol=d["reltol"], abstol=d["abstol"])
        if "nvec" in d:
            kw["nvec"] = d["nvec"][cols]
        b = odemi.solve_arrays(fun, alg, *_cols(d, cols), **kw)
        for kk, idx, r in reqs:
            if kk != k:
                continue
            assert np.array_equal(r.u_final, b["u_final"][:, idx]), k
            assert np.array_equal(r.stats, b["stats"][:, idx]) and np.array_equal(r.retcode, b["retcode"][idx])


@pytest.mark.parametrize("kind", ["lorenz-tsit5", "robertson-rodas", "chain-wg"])
def test_per_request_tolerances_equal_batch_engine(kind):
    """Requests with their own reltol / abstol (odemi_submit_opts) in one pool,
    interleaved: each result equals the batch engine's solve at that request's
    tolerances, bit for bit (the per-slot tolerances live in the pool's state,
    so neighbours with other tolerances change nothing)."""
    if kind == "lorenz-tsit5":
        d, fun, alg, ns = P.lorenz_large(N=600), P.LORENZ, odemi.Tsit5(), None
        tols = [(1e-4, 1e-4), (None, None), (1e-8, 1e-9)]
    elif kind == "robertson-rodas":
        d, fun, alg, ns = P.robertson_ensemble(N=600), P.ROBERTSON, odemi.Rodas5P(), None
        tols = [(1e-4, 1e-8), (None, None), (1e-8, 1e-12)]
    else:
        d = P.chain_ensemble(N=48, seed=61, m_lo=4, m_hi=12)
        d["tf"] = np.full(48, 0.5)
        fun, alg, ns = P.CHAIN, odemi.Rodas5P(), sorted(set(int(x) for x in d["nvec"]))
        tols = [(1e-4, 1e-4), (None, None), (1e-8, 1e-8)]
    n, N = d["u0"].shape
    print(f"[case] serving per-request tolerances {kind}", flush=True)
    chunks = np.array_split(np.arange(N), 6)
    with odemi.EnsembleServer(fun, n, d["p"].shape[0], alg, reltol=d["reltol"], abstol=d["abstol"], ns=ns,
                              tick_attempts=16) as srv:
        reqs = []
        for q, idx in enumerate(chunks):
            rt, at = tols[q % 3]
            nv = d["nvec"][idx] if ns is not None else None
            reqs.append((idx, rt, at, srv.submit(*_cols(d, idx), nvec=nv, reltol=rt, abstol=at)))
        for *_, r in reqs:
            assert r.wait(timeout=300)
    for idx, rt, at, r in reqs:
        kw = dict(reltol=d["reltol"] if rt is None else rt, abstol=d["abstol"] if at is None else at)
        if ns is not None:
            kw["nvec"] = d["nvec"][idx]
        b = odemi.solve_arrays(fun, alg, *_cols(d, idx), **kw)
        assert (b["retcode"] == 1).all()
        assert np.array_equal(r.u_final, b["u_final"]) and np.array_equal(r.stats, b["stats"]), (rt, at)
    # the tolerances did change the work
    steps = [int(r.stats[1].sum()) for *_, r in reqs[:3]]
    assert steps[0] < steps[1] < steps[2]


SAVE_CASES = [c for c in CASES if c[0] in ("tsit5-fp64", "tsit5-mixed", "rodas-fp64")]


@pytest.mark.parametrize("name,alg,prec,fun,pid,dgen", SAVE_CASES, ids=[c[0] for c in SAVE_CASES])
def test_request_saveat_and_tstops_equal_batch(name, alg, prec, fun, pid, dgen):
    """Requests carrying their own save times, stopping times and save flags
    (lane servers with max_saves / max_tstops), in flight together with plain
    requests: each request's saved points, final states and stats are those of
    the batch engine run with the same keywords (bit for bit)."""
    d = dgen()
    n, N = d["u0"].shape
    print(f"[case] serving saveat/tstops {name}", flush=True)
    span = float(np.median(d["tf"] - d["t0"]))
    specs = [   # (trajectories, saveat, tstops, save_start, save_end)
        (np.arange(0, 48), np.linspace(0.1, 0.9, 5) * span, None, False, True),
        (np.arange(48, 120), np.array([0.25 * span]), np.array([0.3 * span, 0.61 * span]), True, True),
        (np.arange(120, 130), None, np.array([0.5 * span]), False, True),
        (np.arange(130, 200), np.linspace(0.05, 1.5, 12) * span, None, True, False),
    ]
    kw = dict(precision=prec, reltol=d["reltol"], abstol=d["abstol"], tick_attempts=16, max_saves=16, max_tstops=4)
    with odemi.EnsembleServer(fun, n, d["p"].shape[0], alg, **kw) as srv:
        reqs = []
        for k, (idx, sa, ts, ss, se) in enumerate(specs):
            if k == 2:   # a plain request in between
                reqs.append(srv.submit(*_cols(d, np.arange(200, 260))))
            reqs.append(srv.submit(*_cols(d, idx), saveat=sa, tstops=ts, save_start=ss, save_end=se))
        for r in reqs:
            assert r.wait(timeout=120)
    reqs.pop(2)
    ea = odemi.EnsembleMI355X(precision=prec)
    for (idx, sa, ts, ss, se), r in zip(specs, reqs):
        b = odemi.solve_arrays(fun, alg, *_cols(d, idx), ea, reltol=d["reltol"], abstol=d["abstol"], saveat=sa,
                               tstops=ts, save_start=ss, save_end=se, max_saves=16)
        assert np.array_equal(r.u_final, b["u_final"]), name
        assert np.array_equal(r.stats, b["stats"]) and np.array_equal(r.retcode, b["retcode"])
        if sa is None and not ss:
            assert r.t_out is None
            continue
        assert np.array_equal(r.nsaved, b["nsaved"])
        assert int(r.nsaved.max()) > (1 if sa is not None else 0)
        for j in range(len(idx)):
            k = int(r.nsaved[j])
            assert np.array_equal(r.t_out[:k, j], b["t_out"][:k, j])
            assert np.array_equal(r.u_out[:k, :, j], b["u_out"][:k, :, j])


def test_chain_server_request_tstops_equal_batch():
    """Uneven-n chain (workgroup) server: requests with their own stopping
    times, among plain requests, give the batch workgroup engine's results with
    the same tstops, bit for bit (FP64)."""
    d = P.chain_ensemble(N=96)
    d["tf"] = np.full(96, 1.0)
    nv = np.asarray(d["nvec"])
    n = d["u0"].shape[0]
    ns = sorted(set(int(x) for x in nv))
    print("[case] chain server per-request tstops", flush=True)
    specs = [(np.arange(0, 24), np.array([0.25, 0.5])), (np.arange(24, 48), None),
             (np.arange(48, 96), np.array([0.1, 0.333, 0.75]))]
    with odemi.EnsembleServer(P.CHAIN, n, d["p"].shape[0], odemi.Rodas5P(), reltol=d["reltol"], abstol=d["abstol"],
                              tick_attempts=16, ns=ns, max_tstops=4) as srv:
        reqs = [srv.submit(*_cols(d, idx), nvec=nv[idx], tstops=ts) for idx, ts in specs]
        for r in reqs:
            assert r.wait(timeout=300)
        with pytest.raises((ValueError, RuntimeError)):   # save times need a lane server
            srv.submit(*_cols(d, np.arange(4)), nvec=nv[:4], saveat=[0.5])
    for (idx, ts), r in zip(specs, reqs):
        b = odemi.solve_arrays(P.CHAIN, odemi.Rodas5P(), *_cols(d, idx), odemi.EnsembleMI355X(), reltol=d["reltol"],
                               abstol=d["abstol"], nvec=nv[idx], tstops=ts)
        for c, j in enumerate(idx):
            assert np.array_equal(r.u_final[:nv[j], c], b["u_final"][:nv[j], c])
        assert np.array_equal(r.stats, b["stats"]) and np.array_equal(r.retcode, b["retcode"])
        if ts is not None:   # every interior stopping time was stepped on: more accepted steps than without
            assert (r.stats[1] >= 1).all()


_DOSES = [2.0, 4.0, 6.0]


def _pk(du, u, p, t):
    du[0] = -p[0] * u[0]
    du[1] = p[0] * u[0] - p[1] * u[1]


def _pk_jac(J, u, p, t):
    J[0, 0] = -p[0]
    J[0, 1] = 0.0
    J[1, 0] = p[0]
    J[1, 1] = -p[1]


def _dose(itg):
    itg.u[0] = itg.u[0] + itg.p[2]


def _stop(itg):
    odemi.terminate(itg)


_PK_CBS = odemi.CallbackSet(odemi.DiscreteCallback(lambda u, t, itg: odemi.sym_in(t, _DOSES), _dose),
                            odemi.DiscreteCallback(lambda u, t, itg: u[1] > itg.p[3], _stop,
                                                   save_positions=(False, True)))


def test_lane_servers_with_callbacks_equal_batch():
    """Servers of problems with callbacks (traced into the served problem):
    bouncing balls with a ContinuousCallback floor (Tsit5), and a dosing model
    whose DiscreteCallbacks dose at stopping times and terminate! (Rodas5P,
    the doses passed as each request's tstops, forced saves kept with saveat)
    give the batch engine's results bit for bit, several requests in flight."""
    print("[case] serving with continuous callbacks (balls, tsit5)", flush=True)
    d = P.ball_ensemble(N=512)
    with odemi.EnsembleServer(P.BALL, 2, 1, odemi.Tsit5(), reltol=d["reltol"], abstol=d["abstol"], tick_attempts=16,
                              callback=P.BALL_BOUNCE) as srv:
        chunks = np.array_split(np.arange(512), 5)
        reqs = [srv.submit(*_cols(d, idx)) for idx in chunks]
        for r in reqs:
            assert r.wait(timeout=120)
    for idx, r in zip(chunks, reqs):
        b = odemi.solve_arrays(P.BALL, odemi.Tsit5(), *_cols(d, idx), reltol=d["reltol"], abstol=d["abstol"],
                               callback=P.BALL_BOUNCE)
        assert np.array_equal(r.u_final, b["u_final"]) and np.array_equal(r.stats, b["stats"])
        assert np.array_equal(r.retcode, b["retcode"])
        assert (r.stats[6] > 0).all()   # conditions were evaluated
    print("[case] serving with discrete callbacks (dosing, rodas5p)", flush=True)
    N = 96
    rng = np.random.default_rng(11)
    p = np.stack([rng.uniform(0.5, 2.0, N), rng.uniform(0.1, 0.5, N), rng.uniform(1.0, 3.0, N),
                  rng.uniform(1.5, 4.0, N)])
    u0 = np.stack([p[2].copy(), np.zeros(N)])
    fun = odemi.ODEFunction(_pk, jac=_pk_jac)
    t0, tf = np.zeros(N), np.full(N, 10.0)
    sa = [1.0, 3.0, 5.0, 7.5]
    with odemi.EnsembleServer(fun, 2, 4, odemi.Rodas5P(), reltol=1e-6, abstol=1e-8, tick_attempts=16,
                              callback=_PK_CBS, max_saves=16, max_tstops=4) as srv:
        halves = (np.arange(0, 48), np.arange(48, 96))
        reqs = [srv.submit(u0[:, h], p[:, h], t0[h], tf[h], tstops=_DOSES, saveat=sa) for h in halves]
        plain = srv.submit(u0, p, t0, tf, tstops=_DOSES)
        for r in reqs + [plain]:
            assert r.wait(timeout=120)
    for h, r in zip(halves, reqs):
        b = odemi.solve_arrays(fun, odemi.Rodas5P(), u0[:, h], p[:, h], t0[h], tf[h], reltol=1e-6, abstol=1e-8,
                               tstops=_DOSES, callback=_PK_CBS, saveat=sa, max_saves=16)
        assert np.array_equal(r.retcode, b["retcode"]) and np.array_equal(r.stats, b["stats"])
        assert np.array_equal(r.nsaved, b["nsaved"])
        for j in range(len(h)):
            k = min(int(r.nsaved[j]), 16)
            assert np.array_equal(r.t_out[:k, j], b["t_out"][:k, j])
            assert np.array_equal(r.u_out[:k, :, j], b["u_out"][:k, :, j])
    b = odemi.solve_arrays(fun, odemi.Rodas5P(), u0, p, t0, tf, reltol=1e-6, abstol=1e-8, tstops=_DOSES,
                           callback=_PK_CBS)
    assert np.array_equal(plain.u_final, b["u_final"]) and np.array_equal(plain.retcode, b["retcode"])
    assert set(int(x) for x in plain.retcode) == {1, 2}   # some trajectories terminate
