"""GPU parity: the HIP engines (through libodemi.so's C ABI) against the CPU
restatement on identical seeded inputs.

Bars: FP64 engines reproduce the oracle's final states to 1e-10 relative and its
accepted/rejected step counts on >= 99% of trajectories (they are expected to be
bit-identical); the mixed engine meets the requested tolerance scale."""
import numpy as np
import pytest

import oracle as O
import odemi
from odemi import problems as P

pytestmark = pytest.mark.gpu


def _oracle_ensemble(prob_id, d, alg, N=None, **kw):
    N = N or d["u0"].shape[1]
    return O.ensemble_builtin(prob_id, d["u0"][:, :N], d["p"][:, :N], d["t0"][:N], d["tf"][:N], alg=alg,
                              reltol=d["reltol"], abstol=d["abstol"], save_start=False, **kw)


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def _run(fun, alg, d, engine, precision, N=None, **kw):
    N = N or d["u0"].shape[1]
    ea = odemi.EnsembleMI355X(engine=engine, precision=precision)
    return odemi.solve_arrays(fun, alg, d["u0"][:, :N], d["p"][:, :N], d["t0"][:N], d["tf"][:N], ea,
                              reltol=d["reltol"], abstol=d["abstol"], save_start=False, save_end=True, **kw)


@pytest.mark.parametrize("engine", ["persistent", "plain"])
def test_lorenz_small_fp64_matches_oracle(engine):
    d = P.lorenz_small()
    print(f"[case] lorenz N=1024 tsit5 fp64 engine={engine}", flush=True)
    ref = _oracle_ensemble(O.PROB_LORENZ, d, O.TSIT5)
    out = _run(P.LORENZ, odemi.Tsit5(), d, engine, "fp64")
    assert (out["retcode"] == 1).all()
    rel = _rel(out["u_final"], ref["u_final"])
    same_steps = (out["stats"][1] == ref["naccept"]) & (out["stats"][2] == ref["nreject"])
    print(f"max rel {rel.max():.3e}  exact {np.mean(out['u_final'] == ref['u_final']):.4f}  "
          f"steps equal {same_steps.mean():.4f}")
    assert rel.max() <= 1e-10
    assert same_steps.mean() >= 0.99
    assert (out["stats"][0] == ref["nf"]).mean() >= 0.99
    # same operation order and the same elementary-function algorithm: bit-identical
    assert np.mean(np.all(out["u_final"] == ref["u_final"], axis=0)) >= 0.99


def test_lorenz_large_subset_fp64_and_mixed():
    N = 8192
    d = P.lorenz_large(N=N)
    print(f"[case] lorenz_large N={N} tsit5 fp64+mixed", flush=True)
    ref = _oracle_ensemble(O.PROB_LORENZ, d, O.TSIT5)
    out = _run(P.LORENZ, odemi.Tsit5(), d, "persistent", "fp64")
    rel = _rel(out["u_final"], ref["u_final"])
    same = (out["stats"][1] == ref["naccept"]) & (out["stats"][2] == ref["nreject"])
    print(f"fp64: max rel {rel.max():.3e} p99 {np.quantile(rel, 0.99):.3e} steps equal {same.mean():.4f}")
    assert np.quantile(rel.max(axis=0), 0.99) <= 1e-10
    assert same.mean() >= 0.99
    mx = _run(P.LORENZ, odemi.Tsit5(), d, "persistent", "mixed")
    assert (mx["retcode"] == 1).all()
    relm = _rel(mx["u_final"], ref["u_final"]).max(axis=0)
    print(f"mixed vs fp64: median rel {np.median(relm):.3e} p90 {np.quantile(relm, 0.9):.3e}")
    # Accuracy gate of the mixed engine: against the FP64 Vern9 solution at
    # reltol 1e-12, its norm-wise final-state error must be as small as the FP64
    # engine's at the same requested tolerance (Lorenz is chaotic over t <= 10,
    # so both carry global errors well above reltol; the FP32 stages must not
    # add to them): median within 1.5x, p99 within 2x.
    S = 2048
    v9 = O.vern9_ensemble(O.PROB_LORENZ, d["u0"][:, :S], d["p"][:, :S], d["t0"][:S], d["tf"][:S])
    assert (v9["status"] == 0).all()
    nrm = np.linalg.norm(v9["u_final"], axis=0)
    e64 = np.linalg.norm(out["u_final"][:, :S] - v9["u_final"], axis=0) / nrm
    emx = np.linalg.norm(mx["u_final"][:, :S] - v9["u_final"], axis=0) / nrm
    print(f"vs Vern9: fp64 median {np.median(e64):.3e} p99 {np.quantile(e64, 0.99):.3e}; "
          f"mixed median {np.median(emx):.3e} p99 {np.quantile(emx, 0.99):.3e}")
    assert np.median(emx) <= 1.5 * np.median(e64)
    assert np.quantile(emx, 0.99) <= 2.0 * np.quantile(e64, 0.99)


@pytest.mark.parametrize("engine,precision", [("persistent", "fp64"), ("plain", "fp64"), ("persistent", "mixed")])
def test_robertson_matches_oracle(engine, precision):
    N = 2048
    d = P.robertson_ensemble(N=N)
    print(f"[case] robertson N={N} rodas5p {engine} {precision}", flush=True)
    ref = _oracle_ensemble(O.PROB_ROBERTSON, d, O.RODAS5P)
    out = _run(P.ROBERTSON, odemi.Rodas5P(), d, engine, precision)
    assert (out["retcode"] == 1).all()
    rel = _rel(out["u_final"], ref["u_final"]).max(axis=0)
    same = (out["stats"][1] == ref["naccept"]) & (out["stats"][2] == ref["nreject"])
    print(f"max rel {rel.max():.3e} steps equal {same.mean():.4f} exact {np.mean(out['u_final'] == ref['u_final']):.4f}")
    if precision == "fp64":
        assert np.quantile(rel, 0.99) <= 1e-10
        assert same.mean() >= 0.99
        assert (out["stats"][0] == ref["nf"]).mean() >= 0.99
    else:
        # FP32 LU + FP64 refinement: the solution must stay within the requested tolerance scale
        err = np.abs(out["u_final"] - ref["u_final"]) / (d["abstol"] + d["reltol"] * np.abs(ref["u_final"]))
        print(f"mixed scaled err max {err.max():.3e}")
        assert np.quantile(err.max(axis=0), 0.99) <= 10.0


def test_host_arrays_with_pinned_outputs_equal_device_arrays():
    """A host-array (numpy) solve allocates its outputs of 1 MiB and more as
    views of pinned torch tensors (the device-to-host copies then run at DMA
    speed); with an odd N and saveat outputs the results equal the same solve
    on device-resident arrays bit for bit, on two solves in a row (the second
    one's pinned blocks may be the first one's, reused)."""
    import torch
    N = 200001
    d = P.lorenz_large(N=N)
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], saveat=np.linspace(0.0, 1.0, 9), save_start=False,
              save_end=True)
    ea = odemi.EnsembleMI355X(engine="persistent", precision="fp64")
    print(f"[case] lorenz N={N} host arrays (pinned outputs) vs device arrays", flush=True)
    dev = torch.device("cuda:0")
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    ref = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), T(d["u0"]), T(d["p"]), T(d["t0"]), T(d["tf"]), ea, **kw)
    ref = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in ref.items()}
    for rep in range(2):
        out = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"], ea, **kw)
        for key in ("u_final", "t_final", "retcode", "stats", "t_out", "u_out", "nsaved"):
            assert isinstance(out[key], np.ndarray), key
            assert np.array_equal(out[key], ref[key], equal_nan=True), (rep, key)
        assert isinstance(out["u_out"].base, torch.Tensor) and out["u_out"].base.is_pinned()
    assert (out["retcode"] == 1).all() and (out["nsaved"] >= 9).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_queue_order_records_equal_in_place(precision, monkeypatch):
    """The final-state-only Tsit5 persistent engine reads its inputs and writes
    its results as queue-order records (prep writes each trajectory's record
    at its queue slot, odemi_rec_unpack restores the SoA outputs); with the
    records switched off (ODEMI_QUEUE_RECORDS=0, in-place SoA access) every
    output is the same bit for bit, from device-resident and from host arrays,
    with and without the saved end point."""
    import torch
    N = 50001
    d = P.lorenz_large(N=N)
    ea = odemi.EnsembleMI355X(engine="persistent", precision=precision)
    dev = torch.device("cuda:0")
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    keys = ("u_final", "t_final", "retcode", "stats", "nsaved", "t_out", "u_out")
    print(f"[case] lorenz N={N} {precision} queue-order records vs in place", flush=True)
    for save_end in (True, False):
        kw = dict(reltol=d["reltol"], abstol=d["abstol"], save_start=False, save_end=save_end)
        res = []
        for flag in ("1", "0"):
            monkeypatch.setenv("ODEMI_QUEUE_RECORDS", flag)
            od = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), T(d["u0"]), T(d["p"]), T(d["t0"]), T(d["tf"]), ea, **kw)
            oh = odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), d["u0"], d["p"], d["t0"], d["tf"], ea, **kw)
            res.append({k: od[k].cpu().numpy() for k in keys})
            res.append({k: oh[k] for k in keys})
        ns = res[1]["nsaved"]
        assert (ns == (1 if save_end else 0)).all()
        for r in res[1:]:
            for k in keys:
                a, b = res[0][k], r[k]
                if k in ("t_out", "u_out"):   # rows past nsaved are not written
                    a, b = a[:1], b[:1]
                    if not save_end:
                        continue
                assert np.array_equal(a, b, equal_nan=True), (precision, save_end, k)
