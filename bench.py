"""Headline benchmark: ensemble trajectories/s at the stated tolerance.

Workload (N=1): 1,048,576 Lorenz trajectories, Tsit5, rho~U[0,28], sigma~U[8,12],
beta=8/3, u0~U[-1,1]^3, t_end~U[1,10], reltol=abstol=1e-6, final state only.
A "step" is one full ensemble solve of that batch with inputs resident in HBM.
`value` is the batched persistent FP64 engine (the reference's arithmetic,
bit-identical to the CPU restatement); the persistent mixed-precision engine
(`persistent_mixed`) and the plain FP64 one-kernel-per-attempt engine
(`plain_fp64`) are timed beside it, the same FP64 solve from host buffers
(`pcie_inclusive_fp64`), the CPU restatement on the host cores
(`cpu_baseline`), a Robertson Rodas5P leg (`robertson`) and a short open-loop
serving window with p50/p99 latency (`serving`).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "ordinarydiffeq.jl_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

# dominant-kernel roofline peaks (MI355X published, dense vector)
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def tsit5_flops_per_attempt(F_f, n):
    """Algorithmic flops of one Tsit5 step attempt: 6 RHS + stage/error/residual/norm
    + controller (FMA = 2, fastpower = 20 each)."""
    return 6 * F_f + 77 * n + 60


def tsit5_flops_init(F_f, n):
    return 3 * F_f + 12 * n


def rodas_flops_per_attempt(F_f, F_J, n, saves=False):
    """Algorithmic flops of one Rodas5P step attempt: 8 RHS + J + LU + 8 solves +
    stage/update vectors + controller (dense rows +48n when saving)."""
    return 8 * F_f + F_J + (2 * n ** 3 / 3 - n ** 2 / 2 - n / 6) + 16 * n ** 2 + 175 * n + 60 + (48 * n if saves else 0)


def _traffic(key):
    """HBM bytes per launch of a kernel on its bench workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes recorded in profiles/traffic.json (scaled by a
    copy-kernel calibration there)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)[key]["bytes"]
    except (OSError, KeyError, ValueError):
        return None


def _accuracy(prob_id, d, outs, sample, nvec=None):
    """Norm-wise relative final-state error |u - u_ref| / |u_ref| of each engine on
    the first `sample` trajectories of the batch, against an FP64 CPU reference
    at reltol 1e-12: Vern9 (the C restatement in oracle/) for non-stiff problems,
    scipy's Radau (rtol 1e-11) for stiff Robertson, where an explicit method is
    impractical."""
    import oracle as O
    S = min(sample, d["u0"].shape[1])
    if prob_id == O.PROB_ROBERTSON:
        from scipy.integrate import solve_ivp
        ref = np.zeros((3, S))
        for j in range(S):
            k = d["p"][:, j]
            r = solve_ivp(lambda t, u: [-k[0] * u[0] + k[2] * u[1] * u[2],
                                        k[0] * u[0] - k[1] * u[1] ** 2 - k[2] * u[1] * u[2], k[1] * u[1] ** 2],
                          (d["t0"][j], d["tf"][j]), d["u0"][:, j], method="Radau", rtol=1e-11, atol=1e-18)
            ref[:, j] = r.y[:, -1]
        kind = "scipy Radau rtol 1e-11 atol 1e-18"
    else:
        r = O.vern9_ensemble(prob_id, d["u0"][:, :S], d["p"][:, :S], d["t0"][:S], d["tf"][:S],
                             nvec=None if nvec is None else nvec[:S], reltol=1e-12, abstol=1e-12)
        ref = r["u_final"]
        kind = "Vern9 (C restatement) reltol = abstol = 1e-12"
    res = {"reference": kind, "sample": f"first {S} trajectories"}
    for name, uf in outs.items():
        e = np.linalg.norm(uf[:, :S] - ref, axis=0) / np.maximum(np.linalg.norm(ref, axis=0), 1e-300)
        res[name] = {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)), "max": float(e.max())}
    return res


def _timed(run, steps, warmup, torch):
    """W untimed warm-up solves, then exactly K solves timed as one region
    bracketed by device synchronisation; returns (last output, ms per solve,
    median kernel ms of the K solves)."""
    out = None
    for _ in range(warmup):
        out = run()
    torch.cuda.synchronize()
    kms = []
    t_a = time.perf_counter()
    for s in range(steps):
        out = run()
        kms.append(out["kernel_ms"])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t_a) * 1e3 / max(steps, 1)
    print(f"[bench] {steps} solves: {ms:.3f} ms per solve (kernel median {float(np.median(kms)):.3f} ms)",
          file=sys.stderr, flush=True)
    return out, ms, float(np.median(kms))


# host cores of one GPU's share of the box (the CPU baseline's thread pool)
CPU_SHARE = 16


def _host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model}


def _cpu_baseline(prob_id, d, alg, sample, sample_one, nvec=None, chunk=64):
    """The CPU restatement (oracle/, FP64, one trajectory per task with dynamic
    chunking like EnsembleThreads) on the host: `sample` trajectories on the
    GPU's CPU share (at most 16 threads), and `sample_one` on one thread."""
    import oracle as O
    N = d["u0"].shape[1]
    threads = max(1, min(len(os.sched_getaffinity(0)), CPU_SHARE))
    kw = dict(alg=alg, reltol=d["reltol"], abstol=d["abstol"], save_start=False)
    S = min(sample, N)
    print(f"[bench] cpu baseline {S} trajectories on {threads} threads", file=sys.stderr, flush=True)
    r = O.ensemble_builtin(prob_id, d["u0"][:, :S], d["p"][:, :S], d["t0"][:S], d["tf"][:S],
                           nvec=None if nvec is None else nvec[:S], threads=threads, chunk=chunk, **kw)
    S1 = min(sample_one, N)
    print(f"[bench] cpu baseline {S1} trajectories on 1 thread", file=sys.stderr, flush=True)
    r1 = O.ensemble_builtin(prob_id, d["u0"][:, :S1], d["p"][:, :S1], d["t0"][:S1], d["tf"][:S1],
                            nvec=None if nvec is None else nvec[:S1], threads=1, chunk=chunk, **kw)
    return {"value": S / r["seconds"], "unit": "trajectories/s", "cores": threads, "kind": "port",
            "sample": f"first {S} trajectories of the same batch, {r['seconds']:.2f} s wall on {threads} threads",
            "one_thread": {"value": S1 / r1["seconds"], "unit": "trajectories/s",
                           "sample": f"first {S1} trajectories, {r1['seconds']:.2f} s wall"},
            "host": _host_info()}


def _rodas_data(workload, N):
    from odemi import problems as P
    import oracle as O
    if workload == "robertson":
        N = N or 262144
        d = P.robertson_ensemble(N=N)
        return N, d, P.ROBERTSON, O.PROB_ROBERTSON, None, (
            f"Robertson ensemble, {N:,} trajectories, Rodas5P, k=(0.04,3e7,1e4)*LogU[0.5,2], u0=(1,0,0), "
            "tspan (0,1e5), reltol 1e-6, abstol 1e-10, analytic 3x3 J")
    N = N or 16384
    d = P.chain_ensemble(N=N)
    return N, d, P.CHAIN, O.PROB_CHAIN, d["nvec"], (
        f"Uneven oscillator chains, {N:,} trajectories, n=2m, m~U{{4..64}}, k~U[1,1e4], c~U[0.01,1], "
        "alpha~U[0,1], x0~N(0,0.1^2), tspan (0,10), Rodas5P, reltol=abstol=1e-6, one launch")


def _rodas_roofline(out, nvec, kms, kernel, peak, bound, traffic=None):
    st = out["stats"].cpu().numpy()
    att = (st[1] + st[2]).astype(np.float64)
    tr = out["traced"]
    if nvec is None:
        fl = att * rodas_flops_per_attempt(tr.flops_f, tr.flops_j, 3)
    else:
        fl = np.array([att[j] * rodas_flops_per_attempt(tr.flops_f[int(nvec[j])], tr.flops_j[int(nvec[j])],
                                                         int(nvec[j])) for j in range(len(att))])
    ach = float(fl.sum()) / (kms / 1e3) / 1e12
    return int(att.sum()), {"bound": bound, "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                            "traffic": traffic, "kernel": kernel, "kernel_ms": kms}


def rodas_legs(args, workload, steps, warmup, cpu=True, plain=True):
    """Rodas5P ensemble (Robertson: lane kernels; chains: uneven n in one
    workgroup-kernel launch): the FP64 engine (the reference's arithmetic,
    bit-identical to the restatement) as `value`, the FP32 LU / Gauss-Jordan +
    FP64 refinement engine and the plain FP64 one-launch-per-attempt engine
    beside it, accuracy against a reltol-1e-12 reference, and the CPU
    restatement on a bounded sample."""
    import torch
    import odemi
    import oracle as O
    dev = torch.device("cuda:0")
    N, d, fun, pid, nvec, desc = _rodas_data(workload, args.N if workload == args.workload else 0)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    u0, p, t0, tf = T(d["u0"]), T(d["p"]), T(d["t0"]), T(d["tf"])
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], save_start=False, save_end=True, nvec=nvec)
    mixed = odemi.EnsembleMI355X(precision="mixed", refine_sweeps=args.refine_sweeps)
    fp64 = odemi.EnsembleMI355X(precision="fp64")
    run = lambda ea: odemi.solve_arrays(fun, odemi.Rodas5P(), u0, p, t0, tf, ea, **kw)  # noqa: E731
    # hipRTC compile (cached on disk) stays outside every timed region
    odemi.compile_problem(fun, d["u0"].shape[0], d["p"].shape[0], algs=("rodas5p",),
                          ns=None if nvec is None else sorted(set(int(x) for x in nvec)))
    lane = nvec is None
    if getattr(args, "no_fp64", False):
        # profiling aid: the mixed engine alone (no FP64 leg, no comparisons)
        om, msm, kmsm = _timed(lambda: run(mixed), max(1, steps), warmup, torch)
        attm, roofm = _rodas_roofline(om, nvec, kmsm, "odemi_rodas_persist_mixed" if lane else "odemi_rodas_wg_mixed",
                                      PEAK_FP64_TFLOPS if lane else PEAK_FP32_TFLOPS,
                                      "valu-fp64" if lane else "lds/valu-fp32", None)
        return {"metric": "ensemble trajectories/s at reltol 1e-6 (Rodas5P), persistent mixed engine only",
                "value": N / (msm / 1e3), "unit": "trajectories/s", "n_gpus": 1, "steps": max(1, steps),
                "warmup": warmup, "ms_per_step": msm, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32 solves + f64 state (mixed)", "data": "synthetic (seeded PCG64)",
                "config": {"workload": desc, "trajectories": N}, "step_attempts": attm,
                "success_fraction": float((om["retcode"].cpu().numpy() == 1).mean()), "roofline": roofm}
    o64, ms64, kms64 = _timed(lambda: run(fp64), steps, warmup, torch)
    att64, roof64 = _rodas_roofline(o64, nvec, kms64, "odemi_rodas_persist_f64" if lane else "odemi_rodas_wg_f64",
                                    PEAK_FP64_TFLOPS, "valu-fp64",
                                    _traffic("odemi_rodas_persist_f64@robertson_262k") if lane and N == 262144
                                    else None)
    res = {
        "metric": "ensemble trajectories/s at reltol 1e-6 (Rodas5P), batched persistent FP64 engine",
        "value": N / (ms64 / 1e3), "unit": "trajectories/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": ms64, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded PCG64)",
        "config": {"workload": desc, "trajectories": N,
                   "parallelism": "one MI355X, " + ("lane per trajectory" if lane else "workgroup per trajectory")
                                  + ", work queue"},
        "step_attempts": att64, "success_fraction": float((o64["retcode"].cpu().numpy() == 1).mean()),
        "roofline": roof64,
    }
    om, msm, kmsm = _timed(lambda: run(mixed), max(1, steps), 1 if warmup else 0, torch)
    attm, roofm = _rodas_roofline(om, nvec, kmsm, "odemi_rodas_persist_mixed" if lane else "odemi_rodas_wg_mixed",
                                  PEAK_FP64_TFLOPS if lane else PEAK_FP32_TFLOPS,
                                  "valu-fp64" if lane else "lds/valu-fp32",
                                  (_traffic("odemi_rodas_persist_mixed@robertson_262k") if lane and N == 262144
                                   else _traffic("odemi_rodas_wg_mixed@chain_2048") if not lane and N == 2048
                                   else None))
    u64, um = o64["u_final"].cpu().numpy(), om["u_final"].cpu().numpy()
    err = (np.abs(um - u64) / (d["abstol"] + d["reltol"] * np.abs(u64))).max(axis=0)
    res["persistent_mixed"] = {
        "trajectories_per_s": N / (msm / 1e3), "ms": msm, "kernel_ms": kmsm, "step_attempts": attm,
        "dtype": "f32 LU (lane) / f32 Gauss-Jordan inverse (workgroup) + f64 refinement, f64 state/RHS/controller",
        "refine_sweeps": args.refine_sweeps, "roofline": roofm,
        "scaled_err_vs_fp64": {"median": float(np.median(err)), "p99": float(np.quantile(err, 0.99)),
                               "max": float(err.max())}}
    res["error_vs_reference"] = _accuracy(pid, d, {"fp64": u64, "mixed": um}, 8 if lane else 64, nvec=nvec)
    if plain and lane:
        op, msp, _ = _timed(lambda: run(odemi.EnsembleMI355X(engine="plain")), 1, 0, torch)
        res["plain_fp64"] = {"trajectories_per_s": N / (msp / 1e3), "ms": msp, "launches": int(op["launches"])}
        res["batched_over_plain"] = {"fp64": msp / ms64, "mixed": msp / msm}
    elif plain:
        # uneven chains: the plain FP64 engine (one launch per step attempt, the
        # workgroup kernels with their state in HBM) against both persistent
        # engines on the same leading sample of the batch
        S = min(N, args.plain_sample)
        kws = dict(kw, nvec=nvec[:S])
        sub = [x[..., :S].contiguous() for x in (u0, p, t0, tf)]
        runs = lambda ea: odemi.solve_arrays(fun, odemi.Rodas5P(), *sub, ea, **kws)  # noqa: E731
        print(f"[bench] plain fp64 workgroup engine on {S} trajectories", file=sys.stderr, flush=True)
        op, msp, _ = _timed(lambda: runs(odemi.EnsembleMI355X(engine="plain")), 1, 0, torch)
        _, mss64, _ = _timed(lambda: runs(fp64), 1, 0, torch)
        _, mssm, _ = _timed(lambda: runs(mixed), 1, 0, torch)
        same = bool(np.array_equal(op["u_final"].cpu().numpy(), u64[:, :S]))
        res["plain_fp64"] = {"sample": S, "trajectories_per_s": S / (msp / 1e3), "ms": msp,
                             "launches": int(op["launches"]), "identical_to_persistent_fp64": same}
        res["batched_over_plain"] = {"sample": S, "fp64": msp / mss64, "mixed": msp / mssm,
                                     "persistent_fp64_ms": mss64, "persistent_mixed_ms": mssm}
    if cpu and not args.no_cpu:
        res["cpu_baseline"] = _cpu_baseline(pid, d, O.RODAS5P, 262144 if lane else 256, 16384 if lane else 32,
                                            nvec=nvec, chunk=64 if lane else 1)
    return res


def bench_rodas(args, workload):
    print(json.dumps(rodas_legs(args, workload, args.steps, args.warmup, plain=not args.no_plain)))


def bench_events(args):
    """Bouncing balls with a ContinuousCallback floor (event search at every
    accepted step, root to adjacent floats, restep, affect_neg!): the callback
    path at scale, on the FP64 persistent engine (bit-identical to the CPU
    restatement).  Free fall is a quadratic that Tsit5 integrates exactly, so
    the FP64 error estimate is near zero and steps are long; the mixed engine's
    FP32 stage rounding puts a floor under its error estimate that caps the
    step at reltol 1e-6, so it takes several times the steps and is reported
    beside, not as the headline.  Algorithmic flops = Tsit5 steps + one
    dense-output evaluation (14 n + 20 flops) per interior condition call."""
    import torch
    import odemi
    import oracle as O
    from odemi import problems as P
    N = args.N or 1048576
    d = P.ball_ensemble(N=N)
    dev = torch.device("cuda:0")
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous()
    u0, p, t0, tf = T(d["u0"]), T(d["p"]), T(d["t0"]), T(d["tf"])
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], save_start=False, save_end=True, callback=P.BALL_BOUNCE)

    def run(ea):
        return odemi.solve_arrays(P.BALL, odemi.Tsit5(), u0, p, t0, tf, ea, **kw)

    def work(o):
        st = o["stats"].cpu().numpy()
        att = int(st[1].sum() + st[2].sum())
        interior = int(st[6].sum() - 2 * st[1].sum())
        F_f = o["traced"].flops_f
        return att, int(st[6].sum()), (att * tsit5_flops_per_attempt(F_f, 2) + N * tsit5_flops_init(F_f, 2)
                                       + interior * (14 * 2 + 20))
    out, ms, kms = _timed(lambda: run(odemi.EnsembleMI355X(precision="fp64")), args.steps, args.warmup, torch)
    attempts, conds, flops = work(out)
    res = {
        "metric": "ensemble trajectories/s at reltol 1e-6 (Tsit5 with a ContinuousCallback), batched FP64 engine",
        "value": N / (ms / 1e3), "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded PCG64)",
        "config": {"workload": f"Bouncing balls, {N:,} trajectories, Tsit5, h0~U[10,50], v0~U[-5,5], "
                               "restitution~U[0.8,0.95], floor event (ContinuousCallback, interp_points 10), "
                               "tspan (0,10), reltol=abstol=1e-6, final state only",
                   "trajectories": N, "parallelism": "one MI355X, lane per trajectory, work queue"},
        "step_attempts": attempts, "condition_calls": conds,
        "success_fraction": float((out["retcode"].cpu().numpy() == 1).mean()),
        "roofline": {"bound": "valu-fp64", "achieved": flops / (kms / 1e3) / 1e12, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / (kms / 1e3) / 1e12 / PEAK_FP64_TFLOPS,
                     "traffic": _traffic("odemi_tsit5_persist_f64@balls_1m") if N == 1048576 else None,
                     "kernel": "odemi_tsit5_persist_f64", "kernel_ms": kms},
    }
    om, msm, kmm = _timed(lambda: run(odemi.EnsembleMI355X(precision="mixed")), max(1, args.steps), 1, torch)
    am, cm, _ = work(om)
    u64, um = out["u_final"].cpu().numpy(), om["u_final"].cpu().numpy()
    err = np.abs(um - u64) / (d["abstol"] + d["reltol"] * np.abs(u64))
    res["persistent_mixed"] = {"trajectories_per_s": N / (msm / 1e3), "kernel_ms": kmm, "step_attempts": am,
                               "condition_calls": cm,
                               "scaled_err_vs_fp64": {"median": float(np.median(err)),
                                                      "p99": float(np.quantile(err, 0.99))}}
    if not args.no_plain:
        _, msp, _ = _timed(lambda: run(odemi.EnsembleMI355X(engine="plain")), 1, 0, torch)
        res["plain_fp64"] = {"trajectories_per_s": N / (msp / 1e3)}
        res["batched_over_plain"] = res["value"] / res["plain_fp64"]["trajectories_per_s"]
    # the FP64 engine against the CPU restatement (built-in floor event) on a sample, bit for bit
    S = min(4096, N)
    threads = min(16, os.cpu_count() or 1)
    ref = O.ensemble_builtin(O.PROB_BALL, d["u0"][:, :S], d["p"][:, :S], d["t0"][:S], d["tf"][:S], alg=O.TSIT5,
                             reltol=d["reltol"], abstol=d["abstol"], save_start=False, threads=threads)
    res["fp64_identical_to_restatement"] = {"sample": S, "identical": bool(np.array_equal(u64[:, :S],
                                                                                          ref["u_final"]))}
    if not args.no_cpu:
        Sc = min(args.cpu_sample, N)
        print(f"[bench] cpu baseline {Sc} trajectories on {threads} threads", file=sys.stderr, flush=True)
        r = O.ensemble_builtin(O.PROB_BALL, d["u0"][:, :Sc], d["p"][:, :Sc], d["t0"][:Sc], d["tf"][:Sc], alg=O.TSIT5,
                               reltol=d["reltol"], abstol=d["abstol"], save_start=False, threads=threads)
        res["cpu_baseline"] = {"value": Sc / r["seconds"], "unit": "trajectories/s", "cores": threads, "kind": "port",
                               "sample": f"first {Sc} trajectories of the same batch, {r['seconds']:.2f} s wall"}
    print(json.dumps(res))


def serving_leg(args, dur):
    """Open-loop Poisson arrivals of ensemble requests of ragged size, served by
    continuous batching (one EnsembleServer per instance type, each with its own
    stream and worker thread).  Latency = completion - scheduled arrival.

    Two merged arrival streams: `rate` req/s over the lane-sized instance types
    (Lorenz at reltol 1e-8 and 1e-6, Robertson; 1..max_request trajectories) and
    `chain_rate` req/s of uneven oscillator chains n = 8..128 (1..chain_max
    trajectories, workgroup kernels), whose per-trajectory cost is ~10^4 times
    a Lorenz trajectory's."""
    import odemi
    from odemi import problems as P
    rate = args.rate
    rng = np.random.Generator(np.random.PCG64(7))
    pools = {
        "lorenz_small": (P.lorenz_small(N=65536), P.LORENZ, odemi.Tsit5(), "fp64"),
        "lorenz_large": (P.lorenz_large(N=262144), P.LORENZ, odemi.Tsit5(), "mixed"),
        "robertson": (P.robertson_ensemble(N=262144), P.ROBERTSON, odemi.Rodas5P(), "mixed"),
    }
    if args.chain_rate > 0:
        pools["chain"] = (P.chain_ensemble(N=16384), P.CHAIN, odemi.Rodas5P(), "mixed")
    kinds = list(pools)
    servers = {}
    budget = args.hbm_budget_gb * (1 << 30) // len(pools)
    for k, (d, fun, alg, prec) in pools.items():
        ns = sorted(set(int(x) for x in d["nvec"])) if "nvec" in d else None
        servers[k] = odemi.EnsembleServer(fun, d["u0"].shape[0], d["p"].shape[0], alg, precision=prec,
                                          reltol=d["reltol"], abstol=d["abstol"], ns=ns,
                                          tick_attempts=args.tick_attempts,
                                          hbm_budget_bytes=budget)
    # schedule: two Poisson streams merged by arrival time
    t_a = np.cumsum(rng.exponential(1.0 / rate, int(rate * dur * 1.5) + 16))
    t_a = t_a[t_a < dur]
    kind_a = rng.integers(0, 3, len(t_a))
    size_a = rng.integers(1, args.max_request + 1, len(t_a))
    if args.chain_rate > 0:
        t_c = np.cumsum(rng.exponential(1.0 / args.chain_rate, int(args.chain_rate * dur * 1.5) + 16))
        t_c = t_c[t_c < dur]
    else:
        t_c = np.zeros(0)
    kind_c = np.full(len(t_c), 3)
    size_c = rng.integers(1, args.chain_max + 1, len(t_c))
    t_arr = np.concatenate([t_a, t_c])
    order = np.argsort(t_arr, kind="stable")
    t_arr, kind, size = t_arr[order], np.concatenate([kind_a, kind_c])[order], np.concatenate([size_a, size_c])[order]
    nreq = len(t_arr)
    off = [int(rng.integers(0, pools[kinds[k]][0]["u0"].shape[1] - s + 1)) for k, s in zip(kind, size)]

    def submit(k, a, b):
        d = pools[k][0]
        nv = d["nvec"][a:b] if "nvec" in d else None
        return servers[k].submit(d["u0"][:, a:b], d["p"][:, a:b], d["t0"][a:b], d["tf"][a:b], stats=False, nvec=nv)

    for k in kinds:   # warm the kernels (hipRTC compile happened at server creation)
        submit(k, 0, 2).wait(timeout=300)
    print(f"[bench] serving {nreq} requests over {dur:.0f} s: {rate:.0f} req/s lane types + "
          f"{args.chain_rate:g} req/s chains", file=sys.stderr, flush=True)
    reqs = [None] * nreq
    lag = np.zeros(nreq)
    t0 = time.monotonic() + 0.05
    for q in range(nreq):
        ts = t0 + t_arr[q]
        while True:
            now = time.monotonic()
            if now >= ts:
                break
            if ts - now > 0.002:
                time.sleep(ts - now - 0.001)
        reqs[q] = submit(kinds[kind[q]], off[q], off[q] + int(size[q]))
        lag[q] = time.monotonic() - ts
        if q % 2000 == 0:
            print(f"[bench] submitted {q}/{nreq}", file=sys.stderr, flush=True)
    for q, r in enumerate(reqs):
        if not r.wait(timeout=900):
            raise RuntimeError(f"request {q} ({kinds[kind[q]]}) did not finish")
    done = np.array([r.times()[2] for r in reqs])
    lat = done - (t0 + t_arr)
    t_end = float(done.max())
    ok = float(np.mean([float((r.retcode == 1).mean()) for r in reqs]))
    info = {k: s_.info() for k, s_ in servers.items()}
    for s_ in servers.values():
        s_.close()
    keep = t_arr >= 0.1 * dur                   # drop the first 10% of the window (warm-up)

    def q(x):
        return {"p50": float(np.quantile(x, 0.5) * 1e3), "p99": float(np.quantile(x, 0.99) * 1e3),
                "max": float(x.max() * 1e3), "n": int(len(x))} if len(x) else None

    lane = keep & (kind < 3)
    traj = int(size.sum())
    res = {
        "metric": "serving: trajectories/s and p50/p99 request latency, open-loop Poisson arrivals",
        "value": traj / (t_end - t0), "unit": "trajectories/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": (t_end - t0) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (Lorenz reltol 1e-8), f32 stages/LU/Gauss-Jordan + f64 state (Lorenz 1e-6, Robertson, chains)",
        "data": "synthetic (seeded PCG64)",
        "config": {"workload": f"open-loop Poisson arrivals for {dur:.0f} s: {rate:.0f} req/s of 1..{args.max_request} "
                               "trajectories (uniform) of one of Lorenz (0,1) reltol 1e-8 Tsit5 / Lorenz "
                               "t_end~U[1,10] reltol 1e-6 Tsit5 / Robertson (0,1e5) Rodas5P, plus "
                               f"{args.chain_rate:g} req/s of 1..{args.chain_max} oscillator chains n=8..128 "
                               "tspan (0,10) Rodas5P",
                   "requests": nreq, "trajectories": traj,
                   "parallelism": "one MI355X, one server (stream + worker thread) per instance type"},
        "latency_ms": q(lat[lane]),
        "latency_ms_by_type": {k: q(lat[keep & (kind == i)]) for i, k in enumerate(kinds)},
        "submit_lag_ms": {"p99": float(np.quantile(lag, 0.99) * 1e3), "max": float(lag.max() * 1e3)},
        "success_fraction": ok,
        "servers": info,
        "roofline": None,
    }
    return res


def bench_serving(args):
    print(json.dumps(serving_leg(args, args.duration)))


def bench_lorenz(args):
    """1,048,576 Lorenz trajectories (Tsit5, reltol = abstol = 1e-6, uneven
    t_end): `value` is the batched persistent FP64 engine (the reference's
    arithmetic; bit-identical to the CPU restatement), the mixed-precision
    persistent engine and the plain FP64 one-launch-per-attempt engine are
    timed beside it on the same device-resident batch."""
    import torch
    import odemi
    import oracle as O
    from odemi import problems as P
    N = args.N or 1048576
    dev = torch.device("cuda:0")
    d = P.lorenz_large(N=N)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous()  # noqa: E731
    u0, p, t0, tf = T(d["u0"]), T(d["p"]), T(d["t0"]), T(d["tf"])
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], save_start=False, save_end=True)
    fp64 = odemi.EnsembleMI355X(engine="persistent", precision="fp64")
    mixed = odemi.EnsembleMI355X(engine="persistent", precision="mixed")
    plain = odemi.EnsembleMI355X(engine="plain", precision="fp64")

    def run(ea, *arrs):
        return odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), *(arrs or (u0, p, t0, tf)), ea, **kw)

    odemi.compile_problem(P.LORENZ, 3, 3, algs=("tsit5",))   # hipRTC compile outside the timed region

    def roof(out, kms, kernel, peak, bound, traffic):
        st = out["stats"].cpu().numpy()
        att = int(st[1].sum() + st[2].sum())
        F_f = out["traced"].flops_f
        fl = att * tsit5_flops_per_attempt(F_f, 3) + N * tsit5_flops_init(F_f, 3)
        ach = fl / (kms / 1e3) / 1e12
        return att, {"bound": bound, "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                     "traffic": traffic, "kernel": kernel, "kernel_ms": kms,
                     "flops_per_attempt": tsit5_flops_per_attempt(F_f, 3)}

    o64, ms64, kms64 = _timed(lambda: run(fp64), args.steps, args.warmup, torch)
    att64, roof64 = roof(o64, kms64, "odemi_tsit5_persist_f64", PEAK_FP64_TFLOPS, "valu-fp64",
                         _traffic("odemi_tsit5_persist_f64@lorenz_1m") if N == 1048576 else None)
    res = {
        "metric": "ensemble trajectories/s at reltol 1e-6 (Tsit5, Lorenz), batched persistent FP64 engine "
                  "vs plain FP64 GPU engine",
        "value": N / (ms64 / 1e3), "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded PCG64)",
        "config": {"workload": "Lorenz ensemble, 1,048,576 trajectories, Tsit5, rho~U[0,28], sigma~U[8,12], "
                               "u0~U[-1,1]^3, t_end~U[1,10], reltol=abstol=1e-6, final state only",
                   "trajectories": N, "parallelism": "one MI355X, lane per trajectory, work queue"},
        "step_attempts": att64, "success_fraction": float((o64["retcode"].cpu().numpy() == 1).mean()),
        "roofline": roof64,
    }
    om, msm, kmsm = _timed(lambda: run(mixed), args.steps, 1, torch)
    attm, roofm = roof(om, kmsm, "odemi_tsit5_persist_mixed", PEAK_FP32_TFLOPS, "valu-fp32",
                       _traffic("odemi_tsit5_persist_mixed@lorenz_1m") if N == 1048576 else None)
    u64, um = o64["u_final"].cpu().numpy(), om["u_final"].cpu().numpy()
    err = (np.abs(um - u64) / (d["abstol"] + d["reltol"] * np.abs(u64))).max(axis=0)
    res["persistent_mixed"] = {
        "trajectories_per_s": N / (msm / 1e3), "ms": msm, "kernel_ms": kmsm, "step_attempts": attm,
        "dtype": "f32 stages and stage sums + f64 state/t/error norm/controller", "roofline": roofm,
        "scaled_err_vs_fp64": {"median": float(np.median(err)), "p99": float(np.quantile(err, 0.99))}}
    # FP64 engine on the first 65,536 trajectories against the CPU restatement, bit for bit
    S = min(65536, N)
    ref = O.ensemble_builtin(O.PROB_LORENZ, d["u0"][:, :S], d["p"][:, :S], d["t0"][:S], d["tf"][:S], alg=O.TSIT5,
                             reltol=d["reltol"], abstol=d["abstol"], save_start=False,
                             threads=max(1, min(len(os.sched_getaffinity(0)), CPU_SHARE)))
    st64 = o64["stats"].cpu().numpy()
    res["fp64_vs_restatement"] = {
        "sample": S, "final_states_identical": float(np.mean(np.all(u64[:, :S] == ref["u_final"], axis=0))),
        "step_counts_identical": float(np.mean((st64[1, :S] == ref["naccept"]) & (st64[2, :S] == ref["nreject"])))}
    res["error_vs_reference"] = _accuracy(O.PROB_LORENZ, d, {"fp64": u64, "mixed": um}, 1024)
    # the same FP64 solve from host arrays: H2D of the inputs through D2H of the outputs
    _, mspc, _ = _timed(lambda: run(fp64, d["u0"], d["p"], d["t0"], d["tf"]), 2, 1, torch)
    res["pcie_inclusive_fp64"] = {"trajectories_per_s": N / (mspc / 1e3), "ms": mspc,
                                  "note": "numpy inputs (pageable), H2D + solve + D2H into the host outputs "
                                          "solve_arrays allocates (pinned from 1 MiB up)"}
    if not args.no_plain:
        print("[bench] plain fp64 engine", file=sys.stderr, flush=True)
        op, msp, _ = _timed(lambda: run(plain), 1, 0, torch)
        res["plain_fp64"] = {"trajectories_per_s": N / (msp / 1e3), "ms": msp, "launches": int(op["launches"])}
        res["batched_over_plain"] = {"fp64": msp / ms64, "mixed": msp / msm}
    if not args.no_cpu:
        res["cpu_baseline"] = _cpu_baseline(O.PROB_LORENZ, d, O.TSIT5, min(args.cpu_sample, N), 131072)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=0)
    ap.add_argument("--workload", default="lorenz", choices=["lorenz", "robertson", "chain", "serving", "events"])
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: requests/s")
    ap.add_argument("--duration", type=float, default=60.0, help="serving: arrival window, s")
    ap.add_argument("--serve-window", type=float, default=5.0,
                    help="default line: arrival window of its serving leg, s (0: no serving leg)")
    ap.add_argument("--max-request", type=int, default=4096, help="serving: trajectories per request, max")
    ap.add_argument("--tick-attempts", type=int, default=32, help="serving: step attempts per tick")
    ap.add_argument("--chain-rate", type=float, default=20.0, help="serving: chain requests/s (0: none)")
    ap.add_argument("--chain-max", type=int, default=16, help="serving: chains per chain request, max")
    ap.add_argument("--hbm-budget-gb", type=int, default=8, help="serving: slot-pool budget, GiB")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="robertson/chain: time the mixed engine alone (a profiling aid; no comparisons)")
    ap.add_argument("--plain-sample", type=int, default=2048,
                    help="chains: trajectories of the plain-vs-batched comparison")
    ap.add_argument("--no-robertson", action="store_true", help="default line: skip the Robertson leg")
    ap.add_argument("--refine-sweeps", type=int, default=1, help="mixed Rodas5P: FP64 refinement sweeps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1048576)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    # one hardware queue per server stream (4 servers + the context stream), so
    # a long workgroup tick on the chain server's stream never sits in front of
    # a lane server's tick in a shared in-order queue; set before HIP starts
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"
    # the engine's library, bound to the HIP runtime torch brings
    from odemi import _lib
    _lib.load()
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "events":
        return bench_events(args)
    if args.workload != "lorenz":
        return bench_rodas(args, args.workload)
    res = bench_lorenz(args)
    if not args.no_robertson:
        print("[bench] robertson leg", file=sys.stderr, flush=True)
        rob = rodas_legs(args, "robertson", max(1, min(args.steps, 5)), 1, plain=not args.no_plain)
        for k in ("metric", "unit", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
            rob.pop(k, None)
        res["robertson"] = rob
    if args.serve_window > 0:
        print("[bench] serving leg", file=sys.stderr, flush=True)
        sv = serving_leg(args, args.serve_window)
        res["serving"] = {k: sv[k] for k in ("value", "unit", "config", "latency_ms", "latency_ms_by_type",
                                             "submit_lag_ms", "success_fraction", "dtype")}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
