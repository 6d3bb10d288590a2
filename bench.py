"""Headline benchmark: ensemble trajectories/s at the stated tolerance.

Workload (N=1): 1,048,576 Lorenz trajectories, Tsit5, rho~U[0,28], sigma~U[8,12],
beta=8/3, u0~U[-1,1]^3, t_end~U[1,10], reltol=abstol=1e-6, final state only.
A "step" is one full ensemble solve of that batch with inputs resident in HBM.
`value` is the batched persistent mixed-precision engine; the plain FP64
one-kernel-per-step engine is timed beside it (`plain_fp64`), and the CPU
restatement runs on a bounded sample on the host cores (`cpu_baseline`).
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=1048576)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1048576)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import odemi
    from odemi import problems as P

    dev = torch.device("cuda:0")
    d = P.lorenz_large(N=args.N)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous()
    u0, p, t0, tf = T(d["u0"]), T(d["p"]), T(d["t0"]), T(d["tf"])
    kw = dict(reltol=d["reltol"], abstol=d["abstol"], save_start=False, save_end=True)
    batched = odemi.EnsembleMI355X(engine="persistent", precision="mixed")
    fp64p = odemi.EnsembleMI355X(engine="persistent", precision="fp64")
    plain = odemi.EnsembleMI355X(engine="plain", precision="fp64")

    def run(ea):
        return odemi.solve_arrays(P.LORENZ, odemi.Tsit5(), u0, p, t0, tf, ea, **kw)

    print(f"[bench] warmup x{args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        out = run(batched)
    torch.cuda.synchronize()
    times, kms = [], []
    for s in range(args.steps):
        torch.cuda.synchronize()
        t_a = time.perf_counter()
        out = run(batched)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t_a)
        kms.append(out["kernel_ms"])
        print(f"[bench] step {s} {times[-1]*1e3:.2f} ms (kernel {kms[-1]:.2f} ms)", file=sys.stderr, flush=True)
    stats = out["stats"].cpu().numpy()
    attempts = int(stats[1].sum() + stats[2].sum())
    tr = out["traced"]
    F_f = tr.flops_f
    flops = attempts * tsit5_flops_per_attempt(F_f, 3) + args.N * tsit5_flops_init(F_f, 3)
    ms_step = float(np.median(times)) * 1e3
    kernel_ms = float(np.median(kms))
    value = args.N / (ms_step / 1e3)
    ok = float((out["retcode"].cpu().numpy() == 1).mean())
    res = {
        "metric": "ensemble trajectories/s at reltol 1e-6 (Tsit5, Lorenz), batched persistent mixed-precision engine",
        "value": value, "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 stages + f64 state/t/error/controller", "data": "synthetic (seeded PCG64)",
        "config": {"workload": "Lorenz ensemble, 1,048,576 trajectories, Tsit5, rho~U[0,28], sigma~U[8,12], "
                               "u0~U[-1,1]^3, t_end~U[1,10], reltol=abstol=1e-6, final state only",
                   "trajectories": args.N, "parallelism": "one MI355X, lane per trajectory, work queue"},
        "step_attempts": attempts, "success_fraction": ok,
        "roofline": {"bound": "valu-fp32", "achieved": flops / (kernel_ms / 1e3) / 1e12,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": flops / (kernel_ms / 1e3) / 1e12 / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "odemi_tsit5_persist_mixed", "kernel_ms": kernel_ms,
                     "flops_per_attempt": tsit5_flops_per_attempt(F_f, 3)},
    }
    # FP64 persistent engine (same batch)
    o64 = run(fp64p)
    torch.cuda.synchronize()
    t_a = time.perf_counter()
    o64 = run(fp64p)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t_a
    res["persistent_fp64"] = {"trajectories_per_s": args.N / t64, "ms": t64 * 1e3, "kernel_ms": o64["kernel_ms"]}
    uref = o64["u_final"].cpu().numpy()
    um = out["u_final"].cpu().numpy()
    relm = (np.abs(um - uref) / (1e-6 + 1e-6 * np.abs(uref))).max(axis=0)
    res["mixed_vs_fp64_scaled_err"] = {"median": float(np.median(relm)), "p99": float(np.quantile(relm, 0.99))}
    if not args.no_plain:
        print("[bench] plain fp64 engine", file=sys.stderr, flush=True)
        t_a = time.perf_counter()
        op = run(plain)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t_a
        res["plain_fp64"] = {"trajectories_per_s": args.N / tp, "ms": tp * 1e3, "launches": int(op["launches"])}
        res["batched_over_plain"] = (args.N / (ms_step / 1e3)) / (args.N / tp)
    if not args.no_cpu:
        import oracle as O
        ncpu = len(os.sched_getaffinity(0))
        threads = min(ncpu, 16)
        S = min(args.cpu_sample, args.N)
        print(f"[bench] cpu baseline {S} trajectories on {threads} threads", file=sys.stderr, flush=True)
        r = O.ensemble_builtin(O.PROB_LORENZ, d["u0"][:, :S], d["p"][:, :S], d["t0"][:S], d["tf"][:S],
                               alg=O.TSIT5, reltol=d["reltol"], abstol=d["abstol"], save_start=False,
                               threads=threads)
        res["cpu_baseline"] = {"value": S / r["seconds"], "unit": "trajectories/s", "cores": threads,
                               "kind": "port", "sample": f"first {S} trajectories of the same batch, "
                                                         f"{r['seconds']:.2f} s wall"}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
