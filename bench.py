"""Benchmark of the VEGAS iteration on one MI355X.

Workload (the headline configuration): VegasFlow, normalised Gaussian
f(x) = (1/(a sqrt(pi)))^8 exp(-sum (x_i - 0.5)^2 / a^2), a = 0.1, d = 8,
50 bins per dimension, 1e7 events per iteration, Philox seed 11.

A step is one training iteration of 1e7 events: sample -> grid map ->
integrand -> bin histograms and sums (the fused kernel) -> reduction, sigma
and grid refinement (the finalize kernel). ``value`` is events/s of the
mixed-precision path (FP32 sampling, map and integrand; FP64 sums,
histograms, sigma, refinement); the FP64 drop-in is timed the same way and
reported beside it, with the integral errors of the full protocol
(5 adaptive + 5 frozen iterations) for both.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

D = 8
N = 10_000_000
A = 0.1
SEED = 11
PEAK = {"mixed": 157.3, "f64": 78.6}  # TFLOP/s, MI355X datasheet (FP32 / FP64 vector)


def flops_per_event(d, train=True):
    """Algorithmic FP ops per event of the fused Gaussian iteration (FMA = 2):
    map 8 per dimension, Gaussian 4 per dimension + 2, weight and sums 5,
    histogram d when training."""
    return 13 * d + 7 if train else 12 * d + 7


def timed_steps(vf, inst, steps, warmup):
    """Run warmup + steps fused training iterations; returns (seconds,
    mean fused-kernel ms, last (res, sigma))."""
    import torch
    from vegasflow_amd import kernels

    stream = torch.cuda.current_stream()
    ws = inst._workspace_for(kernels.partial_stride(inst.n_dim, 1, True))

    def one(ev=None):
        inst._next_draw()
        step = inst._step(inst.n_events, train=True, offset=0)
        if ev is not None:
            ev[0].record(stream)
        ng = kernels.vegas_fused(step, inst._fused, inst.divisions, ws)
        if ev is not None:
            ev[1].record(stream)
        return kernels.vegas_finalize(inst.n_dim, 1, True, True, inst.n_events, ws, ng, grid=inst.divisions)

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    out = None
    for k in range(steps):
        out = one(evs[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kms = sum(a.elapsed_time(b) for a, b in evs) / steps
    return dt, kms, out


def protocol(vf, dtype, fn, d, n, seed):
    """5 adaptive + 5 frozen iterations; returns dict of results."""
    import torch

    inst = vf.VegasFlow(d, n, seed=seed, verbose=False, dtype=dtype)
    inst.compile(fn)
    inst.run_integration(5)
    inst.freeze_grid()
    res, sig = inst.run_integration(5)
    torch.cuda.synchronize()
    return float(res), float(sig)


def bench_vplus(vf, dtype, n_iter=10):
    """VegasFlowPlus, d=4 two-Gaussian mixture, 1e6 events requested,
    adaptive (5e5 events in the first iteration), 10 iterations."""
    import torch

    fn = vf.GaussianMix((0.5, 0.5), (0.05, 0.1), (0.33, 0.67))
    inst = vf.VegasFlowPlus(4, 1_000_000, adaptive=True, seed=SEED, verbose=False, dtype=dtype)
    inst.compile(fn)
    inst.run_integration(1)  # warm-up (code objects, workspace)
    inst = vf.VegasFlowPlus(4, 1_000_000, adaptive=True, seed=SEED, verbose=False, dtype=dtype)
    inst.compile(fn)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, sig = inst.run_integration(n_iter)
    dt = time.perf_counter() - t0
    events = sum(inst.events_history)
    return {"integral": float(res), "sigma": float(sig), "pull_vs_exact": (float(res) - fn.exact(4)) / float(sig),
            "events_per_s": events / dt, "ms_per_iteration": 1e3 * dt / n_iter, "events": events,
            "iterations": n_iter, "events_first_iteration": inst.events_history[0],
            "events_last_iteration": inst.events_history[-1]}


def bench_batched(torch, dtype, n_iter=8):
    """256 Gaussian integrals (d ~ U{2..12}, events ~ logU[1e4, 1e6],
    a ~ U[0.05, 0.3], centres ~ U[0.2, 0.8]^d, seed 23), 8 iterations, each
    round a replay of one captured HIP graph."""
    import numpy as np
    from vegasflow_amd.batch import BatchVegas, random_batch

    specs = random_batch(256, seed=23)
    b = BatchVegas(specs, dtype=dtype)
    b.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(n_iter)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res, err = b.results()
    exact = np.array([fn.exact(d) for fn, d, _, _ in specs])
    pulls = (res - exact) / err
    return {"events_per_s": b.events_per_round() * n_iter / dt, "ms_per_round": 1e3 * dt / n_iter,
            "events_per_round": b.events_per_round(), "items": len(specs),
            "events_per_tile": b.events_per_tile, "workgroups": b.blocks,
            "frac_within_3sigma": float(np.mean(np.abs(pulls) < 3)),
            "median_abs_pull": float(np.median(np.abs(pulls)))}


def bench_serving(seconds, rate=200.0):
    """Poisson stream at ``rate`` req/s for ``seconds``: two-Gaussian mixture
    (d=4) or Breit-Wigner x power (d=6); target 1e-3 (1e6 events/iteration)
    or 1e-4 (2e7 events/iteration); at most 20 iterations."""
    from vegasflow_amd.serve import Server, latency_report, poisson_requests

    reqs = poisson_requests(rate=rate, duration=seconds, seed=5)
    server = Server(memory_budget=256 << 20)
    t0 = time.perf_counter()
    done = server.run_stream(reqs)
    dt = time.perf_counter() - t0
    rep = latency_report(done)
    rep.update({"rate_req_per_s": rate, "duration_s": seconds, "wall_s": dt,
                "events_per_s": server.events_done / dt, "ticks": server.ticks,
                "slots": server.capacity})
    return rep


def cpu_baseline(seconds_budget=20.0):
    """The NumPy FP64 oracle (a restatement of the reference algorithm,
    Philox draws included) timed on the host on one core-bound iteration."""
    import numpy as np
    from oracle import vegas as ora

    n = 200_000
    grid = ora.initial_grid(D)
    fn = lambda x: ora.gaussian(x, A, 0.5)  # noqa: E731
    reps, t_total = 0, 0.0
    while t_total < seconds_budget and reps < 50:
        t0 = time.perf_counter()
        u = ora.uniform_f64(SEED, reps, n, D)
        ora.vegas_iteration(grid, u, fn, n)
        t_total += time.perf_counter() - t0
        reps += 1
    return {
        "value": reps * n / t_total,
        "unit": "events/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{reps} training iterations x {n} events, d=8 Gaussian, NumPy FP64 oracle "
                  f"(Philox draws, grid map, bin histograms, refinement), single-threaded",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-protocol", action="store_true", help="skip the 5+5 iteration error check")
    ap.add_argument("--no-extra", action="store_true", help="skip the VEGAS+, batched and serving runs")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="length of the Poisson stream")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    import vegasflow_amd as vf

    torch.cuda.set_device(0)
    fn = vf.Gaussian(A, 0.5)
    exact = fn.exact(D)

    results = {}
    for prec, dtype in (("mixed", torch.float32), ("f64", torch.float64)):
        inst = vf.VegasFlow(D, N, seed=SEED, verbose=False, dtype=dtype)
        inst.compile(fn)
        dt, kms, out = timed_steps(vf, inst, args.steps, args.warmup)
        results[prec] = {
            "events_per_s": args.steps * N / dt,
            "ms_per_step": 1e3 * dt / args.steps,
            "fused_kernel_ms": kms,
        }

    err = {}
    if not args.no_protocol:
        for prec, dtype in (("mixed", torch.float32), ("f64", torch.float64)):
            res, sig = protocol(vf, dtype, fn, D, N, SEED)
            err[prec] = {"integral": res, "sigma": sig, "pull_vs_exact": (res - exact) / sig}
        prod = vf.Product()
        for prec, dtype in (("mixed", torch.float32), ("f64", torch.float64)):
            res, sig = protocol(vf, dtype, prod, D, N, SEED + 1)
            err[prec + "_product"] = {"integral": res, "sigma": sig,
                                      "pull_vs_exact": (res - 2.0 ** -D) / sig}
        err["mixed_minus_f64_in_sigma"] = (err["mixed"]["integral"] - err["f64"]["integral"]) / err["f64"]["sigma"]

    head = results["mixed"]
    flop = flops_per_event(D, True) * N
    achieved = flop / (head["fused_kernel_ms"] * 1e-3) / 1e12
    line = {
        "metric": "VEGAS events/s, d=8 Gaussian, mixed vs FP64 at equal error",
        "value": head["events_per_s"],
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 eval / f64 accumulate",
        "data": "synthetic (Philox draws, built-in Gaussian integrand)",
        "config": {"workload": "VegasFlow d=8 Gaussian a=0.1, 50 bins, 1e7 events/iteration, "
                               "training iteration (fused kernel + finalize/refine), seed 11",
                   "events_per_iteration": N, "n_dim": D},
        "f64": results["f64"],
        "mixed_over_f64": head["events_per_s"] / results["f64"]["events_per_s"],
        "errors": err,
        "roofline": {
            "bound": "valu",
            "kernel": "vegas_fused_mixed<GAUSSIAN, train>",
            "achieved": achieved,
            "peak": PEAK["mixed"],
            "unit": "TFLOP/s",
            "frac": achieved / PEAK["mixed"],
            "traffic": None,
            "flop_per_event": flops_per_event(D, True),
            "f64_kernel": {
                "achieved": flop / (results["f64"]["fused_kernel_ms"] * 1e-3) / 1e12,
                "peak": PEAK["f64"],
            },
        },
    }
    line["roofline"]["f64_kernel"]["frac"] = line["roofline"]["f64_kernel"]["achieved"] / PEAK["f64"]
    if not args.no_extra:
        line["vegasflowplus_d4"] = {p: bench_vplus(vf, dt_) for p, dt_ in
                                    (("mixed", torch.float32), ("f64", torch.float64))}
        line["batched_256"] = {p: bench_batched(torch, dt_) for p, dt_ in
                               (("mixed", torch.float32), ("f64", torch.float64))}
        line["serving"] = bench_serving(args.serve_seconds)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
