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
# HBM bytes per launch of the fused training kernels at this workload, from the
# FETCH_SIZE and WRITE_SIZE passes of rocprofv3 scaled by a 1 GiB copy
# (profiles/vegas_fused_gauss_d8_1e7_pmc.md): 0 bytes per event; the grid
# reads and the FP64 atomic adds of the partial rows
TRAFFIC = {"mixed": 0.9748e6, "f64": 0.9830e6}
TRAFFIC_SOURCE = "profiles/vegas_fused_gauss_d8_1e7_pmc.md"
# VALU wave instructions per launch (SQ_INSTS_VALU, same profile) and the
# VALU-busy cycles per 64 events: the instruction classes of the counters
# (profiles/vegas_fused_gauss_d8_1e7_valu_mix.md) priced at the issue costs
# `bench.py --only issue` measures on gfx950 (2 cycles per wave64 FP32 add,
# multiply or FMA, 8 per v_exp_f32, 4 for every other kind), set against
# 4 SIMDs per CU at the 2.4 GHz engine clock (MI355X datasheet)
VALU_INSTS = {"mixed": 2.351e7, "f64": 5.107e7}
VALU_BUSY_CYCLES = {"mixed": 549.9, "f64": 1307.3}
VALU_SOURCE = "profiles/vegas_fused_gauss_d8_1e7_valu_mix.md"
CLOCK_HZ = 2.4e9
# host cores allotted to one GPU of the node (CPU worker pools are sized to it)
CPU_SHARE = 16


def flops_per_event(d, train=True):
    """Algorithmic FP ops per event of the fused Gaussian iteration (FMA = 2):
    map 8 per dimension, Gaussian 4 per dimension + 2, weight and sums 5,
    histogram d when training."""
    return 13 * d + 7 if train else 12 * d + 7


def timed_steps(vf, inst, steps, warmup):
    """Run warmup + steps fused training iterations; returns (seconds of the
    steps, mean fused-kernel ms of a second instrumented pass, last (res,
    sigma))."""
    import torch
    from vegasflow_amd import kernels

    stream = torch.cuda.current_stream()
    ws = inst._workspace_for(kernels.partial_stride(inst.n_dim, 1, True))

    def producer():
        inst._next_draw()
        step = inst._step(inst.n_events, train=True, offset=0)
        return kernels.vegas_fused(step, inst._fused, inst.divisions, ws)

    def one():
        ng = producer()
        return kernels.vegas_finalize(inst.n_dim, 1, True, True, inst.n_events, ws, ng, grid=inst.divisions)

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    # throughput: K steps, nothing but the iteration on the stream
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = one()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # kernel time: K producer launches back to back between one pair of HIP
    # events on the launch stream (the pair's own cost spread over K; a spin
    # kernel first lets the host enqueue ahead of the GPU), then one finalize
    # that consumes the summed rows without refining
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda._sleep(5_000_000)
    e0.record(stream)
    rows = 1
    for _ in range(steps):
        rows = producer()
    e1.record(stream)
    kernels.vegas_finalize(inst.n_dim, 1, True, False, inst.n_events, ws, rows)
    torch.cuda.synchronize()
    kms = e0.elapsed_time(e1) / steps
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
    return float(res), float(sig), float(inst.chi2_dof)


def unfused_bytes(d, s):
    """Algorithmic HBM bytes per event of the unfused path (float size s,
    uint8 bins): sample writes x, w, bins; digest reads f, w, bins."""
    return {"sample": s * d + s + d, "digest": 2 * s + d}


def bench_unfused(vf, dtype, steps=10):
    """The torch-hook path at the headline shape: VegasFlow with the Gaussian
    written as a torch function (not a built-in), d=8, 1e7 events per
    iteration in one step; the sample and digest kernels are timed with HIP
    events around each launch and priced against HBM bandwidth."""
    import torch
    from vegasflow_amd import kernels

    s = 8 if dtype == torch.float64 else 4

    def gaussian(x, weight=None):
        pref = (1.0 / (A * math.sqrt(math.pi))) ** x.shape[1]
        return pref * torch.exp(-torch.sum(torch.square(x - 0.5), dim=1) / A ** 2)

    inst = vf.VegasFlow(D, N, seed=SEED, verbose=False, dtype=dtype, events_limit=N)
    inst.compile(gaussian)
    stream = torch.cuda.current_stream()
    ws = inst._workspace_for(kernels.partial_stride(D, 1, True))
    x = torch.empty((N, D), dtype=dtype, device="cuda")
    w = torch.empty((N,), dtype=dtype, device="cuda")
    ind = torch.empty((N, D), dtype=torch.uint8, device="cuda")
    marks = []
    for k in range(steps + 2):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        inst._next_draw()
        step = inst._step(N, train=True, offset=0)
        ev[0].record(stream)
        kernels.vegas_sample(step, inst.divisions, x, w, ind)
        ev[1].record(stream)
        f = gaussian(x)
        ev[2].record(stream)
        ng = kernels.vegas_digest(step, f, w, ind, ws)
        ev[3].record(stream)
        kernels.vegas_finalize(D, 1, True, True, N, ws, ng, grid=inst.divisions)
        if k >= 2:
            marks.append(ev)
    torch.cuda.synchronize()
    ms = [sum(m[i].elapsed_time(m[i + 1]) for m in marks) / steps for i in range(3)]
    b = unfused_bytes(D, s)
    inst.run_integration(1)  # allocations of the integrand's temporaries
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, sig = inst.run_integration(5)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # the whole iteration's algorithmic bytes: x written and read, w and the
    # bins written and read, f written and read (2 s d + 4 s + 2 d per event)
    it_bytes = 2 * s * D + 4 * s + 2 * D
    return {"events_per_s_run_integration": 5 * N / dt,
            "iteration_bytes_per_event": it_bytes,
            "iteration_GB_s": it_bytes * 5 * N / dt / 1e9,
            "iteration_frac_of_hbm_peak": it_bytes * 5 * N / dt / 8e12,
            "integral_5_iterations": float(res),
            "sigma": float(sig),
            "sample_kernel": {"ms": ms[0], "bytes_per_event": b["sample"],
                              "achieved_GB_s": b["sample"] * N / (ms[0] * 1e-3) / 1e9},
            "torch_integrand_ms": ms[1],
            "digest_kernel": {"ms": ms[2], "bytes_per_event": b["digest"],
                              "achieved_GB_s": b["digest"] * N / (ms[2] * 1e-3) / 1e9},
            "hbm_peak_GB_s": 8000.0}


def bench_plain(vf, dtype, n_iter=20):
    """PlainFlow, d=8 Gaussian a=0.1, 1e7 events per iteration: a warm-up run
    of n_iter iterations (the second half replayed from the captured
    iteration), then n_iter timed iterations, all replayed."""
    import torch

    fn = vf.Gaussian(A, 0.5)
    inst = vf.PlainFlow(D, N, seed=SEED, verbose=False, dtype=dtype)
    inst.compile(fn)
    inst.run_integration(n_iter)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, sig = inst.run_integration(n_iter)
    dt = time.perf_counter() - t0
    return {"events_per_s": N * n_iter / dt, "ms_per_iteration": 1e3 * dt / n_iter, "iterations": n_iter,
            "integral": float(res), "sigma": float(sig), "pull_vs_exact": (float(res) - fn.exact(D)) / float(sig)}


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
    out = {"integral": float(res), "sigma": float(sig), "pull_vs_exact": (float(res) - fn.exact(4)) / float(sig),
           "chi2_dof": float(inst.chi2_dof), "events_per_s": events / dt, "ms_per_iteration": 1e3 * dt / n_iter,
           "events": events, "iterations": n_iter, "events_first_iteration": inst.events_history[0],
           "events_last_iteration": inst.events_history[-1]}
    # steady state: 50 further iterations of the same instance, replayed from
    # the HIP graph of one iteration (runs of 16 or more iterations capture
    # one); an untimed run of 50 first takes the capture, and the process's
    # first graph instantiation (about 4 ms), out of the timed run
    inst.run_integration(50)
    first = len(inst.events_history)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    inst.run_integration(50)
    dt = time.perf_counter() - t0
    ev = sum(inst.events_history[first:])
    out["steady_50_iterations"] = {"events_per_s": ev / dt, "ms_per_iteration": 1e3 * dt / 50}
    return out


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
    # the tail beyond 3 sigma: items whose iterations never sampled the peak
    # (result below 1e-3 of the exact value: narrow Gaussians in many
    # dimensions, a peak volume of (a sqrt(pi))^d against N events) and the
    # rest; the oracle gives the same pulls on the same draws
    # (test_c4_batch_full_config_against_the_oracle)
    beyond = np.abs(pulls) >= 3
    missed = res < 1e-3 * exact
    tail = [{"item": int(i), "n_dim": specs[i][1], "events": specs[i][2], "a": round(specs[i][0].a, 4),
             "pull": float(pulls[i]), "result_over_exact": float(res[i] / exact[i])}
            for i in np.flatnonzero(beyond & ~missed)]
    return {"events_per_s": b.events_per_round() * n_iter / dt, "ms_per_round": 1e3 * dt / n_iter,
            "items_beyond_3sigma": int(beyond.sum()), "items_peak_missed": int(missed.sum()),
            "peak_missed_dims": sorted(int(specs[i][1]) for i in np.flatnonzero(missed)),
            "tail_not_missed": tail,
            "events_per_round": b.events_per_round(), "items": len(specs),
            "events_per_tile": b.events_per_tile, "pair_histograms": b.pairs,
            "workgroups": max(b._pair_blocks.values()) if b.pairs else b.blocks,
            "frac_within_3sigma": float(np.mean(np.abs(pulls) < 3)),
            "median_abs_pull": float(np.median(np.abs(pulls))),
            "median_chi2_dof": float(np.nanmedian(b.chi2_dof()))}


def bench_serving(seconds, dtype, rate=200.0):
    """Poisson stream at ``rate`` req/s for ``seconds``: two-Gaussian mixture
    (d=4) or Breit-Wigner x power (d=6); target 1e-3 (1e6 events/iteration)
    or 1e-4 (2e7 events/iteration); at most 20 iterations."""
    import numpy as np
    from vegasflow_amd.serve import Server, latency_report, poisson_requests

    reqs = poisson_requests(rate=rate, duration=seconds, seed=5)
    server = Server(dtype=dtype)  # slab budget from the free device memory
    t0 = time.perf_counter()
    done = server.run_stream(reqs)
    dt = time.perf_counter() - t0
    rep = latency_report(done)
    ticks = np.array(server.tick_seconds) * 1e3
    rep.update({"rate_req_per_s": rate, "duration_s": seconds, "wall_s": dt,
                "events_per_s": server.events_done / dt, "ticks": server.ticks,
                "tick_ms_p50": float(np.percentile(ticks, 50)), "tick_ms_p99": float(np.percentile(ticks, 99)),
                "tick_ms_max": float(ticks.max()), "slots": server.capacity,
                "slab_budget_bytes": server.memory_budget,
                "pair_histograms": server.pairs})
    return rep


def _cpu_model():
    """lscpu's model name (the same line /proc/cpuinfo holds)."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(numpy_seconds=6.0):
    """The reference algorithm on the host cores, FP64.

    Headline: the reference's TF graph restated op for op with torch CPU
    tensors (dense one-hot histograms, 1e6-event steps, serial rebin),
    multi-threaded, timed on one whole training iteration of the benchmark
    workload (d=8 Gaussian, 1e7 events). Beside it: the NumPy oracle,
    single-threaded, on 2e5-event iterations for ``numpy_seconds``, and the
    reference's own CPU case (d=4 Gaussian, 5 x 1e5, seed 7) with its result.
    """
    import numpy as np
    import torch

    from oracle import torch_cpu as tc
    from oracle import vegas as ora

    # the host cores this process may run on, capped at the host cores
    # allotted to one GPU (16)
    affinity = len(os.sched_getaffinity(0))
    threads = max(1, min(CPU_SHARE, affinity))
    torch.set_num_threads(threads)
    gen = torch.Generator().manual_seed(SEED)
    div = torch.from_numpy(np.ascontiguousarray(ora.initial_grid(D)))
    t0 = time.perf_counter()
    _, _, done = tc.iteration(div, N, tc.gaussian, gen)
    t_torch = time.perf_counter() - t0

    gen = torch.Generator().manual_seed(7)
    div4 = torch.from_numpy(np.ascontiguousarray(ora.initial_grid(4)))
    t0 = time.perf_counter()
    its = [tc.iteration(div4, 100_000, tc.gaussian, gen)[:2] for _ in range(5)]
    t_c1 = time.perf_counter() - t0
    c1_res, c1_sig = ora.weighted_average(its)

    n = 200_000
    grid = ora.initial_grid(D)
    fn = lambda x: ora.gaussian(x, A, 0.5)  # noqa: E731
    reps, t_np = 0, 0.0
    while t_np < numpy_seconds and reps < 50:
        t0 = time.perf_counter()
        u = ora.uniform_f64(SEED, reps, n, D)
        ora.vegas_iteration(grid, u, fn, n)
        t_np += time.perf_counter() - t0
        reps += 1
    return {
        "value": done / t_torch,
        "unit": "events/s",
        "cores": threads,
        "cores_in_affinity": affinity,
        "cpu_share_per_gpu": CPU_SHARE,
        "kind": "port",
        "sample": f"one training iteration of {done} events (10 steps of 1e6), d=8 Gaussian, FP64, "
                  f"torch-CPU op-for-op restatement of the reference graph, {threads} threads, "
                  f"{t_torch:.1f} s",
        "cpu_model": _cpu_model(),
        "numpy_oracle_1_thread": {"events_per_s": reps * n / t_np,
                                  "sample": f"{reps} training iterations x {n} events, d=8 Gaussian, "
                                            f"Philox draws included"},
        "reference_cpu_case_d4": {"integral": float(c1_res), "sigma": float(c1_sig),
                                  "pull_vs_exact": (float(c1_res) - 0.99999999999385025) / float(c1_sig),
                                  "events_per_s": 5 * 100_000 / t_c1,
                                  "sample": "5 training iterations x 1e5 events, d=4 Gaussian a=0.1"},
    }


def calibrate_copy(torch, nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of ``nbytes`` (torch's copy kernel):
    the reference point that scales the FETCH_SIZE / WRITE_SIZE counters to
    bytes, and the achievable HBM bandwidth of this GPU."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    return {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": reps + 1,
            "ms_per_copy": ms, "GB_per_s_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}


ISSUE_KINDS = ("v_mad_u64_u32", "v_bitop3_b32", "v_fma_f32", "v_mul_f32", "v_fma_f64", "v_add_f64",
               "v_cvt_f64_f32 + v_mul_f64 + v_cvt_f32_f64 (mean)", "v_exp_f32 + v_mul_f32 (mean)",
               "v_lshl_add_u32", "Philox4x32-10 block (cycles per block)")


def bench_issue_costs(iters=4096):
    """Shader cycles per wave64 instruction per SIMD of the VALU kinds the
    fused kernels are made of (libmvprobe.so: independent chains of one
    kind, 16 waves per SIMD, HIP-event time x the shader clock of the same
    run), and of one Philox4x32-10 block as the kernels compute it."""
    import ctypes

    from vegasflow_amd import _build

    path = _build.probe_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.mvp_issue_cost.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
    lib.mvp_issue_cost.restype = ctypes.c_int
    out = {}
    for k in range(lib.mvp_kinds()):
        ms, ghz, per = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        if lib.mvp_issue_cost(k, iters, ctypes.byref(ms), ctypes.byref(ghz), ctypes.byref(per)) != 0:
            raise RuntimeError(f"issue-cost probe {ISSUE_KINDS[k]} failed")
        out[ISSUE_KINDS[k]] = round(ms.value * 1e-3 * ghz.value * 1e9 / per.value, 2)
    out["clock_GHz_last_probe"] = round(ghz.value, 3)
    return out


LDS_KINDS = ("ds_read_b64, random map entries", "ds_read_b64, lane-consecutive entries",
             "ds_add_f64, random pair cells", "ds_add_f64, lane-consecutive cells",
             "8 random reads + 4 random adds (mean)")


def bench_lds_costs(iters=8192):
    """Shader cycles per LDS wave instruction per CU (libmvprobe.so) at the
    d = 8 pair kernel's launch shape: one 1024-thread workgroup per CU, 83.2
    KB of LDS laid out as its map table and pair cells."""
    import ctypes

    from vegasflow_amd import _build

    lib = ctypes.CDLL(_build.probe_path())
    dp = ctypes.POINTER(ctypes.c_double)
    lib.mvp_lds_cost.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
    lib.mvp_lds_cost.restype = ctypes.c_int
    out = {}
    for k in range(lib.mvp_lds_kinds()):
        ms, ghz, per = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        if lib.mvp_lds_cost(k, iters, ctypes.byref(ms), ctypes.byref(ghz), ctypes.byref(per)) != 0:
            raise RuntimeError(f"LDS probe {LDS_KINDS[k]} failed")
        out[LDS_KINDS[k]] = round(ms.value * 1e-3 * ghz.value * 1e9 / per.value, 2)
    return out


def bench_scaling(torch, clock_hz=CLOCK_HZ, sizes=(2_500_000, 5_000_000, 10_000_000, 20_000_000), reps=20):
    """Fused training kernel time against events per launch, t(N) = a + b N
    (HIP events around each launch, the GPU kept busy while the host
    enqueues it): a is the launch's fixed cost (launch, table build, LDS
    clearing, epilogue), b the event loop's; the loop's VALU-busy fraction
    is the counted busy cycles per 64 events over b in cycles per wave64
    iteration per SIMD."""
    import numpy as np

    from vegasflow_amd import integrands, kernels

    dev = torch.device("cuda", 0)
    fn = integrands.Gaussian(A, 0.5).to_struct(D)
    simds = 4 * torch.cuda.get_device_properties(0).multi_processor_count
    out = {}
    for prec in ("mixed", "f64"):
        ws = kernels.Workspace(dev, kernels.partial_stride(D, 1, True))
        grid = torch.linspace(0, 1, 51, dtype=torch.float64, device=dev).repeat(D, 1).contiguous()
        points = []
        for n in sizes:
            evs = []
            for k in range(reps + 3):
                step = kernels.make_step(D, n, precision=prec, train=True, seed=3, draw_id=k, xjac=1.0 / n)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                torch.cuda._sleep(1_000_000)  # the host's enqueue latency out of the timed region
                e0.record()
                rows = kernels.vegas_fused(step, fn, grid, ws)
                e1.record()
                kernels.vegas_finalize(D, 1, True, False, n, ws, rows)  # consume the rows
                if k >= 3:
                    evs.append((e0, e1))
            torch.cuda.synchronize()
            points.append((n, float(np.median([a.elapsed_time(b) for a, b in evs]))))
        ns = np.array([p[0] for p in points], float)
        ts = np.array([p[1] for p in points], float) * 1e-3
        b, a = np.polyfit(ns, ts, 1)
        cycles = b * clock_hz * 64 * simds  # per wave64 iteration (64 events) per SIMD
        out[prec] = {"points_ms": points, "fixed_us": a * 1e6, "us_per_1e7_events": b * 1e13,
                     "cycles_per_64_events_per_simd": cycles,
                     "loop_valu_busy_frac": VALU_BUSY_CYCLES[prec] / cycles, "clock_GHz": clock_hz / 1e9}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only time a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-protocol", action="store_true", help="skip the 5+5 iteration error check")
    ap.add_argument("--no-extra", action="store_true", help="skip the VEGAS+, batched and serving runs")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson stream (60 s: the serving configuration)")
    ap.add_argument("--only", choices=["unfused", "plain", "vplus", "batch", "serve", "issue", "scaling"],
                    help="run one extra section alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    import vegasflow_amd as vf

    torch.cuda.set_device(0)
    if args.calibrate_copy:
        print(json.dumps(calibrate_copy(torch)))
        return 0
    precisions = (("mixed", torch.float32), ("f64", torch.float64))
    if args.only:
        if args.only == "unfused":
            out = {p: bench_unfused(vf, dt_) for p, dt_ in precisions}
            out["measured_copy_GB_s"] = calibrate_copy(torch)["GB_per_s_read_plus_write"]
        elif args.only == "vplus":
            out = {p: bench_vplus(vf, dt_) for p, dt_ in precisions}
        elif args.only == "plain":
            out = {p: bench_plain(vf, dt_) for p, dt_ in precisions}
        elif args.only == "batch":
            out = {p: bench_batched(torch, dt_) for p, dt_ in precisions}
        elif args.only == "issue":
            out = {"valu": bench_issue_costs(), "lds_per_cu": bench_lds_costs()}
        elif args.only == "scaling":
            out = bench_scaling(torch)
        else:
            out = {p: bench_serving(args.serve_seconds, dt_) for p, dt_ in precisions}
        print(json.dumps({"section": args.only, "result": out}))
        return 0
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
            res, sig, chi2 = protocol(vf, dtype, fn, D, N, SEED)
            err[prec] = {"integral": res, "sigma": sig, "pull_vs_exact": (res - exact) / sig,
                         "chi2_dof_frozen": chi2}
        prod = vf.Product()
        for prec, dtype in (("mixed", torch.float32), ("f64", torch.float64)):
            res, sig, chi2 = protocol(vf, dtype, prod, D, N, SEED + 1)
            err[prec + "_product"] = {"integral": res, "sigma": sig,
                                      "pull_vs_exact": (res - 2.0 ** -D) / sig, "chi2_dof_frozen": chi2}
        err["mixed_minus_f64_in_sigma"] = (err["mixed"]["integral"] - err["f64"]["integral"]) / err["f64"]["sigma"]

    head = results["mixed"]
    flop = flops_per_event(D, True) * N
    achieved = flop / (head["fused_kernel_ms"] * 1e-3) / 1e12
    cus = torch.cuda.get_device_properties(0).multi_processor_count

    def valu_issue(prec):
        simd_cycles = cus * 4 * CLOCK_HZ * results[prec]["fused_kernel_ms"] * 1e-3
        busy = VALU_BUSY_CYCLES[prec] * N / 64
        return {"insts_per_launch": VALU_INSTS[prec], "busy_cycles_per_64_events": VALU_BUSY_CYCLES[prec],
                "busy_frac": busy / simd_cycles, "source": VALU_SOURCE}

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
            "kernel": "vegas_fused_pairs<GAUSSIAN, 8> (mixed training, n_dim 8)",
            "achieved": achieved,
            "peak": PEAK["mixed"],
            "unit": "TFLOP/s",
            "frac": achieved / PEAK["mixed"],
            "traffic": TRAFFIC["mixed"],
            "traffic_unit": "bytes per launch",
            "traffic_source": TRAFFIC_SOURCE,
            "kernel_ms": head["fused_kernel_ms"],
            "flop_per_event": flops_per_event(D, True),
            "valu_issue": valu_issue("mixed"),
            "f64_kernel": {
                "kernel": "vegas_fused_pairs_f64<GAUSSIAN, 8> (FP64 training, n_dim 8)",
                "achieved": flop / (results["f64"]["fused_kernel_ms"] * 1e-3) / 1e12,
                "peak": PEAK["f64"],
                "traffic": TRAFFIC["f64"],
                "valu_issue": valu_issue("f64"),
            },
        },
    }
    line["roofline"]["f64_kernel"]["frac"] = line["roofline"]["f64_kernel"]["achieved"] / PEAK["f64"]
    try:  # the issue costs the busy cycles are priced at, measured on this GPU now
        line["roofline"]["valu_issue_cycles"] = bench_issue_costs()
        line["roofline"]["lds_cycles_per_cu"] = bench_lds_costs()
    except (OSError, RuntimeError) as e:
        line["roofline"]["valu_issue_cycles"] = {"error": str(e)}
    # the event loop apart from the launch's fixed cost (at the 2.4 GHz clock
    # the kernel runs at: GRBM_GUI_ACTIVE over its traced duration)
    line["roofline"]["loop"] = bench_scaling(torch)
    if not args.no_extra:
        line["unfused_torch_integrand"] = {p: bench_unfused(vf, dt_) for p, dt_ in
                                           (("mixed", torch.float32), ("f64", torch.float64))}
        # the achievable HBM rate next to the 8 TB/s datasheet figure
        line["unfused_torch_integrand"]["measured_copy_GB_s"] = \
            calibrate_copy(torch)["GB_per_s_read_plus_write"]
        line["plainflow_d8"] = {p: bench_plain(vf, dt_) for p, dt_ in
                                (("mixed", torch.float32), ("f64", torch.float64))}
        line["vegasflowplus_d4"] = {p: bench_vplus(vf, dt_) for p, dt_ in
                                    (("mixed", torch.float32), ("f64", torch.float64))}
        line["batched_256"] = {p: bench_batched(torch, dt_) for p, dt_ in
                               (("mixed", torch.float32), ("f64", torch.float64))}
        line["serving"] = {p: bench_serving(args.serve_seconds, dt_) for p, dt_ in
                           (("mixed", torch.float32), ("f64", torch.float64))}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
