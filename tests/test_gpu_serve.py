"""Serving loop: continuous batching, admission by memory, retirement at the
error target; each result equals the request run alone (same draws) up to
summation order."""
import numpy as np
import pytest
import torch

import vegasflow_amd as vf
from vegasflow_amd.serve import Request, Server, latency_report, poisson_requests, run_alone, slot_bytes

pytestmark = pytest.mark.gpu


def mixed_requests(n, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = i % 3
        if kind == 0:
            fn, d = vf.GaussianMix(), 4
        elif kind == 1:
            fn, d = vf.BreitWignerPower(), 6
        else:
            fn, d = vf.Gaussian(0.2, 0.4), 5
        out.append(Request(fn, d, int(rng.integers(20_000, 120_000)), target=float(rng.choice([1e-2, 3e-3])),
                           max_iter=12, seed=500 + i, arrival=0.002 * i, rid=i))
    return out


@pytest.mark.parametrize("dtype,pairs", [(torch.float64, None), (torch.float32, None), (torch.float64, True)])
def test_results_same_alone_and_in_the_mix(dtype, pairs):
    """FP64, mixed precision (through the pair-histogram batch kernel by
    default), and FP64 through its pair-histogram batch kernel on request."""
    server = Server(memory_budget=64 << 20, dtype=dtype, pair_histograms=pairs)
    assert server.pairs == (dtype == torch.float32 if pairs is None else pairs)
    done = server.run_stream(mixed_requests(18))
    assert len(done) == 18
    for req in done:
        assert req.iterations >= server.min_iter
        res, err = run_alone(req, dtype=dtype)
        assert abs(req.result - res) <= 1e-9 * abs(res), req.rid
        assert abs(req.error - err) <= 1e-6 * abs(err), req.rid


def test_admission_respects_the_memory_budget():
    server = Server(memory_budget=3 * slot_bytes() + 10)
    assert server.capacity == 3
    reqs = mixed_requests(8, seed=1)
    for r in reqs:
        r.arrival = 0.0
    seen = []
    orig_tick = server.tick

    def tick():
        seen.append(len(server.live))
        return orig_tick()

    server.tick = tick
    done = server.run_stream(reqs)
    assert len(done) == 8 and max(seen) <= 3 and min(seen) >= 1


def test_loose_targets_retire_early_and_report_latency():
    server = Server(memory_budget=16 << 20)
    reqs = poisson_requests(rate=400.0, duration=0.1, seed=3, events_gauss=(20_000, 50_000),
                            events_bw=(20_000, 50_000), targets=(5e-2, 2e-2), max_iter=20)
    done = server.run_stream(reqs)
    assert len(done) == len(reqs) > 10
    rep = latency_report(done)
    assert rep["p99_ms"] >= rep["p50_ms"] > 0
    assert all(r.iterations < 20 for r in done)
    assert rep["met_target"] == 1.0
