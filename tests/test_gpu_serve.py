"""Serving loop: continuous batching, admission by memory, retirement at the
error target; each result equals the request run alone (same draws) up to
summation order, VegasFlow and VegasFlowPlus requests alike."""
import numpy as np
import pytest
import torch

import vegasflow_amd as vf
from oracle import vegas as ora
from tests.parity import RTOL, assert_sigma, kappa_vplus
from vegasflow_amd.serve import Request, Server, latency_report, poisson_requests, run_alone, slot_bytes

pytestmark = pytest.mark.gpu


def mixed_requests(n, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = i % 3
        alg = "vegas"
        if kind == 0:
            fn, d, alg = vf.GaussianMix(), 4, "vegas+"
        elif kind == 1:
            fn, d = vf.BreitWignerPower(), 6
        else:
            fn, d = vf.Gaussian(0.2, 0.4), 5
        out.append(Request(fn, d, int(rng.integers(20_000, 120_000)), target=float(rng.choice([1e-2, 3e-3])),
                           max_iter=12, seed=500 + i, arrival=0.002 * i, rid=i, algorithm=alg))
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


def mix_np(x):
    return ora.gaussian_mix(x, (0.5, 0.5), (0.05, 0.1), (0.33, 0.67))


@pytest.mark.parametrize("d,n,adaptive", [(4, 100_000, True), (3, 30_000, False), (2, 50_000, True)])
def test_vplus_slot_first_iteration_matches_oracle(d, n, adaptive):
    """One tick of a VegasFlowPlus slot against the oracle's VEGAS+ iteration
    on the same Philox draws (vflowplus.py:37-80, 187-242): integral and
    sigma at the 1e-12 bar, refined grid, and the reallocated events of the
    next iteration; a VegasFlow request shares the ticks."""
    server = Server(memory_budget=8 * slot_bytes())
    server.min_iter = 99
    req = Request(vf.GaussianMix(), d, n, target=1e-12, max_iter=3, seed=77 + d, algorithm="vegas+",
                  adaptive=adaptive)
    other = Request(vf.Gaussian(0.2, 0.4), 5, 40_000, target=1e-12, max_iter=3, seed=5)
    server.submit(req)
    server.submit(other)
    server.admit_waiting(0.0)
    server.tick()
    st = ora.vplus_setup(d, n, adaptive)
    total = int(st["n_ev"].sum())
    u = ora.uniform_f64(req.seed, 0, total, d)
    grid0 = ora.initial_grid(d)
    res, sig, new_div, new_n_ev, o_var, _ = ora.vplus_iteration(grid0, st, u, mix_np)
    s = req.slot
    stats = server.stats[4 * s:4 * s + 4].cpu().numpy()
    assert abs(stats[0] - res) <= RTOL * abs(res)
    x, w, ind, seg = ora.generate_samples_in_hypercubes(u, st["n_strat"], st["n_ev"],
                                                        ora.hypercubes(st["n_strat"], d), grid0)
    x, w = ora.apply_xjac_and_limits(x, w, 1.0 / st["n_cubes"])
    o_ress, _, _ = ora.vplus_event(x, w, ind, seg, st["n_ev"], mix_np(x), False)
    assert_sigma(stats[1], sig, kappa_vplus(o_ress, o_var, st["n_ev"], sig))
    grid = server.grids[s * 816:s * 816 + d * 51].view(d, 51).cpu().numpy()
    np.testing.assert_allclose(grid, new_div, rtol=RTOL, atol=1e-15)
    n_ev = server.n_ev[s * 10_000:s * 10_000 + st["n_cubes"]].cpu().numpy()
    diff = np.abs(n_ev.astype(np.int64) - new_n_ev.astype(np.int64))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
    assert req._vp["events"] == int(n_ev.sum())
    while server.live:
        server.tick()
    assert req.iterations == other.iterations == 3


def test_vplus_requests_converge_where_the_grid_alone_does_not():
    """The two-Gaussian mixture has its peaks on the diagonal, which a
    separable grid cannot follow; stratification with adaptive allocation
    reaches a 1e-3 target that plain VEGAS misses in the same budget."""
    server = Server(memory_budget=8 * slot_bytes())
    vp = Request(vf.GaussianMix(), 4, 200_000, target=1e-3, max_iter=20, seed=3, algorithm="vegas+")
    vg = Request(vf.GaussianMix(), 4, 200_000, target=1e-3, max_iter=20, seed=3, algorithm="vegas")
    done = server.run_stream([vp, vg], warm_up=False)
    assert len(done) == 2
    assert vp.status == "converged" and vp.iterations < vg.iterations
    exact = vf.GaussianMix().exact(4)
    assert abs(vp.result - exact) < 4 * vp.error


def test_tick_event_cap_bounds_every_tick():
    """With max_tick_events, a tick's live requests hold at most that many
    events per iteration (or one request alone), and admission order leaves
    each result as it is alone."""
    cap = 150_000
    server = Server(memory_budget=32 * slot_bytes(), max_tick_events=cap)
    reqs = mixed_requests(12, seed=4)
    for r in reqs:
        r.arrival = 0.0
    seen = []
    orig_tick = server.tick

    def tick():
        seen.append((sum(r.n_events for r in server.live.values()), len(server.live)))
        assert server.live_events == seen[-1][0]
        return orig_tick()

    server.tick = tick
    done = server.run_stream(reqs, warm_up=False)
    assert len(done) == 12 and server.live_events == 0
    assert all(ev <= cap or n == 1 for ev, n in seen)
    assert max(n for _, n in seen) >= 2  # the cap leaves room for several
    assert any(r.admitted - r.arrival > 0 for r in done)  # and it held some back
    for req in done[:4]:
        res, err = run_alone(req)
        assert abs(req.result - res) <= 1e-9 * abs(res), req.rid


def test_default_budget_follows_free_device_memory():
    before, _ = torch.cuda.mem_get_info()
    server = Server()
    after, _ = torch.cuda.mem_get_info()
    assert min(after // 16, 1 << 30) <= server.memory_budget <= min(before // 16, 1 << 30)
    assert server.capacity == server.memory_budget // slot_bytes() >= 1
