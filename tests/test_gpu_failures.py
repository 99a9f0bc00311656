"""Failure paths, forced from the host with inputs that are valid for the
kernels: integrands that return NaN or Inf (the finalize kernels' status
word, a RuntimeWarning, a grid left unrefined), and serving requests that
are cancelled, time out or go non-finite while other requests run on.

The nearest reference behaviour is the output-shape check of the integrand
(monte_carlo.py:626-633); the reference passes NaN results through without
comment, and so do the results here, with the warning added."""
import math
import warnings

import numpy as np
import pytest
import torch

import vegasflow_amd as vf
from vegasflow_amd import PlainFlow, VegasFlow, VegasFlowPlus, kernels, serve

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def nan_tail(xarr, weight=None):
    """A Gaussian that returns NaN on the slice x_0 > 0.999."""
    f = vf.Gaussian(0.3, 0.5)(xarr)
    return torch.where(xarr[:, 0] > 0.999, torch.full_like(f, math.nan), f)


@pytest.mark.parametrize("alg", [VegasFlow, PlainFlow, VegasFlowPlus])
def test_nan_torch_integrand_warns_and_keeps_the_grid(alg):
    inst = alg(3, 20_000, seed=3, verbose=False)
    inst.compile(nan_tail)
    grid0 = inst.divisions.cpu().numpy().copy() if hasattr(inst, "divisions") else None
    with pytest.warns(RuntimeWarning, match="NaN or Inf"):
        res, err = inst.run_integration(2)
    assert inst.nonfinite_iterations == [0, 1]
    assert math.isnan(res)
    if grid0 is not None:  # no refinement from a non-finite histogram
        np.testing.assert_array_equal(inst.divisions.cpu().numpy(), grid0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n_iter", [2, 20])  # eager, and replayed from a graph
def test_inf_builtin_integrand_in_the_fused_kernels(dtype, n_iter):
    inst = VegasFlow(4, 10_000, seed=4, verbose=False, dtype=dtype)
    inst.compile(vf.Constant(math.inf))
    assert inst._fused is not None
    grid0 = inst.divisions.cpu().numpy().copy()
    with pytest.warns(RuntimeWarning, match="NaN or Inf"):
        inst.run_integration(n_iter)
    assert inst.nonfinite_iterations == list(range(n_iter))
    assert all(not np.isfinite(r) for r, _, _ in inst.history)
    np.testing.assert_array_equal(inst.divisions.cpu().numpy(), grid0)
    # the status word is reset per run: a finite integrand after it is quiet
    inst.compile(vf.Gaussian(0.2, 0.5))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        res, err = inst.run_integration(2)
    assert np.isfinite(res) and inst.nonfinite_iterations == []


def test_vplus_fused_inf_keeps_the_allocation():
    inst = VegasFlowPlus(3, 30_000, adaptive=True, seed=5, verbose=False)
    inst.compile(vf.Constant(math.inf))
    n_ev0 = inst.n_ev.cpu().numpy().copy()
    with pytest.warns(RuntimeWarning, match="NaN or Inf"):
        inst.run_integration(2)
    np.testing.assert_array_equal(inst.n_ev.cpu().numpy(), n_ev0)


def test_status_word_through_the_c_abi():
    """mv_vegas_finalize ORs MV_STATUS_NONFINITE into the caller's word only
    for non-finite sums."""
    d, n = 3, 5000
    for value, want in [(math.inf, 1), (2.0, 0)]:
        step = kernels.make_step(d, n, precision="f64", train=True, seed=1, xjac=1.0 / n)
        ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
        g = torch.linspace(0, 1, 51, dtype=torch.float64, device=DEV).repeat(d, 1).contiguous()
        status = torch.zeros(1, dtype=torch.int32, device=DEV)
        ng = kernels.vegas_fused(step, vf.Constant(value).to_struct(d), g, ws)
        kernels.vegas_finalize(d, 1, True, True, n, ws, ng, grid=g, status=status)
        assert int(status.item()) == want


# ---------------------------------------------------------------- serving
def _req(fn, d, n, seed, **kw):
    return serve.Request(fn, d, n, target=1e-9, max_iter=6, seed=seed, **kw)


def test_server_cancel_timeout_and_nonfinite():
    srv = serve.Server(memory_budget=16 * serve.slot_bytes())
    srv.min_iter = 1
    keep = _req(vf.Gaussian(0.2, 0.5), 3, 20_000, seed=11)
    running = _req(vf.Gaussian(0.2, 0.5), 4, 20_000, seed=12)
    bad = _req(vf.Constant(math.inf), 2, 10_000, seed=13)
    slow = _req(vf.Gaussian(0.2, 0.5), 2, 10_000, seed=14, timeout=0.0)
    waiting = _req(vf.Gaussian(0.2, 0.5), 2, 10_000, seed=15, arrival=1e9)
    for r in (keep, running, bad, slow, waiting):
        srv.submit(r)
    srv.admit_waiting(1.0)
    assert slow.status == "timeout" and slow.slot == -1  # expired while waiting
    assert {keep.status, running.status, bad.status} == {"running"}
    assert waiting.status == "waiting"
    retired = srv.tick()
    assert bad in retired and bad.status == "nonfinite"
    assert srv.cancel(running) and running.status == "cancelled"
    assert srv.cancel(waiting) and waiting.status == "cancelled" and not srv.queue
    assert not srv.cancel(running)  # already finished
    while srv.live:
        srv.tick()
    assert keep.status == "max_iter" and keep.iterations == 6
    # the survivor is unaffected by its neighbours' failures
    res, err = serve.run_alone(keep)
    assert abs(keep.result - res) <= 1e-9 * abs(res)
    assert len(srv.free) == srv.capacity
    rep = serve.latency_report(srv.finished)
    assert rep["status"] == {"timeout": 1, "nonfinite": 1, "cancelled": 2, "max_iter": 1}
