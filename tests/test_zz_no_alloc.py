"""mv_vegas_iteration_static (no allocation: static partial rows) against the
existing entry points mv_vegas_fused + mv_vegas_finalize with a caller
workspace, on the same step: integral and sigma within 1e-12 (sigma scaled
by its cancellation factor), histogram within 1e-12, refined grid within
1e-12. Two calls in a row, the second with fewer events and dimensions, in
both precisions; the library's array keeps no state between calls."""
import numpy as np
import pytest
import torch

from oracle import vegas as ora
from tests.parity import RTOL, assert_sigma, kappa_vegas
from vegasflow_amd import integrands as vfi
from vegasflow_amd import kernels

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _grid(d, seed):
    rng = np.random.default_rng(seed)
    g = np.empty((d, 51))
    for j in range(d):
        w = rng.gamma(0.7, size=50) + 1e-3
        c = np.concatenate([[0.0], np.cumsum(w)])
        g[j] = c / c[-1]
        g[j, 0], g[j, -1] = 0.0, 1.0
    return torch.tensor(g, dtype=torch.float64, device=DEV)


def _pair(d, n, precision, seed, draw):
    fn = vfi.Gaussian(0.2, 0.45).to_struct(d)
    step = kernels.make_step(d, n, precision=precision, train=True, seed=seed, draw_id=draw, xjac=1.0 / n)
    base = _grid(d, seed)
    # existing entry points, caller workspace
    g1 = base.clone()
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    h1 = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    ng = kernels.vegas_fused(step, fn, g1, ws)
    o1 = kernels.vegas_finalize(d, 1, True, True, n, ws, ng, grid=g1, hist_out=h1)
    # no-allocation entry point
    g2 = base.clone()
    h2 = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    o2 = torch.empty(2, dtype=torch.float64, device=DEV)
    kernels.vegas_iteration_static(step, fn, g2, True, o2, h2)
    torch.cuda.synchronize()
    return (o1.cpu().numpy(), h1.cpu().numpy(), g1.cpu().numpy()), (o2.cpu().numpy(), h2.cpu().numpy(),
                                                                      g2.cpu().numpy())


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_static_iteration_equals_workspace_entry_points(precision):
    for d, n, seed, draw in [(8, 1_000_000, 5, 3), (3, 20_011, 6, 0)]:
        (o1, h1, g1), (o2, h2, g2) = _pair(d, n, precision, seed, draw)
        assert abs(o2[0] - o1[0]) <= RTOL * abs(o1[0]), (d, n, o1, o2)
        assert_sigma(o2[1], o1[1], kappa_vegas(o1[0], o1[1], n))
        np.testing.assert_allclose(h2, h1, rtol=RTOL, atol=1e-300)
        np.testing.assert_allclose(g2, g1, rtol=RTOL, atol=1e-15)
        if precision == "f64":  # and both equal the oracle on the same draws
            grid = _grid(d, seed).cpu().numpy()
            res, sig, *_ = ora.vegas_iteration(grid, ora.uniform_f64(seed, draw, n, d),
                                               lambda x: ora.gaussian(x, 0.2, 0.45), n)
            assert abs(o2[0] - res) <= RTOL * abs(res)
            assert_sigma(o2[1], sig, kappa_vegas(res, sig, n))


def test_static_iteration_frozen_grid_and_limits():
    d, n = 5, 50_000
    fn = vfi.Product().to_struct(d)
    xmin = torch.tensor([0.1, -1.0, 0.0, 2.0, 0.5], dtype=torch.float64, device=DEV)
    xmax = torch.tensor([0.9, 1.0, 3.0, 2.5, 0.75], dtype=torch.float64, device=DEV)
    xd = xmax - xmin
    step = kernels.make_step(d, n, precision="f64", train=False, seed=9, xjac=1.0 / n, xmin=xmin, xdelta=xd,
                             xdeltajac=float(torch.prod(xd)))
    g = _grid(d, 9)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, False))
    o1 = kernels.vegas_finalize(d, 1, False, False, n, ws, kernels.vegas_fused(step, fn, g, ws))
    o2 = torch.empty(2, dtype=torch.float64, device=DEV)
    kernels.vegas_iteration_static(step, fn, g, False, o2)
    o1, o2 = o1.cpu().numpy(), o2.cpu().numpy()
    assert abs(o2[0] - o1[0]) <= RTOL * abs(o1[0])
    assert_sigma(o2[1], o1[1], kappa_vegas(o1[0], o1[1], n))


def test_static_iteration_rejects_sizes_beyond_its_array():
    d = 17
    step = kernels.make_step(d, 1000, precision="f64", train=True, xjac=1e-3)
    g = torch.zeros((d, 51), dtype=torch.float64, device=DEV)
    out = torch.empty(2, dtype=torch.float64, device=DEV)
    with pytest.raises(NotImplementedError, match="exceeds 16"):
        kernels.vegas_iteration_static(step, vfi.Product().to_struct(d), g, True, out)
    step = kernels.make_step(4, 1000, precision="f64", train=False, xjac=1e-3)
    with pytest.raises(ValueError, match="training step"):
        kernels.vegas_iteration_static(step, vfi.Product().to_struct(4), g[:4].contiguous(), True, out)
