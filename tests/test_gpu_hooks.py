"""Subclass hooks, multi-step iterations and the sampler API against the
oracle on the same Philox draws.

- Hooks (monte_carlo.py:277-310, vflow.py:349-442, vflowplus.py:153-247):
  a subclass that overrides ``_run_event``, ``_digest_random_generation``
  (the reference's (xrand, *args) contract), ``_importance_sampling_array_
  filling``, ``refine_grid`` or ``_run_iteration`` has its override called
  every iteration, fused or replayed, and still matches the oracle.
- Steps (monte_carlo.py:420-480 and _accumulate 72-92): an iteration of
  more events than ``events_limit`` runs in ragged steps whose sums equal
  the oracle's single pass.
- Sampler (monte_carlo.py:229-247, 765-794; vflowplus.py:173-185):
  ``generate_random_array`` returns the oracle's points and densities.
"""
import numpy as np
import pytest
import torch

import vegasflow_amd as vf
from oracle import vegas as ora
from tests.parity import RTOL, assert_sigma, kappa_vegas
from vegasflow_amd import PlainFlow, VegasFlow, VegasFlowPlus

pytestmark = pytest.mark.gpu

A, C = 0.2, 0.45


def gauss_np(x):
    return ora.gaussian(x, A, C)


def gauss_torch(x):
    return vf.Gaussian(A, C)(x)


def _check_iterations(inst, seed, n, d, n_iter, u_of=None):
    """Iteration i of ``inst`` equals the oracle's iteration i from the
    engine's grid i; ``u_of(i)`` gives the uniforms the oracle maps."""
    grids = inst._test_grids
    for i in range(n_iter):
        r, s, _ = inst.history[i]
        u = ora.uniform_f64(seed, i, n, d) if u_of is None else u_of(i)
        res, sig, _, o_grid, *_ = ora.vegas_iteration(grids[i], u, gauss_np, n)
        assert abs(r - res) <= RTOL * abs(res), (i, r, res)
        assert_sigma(s, sig, kappa_vegas(res, sig, n))
        np.testing.assert_allclose(grids[i + 1], o_grid, rtol=RTOL, atol=1e-15)


def _run_recording_grids(inst, n_iter):
    inst._test_grids = [inst.divisions.cpu().numpy()]
    for _ in range(n_iter):
        inst.run_integration(1)
        inst._test_grids.append(inst.divisions.cpu().numpy())


# ----------------------------------------------------------------- steps
@pytest.mark.parametrize("limit", [10_000, 7_777])
def test_vegasflow_ragged_steps_match_oracle(limit):
    """25 000 events per iteration in steps of ``limit`` (10 000 + 10 000 +
    5 000, or 7 777 x 3 + 1 669) through the torch-callable path."""
    d, n, seed = 4, 25_000, 61
    inst = VegasFlow(d, n, events_limit=limit, seed=seed, verbose=False)
    inst.compile(gauss_torch)
    assert inst._fused is None and inst.events_per_run == limit
    _run_recording_grids(inst, 3)
    _check_iterations(inst, seed, n, d, 3)


def test_plainflow_ragged_steps_match_oracle():
    d, n, seed = 3, 30_001, 62
    inst = PlainFlow(d, n, events_limit=8_000, seed=seed, verbose=False)
    inst.compile(gauss_torch)
    inst.run_integration(2)
    for i, (r, s, _) in enumerate(inst.history):
        res, sig = ora.plain_iteration(ora.uniform_f64(seed, i, n, d), gauss_np, n)
        assert abs(r - res) <= RTOL * abs(res)
        assert_sigma(s, sig, kappa_vegas(res, sig, n))


# ----------------------------------------------------------------- hooks
class CountingRunEvent(VegasFlow):
    calls = 0

    def _run_event(self, integrand, ncalls=None):
        type(self).calls += 1
        return super()._run_event(integrand, ncalls=ncalls)


def test_run_event_override_is_called_and_matches_oracle():
    d, n, seed = 3, 20_000, 63
    CountingRunEvent.calls = 0
    inst = CountingRunEvent(d, n, events_limit=8_000, seed=seed, verbose=False)
    inst.compile(vf.Gaussian(A, C))  # a built-in integrand, run through the hook
    assert inst._fused is None
    _run_recording_grids(inst, 2)
    assert CountingRunEvent.calls == 2 * 3  # three steps per iteration
    _check_iterations(inst, seed, n, d, 2)


class ReflectedDraws(VegasFlow):
    """The reference's contract: uniforms in, (x, w, ind) out."""

    def _digest_random_generation(self, xrand):
        return super()._digest_random_generation(1.0 - xrand)


def test_digest_override_gets_uniforms_and_matches_oracle():
    d, n, seed = 3, 20_000, 64
    inst = ReflectedDraws(d, n, seed=seed, verbose=False)
    inst.compile(vf.Gaussian(A, C))
    assert inst._fused is None
    _run_recording_grids(inst, 2)
    _check_iterations(inst, seed, n, d, 2, u_of=lambda i: 1.0 - ora.uniform_f64(seed, i, n, d))


class SquaredPoints(PlainFlow):
    """x = u^2 with Jacobian prod 2u: the integral is unchanged."""

    def _digest_random_generation(self, xrand):
        return xrand * xrand, torch.prod(2.0 * xrand, dim=1)


def test_plain_digest_override_applies_xjac_and_limits():
    d, n, seed = 2, 50_000, 65
    xmin, xmax = np.array([0.5, -1.0]), np.array([2.0, 1.0])
    inst = SquaredPoints(d, n, seed=seed, verbose=False, xmin=xmin, xmax=xmax)
    inst.compile(lambda x: torch.prod(x, dim=1) ** 2)
    assert inst._fused is None
    inst.run_integration(1)
    r, s, _ = inst.history[0]
    u = ora.uniform_f64(seed, 0, n, d)
    x = xmin + u * u * (xmax - xmin)
    tmp = np.prod(x, axis=1) ** 2 * np.prod(2.0 * u, axis=1) * (1.0 / n) * np.prod(xmax - xmin)
    res = tmp.sum()
    assert abs(r - res) <= RTOL * abs(res)
    exact = np.prod((xmax ** 3 - xmin ** 3) / 3.0)
    assert abs(r - exact) < 4 * s


class OwnHistogram(VegasFlow):
    calls = 0

    def _importance_sampling_array_filling(self, results2, indices):
        type(self).calls += 1
        cols = [np.bincount(indices[:, j].cpu().numpy().astype(np.int64),
                            weights=results2.cpu().numpy().astype(np.float64), minlength=50)
                for j in range(self.n_dim)]
        return torch.tensor(np.stack(cols), dtype=torch.float64, device=self.device)


def test_histogram_override_feeds_refinement():
    d, n, seed = 3, 20_000, 66
    OwnHistogram.calls = 0
    inst = OwnHistogram(d, n, seed=seed, verbose=False)
    inst.compile(vf.Gaussian(A, C))
    assert inst._fused is None
    _run_recording_grids(inst, 2)
    assert OwnHistogram.calls == 2
    _check_iterations(inst, seed, n, d, 2)


class CountingRefine(VegasFlow):
    calls = 0

    def refine_grid(self, arr_res2):
        type(self).calls += 1
        super().refine_grid(arr_res2)


class CountingIteration(VegasFlow):
    calls = 0

    def _run_iteration(self):
        type(self).calls += 1
        return super()._run_iteration()


@pytest.mark.parametrize("cls", [CountingRefine, CountingIteration])
def test_per_iteration_overrides_run_every_iteration(cls):
    """Runs long enough to be replayed from a graph (20 >= 16 iterations)
    still call the override each iteration, on the fused kernels, and
    equal the plain class's results on the same seed."""
    d, n, seed = 4, 20_000, 67
    cls.calls = 0
    a = cls(d, n, seed=seed, verbose=False)
    a.compile(vf.Gaussian(A, C))
    b = VegasFlow(d, n, seed=seed, verbose=False)
    b.compile(vf.Gaussian(A, C))
    assert a._fused is not None and not a._replay_ok() and b._replay_ok()
    ra = a.run_integration(20)
    rb = b.run_integration(20)
    assert cls.calls == 20
    for (r1, s1, _), (r2, s2, _) in zip(a.history, b.history):
        assert abs(r1 - r2) <= 1e-10 * abs(r2) and abs(s1 - s2) <= 1e-8 * abs(s2)
    assert abs(ra[0] - rb[0]) <= 1e-10 * abs(rb[0])


class PlusCountingRefine(VegasFlowPlus):
    calls = 0

    def refine_grid(self, arr_res2):
        type(self).calls += 1
        super().refine_grid(arr_res2)


def test_vplus_refine_override_is_called():
    """Fixed allocation: the adaptive one takes int() of sums whose last
    bits depend on the order of the atomic adds, so two runs may place an
    event differently."""
    d, n, seed = 3, 40_000, 68
    PlusCountingRefine.calls = 0
    a = PlusCountingRefine(d, n, adaptive=False, seed=seed, verbose=False)
    a.compile(vf.GaussianMix())
    b = VegasFlowPlus(d, n, adaptive=False, seed=seed, verbose=False)
    b.compile(vf.GaussianMix())
    ra = a.run_integration(18)
    rb = b.run_integration(18)
    assert PlusCountingRefine.calls == 18
    assert abs(ra[0] - rb[0]) <= 1e-9 * abs(rb[0])


# --------------------------------------------------------------- sampler
def test_vegas_generate_random_array_matches_oracle():
    d, n, seed, n_draw = 4, 20_000, 69, 3000
    inst = VegasFlow(d, n, seed=seed, verbose=False)
    inst.compile(vf.Gaussian(A, C))
    inst.run_integration(2)
    grid = inst.divisions.cpu().numpy()
    draw = inst._draw_count  # the draw the sampler takes next
    x, px = inst.generate_random_array(n_draw)
    ox, ow, _ = ora.vegas_map(ora.uniform_f64(seed, draw, n_draw, d), grid)
    np.testing.assert_array_equal(x.cpu().numpy(), ox)
    # p(x) = w xjac / (xjac n) with xjac = 1/N (monte_carlo.py:245-246)
    want = (ow * (1.0 / n)) / ((1.0 / n) * n_draw)
    np.testing.assert_allclose(px.cpu().numpy(), want, rtol=1e-15)


def test_vplus_generate_random_array_concatenates_iterations():
    """n_events beyond one iteration: whole iterations of draws, one draw id
    each, concatenated and cut (vflowplus.py:173-185)."""
    d, n, seed = 2, 4000, 70
    inst = VegasFlowPlus(d, n, seed=seed, verbose=False)
    inst.compile(vf.GaussianMix())
    inst.run_integration(2)
    grid = inst.divisions.cpu().numpy()
    n_ev = inst.n_ev.cpu().numpy()
    per_it = int(n_ev.sum())
    n_draw = 2 * per_it + 123
    draw = inst._draw_count
    x, px = inst.generate_random_array(n_draw)
    st = ora.vplus_setup(d, n, False)
    cubes = ora.hypercubes(st["n_strat"], d)
    xs, ws = [], []
    for k in range(n_draw // per_it + 1):
        u = ora.uniform_f64(seed, draw + k, per_it, d)
        ox, ow, _, _ = ora.generate_samples_in_hypercubes(u, st["n_strat"], n_ev, cubes, grid)
        xjac = 1.0 / st["n_cubes"]
        xs.append(ox)
        ws.append((ow * xjac) / (xjac * per_it))
    ox = np.concatenate(xs)[:n_draw]
    opx = np.concatenate(ws)[:n_draw] * per_it / n_draw
    np.testing.assert_array_equal(x.cpu().numpy(), ox)
    np.testing.assert_allclose(px.cpu().numpy(), opx, rtol=1e-14)


# -------------------------------------------------- mixed, narrow limits
@pytest.mark.parametrize("train", [True, False])
def test_mixed_narrow_limits_stay_finite(train):
    """[0, 0.1]^16 with the product integrand: the per-dimension Jacobian
    constants are 500 each (500^16 overflows FP32). The mixed result equals
    the FP64 oracle on the mixed positions to FP32 accuracy, and the FP64
    kernel matches the closed form prod (0.1^2 / 2)."""
    d, n, seed = 16, 200_000, 71
    xmin, xmax = np.zeros(d), np.full(d, 0.1)
    out = {}
    for dtype in (torch.float64, torch.float32):
        inst = VegasFlow(d, n, seed=seed, verbose=False, xmin=xmin, xmax=xmax, dtype=dtype, train=train)
        inst.compile(vf.Product())
        assert inst._fused is not None
        inst.run_integration(1)
        out[dtype] = inst.history[0][:2]
    r32, s32 = out[torch.float32]
    assert np.isfinite(r32) and np.isfinite(s32)
    words = ora.hi_words(seed, 0, n, d)
    x, w, _ = ora.mixed_map(words, ora.initial_grid(d))
    x, w = ora.apply_xjac_and_limits(x, w, 1.0 / n, xmin, xmax)
    res = np.sum(w * ora.product(x))
    assert abs(r32 - res) <= 1e-4 * abs(res)
    r64, s64 = out[torch.float64]
    exact = (0.1 ** 2 / 2) ** d
    assert abs(r64 - exact) < 4 * s64


def test_zero_width_dimension_gives_zero_in_both_precisions():
    d, n = 5, 50_000
    xmin = np.array([0.0, 0.2, 0.3, 0.0, 0.0])
    xmax = np.array([1.0, 0.2, 0.9, 1.0, 1.0])  # dimension 1 has zero width
    for dtype in (torch.float64, torch.float32):
        inst = VegasFlow(d, n, seed=72, verbose=False, xmin=xmin, xmax=xmax, dtype=dtype)
        inst.compile(vf.Gaussian(0.3, 0.5))
        inst.run_integration(1)
        r, s, _ = inst.history[0]
        assert r == 0.0 and s == 0.0, (dtype, r, s)
