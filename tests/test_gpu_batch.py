"""Batched integrals (mv_batch_*): each item must reproduce a lone VegasFlow
run with the same seed (the same Philox draws) up to summation order, and a
captured HIP graph must reproduce eager rounds replay after replay.
"""
import os
import tempfile

import numpy as np
import pytest
import torch

import vegasflow_amd as vf
from oracle import vegas as ora
from tests.parity import RTOL, assert_history, assert_sigma, kappa_vegas
from vegasflow_amd import autotune, kernels
from vegasflow_amd.batch import BatchVegas, random_batch

pytestmark = pytest.mark.gpu


def small_specs():
    return [
        (vf.Gaussian(0.2, [0.3, 0.5, 0.6]), 3, 20000, 101),
        (vf.Product(), 2, 5000, 102),
        (vf.GaussianMix(), 4, 30011, 103),
        (vf.BreitWignerPower(), 6, 10000, 104),
        (vf.Constant(3.0), 1, 1000, 105),
        (vf.Gaussian(0.1, 0.5), 12, 70000, 106),
        (vf.Gaussian(0.15, 0.45), 16, 9000, 107),
    ]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_items_match_lone_runs(dtype):
    specs = small_specs()
    b = BatchVegas(specs, dtype=dtype, events_per_tile=2048, blocks_per_cu=2)
    b.run(3, graph=False)
    hist = b.history()
    assert hist.shape == (len(specs), 3, 2)
    for i, (fn, d, n, seed) in enumerate(specs):
        lone = vf.VegasFlow(d, n, seed=seed, verbose=False, dtype=dtype)
        lone.compile(fn)
        lone.run_integration(3)
        want = np.array([[(r, s) for r, s, _ in lone.history]])
        assert_history(hist[i:i + 1], want, [n], rtol=1e-10)
        np.testing.assert_allclose(b.grid(i).cpu().numpy(), lone.divisions.cpu().numpy(),
                                   rtol=1e-10, atol=1e-13)
    np.testing.assert_array_equal(b.iterations(), 3)


def pair_specs():
    return [
        (vf.Gaussian(0.2, [0.3, 0.5, 0.6]), 3, 20000, 101),
        (vf.Product(), 2, 5000, 102),
        (vf.GaussianMix(), 4, 30011, 103),
        (vf.BreitWignerPower(), 6, 10000, 104),
        (vf.Constant(3.0), 1, 1000, 105),
        (vf.Gaussian(0.1, 0.5), 12, 70000, 106),
        (vf.Gaussian(0.15, 0.45), 13, 9000, 107),
        (vf.Gaussian(0.12, 0.55), 14, 40000, 108),
    ]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_pair_histograms_match_lone_runs(dtype):
    """Batches through mv_batch_fused_pairs / mv_batch_fused_pairs_f64 (tiles
    of 16384 events over 1024-thread workgroups), replayed from a graph: every
    item equals its lone run (odd n_dim, n_dim = 1 and 14 included)."""
    specs = pair_specs()
    b = BatchVegas(specs, dtype=dtype, events_per_tile=16384, pair_histograms=True)
    assert b.pairs and all(v > 0 for v in b._pair_blocks.values())
    b.run(3, graph=True)
    hist = b.history()
    for i, (fn, d, n, seed) in enumerate(specs):
        lone = vf.VegasFlow(d, n, seed=seed, verbose=False, dtype=dtype)
        lone.compile(fn)
        lone.run_integration(3)
        want = np.array([[(r, s) for r, s, _ in lone.history]])
        assert_history(hist[i:i + 1], want, [n], rtol=1e-10)
        np.testing.assert_allclose(b.grid(i).cpu().numpy(), lone.divisions.cpu().numpy(),
                                   rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_pair_kernel_frozen_items_match_per_wave_kernel(dtype):
    """Items that do not train run the plain event body of their precision
    inside the pair kernel: same sums as the per-wave kernel on the same
    draws."""
    specs = pair_specs()
    a = BatchVegas(specs, dtype=dtype, train=False, events_per_tile=4096, blocks_per_cu=4)
    b = BatchVegas(specs, dtype=dtype, train=False, events_per_tile=65536, pair_histograms=True)
    a.run(2, graph=False)
    b.run(2, graph=False)
    assert_history(b.history(), a.history(), [n for _, _, n, _ in specs], rtol=1e-10)


def test_batch_graph_replays_match_eager_rounds():
    specs = small_specs()
    eager = BatchVegas(specs, events_per_tile=4096, blocks_per_cu=4)
    graph = BatchVegas(specs, events_per_tile=4096, blocks_per_cu=4)
    graph.capture()
    for k in range(4):
        eager.step(graph=False)
        graph.step(graph=True)
        torch.cuda.synchronize()
        # eager work and a synchronisation between replays
        scratch = torch.ones(1000, device="cuda").sum()
        assert float(scratch) == 1000.0
        np.testing.assert_allclose(graph.last()[0], eager.last()[0], rtol=1e-12)
        np.testing.assert_allclose(graph.last()[1], eager.last()[1], rtol=1e-9)
        np.testing.assert_allclose(graph.grids.cpu().numpy(), eager.grids.cpu().numpy(), rtol=1e-11,
                                   atol=1e-14)
    np.testing.assert_array_equal(graph.iterations(), 4)
    np.testing.assert_allclose(graph.results()[0], eager.results()[0], rtol=1e-12)


def test_autotune_picks_and_caches_a_shape():
    path = tempfile.mktemp(suffix=".json")
    os.environ["MIVEGAS_TUNE_CACHE"] = path
    try:
        specs = random_batch(16, seed=5)
        b1 = BatchVegas(specs)
        assert b1.events_per_tile in autotune.EVENTS_PER_TILE + autotune.PAIR_EVENTS_PER_TILE
        assert os.path.exists(path)
        autotune._memory.clear()
        b2 = BatchVegas(specs)
        assert (b2.events_per_tile, b2.blocks, b2.pairs) == (b1.events_per_tile, b1.blocks, b1.pairs)
        # mixed precision: the pair-histogram kernel competes too
        b3 = BatchVegas(specs, dtype=torch.float32)
        assert b3.events_per_tile in autotune.EVENTS_PER_TILE + autotune.PAIR_EVENTS_PER_TILE
    finally:
        os.environ.pop("MIVEGAS_TUNE_CACHE", None)
        if os.path.exists(path):
            os.remove(path)


def test_random_batch_integrals_agree_with_closed_forms():
    specs = random_batch(48, seed=23, d_range=(2, 6), a_range=(0.1, 0.3))
    b = BatchVegas(specs, events_per_tile=4096, blocks_per_cu=4)
    b.run(8)
    res, err = b.results()
    exact = np.array([fn.exact(d) for fn, d, _, _ in specs])
    pulls = (res - exact) / err
    assert np.all(np.isfinite(pulls))
    assert np.mean(np.abs(pulls) < 4) > 0.9


def _philox_uniforms(seed, draw, n, d):
    """The oracle's uniforms (oracle.uniform_f64) drawn by the plain sampling
    kernel, which reproduces them bit for bit (test_gpu_kernels.py checks
    it); the oracle's map, sums and refinement then run on the host."""
    step = kernels.make_step(d, n, precision="f64", train=False, seed=seed, draw_id=draw)
    u = torch.empty((n, d), dtype=torch.float64, device="cuda")
    kernels.plain_sample(step, u)
    return u.cpu().numpy()


def test_c4_batch_full_config_against_the_oracle():
    """The batched configuration in full: random_batch(256, seed=23) (d 2..12,
    1e4..1e6 events, 8 iterations) in FP64, rounds replayed from the captured
    graph. A sample of items, the five with the largest |pull| against their
    closed form and the three with the fewest events, is checked against the
    oracle on the same draws:

    * every iteration at the 1e-12 bar against the oracle's iteration started
      from the engine's own grid (integral, sigma with its cancellation
      factor, refined edges);
    * the whole run against the oracle's own eight iterations from the
      linspace grid within 1e-7, and the same pull. The two runs' grids
      drift apart by the rounding of the histogram sums (atomic arrival
      order on the GPU, left to right in the oracle), which every rebinning
      carries into the next iteration and can amplify where a cut falls in
      a bin of small weight: measured up to 1.0e-10 (iteration 6) and
      1.4e-9 (iteration 7) on two runs whose autotuner picked different
      kernels, i.e. different addition orders.

    The pull tail therefore belongs to the estimator (the inverse-variance
    average of monte_carlo.py:713-732 over iterations whose grids are still
    adapting), not to a kernel."""
    specs = random_batch(256, seed=23)
    b = BatchVegas(specs, dtype=torch.float64)
    snaps = [b.grids.cpu().numpy().copy()]
    for _ in range(8):
        b.step(graph=True)
        snaps.append(b.grids.cpu().numpy().copy())
    res, err = b.results()
    hist = b.history()
    assert hist.shape == (256, 8, 2)
    exact = np.array([fn.exact(d) for fn, d, _, _ in specs])
    pulls = (res - exact) / err
    worst = [int(i) for i in np.argsort(-np.abs(pulls))[:5]]
    small = [int(i) for i in np.argsort([n for _, _, n, _ in specs]) if int(i) not in worst][:3]
    for i in worst + small:
        fn, d, n, seed = specs[i]
        off = b._grid_offsets[i]
        engine_grids = [g[off:off + d * 51].reshape(d, 51) for g in snaps]
        f = lambda x: ora.gaussian(x, fn.a, fn.centre)  # noqa: E731
        grid = ora.initial_grid(d)
        o_hist = []
        for k in range(8):
            u = _philox_uniforms(seed, k, n, d)
            r, s, _, g_next, *_ = ora.vegas_iteration(engine_grids[k], u, f, n)
            assert abs(hist[i, k, 0] - r) <= RTOL * abs(r), (i, k)
            assert_sigma(hist[i, k, 1], s, kappa_vegas(r, s, n))
            np.testing.assert_allclose(engine_grids[k + 1], g_next, rtol=RTOL, atol=1e-15)
            r, s, _, grid, *_ = ora.vegas_iteration(grid, u, f, n)
            o_hist.append((r, s))
            assert abs(hist[i, k, 0] - r) <= 1e-7 * abs(r), (i, k)
        o_res, o_err = ora.weighted_average(o_hist)
        assert abs(res[i] - o_res) <= 1e-7 * abs(o_res), i
        o_pull = (o_res - exact[i]) / o_err
        assert abs(o_pull - pulls[i]) <= 1e-6 * max(1.0, abs(pulls[i])), i
