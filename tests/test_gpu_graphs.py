"""run_integration replays a captured one-iteration HIP graph for fused
built-in integrands, with the draw id in a device counter that every
finalize advances. The replayed run must draw exactly what the eager loop
draws: same results (to summation order), grids, events and history, across
several run_integration calls, a frozen grid, both precisions and the three
algorithms."""
import numpy as np
import pytest
import torch

import vegasflow_amd as vf

pytestmark = pytest.mark.gpu


def _pair(cls, *args, **kwargs):
    out = []
    for graphs in (False, True):
        inst = cls(*args, verbose=False, **kwargs)
        inst.use_graphs = graphs
        inst.graph_min_iterations = 2  # capture even short runs here
        out.append(inst)
    return out


def _same_history(a, b, rtol=1e-12):
    assert len(a.history) == len(b.history)
    for (ra, ea, _), (rb, eb, _) in zip(a.history, b.history):
        assert abs(ra - rb) <= rtol * abs(ra) + 1e-300
        assert abs(ea - eb) <= 1e-9 * abs(ea) + 1e-300


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_vegasflow_replayed_iterations_match_eager(dtype):
    eager, graph = _pair(vf.VegasFlow, 8, 200_000, seed=11, dtype=dtype)
    fn = vf.Gaussian(0.1, 0.5)
    for inst in (eager, graph):
        inst.compile(fn)
        inst.run_integration(3)
        inst.run_integration(2)  # a second run replays the captured graph
        inst.freeze_grid()       # a new shape: frozen-grid iteration
        inst.run_integration(3)
    assert graph._graphs, "the replayed path was not taken"
    _same_history(eager, graph)
    np.testing.assert_allclose(graph.divisions.cpu().numpy(), eager.divisions.cpu().numpy(),
                               rtol=1e-11, atol=1e-14)
    assert graph._draw_count == eager._draw_count == 8


def test_long_runs_replay_one_graph():
    """A long run replays the one captured iteration many times."""
    eager, graph = _pair(vf.VegasFlow, 6, 100_000, seed=19, dtype=torch.float32)
    fn = vf.Gaussian(0.15, 0.45)
    for inst in (eager, graph):
        inst.compile(fn)
        inst.run_integration(19)
        inst.run_integration(7)
    assert len(graph._graphs) == 1
    _same_history(eager, graph)
    assert graph._draw_count == eager._draw_count == 26


def test_short_first_runs_stay_eager():
    """Below graph_min_iterations a new shape runs eagerly (a capture would
    cost more than it saves); a long run captures and later short runs of
    the same shape replay."""
    inst = vf.VegasFlow(4, 50_000, seed=2, verbose=False)
    inst.compile(vf.Gaussian(0.2, 0.5))
    inst.run_integration(3)
    assert not inst._graphs
    inst.run_integration(inst.graph_min_iterations)
    assert len(inst._graphs) == 1
    inst.run_integration(2)
    assert len(inst._graphs) == 1 and inst._draw_count == 5 + inst.graph_min_iterations


def test_large_iterations_stay_eager():
    """Iterations above graph_max_events run eagerly even in long runs, and
    give the same results as a replayed instance with the limit raised."""
    big = vf.VegasFlow(4, 50_000, seed=6, verbose=False)
    big.graph_max_events = 10_000
    big.compile(vf.Gaussian(0.2, 0.5))
    r1 = big.run_integration(big.graph_min_iterations)
    assert not big._graphs
    small = vf.VegasFlow(4, 50_000, seed=6, verbose=False)
    small.compile(vf.Gaussian(0.2, 0.5))
    r2 = small.run_integration(small.graph_min_iterations)
    assert small._graphs
    assert abs(r1[0] - r2[0]) <= 1e-12 * abs(r2[0]) and abs(r1[1] - r2[1]) <= 1e-9 * abs(r2[1])


def test_vegasflow_replay_with_limits_and_single_iterations():
    eager, graph = _pair(vf.VegasFlow, 3, 50_000, seed=3, xmin=[0.0, -1.0, 2.0], xmax=[1.0, 1.0, 3.5])
    fn = vf.Product()
    for inst in (eager, graph):
        inst.compile(fn)
        inst.run_integration(2)  # captures
        for _ in range(3):
            inst.run_integration(1)  # one replay each
    assert len(graph._graphs) == 1
    _same_history(eager, graph)


def test_plainflow_replayed_iterations_match_eager():
    eager, graph = _pair(vf.PlainFlow, 5, 100_000, seed=5)
    fn = vf.Gaussian(0.2, 0.4)
    for inst in (eager, graph):
        inst.compile(fn)
        inst.run_integration(4)
        inst.run_integration(2)
    assert graph._graphs
    _same_history(eager, graph)


@pytest.mark.parametrize("adaptive", [False, True])
def test_vegasflowplus_replayed_iterations_match_eager(adaptive):
    eager, graph = _pair(vf.VegasFlowPlus, 4, 100_000, seed=7, adaptive=adaptive)
    fn = vf.GaussianMix((0.5, 0.5), (0.05, 0.1), (0.33, 0.67))
    for inst in (eager, graph):
        inst.compile(fn)
        inst.run_integration(3)
        inst.run_integration(3)
    assert graph._graphs
    _same_history(eager, graph, rtol=1e-11)
    assert graph.events_history == eager.events_history
    assert graph.n_events == eager.n_events
    np.testing.assert_array_equal(graph._n_ev.cpu().numpy(), eager._n_ev.cpu().numpy())
    np.testing.assert_allclose(graph.divisions.cpu().numpy(), eager.divisions.cpu().numpy(),
                               rtol=1e-10, atol=1e-13)


def test_stream_handle_follows_torch_current_stream():
    """The kernels launch on torch's current stream, also inside a stream
    context (as during a graph capture)."""
    from vegasflow_amd import kernels

    assert (kernels._stream().value or 0) == torch.cuda.current_stream().cuda_stream
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        assert (kernels._stream().value or 0) == side.cuda_stream
    assert (kernels._stream().value or 0) == torch.cuda.current_stream().cuda_stream
