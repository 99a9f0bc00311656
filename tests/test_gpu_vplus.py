"""VegasFlowPlus kernels (mv_vplus_*) against the oracle's restatement of
vflowplus.py on the same draws.

Cube layout, per-cube sums (segmented wave scans + atomics), sigma over the
cubes, the adaptive reallocation of events (beta = 0.75) and the grid
refinement are compared with oracle/vegas.py. Bins and cube ids are
bit-exact; sums and edges within 1e-12 relative; the reallocated event counts
come from int() of a float whose denominator is summed in a different order,
so they must match exactly except for rare +-1 differences.
"""
import numpy as np
import pytest
import torch

from oracle import vegas as ora
from tests.parity import assert_sigma, kappa_vplus
from vegasflow_amd import kernels
from vegasflow_amd import integrands as vfi

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def random_grid(rng, d):
    g = np.empty((d, 51))
    for j in range(d):
        w = rng.gamma(0.7, size=50) + 1e-3
        c = np.concatenate([[0.0], np.cumsum(w)])
        g[j] = c / c[-1]
        g[j, 0], g[j, -1] = 0.0, 1.0
    return g


MIX = vfi.GaussianMix((0.5, 0.5), (0.05, 0.1), (0.33, 0.67))


def mix_np(x):
    return ora.gaussian_mix(x, (0.5, 0.5), (0.05, 0.1), (0.33, 0.67))


def _oracle_cube_sums(u, st, grid, d):
    """The oracle's per-cube sums ress and arr_var (vflowplus.py:209-217)."""
    x, w, ind, seg = ora.generate_samples_in_hypercubes(
        u, st["n_strat"], st["n_ev"], ora.hypercubes(st["n_strat"], d), grid)
    x, w = ora.apply_xjac_and_limits(x, w, 1.0 / st["n_cubes"])
    ress, arr_var, _ = ora.vplus_event(x, w, ind, seg, st["n_ev"], mix_np(x), False)
    return ress, arr_var


def _setup(d, n, adaptive, rng, ragged=True, n_strat=None):
    st = ora.vplus_setup(d, n, adaptive)
    if n_strat is not None:  # more cubes than the reference's 1e4 cap, or (n_strat, min_neval)
        n_strat, min_neval = n_strat if isinstance(n_strat, tuple) else (n_strat, 2)
        st["n_strat"], st["n_cubes"] = n_strat, n_strat**d
        st["min_neval"] = min_neval
    if ragged:
        st["n_ev"] = rng.integers(st["min_neval"], st["min_neval"] + 40, size=st["n_cubes"]).astype(np.int32)
    return st


def _device_strata(st):
    n_ev = dev(st["n_ev"], torch.int32)
    off = np.zeros(st["n_cubes"] + 1, dtype=np.int64)
    off[1:] = np.cumsum(st["n_ev"])
    offsets = dev(off, torch.int32)
    return n_ev, offsets, kernels.make_strata(st["n_strat"], st["n_cubes"], n_ev, offsets)


@pytest.mark.parametrize("d,n", [(2, 10000), (4, 100000), (3, 50000)])
def test_vplus_sample_bitwise(d, n):
    rng = np.random.default_rng(d)
    st = _setup(d, n, True, rng)
    grid = random_grid(rng, d)
    n_ev, offsets, strata = _device_strata(st)
    total = int(st["n_ev"].sum())
    step = kernels.make_step(d, total, precision="f64", train=True, seed=5, draw_id=2, xjac=1.0 / st["n_cubes"])
    x = torch.empty((total, d), dtype=torch.float64, device=DEV)
    w = torch.empty(total, dtype=torch.float64, device=DEV)
    idx = torch.empty((total, d), dtype=torch.uint8, device=DEV)
    seg = torch.empty(total, dtype=torch.int32, device=DEV)
    kernels.vplus_sample(step, strata, dev(grid), x, w, idx, seg)
    u = ora.uniform_f64(5, 2, total, d)
    cubes = ora.hypercubes(st["n_strat"], d)
    ox, ow, oind, oseg = ora.generate_samples_in_hypercubes(u, st["n_strat"], st["n_ev"], cubes, grid)
    ox, ow = ora.apply_xjac_and_limits(ox, ow, 1.0 / st["n_cubes"])
    np.testing.assert_array_equal(seg.cpu().numpy(), oseg)
    np.testing.assert_array_equal(idx.cpu().numpy(), oind)
    np.testing.assert_array_equal(x.cpu().numpy(), ox)
    np.testing.assert_array_equal(w.cpu().numpy(), ow)


@pytest.mark.parametrize("adaptive", [True, False])
@pytest.mark.parametrize("d,n,n_strat", [(2, 20000, None), (4, 200000, None), (5, 100000, None),
                                         (2, 200000, 150), (1, 30000, None), (15, 40000, None),
                                         (3, 400000, (8, 350))])
def test_vplus_fused_iteration_matches_oracle(d, n, n_strat, adaptive):
    """n_strat = 150 at d = 2 gives 22500 cubes, past the reference's cap of
    1e4 (the engine takes up to 2^24 - 1). d = 2..14 train through pair
    tables, d = 1 and 15 through per-dimension histograms. Cube sizes cover
    every path of the run loop: cubes of a few events (64-event trips across
    many cubes), of 20 to 60 (trips across one boundary), and of several
    hundred (n_strat = 8 at d = 3, 350 to 389 events per cube: trips inside
    one cube), and one cube of all events at d = 15."""
    rng = np.random.default_rng(10 * d + adaptive)
    st = _setup(d, n, adaptive, rng, n_strat=n_strat)
    grid = random_grid(rng, d)
    n_ev, offsets, strata = _device_strata(st)
    total = int(st["n_ev"].sum())
    ress = torch.zeros(st["n_cubes"], dtype=torch.float64, device=DEV)
    ress2 = torch.zeros_like(ress)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    g = dev(grid)
    step = kernels.make_step(d, total + 1000, precision="f64", train=True, seed=17, draw_id=3,
                             xjac=1.0 / st["n_cubes"])
    ng = kernels.vplus_fused(step, strata, MIX.to_struct(d), g, ress, ress2, ws)
    arr_var = torch.empty(st["n_cubes"], dtype=torch.float64, device=DEV)
    hist = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    out = kernels.vplus_finalize(d, strata, n_ev, offsets, ress, ress2, st["adaptive"], st["min_neval"],
                                 st["init_calls"], ora.BETA, ws, ng, True, True, g, arr_var, hist)
    torch.cuda.synchronize()
    u = ora.uniform_f64(17, 3, total, d)
    res, sig, new_div, new_n_ev, o_var, o_arr = ora.vplus_iteration(grid, st, u, mix_np)
    o = out.cpu().numpy()
    assert rel(o[0], res) < 1e-12
    assert rel(hist.cpu().numpy(), o_arr) < 1e-12
    np.testing.assert_allclose(g.cpu().numpy(), new_div, rtol=1e-12, atol=1e-15)
    # arr_var = ress2 * n_ev - ress^2 cancels where f is nearly flat over a
    # cube (narrow cubes, e.g. 7500 of them at d = 1): the summation-order
    # rounding of ress (a few ulp) then scales with ress^2, not with arr_var
    o_ress, _ = _oracle_cube_sums(u, st, grid, d)
    assert_sigma(o[1], sig, kappa_vplus(o_ress, o_var, st["n_ev"], sig))
    cancel = np.square(o_ress)
    scale = np.abs(o_var).max()
    got_var = arr_var.cpu().numpy()
    assert np.all(np.abs(got_var - o_var) <= 1e-9 * np.abs(o_var) + 1e-12 * scale + 1e-13 * cancel)
    got_n_ev = n_ev.cpu().numpy()
    diff = np.abs(got_n_ev.astype(np.int64) - new_n_ev.astype(np.int64))
    assert diff.max() <= 1
    well = np.abs(o_var) > 1e-6 * cancel  # variances not dominated by that rounding
    if well.any():
        assert (diff[well] > 0).mean() < 1e-3
    off = offsets.cpu().numpy()
    np.testing.assert_array_equal(off[1:], np.cumsum(got_n_ev))
    assert int(o[2]) == int(got_n_ev.sum())
    # per-cube accumulators are zeroed for the next iteration
    assert float(ress.abs().max()) == 0.0 and float(ress2.abs().max()) == 0.0


def test_vplus_digest_path_matches_fused():
    d, n = 4, 100000
    rng = np.random.default_rng(3)
    st = _setup(d, n, True, rng)
    grid = random_grid(rng, d)
    total = int(st["n_ev"].sum())
    outs = []
    for path in ("fused", "unfused"):
        n_ev, offsets, strata = _device_strata(st)
        g = dev(grid)
        ress = torch.zeros(st["n_cubes"], dtype=torch.float64, device=DEV)
        ress2 = torch.zeros_like(ress)
        ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
        step = kernels.make_step(d, total, precision="f64", train=True, seed=8, xjac=1.0 / st["n_cubes"])
        if path == "fused":
            ng = kernels.vplus_fused(step, strata, MIX.to_struct(d), g, ress, ress2, ws)
        else:
            x = torch.empty((total, d), dtype=torch.float64, device=DEV)
            w = torch.empty(total, dtype=torch.float64, device=DEV)
            idx = torch.empty((total, d), dtype=torch.uint8, device=DEV)
            seg = torch.empty(total, dtype=torch.int32, device=DEV)
            kernels.vplus_sample(step, strata, g, x, w, idx, seg)
            f = MIX(x)
            ng = kernels.vplus_digest(step, strata, f, w, idx, seg, ress, ress2, ws)
        out = kernels.vplus_finalize(d, strata, n_ev, offsets, ress, ress2, True, st["min_neval"],
                                     st["init_calls"], ora.BETA, ws, ng, True, True, g)
        outs.append((out.cpu().numpy(), g.cpu().numpy(), n_ev.cpu().numpy()))
    (o1, g1, e1), (o2, g2, e2) = outs
    assert rel(o1[0], o2[0]) < 1e-12
    assert rel(o1[1], o2[1]) < 1e-9
    np.testing.assert_allclose(g1, g2, rtol=1e-12, atol=1e-15)
    assert np.abs(e1.astype(int) - e2.astype(int)).max() <= 1


@pytest.mark.parametrize("pattern", ["one_event_each", "empty_runs"])
def test_vplus_fused_cube_search_tiny_and_empty_cubes(pattern):
    """Per-cube sums of the fused kernel (one search per wave, then a
    lane-parallel search over the next 64 offsets) equal those of the sample
    + digest path (a whole-array search per event) when a wave spans 64 cubes
    of one event each, and when runs of empty cubes push events past the
    64-offset window."""
    d, n_strat = 3, 12
    n_cubes = n_strat ** d
    rng = np.random.default_rng(5)
    if pattern == "one_event_each":
        n_ev = np.ones(n_cubes, dtype=np.int32)
    else:
        n_ev = rng.integers(0, 3, size=n_cubes).astype(np.int32)
        n_ev[100:300] = 0
        n_ev[700:720] = 0
        n_ev[0] = 1
    st = {"n_strat": n_strat, "n_cubes": n_cubes, "n_ev": n_ev}
    total = int(n_ev.sum())
    grid = random_grid(rng, d)
    sums = []
    for path in ("fused", "unfused"):
        _, _, strata = _device_strata(st)
        g = dev(grid)
        ress = torch.zeros(n_cubes, dtype=torch.float64, device=DEV)
        ress2 = torch.zeros_like(ress)
        ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, False))
        step = kernels.make_step(d, total, precision="f64", train=False, seed=4, xjac=1.0 / n_cubes)
        if path == "fused":
            kernels.vplus_fused(step, strata, MIX.to_struct(d), g, ress, ress2, ws)
        else:
            x = torch.empty((total, d), dtype=torch.float64, device=DEV)
            w = torch.empty(total, dtype=torch.float64, device=DEV)
            idx = torch.empty((total, d), dtype=torch.uint8, device=DEV)
            seg = torch.empty(total, dtype=torch.int32, device=DEV)
            kernels.vplus_sample(step, strata, g, x, w, idx, seg)
            kernels.vplus_digest(step, strata, MIX(x), w, idx, seg, ress, ress2, ws)
        sums.append((ress.cpu().numpy(), ress2.cpu().numpy()))
    (r1, q1), (r2, q2) = sums
    np.testing.assert_array_equal(r1 == 0.0, n_ev == 0)
    np.testing.assert_allclose(r1, r2, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(q1, q2, rtol=1e-12, atol=1e-300)


def test_vplus_mixed_tracks_f64():
    d, n = 4, 1_000_000
    st = ora.vplus_setup(d, n, True)
    grid = ora.initial_grid(d)
    res = {}
    for prec in ("f64", "mixed"):
        n_ev, offsets, strata = _device_strata(st)
        g = dev(grid)
        ress = torch.zeros(st["n_cubes"], dtype=torch.float64, device=DEV)
        ress2 = torch.zeros_like(ress)
        ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
        total = int(st["n_ev"].sum())
        step = kernels.make_step(d, total, precision=prec, train=True, seed=23, xjac=1.0 / st["n_cubes"])
        ng = kernels.vplus_fused(step, strata, MIX.to_struct(d), g, ress, ress2, ws)
        out = kernels.vplus_finalize(d, strata, n_ev, offsets, ress, ress2, True, st["min_neval"],
                                     st["init_calls"], ora.BETA, ws, ng, True, True, g)
        res[prec] = out.cpu().numpy()
    assert abs(res["mixed"][0] - res["f64"][0]) < 0.05 * res["f64"][1]
    assert abs(res["mixed"][1] - res["f64"][1]) < 0.01 * res["f64"][1]


@pytest.mark.parametrize("n_cubes", [7, 10000, 14641])
def test_vplus_finalize_matches_closed_form(n_cubes):
    """mv_vplus_finalize alone on synthetic per-cube sums (beyond 10240
    cubes a thread takes several chunks): sigma, the reallocation
    (vflowplus.py:153-163), offsets and zeroed accumulators."""
    rng = np.random.default_rng(n_cubes)
    n_strat = {7: 7, 10000: 10, 14641: 11}[n_cubes]
    d = {7: 1, 10000: 4, 14641: 4}[n_cubes]
    ne = rng.integers(2, 90, size=n_cubes).astype(np.int32)
    r = rng.random(n_cubes) * 1e-4
    r2 = (r ** 2 / ne) * (1.0 + rng.random(n_cubes))
    r2[::13] = r[::13] ** 2 / ne[::13] * 0.999  # some negative variances (clamped)
    n_ev = dev(ne, torch.int32)
    off = np.zeros(n_cubes + 1, dtype=np.int64)
    off[1:] = np.cumsum(ne)
    offsets = dev(off, torch.int32)
    strata = kernels.make_strata(n_strat, n_cubes, n_ev, offsets)
    ress, ress2 = dev(r), dev(r2)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    arr_var = torch.empty(n_cubes, dtype=torch.float64, device=DEV)
    init_calls, min_neval = 1_000_000, 2
    out = kernels.vplus_finalize(d, strata, n_ev, offsets, ress, ress2, True, min_neval, init_calls, ora.BETA,
                                 ws, 1, False, False, None, arr_var).cpu().numpy()
    var = r2 * ne - r ** 2
    np.testing.assert_allclose(arr_var.cpu().numpy(), var, rtol=1e-12, atol=1e-30)
    assert rel(out[0], r.sum()) < 1e-12
    assert rel(out[1], np.sqrt(np.sum(np.maximum(var, 0.0) / (ne - 1.0)))) < 1e-12
    want = ora.redistribute_samples(var, min_neval, init_calls)
    got = n_ev.cpu().numpy()
    diff = np.abs(got.astype(np.int64) - want.astype(np.int64))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
    o = offsets.cpu().numpy()
    assert o[0] == 0
    np.testing.assert_array_equal(o[1:], np.cumsum(got))
    assert int(out[2]) == int(got.sum())
    assert float(ress.abs().max()) == 0.0 and float(ress2.abs().max()) == 0.0


def test_vplus_fused_counter_crosses_2_32():
    """Events whose Philox counter crosses a multiple of 2^32 inside one
    launch (event_offset near 2^32) draw the oracle's uniforms."""
    d, n = 3, 30000
    off = (1 << 32) - 5001
    rng = np.random.default_rng(44)
    st = _setup(d, n, False, rng)
    grid = random_grid(rng, d)
    n_ev, offsets, strata = _device_strata(st)
    total = int(st["n_ev"].sum())
    ress = torch.zeros(st["n_cubes"], dtype=torch.float64, device=DEV)
    ress2 = torch.zeros_like(ress)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    g = dev(grid)
    step = kernels.make_step(d, total, precision="f64", train=True, event_offset=off, seed=6, draw_id=1,
                             xjac=1.0 / st["n_cubes"])
    ng = kernels.vplus_fused(step, strata, MIX.to_struct(d), g, ress, ress2, ws)
    out = kernels.vplus_finalize(d, strata, n_ev, offsets, ress, ress2, False, st["min_neval"],
                                 st["init_calls"], ora.BETA, ws, ng, True, True, g)
    u = ora.uniform_f64(6, 1, total, d, event_offset=off)
    res, sig, new_div, _, _, _ = ora.vplus_iteration(grid, st, u, mix_np)
    o = out.cpu().numpy()
    assert rel(o[0], res) < 1e-12
    o_ress, o_var = _oracle_cube_sums(u, st, grid, d)
    assert_sigma(o[1], sig, kappa_vplus(o_ress, o_var, st["n_ev"], sig))
    np.testing.assert_allclose(g.cpu().numpy(), new_div, rtol=1e-12, atol=1e-15)


def _mixed_words_u(seed, draw, total, d):
    # the mixed kernels sample u = (Philox high word) * 2^-32
    return ora.hi_words(seed, draw, total, d).astype(np.float64) * 2.0**-32


@pytest.mark.parametrize("d", [1, 4, 15])
def test_vplus_mixed_bins_match_oracle(d):
    """Mixed VEGAS+: the bins are those the oracle computes from the same
    words (u = hi * 2^-32, FP64 index arithmetic). A constant integrand on the
    initial grid gives every event the same tmp^2, so the histogram divided by
    it counts events per bin exactly. d = 4 trains through pair tables, d = 1
    and 15 through per-dimension histograms."""
    n = 60000
    st = ora.vplus_setup(d, n, False)
    grid = ora.initial_grid(d)
    n_ev, offsets, strata = _device_strata(st)
    total = int(st["n_ev"].sum())
    ress = torch.zeros(st["n_cubes"], dtype=torch.float64, device=DEV)
    ress2 = torch.zeros_like(ress)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    step = kernels.make_step(d, total, precision="mixed", train=True, seed=12, draw_id=2,
                             xjac=1.0 / st["n_cubes"])
    ng = kernels.vplus_fused(step, strata, vfi.Constant(2.5).to_struct(d), dev(grid), ress, ress2, ws)
    hist = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    kernels.vplus_finalize(d, strata, n_ev, offsets, ress, ress2, False, st["min_neval"], st["init_calls"],
                           ora.BETA, ws, ng, True, False, dev(grid), None, hist)
    u = _mixed_words_u(12, 2, total, d)
    _, _, ind, _ = ora.generate_samples_in_hypercubes(u, st["n_strat"], st["n_ev"],
                                                      ora.hypercubes(st["n_strat"], d), grid)
    tmp2 = (2.5 / st["n_cubes"] / st["min_neval"]) ** 2
    h = hist.cpu().numpy()
    for j in range(d):
        counts = np.bincount(ind[:, j].astype(np.int64), minlength=50)
        np.testing.assert_allclose(h[j] / tmp2, counts, rtol=1e-6)


@pytest.mark.parametrize("d", [4, 15])
def test_vplus_mixed_matches_oracle_on_words(d):
    """Mixed VEGAS+ on a random grid and ragged cubes: the FP64 oracle on the
    same words agrees to FP32 accuracy (map and integrand in FP32)."""
    n = 200000
    rng = np.random.default_rng(70 + d)
    st = _setup(d, n, False, rng)
    grid = random_grid(rng, d)
    n_ev, offsets, strata = _device_strata(st)
    total = int(st["n_ev"].sum())
    ress = torch.zeros(st["n_cubes"], dtype=torch.float64, device=DEV)
    ress2 = torch.zeros_like(ress)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    step = kernels.make_step(d, total, precision="mixed", train=True, seed=31, draw_id=4,
                             xjac=1.0 / st["n_cubes"])
    ng = kernels.vplus_fused(step, strata, MIX.to_struct(d), dev(grid), ress, ress2, ws)
    hist = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    out = kernels.vplus_finalize(d, strata, n_ev, offsets, ress, ress2, False, st["min_neval"],
                                 st["init_calls"], ora.BETA, ws, ng, True, False, dev(grid), None, hist)
    u = _mixed_words_u(31, 4, total, d)
    res, sig, _, _, _, o_arr = ora.vplus_iteration(grid, st, u, mix_np)
    o = out.cpu().numpy()
    assert rel(o[0], res) < 1e-5
    assert rel(o[1], sig) < 1e-4
    np.testing.assert_allclose(hist.cpu().numpy(), o_arr, rtol=1e-4, atol=1e-6 * o_arr.max())
