"""Kernel-level parity of libmivegas.so against the CPU oracle.

Every test runs the HIP kernel through the C ABI (vegasflow_amd.kernels)
and the oracle (oracle/vegas.py) on identical inputs: the same Philox draws
(seed, draw id, event index) or injected uniforms. Bar: bin indices and
Philox words bit-exact; FP64 sums, sigma, histograms and refined edges
within 1e-12 relative (the fused kernels may use FMAs; the sampling kernel
reproduces the reference's op order and is compared at 1e-15).
"""

import numpy as np
import pytest
import torch

from oracle import vegas as ora
from tests.parity import assert_sigma, kappa_vegas
from vegasflow_amd import kernels
from vegasflow_amd import integrands as vfi

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
RTOL = 1e-12


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def random_grid(rng, d):
    """Monotone grids with uneven bins, edges 0 and 1 exact."""
    g = np.empty((d, 51))
    for j in range(d):
        w = rng.gamma(0.7, size=50) + 1e-3
        c = np.concatenate([[0.0], np.cumsum(w)])
        g[j] = c / c[-1]
        g[j, 0], g[j, -1] = 0.0, 1.0
    return g


FAMILIES = {
    "gauss": (vfi.Gaussian(0.1, 0.5), lambda x: ora.gaussian(x, 0.1, 0.5)),
    "gauss_offc": (vfi.Gaussian(0.2, 0.37), lambda x: ora.gaussian(x, 0.2, 0.37)),
    "product": (vfi.Product(), ora.product),
    "mix": (vfi.GaussianMix((0.5, 0.5), (0.05, 0.1), (0.33, 0.67)),
            lambda x: ora.gaussian_mix(x, (0.5, 0.5), (0.05, 0.1), (0.33, 0.67))),
    "bw": (vfi.BreitWignerPower(0.5, 0.02, -0.7), lambda x: ora.breit_wigner_power(x, 0.5, 0.02, -0.7)),
    "const": (vfi.Constant(2.5), lambda x: ora.constant(x, 2.5)),
}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


# ---------------------------------------------------------------- draws
@pytest.mark.parametrize("d", [1, 3, 4, 8, 13])
def test_plain_sample_f64_matches_philox_oracle(d):
    n, seed, draw = 4099, 0x1234_5678_9ABC, 5
    step = kernels.make_step(d, n, precision="f64", train=False, event_offset=17, seed=seed, draw_id=draw)
    x = torch.empty((n, d), dtype=torch.float64, device=DEV)
    kernels.plain_sample(step, x)
    u = ora.uniform_f64(seed, draw, n, d, event_offset=17)
    np.testing.assert_array_equal(x.cpu().numpy(), u)


@pytest.mark.parametrize("d", [2, 5, 16])
def test_plain_sample_mixed_words_match_philox_oracle(d):
    n, seed, draw = 3000, 99, 2
    step = kernels.make_step(d, n, precision="mixed", train=False, seed=seed, draw_id=draw)
    x = torch.empty((n, d), dtype=torch.float32, device=DEV)
    kernels.plain_sample(step, x)
    words = ora.hi_words(seed, draw, n, d)
    # x = (hw >> 9) * 2^-23 + 2^-24 exactly
    got = np.round((x.cpu().numpy().astype(np.float64) - 2.0 ** -24) * 2.0 ** 23).astype(np.int64)
    np.testing.assert_array_equal(got, (words >> 9).astype(np.int64))


# ------------------------------------------------------------- sampling
@pytest.mark.parametrize("d", [1, 4, 8, 17])
def test_vegas_sample_f64_bitwise(d):
    rng = np.random.default_rng(d)
    grid = random_grid(rng, d)
    n, seed = 5000, 42
    step = kernels.make_step(d, n, precision="f64", train=True, seed=seed, draw_id=3, xjac=1.0 / n)
    x = torch.empty((n, d), dtype=torch.float64, device=DEV)
    w = torch.empty(n, dtype=torch.float64, device=DEV)
    idx = torch.empty((n, d), dtype=torch.uint8, device=DEV)
    kernels.vegas_sample(step, dev(grid), x, w, idx)
    u = ora.uniform_f64(seed, 3, n, d)
    ox, ow, oind = ora.vegas_map(u, grid)
    ox, ow = ora.apply_xjac_and_limits(ox, ow, 1.0 / n)
    np.testing.assert_array_equal(idx.cpu().numpy(), oind)
    np.testing.assert_array_equal(x.cpu().numpy(), ox)
    np.testing.assert_array_equal(w.cpu().numpy(), ow)


def test_vegas_sample_injected_edge_draws():
    """Draws at TECH_CUT, 1 - TECH_CUT and within an ulp of every bin edge."""
    d = 3
    edges = []
    for k in range(51):
        b = 1.0 - k / 50.0
        for v in (np.nextafter(b, 0.0), b, np.nextafter(b, 1.0)):
            if ora.TECH_CUT <= v <= 1.0 - ora.TECH_CUT:
                edges.append(v)
    edges += [ora.TECH_CUT, 1.0 - ora.TECH_CUT, 0.5, 0.02, 0.98]
    vals = np.array(edges)
    u = np.stack([vals, vals[::-1], np.roll(vals, 7)], axis=1)
    n = u.shape[0]
    grid = random_grid(np.random.default_rng(5), d)
    step = kernels.make_step(d, n, precision="f64", train=True, xjac=1.0 / n, inject=dev(u))
    x = torch.empty((n, d), dtype=torch.float64, device=DEV)
    w = torch.empty(n, dtype=torch.float64, device=DEV)
    idx = torch.empty((n, d), dtype=torch.uint8, device=DEV)
    kernels.vegas_sample(step, dev(grid), x, w, idx)
    ox, ow, oind = ora.vegas_map(u, grid)
    ox, ow = ora.apply_xjac_and_limits(ox, ow, 1.0 / n)
    np.testing.assert_array_equal(idx.cpu().numpy(), oind)
    np.testing.assert_array_equal(x.cpu().numpy(), ox)
    np.testing.assert_array_equal(w.cpu().numpy(), ow)


def test_vegas_sample_limits():
    d, n = 3, 2000
    xmin = np.array([-1.0, 0.5, 2.0])
    xmax = np.array([3.0, 0.75, 2.5])
    grid = random_grid(np.random.default_rng(11), d)
    step = kernels.make_step(d, n, precision="f64", train=True, seed=9, xjac=1.0 / n, xmin=dev(xmin),
                             xdelta=dev(xmax - xmin), xdeltajac=float(np.prod(xmax - xmin)))
    x = torch.empty((n, d), dtype=torch.float64, device=DEV)
    w = torch.empty(n, dtype=torch.float64, device=DEV)
    kernels.vegas_sample(step, dev(grid), x, w, None)
    u = ora.uniform_f64(9, 0, n, d)
    ox, ow, _ = ora.vegas_map(u, grid)
    ox, ow = ora.apply_xjac_and_limits(ox, ow, 1.0 / n, xmin, xmax)
    np.testing.assert_array_equal(x.cpu().numpy(), ox)
    assert rel(w.cpu().numpy(), ow) < 1e-15


# ---------------------------------------------------------------- fused
def _fused_once(fam, d, n, grid, train, seed=7, draw=0, limits=None, precision="f64", inject=None):
    fn, _ = FAMILIES[fam]
    kw = {}
    if limits is not None:
        xmin, xmax = limits
        kw = dict(xmin=dev(xmin), xdelta=dev(xmax - xmin), xdeltajac=float(np.prod(xmax - xmin)))
    step = kernels.make_step(d, n, precision=precision, train=train, seed=seed, draw_id=draw,
                             xjac=1.0 / n, inject=inject, **kw)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, train))
    g = dev(grid)
    ng = kernels.vegas_fused(step, fn.to_struct(d), g, ws)
    hist = torch.empty((d, 50), dtype=torch.float64, device=DEV) if train else None
    out = kernels.vegas_finalize(d, 1, train, train, n, ws, ng, grid=g, hist_out=hist)
    torch.cuda.synchronize()
    return out.cpu().numpy(), (hist.cpu().numpy() if train else None), g.cpu().numpy()


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_fused_event_range_across_a_2_32_counter_boundary(precision):
    """A step whose event counters cross a multiple of 2^32 is cut into
    launches with a uniform high counter word; the draws stay those of the
    oracle's 64-bit event counter."""
    d, n, off = 5, 20000, (1 << 32) - 7001
    grid = random_grid(np.random.default_rng(7), d)
    fn, ofn = FAMILIES["gauss_offc"]
    step = kernels.make_step(d, n, precision=precision, train=True, event_offset=off, seed=5, draw_id=9,
                             xjac=1.0 / n)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    ng = kernels.vegas_fused(step, fn.to_struct(d), dev(grid), ws)
    hist = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    out = kernels.vegas_finalize(d, 1, True, False, n, ws, ng, hist_out=hist).cpu().numpy()
    if precision == "f64":
        x, w, ind = ora.vegas_map(ora.uniform_f64(5, 9, n, d, event_offset=off), grid)
        tol = RTOL
    else:
        x, w, ind = ora.mixed_map(ora.hi_words(5, 9, n, d, event_offset=off), grid)
        tol = 1e-5
    x, w = ora.apply_xjac_and_limits(x, w, 1.0 / n)
    res, res2, arr = ora.vegas_event(x, w, ind, ofn(x))
    assert rel(out[0], res) < tol
    np.testing.assert_allclose(hist.cpu().numpy(), arr, rtol=max(tol, 1e-12) * 10, atol=1e-6 * arr.max())
    # plain MC across the same boundary
    pstep = kernels.make_step(d, n, precision=precision, train=False, event_offset=off, seed=5, draw_id=9,
                              xjac=1.0 / n)
    pws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, False))
    png = kernels.plain_fused(pstep, fn.to_struct(d), pws)
    pout = kernels.vegas_finalize(d, 1, False, False, n, pws, png).cpu().numpy()
    if precision == "f64":
        pres, _ = ora.plain_iteration(ora.uniform_f64(5, 9, n, d, event_offset=off), ofn, n)
        assert rel(pout[0], pres) < RTOL
        # the unfused sampler across the boundary, bit for bit
        xs = torch.empty((n, d), dtype=torch.float64, device=DEV)
        ws_ = torch.empty(n, dtype=torch.float64, device=DEV)
        idx = torch.empty((n, d), dtype=torch.uint8, device=DEV)
        kernels.vegas_sample(step, dev(grid), xs, ws_, idx)
        np.testing.assert_array_equal(idx.cpu().numpy(), ind)
        np.testing.assert_array_equal(xs.cpu().numpy(), x)
        np.testing.assert_array_equal(ws_.cpu().numpy(), w)
    else:
        u = (ora.hi_words(5, 9, n, d, event_offset=off) >> 9).astype(np.float64) * 2.0 ** -23 + 2.0 ** -24
        pres, _ = ora.plain_iteration(u, ofn, n)
        assert rel(pout[0], pres) < 1e-5


@pytest.mark.parametrize("fam", list(FAMILIES))
@pytest.mark.parametrize("d", [2, 4, 8])
def test_fused_f64_iteration_matches_oracle(fam, d):
    n = 20011
    grid = random_grid(np.random.default_rng(100 + d), d)
    out, hist, new_grid = _fused_once(fam, d, n, grid, True, seed=31, draw=4)
    u = ora.uniform_f64(31, 4, n, d)
    res, sig, arr, o_grid, *_ = ora.vegas_iteration(grid, u, FAMILIES[fam][1], n)
    assert rel(out[0], res) < RTOL
    if sig > 0:
        assert_sigma(out[1], sig, kappa_vegas(res, sig, n))
    assert rel(hist, arr) < RTOL
    np.testing.assert_allclose(new_grid, o_grid, rtol=RTOL, atol=1e-15)


@pytest.mark.parametrize("d", [1, 3, 13, 16])
def test_fused_f64_odd_dims_and_frozen(d):
    n = 7777
    grid = random_grid(np.random.default_rng(d), d)
    for train in (True, False):
        out, hist, new_grid = _fused_once("gauss_offc", d, n, grid, train, seed=5, draw=1)
        u = ora.uniform_f64(5, 1, n, d)
        res, sig, arr, o_grid, *_ = ora.vegas_iteration(grid, u, FAMILIES["gauss_offc"][1], n, train=train)
        assert rel(out[0], res) < RTOL
        if train:
            assert rel(hist, arr) < RTOL
            np.testing.assert_allclose(new_grid, o_grid, rtol=RTOL, atol=1e-15)
        else:
            np.testing.assert_array_equal(new_grid, grid)


def test_fused_f64_limits():
    d, n = 3, 30000
    xmin = np.array([-1.0, 0.5, 2.0])
    xmax = np.array([3.0, 0.75, 2.5])
    grid = random_grid(np.random.default_rng(3), d)
    out, hist, new_grid = _fused_once("product", d, n, grid, True, seed=12, limits=(xmin, xmax))
    u = ora.uniform_f64(12, 0, n, d)
    res, sig, arr, o_grid, *_ = ora.vegas_iteration(grid, u, ora.product, n, xmin=xmin, xmax=xmax)
    assert rel(out[0], res) < RTOL
    assert rel(hist, arr) < RTOL
    np.testing.assert_allclose(new_grid, o_grid, rtol=RTOL, atol=1e-15)


def test_fused_f64_injected_matches_oracle():
    d, n = 4, 3001
    rng = np.random.default_rng(8)
    u = rng.uniform(ora.TECH_CUT, 1 - ora.TECH_CUT, size=(n, d))
    grid = random_grid(rng, d)
    out, hist, new_grid = _fused_once("gauss", d, n, grid, True, inject=dev(u))
    res, sig, arr, o_grid, *_ = ora.vegas_iteration(grid, u, FAMILIES["gauss"][1], n)
    assert rel(out[0], res) < RTOL
    assert rel(hist, arr) < RTOL
    np.testing.assert_allclose(new_grid, o_grid, rtol=RTOL, atol=1e-15)


def test_fused_constant_has_zero_sigma():
    """Constant integrand on the initial grid: every tmp equals c * w/N."""
    d, n = 5, 100000
    out, hist, _ = _fused_once("const", d, n, ora.initial_grid(d), True)
    assert abs(out[0] - 2.5) < 1e-12
    assert out[1] < 1e-6
    # histogram rows all carry the same total mass sum(tmp^2)
    np.testing.assert_allclose(hist.sum(axis=1), hist.sum(axis=1)[0], rtol=1e-12)


def test_fused_large_event_count_multi_block():
    d, n = 8, 3_000_017
    grid = ora.initial_grid(d)
    out, hist, _ = _fused_once("gauss", d, n, grid, True, seed=11)
    u = ora.uniform_f64(11, 0, n, d)
    res, sig, arr, *_ = ora.vegas_iteration(grid, u, FAMILIES["gauss"][1], n)
    assert rel(out[0], res) < RTOL
    assert rel(hist, arr) < RTOL


# ------------------------------------------------------- unfused digest
@pytest.mark.parametrize("d", [3, 8, 12, 14, 20])
def test_sample_digest_path_matches_oracle(d):
    """Digest through pair tables read four bins per load (d = 8, 12), one
    bin per load (3, 14), and per-dimension histograms (20)."""
    n = 10007
    grid = random_grid(np.random.default_rng(40 + d), d)
    step = kernels.make_step(d, n, precision="f64", train=True, seed=2, draw_id=6, xjac=1.0 / n)
    x = torch.empty((n, d), dtype=torch.float64, device=DEV)
    w = torch.empty(n, dtype=torch.float64, device=DEV)
    idx = torch.empty((n, d), dtype=torch.uint8, device=DEV)
    g = dev(grid)
    kernels.vegas_sample(step, g, x, w, idx)
    f = vfi.Gaussian(0.3, 0.4)(x)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, True))
    ng = kernels.vegas_digest(step, f, w, idx, ws)
    hist = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    out = kernels.vegas_finalize(d, 1, True, True, n, ws, ng, grid=g, hist_out=hist)
    u = ora.uniform_f64(2, 6, n, d)
    res, sig, arr, o_grid, *_ = ora.vegas_iteration(grid, u, lambda xx: ora.gaussian(xx, 0.3, 0.4), n)
    o = out.cpu().numpy()
    assert rel(o[0], res) < RTOL
    assert_sigma(o[1], sig, kappa_vegas(res, sig, n))
    assert rel(hist.cpu().numpy(), arr) < RTOL
    np.testing.assert_allclose(g.cpu().numpy(), o_grid, rtol=RTOL, atol=1e-15)


def test_digest_vector_integrand_and_accumulate():
    """(n, 3) integrand, histogram of column 1, two accumulated steps."""
    d, n = 4, 6000
    rng = np.random.default_rng(1)
    f = rng.normal(size=(2 * n, 3))
    w = rng.uniform(size=2 * n)
    ind = rng.integers(0, 50, size=(2 * n, d)).astype(np.uint8)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 3, True))
    ngs = []
    for s in range(2):
        sl = slice(s * n, (s + 1) * n)
        step = kernels.make_step(d, n, precision="f64", train=True)
        ngs.append(kernels.vegas_digest(step, dev(f[sl]), dev(w[sl]), dev(ind[sl], torch.uint8), ws,
                                        main_dim=1))
    hist = torch.empty((d, 50), dtype=torch.float64, device=DEV)
    out = kernels.vegas_finalize(d, 3, True, False, 2 * n, ws, max(ngs), hist_out=hist, raw=True)
    o = out.cpu().numpy()
    tmp = w[:, None] * f
    assert rel(o[:3], tmp.sum(0)) < RTOL
    assert rel(o[3:], (tmp ** 2).sum(0)) < RTOL
    assert rel(hist.cpu().numpy(), ora.bin_histogram(tmp[:, 1] ** 2, ind.astype(np.int64), d)) < RTOL


# ----------------------------------------------------------------- refine
def test_refine_kernel_matches_oracle_on_varied_histograms():
    rng = np.random.default_rng(17)
    cases = [
        np.ones(50),
        rng.uniform(size=50),
        rng.exponential(size=50) ** 4,
        np.where(np.arange(50) == 20, 1.0, 0.0),
        np.where(np.arange(50) < 3, 5.0, 1e-40),
        np.zeros(50),
        np.linspace(1.0, 100.0, 50) ** 3,
    ]
    hist = np.stack(cases)
    grid = np.stack([random_grid(rng, 1)[0] if k % 2 else ora.initial_grid(1)[0] for k in range(len(cases))])
    g = dev(grid)
    kernels.refine_grid(dev(hist), g)
    got = g.cpu().numpy()
    for k in range(len(cases)):
        want = ora.refine_grid_per_dimension(hist[k], grid[k])
        np.testing.assert_allclose(got[k], want, rtol=RTOL, atol=1e-15)
        assert got[k][0] == 0.0 and got[k][-1] == 1.0


def test_consume_array_into_indices_kernel():
    rng = np.random.default_rng(4)
    for size_in, size_out in [(97, 13), (100000, 50), (5000, 5000)]:
        inp = rng.uniform(size=size_in)
        ind = rng.integers(0, size_out, size=size_in)
        got = kernels.consume_array_into_indices(dev(inp), dev(ind, torch.int32), size_out).cpu().numpy()
        want = ora.consume_array_into_indices(inp, ind, size_out) if size_in * size_out < 1e7 else \
            np.bincount(ind, weights=inp, minlength=size_out)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
        assert abs(got.sum() - inp.sum()) < 1e-9 * size_in


# ----------------------------------------------------------------- plain
@pytest.mark.parametrize("fam,d", [("gauss", 4), ("gauss", 1), ("gauss", 13), ("product", 4),
                                   ("product", 20), ("bw", 4), ("bw", 16)])
def test_plain_fused_matches_oracle(fam, d):
    """Kernels compiled for n_dim 1..16, and the runtime-n_dim one (d = 20)."""
    n = 50021
    fn, ofn = FAMILIES[fam]
    step = kernels.make_step(d, n, precision="f64", train=False, seed=21, draw_id=2, xjac=1.0 / n)
    ws = kernels.Workspace(DEV, kernels.partial_stride(d, 1, False))
    ng = kernels.plain_fused(step, fn.to_struct(d), ws)
    out = kernels.vegas_finalize(d, 1, False, False, n, ws, ng).cpu().numpy()
    res, sig = ora.plain_iteration(ora.uniform_f64(21, 2, n, d), ofn, n)
    assert rel(out[0], res) < RTOL
    assert_sigma(out[1], sig, kappa_vegas(res, sig, n))


# ----------------------------------------------------------------- mixed
@pytest.mark.parametrize("d", [1, 5, 8, 17])
def test_vegas_sample_mixed_matches_oracle_positions(d):
    """Mixed sampling: bins bit-exact with oracle.mixed_position on the same
    Philox words; x and w equal the FP64 evaluation of the same positions to
    FP32 rounding."""
    rng = np.random.default_rng(100 + d)
    grid = random_grid(rng, d)
    n, seed = 6000, 77
    step = kernels.make_step(d, n, precision="mixed", train=True, seed=seed, draw_id=1, xjac=1.0 / n)
    x = torch.empty((n, d), dtype=torch.float32, device=DEV)
    w = torch.empty(n, dtype=torch.float32, device=DEV)
    idx = torch.empty((n, d), dtype=torch.uint8, device=DEV)
    kernels.vegas_sample(step, dev(grid), x, w, idx)
    ox, ow, oind = ora.mixed_map(ora.hi_words(seed, 1, n, d), grid)
    np.testing.assert_array_equal(idx.cpu().numpy(), oind)
    np.testing.assert_allclose(x.cpu().numpy(), ox, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(w.cpu().numpy(), ow / n, rtol=4e-6 * d)
    # never exactly at the lower edge; the top edge can round to 1.0f in FP32
    assert x.min().item() > 0.0 and x.max().item() <= 1.0


@pytest.mark.parametrize("d", [1, 6, 9, 14, 16])
def test_mixed_fused_matches_oracle_positions(d):
    """The fused mixed iteration equals the FP64 oracle evaluated on the
    mixed positions (same words) to FP32 accuracy; bins are exact, so the
    histogram counts agree. d = 6, 9, 14 train through pair histograms,
    d = 1 and 16 through per-dimension ones."""
    n = 300000
    rng = np.random.default_rng(12)
    grid = random_grid(rng, d)
    out, hist, _ = _fused_once("gauss_offc", d, n, grid, True, seed=31, precision="mixed")
    x, w, ind = ora.mixed_map(ora.hi_words(31, 0, n, d), grid)
    x, w = ora.apply_xjac_and_limits(x, w, 1.0 / n)
    res, res2, arr = ora.vegas_event(x, w, ind, ora.gaussian(x, 0.2, 0.37))
    sig = ora.vegas_sigma(res, res2, n)
    assert rel(out[0], res) < 1e-5
    assert rel(out[1], sig) < 1e-4
    np.testing.assert_allclose(hist, arr, rtol=1e-4, atol=1e-6 * arr.max())


def test_mixed_fused_tracks_f64_on_same_draws():
    """The mixed kernel samples the leading 32 bits of each FP64 draw: its
    integral differs from the FP64 one by FP32 rounding only."""
    d, n = 8, 1_000_000
    grid = ora.initial_grid(d)
    o64, h64, _ = _fused_once("gauss", d, n, grid, True, seed=11)
    o32, h32, _ = _fused_once("gauss", d, n, grid, True, seed=11, precision="mixed")
    assert abs(o32[0] - o64[0]) < 0.02 * o64[1]
    assert abs(o32[1] - o64[1]) < 1e-3 * o64[1]
    assert rel(h32.sum(), h64.sum()) < 1e-4


@pytest.mark.parametrize("d,n", [(8, 37), (8, 64), (5, 262145), (8, 600037), (14, 524288 + 70)])
def test_mixed_pair_kernel_counts_every_event_once(d, n):
    """The pair kernel's event loop (trips in which a whole wave has events,
    then one trip tested per lane) takes every event of the step exactly
    once: per-bin counts of a constant integrand equal the oracle's integer
    bins for fewer events than a wave, exactly one wave, one event more than
    a full round of the grid (262144 threads), and several rounds with a
    partial wave at the end."""
    out, hist, _ = _fused_once("const", d, n, ora.initial_grid(d), True, seed=9, precision="mixed")
    bins, _ = ora.mixed_position(ora.hi_words(9, 0, n, d))
    tmp2 = (2.5 / n) ** 2
    assert abs(out[0] - 2.5) < 1e-5
    for j in range(d):
        counts = np.bincount(bins[:, j].astype(np.int64), minlength=50)
        np.testing.assert_allclose(hist[j] / tmp2, counts, rtol=1e-6)


@pytest.mark.parametrize("d", [2, 3, 11, 15])
def test_mixed_fused_bins_match_integer_oracle(d):
    """Mixed bins are those of oracle.mixed_position: compare the histogram
    of a constant integrand (counts per bin) with the oracle's integer bins
    (pair-histogram marginals for d <= 14, odd d with an unpaired last
    dimension included)."""
    n = 200000
    out, hist, _ = _fused_once("const", d, n, ora.initial_grid(d), True, seed=4, precision="mixed")
    bins, _ = ora.mixed_position(ora.hi_words(4, 0, n, d))
    tmp2 = (2.5 / n) ** 2
    for j in range(d):
        counts = np.bincount(bins[:, j].astype(np.int64), minlength=50)
        np.testing.assert_allclose(hist[j] / tmp2, counts, rtol=1e-6)
