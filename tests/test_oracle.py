"""Pins the CPU oracle (oracle/) before it is trusted as the parity reference.

Each part of the restatement is checked against something it does not
contain: published known-answer vectors (Philox), an independent
implementation (torch's Philox engine), brute force, closed forms and
invariants, and an independent vectorised formulation of Lepage's rebinning.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import vegas as ora
from oracle.philox import philox4x32_10

HERE = os.path.dirname(os.path.abspath(__file__))


def load_golden(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("source", ["random123_kat", "torch_engine"])
def test_philox_known_answers(source):
    doc = load_golden("philox_vectors.json")
    for case in doc[source]:
        ctr = [np.uint32(int(v, 16)) for v in case["ctr"]]
        key = [np.uint32(int(v, 16)) for v in case["key"]]
        got = ["%08x" % int(v) for v in philox4x32_10(*ctr, *key)]
        assert got == case["out"], case


def test_draws_range_moments_and_chunking():
    u = ora.uniform_f64(11, 0, 200_000, 5)
    assert u.min() >= ora.TECH_CUT and u.max() <= 1 - ora.TECH_CUT
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.var() - 1 / 12) < 2e-3
    counts = np.histogram(u[:, 2], bins=50, range=(0, 1))[0]
    chi2 = ((counts - counts.mean()) ** 2 / counts.mean()).sum()
    assert chi2 < 100  # 49 dof
    # event e of a draw is the same however the draw is split
    a = ora.uniform_f64(11, 3, 1000, 5)
    b = ora.uniform_f64(11, 3, 400, 5, event_offset=600)
    np.testing.assert_array_equal(a[600:], b)
    # draws and seeds give independent streams
    assert not np.array_equal(ora.uniform_f64(11, 4, 10, 5), ora.uniform_f64(11, 3, 10, 5))
    assert not np.array_equal(ora.uniform_f64(12, 3, 10, 5), ora.uniform_f64(11, 3, 10, 5))


def test_mixed_words_are_the_leading_bits_of_the_f64_draws():
    n, d = 5000, 7
    u = ora.uniform_f64(2, 1, n, d)
    hi = ora.hi_words(2, 1, n, d).astype(np.float64)
    r = (u - ora.TECH_CUT) / ((1 - ora.TECH_CUT) - ora.TECH_CUT)
    assert np.abs(np.floor(r * 2.0 ** 32) - hi).max() <= 1


# ------------------------------------------------------------- grid map
def test_identity_grid_map_brute_force():
    rng = np.random.default_rng(0)
    u = rng.uniform(ora.TECH_CUT, 1 - ora.TECH_CUT, size=(300, 3))
    x, w, ind = ora.vegas_map(u, ora.initial_grid(3))
    for e in range(u.shape[0]):
        for j in range(3):
            assert ind[e, j] == int(50.0 * (1.0 - float(u[e, j])))
    assert np.abs(x - (1.0 - u)).max() < 1e-15
    assert np.abs(w - 1.0).max() < 1e-14


def test_map_jacobian_integrates_to_one_on_any_grid():
    """One draw per bin: sum of the weights / 50 is sum of the bin widths."""
    rng = np.random.default_rng(1)
    w_ = rng.gamma(0.5, size=50) + 1e-4
    edges = np.concatenate([[0.0], np.cumsum(w_)]) / w_.sum()
    edges[-1] = 1.0
    u = (1.0 - (np.arange(50) + 0.5) / 50.0).reshape(-1, 1)
    x, w, ind = ora.vegas_map(u, edges.reshape(1, -1))
    np.testing.assert_array_equal(ind[:, 0], np.arange(50))
    assert abs(w.sum() / 50.0 - 1.0) < 1e-13
    # each point is the bin centre of its bin
    np.testing.assert_allclose(x[:, 0], 0.5 * (edges[:-1] + edges[1:]), rtol=1e-13)


def test_importance_sampling_is_unbiased_on_a_skewed_grid():
    rng = np.random.default_rng(2)
    d, n = 2, 400_000
    grid = np.stack([np.linspace(0, 1, 51) ** 2, np.sqrt(np.linspace(0, 1, 51))])
    u = rng.uniform(ora.TECH_CUT, 1 - ora.TECH_CUT, size=(n, d))
    x, w, _ = ora.vegas_map(u, grid)
    est = np.mean(w * ora.product(x))
    assert abs(est - 0.25) < 4 * np.std(w * ora.product(x)) / math.sqrt(n)


# --------------------------------------------------------- histogram
def test_consume_array_into_indices_brute_force():
    """test_utils.py:11-30 restated."""
    rng = np.random.default_rng(3)
    for _ in range(5):
        size_in = int(rng.integers(5, 100))
        size_out = int(rng.integers(1, size_in - 3))
        inp = rng.random(size_in)
        ind = rng.integers(0, size_out, size=size_in)
        res = ora.consume_array_into_indices(inp, ind.reshape(-1, 1), size_out)
        assert abs(inp.sum() - res.sum()) < 1e-12
        check = np.zeros(size_out)
        for val, i in zip(inp, ind):
            check[i] += val
        np.testing.assert_allclose(check, res)
        if size_out <= ora.BINS_MAX:
            hist = ora.bin_histogram(inp, ind.reshape(-1, 1), 1)[0]
            np.testing.assert_allclose(hist[:size_out], check)
            assert not hist[size_out:].any()


# ------------------------------------------------------------- refine
def rebin_by_interpolation(wei, div):
    """Lepage's rebinning written independently: new edge k is where the
    piecewise-linear cumulative weight reaches k * mean(wei)."""
    cum = np.concatenate([[0.0], np.cumsum(wei)])
    targets = np.arange(1, 50) * (cum[-1] / 50.0)
    inner = np.interp(targets, cum, div)
    return np.concatenate([[0.0], inner, [1.0]])


def smoothed_weights(r):
    """Independent statement of vflow.py:152-166: 3-point average with the
    end bins averaged over 2, floor 1e-30, then ((1-s/S)/ln(S/s))^1.5."""
    s = np.empty(50)
    for i in range(50):
        lo, hi = max(i - 1, 0), min(i + 1, 49)
        s[i] = r[lo:hi + 1].sum() / (hi - lo + 1)
    s = np.maximum(s, 1e-30)
    S = s.sum()
    return ((1 - s / S) / np.log(S / s)) ** 1.5


def test_refine_uniform_histogram_is_a_fixed_point():
    for div in (np.linspace(0, 1, 51), np.linspace(0, 1, 51) ** 2):
        new = ora.refine_grid_per_dimension(np.ones(50) * 3.7, div)
        np.testing.assert_allclose(new, div, rtol=1e-13, atol=1e-15)
        assert new[0] == 0.0 and new[-1] == 1.0


@pytest.mark.parametrize("seed", range(6))
def test_refine_matches_independent_rebin(seed):
    rng = np.random.default_rng(seed)
    r = rng.exponential(size=50) ** 3
    if seed % 2:
        r[rng.integers(0, 50, size=20)] = 0.0
    w_ = rng.gamma(1.0, size=50) + 1e-3
    div = np.concatenate([[0.0], np.cumsum(w_)]) / w_.sum()
    div[-1] = 1.0
    new = ora.refine_grid_per_dimension(r, div)
    expect = rebin_by_interpolation(smoothed_weights(r), div)
    np.testing.assert_allclose(new, expect, rtol=1e-11, atol=1e-14)
    assert np.all(np.diff(new) > 0)


def test_refine_concentrates_edges_around_a_spike():
    r = np.zeros(50)
    r[30] = 1.0
    new = ora.refine_grid_per_dimension(r, np.linspace(0, 1, 51))
    # the smoothed spike covers old bins 29..31, i.e. x in [0.58, 0.64]
    inside = np.sum((new > 0.58) & (new < 0.64))
    assert inside >= 40
    assert np.all(np.diff(new) > 0)


# ------------------------------------------------------ sums and sigma
def test_constant_integrand_sigma_is_zero():
    n = 10_000
    u = ora.uniform_f64(1, 0, n, 3)
    res, sig, *_ = ora.vegas_iteration(ora.initial_grid(3), u, lambda x: ora.constant(x, 2.0), n, train=False)
    assert abs(res - 2.0) < 1e-12
    assert sig < 1e-6
    pres, psig = ora.plain_iteration(u, lambda x: ora.constant(x, 2.0), n)
    assert abs(pres - 2.0) < 1e-12 and psig < 1e-6


def test_weighted_average_closed_form():
    res, sig = ora.weighted_average([(1.0, 0.1), (2.0, 0.2)])
    assert abs(res - (1 / 0.01 + 2 / 0.04) / (1 / 0.01 + 1 / 0.04)) < 1e-15
    assert abs(sig - math.sqrt(1 / (1 / 0.01 + 1 / 0.04))) < 1e-15


def test_gaussian_integral_within_errors():
    d, n = 4, 100_000
    hist, _ = ora.run_vegas(lambda x: ora.gaussian(x, 0.1, 0.5), d, n, 5, seed=7)
    res, sig = ora.weighted_average(hist[2:])
    exact = math.erf(5.0) ** 4  # 0.99999999999385025 (closed form)
    assert abs(exact - 0.99999999999385025) < 1e-15
    assert abs(res - exact) < 4 * sig


def test_limits_closed_form():
    rng = np.random.default_rng(4)
    d, n = 3, 200_000
    xmin = -1.0 + rng.random(d) * 2
    xmax = 3.0 + rng.random(d)
    res, sig, *_ = ora.vegas_iteration(ora.initial_grid(d), ora.uniform_f64(3, 0, n, d), ora.product, n,
                                       train=False, xmin=xmin, xmax=xmax)
    exact = np.prod(xmax ** 2 / 2 - xmin ** 2 / 2)
    assert abs(res - exact) < 4 * sig


# ------------------------------------------------------------- VEGAS+
def test_stratification_table():
    """n_strat / n_cubes / n_ev / sum for the documented cases."""
    for row in load_golden("vplus_strata.json")["cases"]:
        st = ora.vplus_setup(row["n_dim"], row["n_events"], row["adaptive"])
        assert (st["n_strat"], st["n_cubes"], st["min_neval"], int(st["n_ev"].sum())) == \
            (row["n_strat"], row["n_cubes"], row["n_ev"], row["total"]), row


def test_hypercube_order_and_digits():
    for n_strat, d in [(3, 4), (10, 2), (2, 5)]:
        cubes = ora.hypercubes(n_strat, d)
        assert cubes.tolist() == [list(p) for p in itertools.product(range(n_strat), repeat=d)]
        for c in range(cubes.shape[0]):
            digits = [(c // n_strat ** (d - 1 - j)) % n_strat for j in range(d)]
            assert digits == cubes[c].tolist()


def test_stratified_points_stay_in_their_cube():
    d, n = 3, 20000
    st = ora.vplus_setup(d, n, True)
    rng = np.random.default_rng(5)
    st["n_ev"] = rng.integers(2, 9, size=st["n_cubes"]).astype(np.int32)
    tot = int(st["n_ev"].sum())
    u = ora.uniform_f64(4, 0, tot, d)
    cubes = ora.hypercubes(st["n_strat"], d)
    x, w, ind, seg = ora.generate_samples_in_hypercubes(u, st["n_strat"], st["n_ev"], cubes, ora.initial_grid(d))
    assert np.all(np.diff(seg) >= 0)
    np.testing.assert_array_equal(np.bincount(seg, minlength=st["n_cubes"]), st["n_ev"])
    h = cubes[seg]
    lo, hi = h / st["n_strat"], (h + 1) / st["n_strat"]
    assert np.all(x >= lo - 1e-12) and np.all(x <= hi + 1e-12)
    f = ora.gaussian(x, 0.2, 0.4)
    ress, arr_var, _ = ora.vplus_event(x, w, ind, seg, st["n_ev"], f, train=False)
    assert abs(ress.sum() - np.sum(w * f)) < 1e-12 * abs(ress.sum())


def test_redistribution_closed_form():
    arr_var = np.array([1.0, 16.0, 0.0, -1e-20])
    got = ora.redistribute_samples(arr_var, min_neval=2, init_calls=1000)
    damped = np.array([1.0, 16.0 ** 0.375, 0.0, 0.0])
    expect = np.maximum(2, damped * 1000 / 2 / damped.sum()).astype(np.int32)
    np.testing.assert_array_equal(got, expect)
    assert got.sum() <= 4 * 2 + 500


def test_torch_cpu_baseline_matches_the_oracle_on_the_same_draws():
    """The timed CPU baseline (torch op-for-op restatement with the dense
    one-hot histogram) computes what the NumPy oracle computes."""
    import torch

    from oracle import torch_cpu as tc

    d, n = 5, 3000
    rng = np.random.default_rng(4)
    grid = ora.initial_grid(d)
    grid = ora.refine_grid(rng.random((d, 50)) + 0.1, grid)  # a non-uniform grid
    u = ora.uniform_f64(3, 0, n, d)
    res, res2, arr = tc.run_step(torch.from_numpy(np.ascontiguousarray(grid)), n, n, tc.gaussian, None,
                                 rnds=torch.from_numpy(u))
    x, w, ind = ora.vegas_map(u, grid)
    x, w = ora.apply_xjac_and_limits(x, w, 1.0 / n)
    ores, ores2, oarr = ora.vegas_event(x, w, ind, ora.gaussian(x, 0.1, 0.5))
    assert abs(float(res) - ores) <= 1e-12 * abs(ores)
    assert abs(float(res2) - ores2) <= 1e-12 * abs(ores2)
    np.testing.assert_allclose(arr.numpy(), oarr, rtol=1e-12, atol=1e-300)


def test_mixed_position_exact_and_tracks_the_f64_draws():
    """The mixed path's (bin, frac) is xn = 50 - q / 2^23 with
    q = floor(((50 2^23 - 10) (w + 1/2) + 5 2^32) / 2^32) exactly (rational
    brute force); u = q / (50 2^23) lies inside [TECH_CUT, 1 - TECH_CUT] like
    the FP64 draws; xn is within 2^-21 of the FP64 position of the same word
    (50 (1 - u) with u = TECH_CUT + (w + 1/2) / 2^32 (1 - 2 TECH_CUT)); and
    the bin is the FP64 draw's but for points within ~2^-21 of an edge."""
    from fractions import Fraction

    words = ora.hi_words(9, 2, 4000, 3).reshape(-1)
    extreme = np.array([0, 1, 9, 10, 511, 512, 0x7FFFFFFF, 0xFFFFFDFF, 0xFFFFFFFF], dtype=np.uint32)
    words = np.concatenate([words, extreme])
    ind, frac = ora.mixed_position(words)
    K = 50 * 2 ** 23 - 10
    tc = Fraction(ora.TECH_CUT)
    for k in list(range(0, words.size, 37)) + list(range(words.size - extreme.size, words.size)):
        w = int(words[k])
        q = (K * (2 * w + 1) + 10 * 2 ** 32) // 2 ** 33
        xn = 50 - Fraction(q, 2 ** 23)
        assert ind[k] == 49 - q // 2 ** 23 and Fraction(float(frac[k])) == xn - ind[k]
        u = Fraction(q, 50 * 2 ** 23)
        assert tc <= u <= 1 - tc
        u64 = tc + Fraction(2 * w + 1, 2 ** 33) * (1 - 2 * tc)
        assert abs(xn - 50 * (1 - u64)) < Fraction(1, 2 ** 21)
    assert ind.min() >= 0 and ind.max() <= 49
    assert frac.min() > 0 and frac.max() < 1
    i0, f0 = ora.mixed_position(np.array([0, 0xFFFFFFFF], dtype=np.uint32))
    assert (i0[0], f0[0]) == (49, 1.0 - 5 * 2.0 ** -23) and (i0[1], f0[1]) == (0, 6 * 2.0 ** -23)
    n, d = 20000, 4
    _, _, ind64 = ora.vegas_map(ora.uniform_f64(5, 1, n, d), ora.initial_grid(d))
    ind32, _ = ora.mixed_position(ora.hi_words(5, 1, n, d))
    assert np.mean(ind32 != ind64) < 1e-3


def _cuts_by_prefix_counts(t_res_sq, div):
    """The rebin as the GPU computes it: cut k independently from the prefix
    sums of the weights (count of prefix sums below k * ave)."""
    r = np.asarray(t_res_sq, dtype=np.float64)
    pad = np.concatenate([[0.0], r, [0.0]])
    s = np.maximum(((pad[1:-1] + pad[2:]) + pad[:-2]) / np.array([2.0] + [3.0] * 48 + [2.0]), 1e-30)
    S = np.sum(s)
    wei = np.power((1.0 - s / S) / (np.log(S) - np.log(s)), 1.5)
    ave = np.sum(wei) / 50
    c = np.cumsum(wei)
    out = [0.0]
    for k in range(1, 50):
        t = k * ave
        n = min(int(np.sum(c < t)), 49)
        cur, prev = div[n + 1], (0.0 if n == 0 else div[n])
        out.append(cur - (cur - prev) * (c[n] - t) / wei[n])
    return np.array(out + [1.0])


def test_parallel_cut_form_equals_serial_rebin():
    """Cut k of the serial 49-cut loop (vflow.py:194-207) stops in the first
    bin whose running weight reaches k * ave: the per-cut form agrees with
    the serial oracle to rounding on varied histograms."""
    rng = np.random.default_rng(8)
    cases = [np.ones(50), np.zeros(50), np.eye(50)[17] * 3.0, np.eye(50)[0], np.eye(50)[49],
             rng.gamma(0.3, size=50), rng.random(50) ** 8, np.exp(-np.linspace(-4, 4, 50) ** 2 * 3)]
    for k, h in enumerate(cases):
        div = ora.initial_grid(1)[0] if k % 2 == 0 else np.sort(np.concatenate([[0.0, 1.0], rng.random(49)]))
        a = ora.refine_grid_per_dimension(h, div)
        b = _cuts_by_prefix_counts(h, div)
        np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-16)


# ------------------------------------------------- integrand closed forms
def test_breit_wigner_power_separable_closed_forms():
    """oracle.breit_wigner_power against integrals it does not contain:
    quadrature of its x_0 factor (at x_1 = 1) is the normalised Cauchy mass
    on [0, 1], (atan(0.5/G2) - atan(-0.5/G2))/pi with G2 = width/2 = 0.01;
    quadrature of its x_1 factor (at x_0 = M, divided by the peak value) is
    int_0^1 x^-0.7 = 10/3; the function is flat in the other dimensions and
    factorises, so its integral is the product 3.2908976726601815."""
    from scipy import integrate

    f = ora.breit_wigner_power
    g2 = 0.01
    bw_mass = (math.atan(0.5 / g2) - math.atan(-0.5 / g2)) / math.pi
    assert abs(bw_mass - 0.98726930179805439) < 1e-15
    i0, _ = integrate.quad(lambda t: f(np.array([[t, 1.0, 0.3]]))[0], 0.0, 1.0, points=[0.5],
                           epsabs=0.0, epsrel=1e-13, limit=200)
    assert abs(i0 - bw_mass) < 1e-13
    peak = f(np.array([[0.5, 1.0, 0.3]]))[0]
    assert abs(peak - 1.0 / (math.pi * g2)) < 1e-12 * peak
    i1, _ = integrate.quad(lambda t: f(np.array([[0.5, t, 0.3]]))[0] / peak, 0.0, 1.0,
                           epsabs=0.0, epsrel=1e-12, limit=200)
    assert abs(i1 - 10.0 / 3.0) < 1e-11
    assert abs(bw_mass * 10.0 / 3.0 - 3.2908976726601815) < 1e-15
    # factorisation and flatness at random points
    rng = np.random.default_rng(6)
    x = rng.uniform(1e-6, 1.0, size=(64, 5))
    ref = np.array([[0.5, 1.0, 0.3, 0.3, 0.3]])
    fx = f(x)
    x0_only = np.repeat(ref, 64, axis=0)
    x0_only[:, 0] = x[:, 0]
    x1_only = np.repeat(ref, 64, axis=0)
    x1_only[:, 1] = x[:, 1]
    np.testing.assert_allclose(fx, f(x0_only) * f(x1_only) / peak, rtol=1e-13)
    # another power: int x^p = 1/(1 + p)
    i1b, _ = integrate.quad(lambda t: ora.breit_wigner_power(np.array([[0.5, t]]), power=-0.3)[0] / peak,
                            0.0, 1.0, epsabs=0.0, epsrel=1e-12, limit=200)
    assert abs(i1b - 1.0 / 0.7) < 1e-11


def test_gaussian_and_mixture_closed_forms():
    """oracle.gaussian / gaussian_mix against products of 1-d quadratures
    (the Gaussian factorises over dimensions)."""
    from scipy import integrate

    for a, c in [(0.1, 0.5), (0.2, 0.37), (0.05, 0.33)]:
        one, _ = integrate.quad(lambda t: ora.gaussian(np.array([[t]]), a, c)[0], 0.0, 1.0,
                                points=[c], epsabs=0.0, epsrel=1e-13, limit=200)
        assert abs(one - 0.5 * (math.erf((1 - c) / a) + math.erf(c / a))) < 1e-13
        # d = 3: product of the 1-d factors at a random point
        x = np.array([[0.41, 0.52, 0.6]])
        got = ora.gaussian(x, a, c)[0]
        want = np.prod([ora.gaussian(np.array([[t]]), a, c)[0] for t in x[0]])
        assert abs(got - want) <= 1e-13 * want
    # C3 mixture, normalised like C1: 0.9999969422972157
    exact = 0.0
    for wk, ak, ck in [(0.5, 0.05, 0.33), (0.5, 0.1, 0.67)]:
        exact += wk * (0.5 * (math.erf((1 - ck) / ak) + math.erf(ck / ak))) ** 4
    assert abs(exact - 0.9999969422972157) < 1e-15


# ------------------------------------------------------ sigma pull tests
def _pull_moments(pulls):
    p = np.asarray(pulls)
    return float(p.mean()), float(p.std())


def _assert_unit_pulls(res, sig, exact):
    """(res - exact)/sigma over independent seeds has mean ~0 and standard
    deviation ~1 when sigma estimates the error of res; a sigma off by a
    factor (the wrong N - 1, a missing square) fails the band."""
    res = np.asarray(res)
    sig = np.asarray(sig)
    mean, std = _pull_moments((res - exact) / sig)
    n = res.size
    assert abs(mean) < 4.0 / math.sqrt(n), mean
    assert 0.85 <= std <= 1.15, std
    # the band rejects a sigma too large or too small by 1.5x
    for factor in (1.5, 1.0 / 1.5):
        assert not 0.85 <= _pull_moments((res - exact) / (sig * factor))[1] <= 1.15


def test_vegas_sigma_pull_distribution():
    """vflow.py:437-438 on an adapted (non-uniform) grid, 400 seeds."""
    fn = lambda x: ora.gaussian(x, 0.3, 0.5)  # noqa: E731
    exact = math.erf(0.5 / 0.3) ** 2
    _, grid = ora.run_vegas(fn, 2, 20000, 3, seed=0)
    assert np.abs(grid - ora.initial_grid(2)).max() > 1e-3  # the grid did adapt
    n = 2000
    out = [ora.vegas_iteration(grid, ora.uniform_f64(1000 + s, 0, n, 2), fn, n, train=False)[:2]
           for s in range(400)]
    _assert_unit_pulls([o[0] for o in out], [o[1] for o in out], exact)


def test_plain_sigma_pull_distribution():
    """plain.py:37-43, 400 seeds."""
    fn = lambda x: ora.gaussian(x, 0.3, 0.45)  # noqa: E731
    exact = (0.5 * (math.erf(0.55 / 0.3) + math.erf(0.45 / 0.3))) ** 3
    n = 2000
    out = [ora.plain_iteration(ora.uniform_f64(2000 + s, 0, n, 3), fn, n) for s in range(400)]
    _assert_unit_pulls([o[0] for o in out], [o[1] for o in out], exact)


def test_vplus_sigma_pull_distribution():
    """vflowplus.py:222-233 with adapted per-cube event counts (first
    iteration's reallocation), 300 seeds."""
    fn = lambda x: ora.gaussian(x, 0.3, 0.5)  # noqa: E731
    exact = math.erf(0.5 / 0.3) ** 2
    st = ora.vplus_setup(2, 8000, adaptive=True)
    grid = ora.initial_grid(2)
    u = ora.uniform_f64(1, 0, int(st["n_ev"].sum()), 2)
    _, _, grid, new_n_ev, *_ = ora.vplus_iteration(grid, st, u, fn)
    st["n_ev"] = new_n_ev
    assert new_n_ev.min() != new_n_ev.max()  # uneven allocation
    tot = int(new_n_ev.sum())
    out = [ora.vplus_iteration(grid, st, ora.uniform_f64(3000 + s, 0, tot, 2), fn, train=False)[:2]
           for s in range(300)]
    _assert_unit_pulls([o[0] for o in out], [o[1] for o in out], exact)
