"""NumPy restatement of the reference VEGAS / VEGAS+ / plain MC hot path.

Test infrastructure only (see ``oracle/__init__.py``): the product package
never imports this module.

Every function cites the reference file:line it restates (paths relative to
the vegasflow 1.4.0 source tree). Arithmetic is FP64 and follows the
reference's operation order wherever a float feeds an integer decision (bin
indices, the float32 stratification count); elsewhere the HIP kernels are
compared to this module within a relative tolerance of 1e-12.

Uniform draws are injected (``u`` arguments) or produced by this module's
restatement of the engine's Philox stream (``uniform_f64`` / ``hi_words``),
because the reference's ``tf.random.uniform`` stream is not reproducible
without TensorFlow.
"""
import itertools
import math

import numpy as np

from .philox import event_words

# src/vegasflow/configflow.py:13-24
BINS_MAX = 50
ALPHA = 1.5
BETA = 0.75
TECH_CUT = 1e-8
MAX_EVENTS_LIMIT = int(1e6)
MAX_NEVAL_HCUBE = int(1e4)
FBINS = float(BINS_MAX)


# ----------------------------------------------------------------------------
# draws
# ----------------------------------------------------------------------------
def n_groups(n_dim):
    """Number of 4-dimension groups (one Philox block of high words each)."""
    return (n_dim + 3) // 4


def uniform_f64(seed, draw_id, n, n_dim, event_offset=0):
    """FP64 uniforms in [TECH_CUT, 1 - TECH_CUT] for events
    ``event_offset .. event_offset + n - 1``.

    Dimension j of an event takes the high word of its 64 random bits from
    word j of Philox blocks 0..G-1 and the low word from word j of blocks
    G..2G-1 (G = ceil(n_dim / 4)); r = (hi:lo >> 12) * 2^-52 in [0, 1), the
    mantissa of a double in [1, 2) minus one.
    The affine step follows tf.random.uniform's ``r * (max - min) + min``
    with max = 1 - TECH_CUT, min = TECH_CUT (monte_carlo.py:264-266).
    """
    g = n_groups(n_dim)
    ev = np.arange(event_offset, event_offset + n, dtype=np.uint64)
    words = event_words(seed, draw_id, ev, 2 * g)
    hi = words[:, : 4 * g][:, :n_dim].astype(np.uint64)
    lo = words[:, 4 * g :][:, :n_dim].astype(np.uint64)
    bits = (hi << np.uint64(20)) | (lo >> np.uint64(12))
    r = (bits | np.uint64(0x3FF0000000000000)).view(np.float64) - 1.0
    span = (1.0 - TECH_CUT) - TECH_CUT
    return r * span + TECH_CUT


def hi_words(seed, draw_id, n, n_dim, event_offset=0):
    """The 32 high bits of each FP64 draw (uint32 (n, n_dim)): the whole
    random input of the mixed-precision path."""
    g = n_groups(n_dim)
    ev = np.arange(event_offset, event_offset + n, dtype=np.uint64)
    words = event_words(seed, draw_id, ev, g)
    return words[:, :n_dim].copy()


MIXED_MUL = 50 * 2 ** 23 - 10
MIXED_ADD = 5 * 2 ** 32 + MIXED_MUL // 2


def mixed_position(words):
    """Bin and fraction of the mixed-precision path (the engine's own form;
    the reference has only FP64/FP32 uniforms). The mirrored coordinate
    m = 50 - xn = 50 u is held to 23 fraction bits as
    q = floor((K (word + 1/2) + 5 2^32) / 2^32), K = 50 2^23 - 10, which runs
    from 5 to 50 2^23 - 6: u = q / (50 2^23) stays inside
    [TECH_CUT, 1 - TECH_CUT] (monte_carlo.py:264-266) and the point is
    xn = 50 - q / 2^23 exactly. Returns (bin = 49 - (q >> 23),
    frac = 1 - (q mod 2^23) / 2^23), i.e. xn = bin + frac exactly, with
    frac in (0, 1] (frac = 1: the bin's upper edge)."""
    q = (np.asarray(words).astype(np.int64) * MIXED_MUL + MIXED_ADD) >> 32
    mfrac = (q & (2 ** 23 - 1)).astype(np.float64) * 2.0 ** -23
    return (49 - (q >> 23)).astype(np.int32), 1.0 - mfrac


def mixed_map(words, divisions):
    """x (n, d), w (n,), ind (n, d) of the mixed path on ``words``, evaluated
    in FP64 (the GPU evaluates the same positions in FP32)."""
    ind, frac = mixed_position(words)
    d = divisions.shape[0]
    cols = np.arange(d)[None, :]
    lo = divisions[cols, ind]
    dx = divisions[cols, ind + 1] - lo
    x = lo + dx * frac
    w = np.prod(dx * FBINS, axis=1)
    return x, w, ind


# ----------------------------------------------------------------------------
# grid
# ----------------------------------------------------------------------------
def initial_grid(n_dim):
    """vflow.py:238-242: every dimension starts from np.linspace(0, 1, 51)."""
    sub = np.linspace(0, 1, BINS_MAX + 1)
    return sub.repeat(n_dim).reshape(-1, n_dim).T.copy()


def importance_sampling_digest(xn, divisions):
    """vflow.py:33-83. ``xn`` is (n_dim, n) FP64, ``divisions`` (n_dim, 51).

    Returns ind (n, n_dim) int32, x (n, n_dim), w (n,).
    """
    ind_i = xn.astype(np.int32)  # int_me truncates (vflow.py:67)
    ind_f = ind_i + 1
    x_ini = np.take_along_axis(divisions, ind_i, axis=1)  # gather batch_dims=1 (70)
    x_fin = np.take_along_axis(divisions, ind_f, axis=1)  # (71)
    xdelta = x_fin - x_ini  # (73)
    aux_rand = xn - np.floor(xn)  # (75)
    x = x_ini + xdelta * aux_rand  # (76)
    weights = np.prod(xdelta * FBINS, axis=0)  # (78)
    return ind_i.T.copy(), x.T.copy(), weights


def vegas_map(u, divisions):
    """vflow.py:87-126: xn = FBINS * (1 - u^T), then the shared digest.

    Returns x (n, n_dim), w (n,), ind (n, n_dim) like _generate_random_array.
    """
    xn = FBINS * (1.0 - u.T)  # (117)
    ind, x, w = importance_sampling_digest(xn, divisions)
    return x, w, ind


def apply_xjac_and_limits(x, w, xjac, xmin=None, xmax=None):
    """monte_carlo.py:270-274: w *= xjac, then optional integration limits."""
    w = w * xjac
    if xmin is not None:
        xmin = np.asarray(xmin, dtype=np.float64)
        xdelta = np.asarray(xmax, dtype=np.float64) - xmin
        x = xmin + x * xdelta
        w = w * np.prod(xdelta)
    return x, w


# ----------------------------------------------------------------------------
# histogram primitive
# ----------------------------------------------------------------------------
def consume_array_into_indices(input_arr, indices, result_size):
    """utils.py:10-44: out[b] = sum of input_arr where indices == b.

    Restated through the same dense one-hot comparison the reference builds
    (utils.py:40-43); ``indices`` is (n, 1) or (n,).
    """
    idx = np.asarray(indices).reshape(-1)
    all_bins = np.arange(result_size)
    eq = idx[None, :] == all_bins[:, None]
    return np.where(eq, np.asarray(input_arr, dtype=np.float64)[None, :], 0.0).sum(axis=1)


def bin_histogram(tmp2, ind, n_dim):
    """vflow.py:370-387 via np.bincount (same sums as the one-hot form)."""
    out = np.empty((n_dim, BINS_MAX), dtype=np.float64)
    for j in range(n_dim):
        out[j] = np.bincount(ind[:, j], weights=tmp2, minlength=BINS_MAX)[:BINS_MAX]
    return out


# ----------------------------------------------------------------------------
# refinement
# ----------------------------------------------------------------------------
def refine_grid_per_dimension(t_res_sq, subdivisions):
    """vflow.py:129-211 (Lepage, J. Comput. Phys. 27 (1978) 192).

    Smoothing evaluated as (r_i + r_{i+1}) + r_{i-1} (vflow.py:158), floor at
    1e-30 (159), weights ((1 - s/S) / (log S - log s))^ALPHA (162-166) and the
    serial 49-cut rebin with prev/cur starting at 0.0 (194-208).

    One addition: the cut loop stops at the last bin (n_bin <= 49); the
    reference has no such guard and would index past the weight array if
    rounding left the running weight short of the average at the end.
    """
    r = np.asarray(t_res_sq, dtype=np.float64)
    sub = np.asarray(subdivisions, dtype=np.float64)
    padded = np.concatenate([[0.0], r, [0.0]])
    meaner = np.full(BINS_MAX, 3.0)
    meaner[0] = meaner[-1] = 2.0
    smeared_tmp = padded[1:-1] + padded[2:] + padded[:-2]
    smeared = np.maximum(smeared_tmp / meaner, 1e-30)
    sum_t = float(np.sum(smeared))
    log_t = np.log(smeared)
    aux_t = (1.0 - smeared / sum_t) / (math.log(sum_t) - log_t)
    wei_t = np.power(aux_t, ALPHA)
    ave_t = float(np.sum(wei_t)) / BINS_MAX

    new_bins = [0.0]
    bin_weight = 0.0
    n_bin = -1
    cur = 0.0
    prev = 0.0
    for _ in range(BINS_MAX - 1):
        while bin_weight < ave_t and n_bin < BINS_MAX - 1:
            n_bin += 1
            bin_weight += wei_t[n_bin]
            prev = cur
            cur = sub[n_bin + 1]
        bin_weight -= ave_t
        delta = (cur - prev) * bin_weight / wei_t[n_bin]
        new_bins.append(cur - delta)
    new_bins.append(1.0)
    return np.array(new_bins)


def refine_grid(arr_res2, divisions):
    """vflow.py:349-362: refine every dimension."""
    out = divisions.copy()
    for j in range(divisions.shape[0]):
        out[j] = refine_grid_per_dimension(arr_res2[j], divisions[j])
    return out


# ----------------------------------------------------------------------------
# integrands (the built-in synthetic families, FP64)
# ----------------------------------------------------------------------------
def gaussian(x, a=0.1, centre=0.5):
    """(1/(a sqrt(pi)))^d exp(-sum((x - c)/a)^2); integrates to ~1 on [0,1]^d
    for centre 0.5 (the test integrand of test_algs.py:27-36 without its
    zero-valued range term)."""
    d = x.shape[1]
    c = np.broadcast_to(np.asarray(centre, dtype=np.float64), (d,))
    pref = (1.0 / (a * math.sqrt(math.pi))) ** d
    coef = np.sum(np.square((x - c) / a), axis=1)
    return pref * np.exp(-coef)


def product(x):
    """prod_i x_i (test_misc.py:46-48); integral 2^-d on [0,1]^d."""
    return np.prod(x, axis=1)


def gaussian_mix(x, weights, widths, centres):
    """sum_k w_k * gaussian(x, a_k, c_k)."""
    out = np.zeros(x.shape[0])
    for wk, ak, ck in zip(weights, widths, centres):
        out = out + wk * gaussian(x, ak, ck)
    return out


def breit_wigner_power(x, mass=0.5, width=0.02, power=-0.7):
    """Normalised Cauchy in x_0 (FWHM = width) times x_1^power; the other
    dimensions are flat."""
    g2 = 0.5 * width
    bw = (g2 / math.pi) / ((x[:, 0] - mass) ** 2 + g2 * g2)
    return bw * np.power(x[:, 1], power)


def constant(x, value=1.0):
    return np.full(x.shape[0], float(value))


# ----------------------------------------------------------------------------
# VegasFlow
# ----------------------------------------------------------------------------
def vegas_event(x, w, ind, f, train=True):
    """vflow.py:389-430 for a scalar integrand: returns res, res2, arr_res2."""
    tmp = w * f
    tmp2 = np.square(tmp)
    res = np.sum(tmp)
    res2 = np.sum(tmp2)
    arr = bin_histogram(tmp2, ind, x.shape[1]) if train else None
    return res, res2, arr


def vegas_sigma(res, res2, n_events):
    """vflow.py:437-438."""
    err_tmp2 = (n_events * res2 - np.square(res)) / (n_events - 1.0)
    return np.sqrt(np.maximum(err_tmp2, 0.0))


def weighted_average(results):
    """monte_carlo.py:713-732 over a list of (res, sigma)."""
    aux_res = 0.0
    weight_sum = 0.0
    for res, sigma in results:
        wgt_tmp = 1.0 / pow(sigma, 2)
        aux_res += res * wgt_tmp
        weight_sum += wgt_tmp
    return aux_res / weight_sum, np.sqrt(1.0 / weight_sum)


def vegas_iteration(divisions, u, integrand, n_events, train=True, xmin=None, xmax=None):
    """One VegasFlow iteration on injected uniforms ``u`` (n_events, d).

    monte_carlo.py:249-275 + vflow.py:389-442. Returns
    (res, sigma, arr_res2, new_divisions, x, w, ind).
    """
    x, w, ind = vegas_map(u, divisions)
    x, w = apply_xjac_and_limits(x, w, 1.0 / n_events, xmin, xmax)
    f = integrand(x)
    res, res2, arr = vegas_event(x, w, ind, f, train)
    sigma = vegas_sigma(res, res2, n_events)
    new_div = refine_grid(arr, divisions) if train else divisions.copy()
    return res, sigma, arr, new_div, x, w, ind


def run_vegas(integrand, n_dim, n_events, n_iter, seed, draw_id0=0, divisions=None,
              train=True, xmin=None, xmax=None):
    """Iterations of VegasFlow with the engine's Philox draws; draw ``i`` of
    the run uses draw_id = draw_id0 + i. Returns (history, divisions)."""
    div = initial_grid(n_dim) if divisions is None else divisions.copy()
    hist = []
    for i in range(n_iter):
        u = uniform_f64(seed, draw_id0 + i, n_events, n_dim)
        res, sigma, _, div, *_ = vegas_iteration(div, u, integrand, n_events, train, xmin, xmax)
        hist.append((res, sigma))
    return hist, div


# ----------------------------------------------------------------------------
# PlainFlow
# ----------------------------------------------------------------------------
def plain_iteration(u, integrand, n_events, xmin=None, xmax=None):
    """plain.py:18-43 on injected uniforms: returns (res, sigma)."""
    x, w = apply_xjac_and_limits(u, np.ones(1), 1.0 / n_events, xmin, xmax)
    tmp = integrand(x) * w
    res = np.sum(tmp)
    raw_res2 = np.sum(np.square(tmp))
    res2 = raw_res2 * n_events
    err_tmp2 = (res2 - np.square(res)) / (n_events - 1.0)
    return res, np.sqrt(np.maximum(err_tmp2, 0.0))


# ----------------------------------------------------------------------------
# VegasFlowPlus
# ----------------------------------------------------------------------------
def vplus_setup(n_dim, n_events, adaptive=False):
    """vflowplus.py:105-139: stratification in float32 as the reference's
    tf.math.pow on Python floats does.

    Returns dict(n_strat, n_cubes, min_neval, n_ev (int32 array), adaptive).
    """
    if n_dim > 13 and adaptive:
        adaptive = False
    if adaptive:
        neval_eff = int(n_events / 2)
    else:
        neval_eff = n_events
    n_strat = np.floor(np.power(np.float32(neval_eff / 2), np.float32(1 / n_dim)))
    if np.power(np.float32(n_strat), np.float32(n_dim)) > np.float32(MAX_NEVAL_HCUBE):
        n_strat = np.floor(np.power(np.float32(1e4), np.float32(1 / n_dim)))
    n_strat = int(np.int32(n_strat))
    n_cubes = n_strat ** n_dim
    min_neval = max(int(neval_eff // n_cubes), 2)
    n_ev = np.full(n_cubes, min_neval, dtype=np.int32)
    return {
        "n_strat": n_strat,
        "n_cubes": n_cubes,
        "min_neval": min_neval,
        "n_ev": n_ev,
        "adaptive": adaptive,
        "init_calls": n_events,
    }


def hypercubes(n_strat, n_dim):
    """vflowplus.py:126-128 (itertools.product order: last dimension fastest)."""
    return np.array(list(itertools.product(range(n_strat), repeat=n_dim)), dtype=np.int32).reshape(-1, n_dim)


def generate_samples_in_hypercubes(u, n_strat, n_ev, cubes, divisions):
    """vflowplus.py:37-80: returns x, w (before xjac), ind, segm."""
    indices = np.repeat(np.arange(cubes.shape[0]), n_ev)
    points = cubes[indices].astype(np.float64)
    n_evs = n_ev[indices].astype(np.float64)
    xn = (points + u).T * FBINS / float(n_strat)
    ind, x, w = importance_sampling_digest(xn, divisions)
    return x, w / n_evs, ind, indices


def vplus_event(x, w, ind, segm, n_ev, f, train=True):
    """vflowplus.py:187-220: returns ress, arr_var, arr_res2."""
    tmp = w * f
    tmp2 = np.square(tmp)
    n_cubes = n_ev.shape[0]
    ress = np.bincount(segm, weights=tmp, minlength=n_cubes)
    ress2 = np.bincount(segm, weights=tmp2, minlength=n_cubes)
    arr_var = ress2 * n_ev.astype(np.float64) - np.square(ress)
    arr = bin_histogram(tmp2, ind, x.shape[1]) if train else None
    return ress, arr_var, arr


def redistribute_samples(arr_var, min_neval, init_calls):
    """vflowplus.py:153-163. Negative variances (rounding) are clamped to 0
    before the power; the reference would produce NaN there."""
    damped = np.power(np.maximum(arr_var, 0.0), BETA / 2)
    total = np.sum(damped)
    if not total > 0.0:
        return None
    new_n_ev = np.maximum(float(min_neval), damped * init_calls / 2 / total)
    return new_n_ev.astype(np.int32)


def vplus_iteration(divisions, setup, u, integrand, train=True, xmin=None, xmax=None):
    """vflowplus.py:222-242 on injected uniforms ``u`` (sum(n_ev), d).

    Returns (res, sigma, new_divisions, new_n_ev, arr_var, arr_res2).
    """
    n_dim = divisions.shape[0]
    n_ev = setup["n_ev"]
    cubes = hypercubes(setup["n_strat"], n_dim)
    x, w, ind, segm = generate_samples_in_hypercubes(u, setup["n_strat"], n_ev, cubes, divisions)
    x, w = apply_xjac_and_limits(x, w, 1.0 / setup["n_cubes"], xmin, xmax)
    f = integrand(x)
    ress, arr_var, arr = vplus_event(x, w, ind, segm, n_ev, f, train)
    sigmas2 = np.maximum(arr_var, 0.0)
    res = np.sum(ress)
    sigma = np.sqrt(np.sum(sigmas2 / (n_ev.astype(np.float64) - 1.0)))
    new_n_ev = n_ev
    if setup["adaptive"]:
        r = redistribute_samples(arr_var, setup["min_neval"], setup["init_calls"])
        if r is not None:
            new_n_ev = r
    new_div = refine_grid(arr, divisions) if train else divisions.copy()
    return res, sigma, new_div, new_n_ev, arr_var, arr
