"""Shared parity bars of the GPU tests.

sigma subtracts two sums: sigma^2 = (N S2 - S1^2) / (N - 1) with S1 = sum tmp
and S2 = sum tmp^2 (vflow.py:437-438, plain.py:37-43). Rounding of S1 and S2
at relative eps (summation order: the kernels add with atomics, the oracle
left to right) moves sigma^2 by up to eps (N S2 + S1^2), so sigma's relative
error is bounded by eps times the cancellation factor

    kappa = (N S2 + S1^2) / (N S2 - S1^2) = 1 + 2 S1^2 / ((N - 1) sigma^2).

The bar for sigma is therefore RTOL * kappa with the same RTOL = 1e-12 as
the integral; kappa is computed from the oracle's own sums. For VEGAS+
(vflowplus.py:216-233) every cube c cancels on its own:
kappa = sum_c (|arr_var_c| + 2 ress_c^2) / (n_c - 1) / sigma^2.
"""
import math

import numpy as np

RTOL = 1e-12


def kappa_vegas(res, sigma, n_events):
    """Cancellation factor of the VegasFlow / PlainFlow sigma."""
    if sigma <= 0.0:
        return math.inf
    return 1.0 + 2.0 * float(res) ** 2 / ((float(n_events) - 1.0) * float(sigma) ** 2)


def kappa_vplus(ress, arr_var, n_ev, sigma):
    """Cancellation factor of the VEGAS+ sigma from per-cube sums."""
    if sigma <= 0.0:
        return math.inf
    n = np.asarray(n_ev, dtype=np.float64)
    num = np.sum((np.abs(arr_var) + 2.0 * np.square(ress)) / (n - 1.0))
    return float(num) / float(sigma) ** 2


def assert_sigma(got, want, kappa, rtol=RTOL):
    """|got - want| <= rtol * kappa * |want|."""
    got, want = float(got), float(want)
    assert abs(got - want) <= rtol * kappa * abs(want), (got, want, kappa, abs(got - want) / abs(want))


def sigma_noise_floor(res, n_events):
    """The sigma that rounding alone leaves when every tmp is equal (true
    sigma 0, e.g. a constant integrand): N S2 + S1^2 = 2 S1^2 rounded at
    2^-52 twice gives sigma^2 <= 4 eps S1^2 / (N - 1); twice that bound."""
    return 2.0 * math.sqrt(4.0 * 2.0 ** -52 / (float(n_events) - 1.0)) * abs(float(res))


def assert_history(got, want, events, rtol):
    """Per-iteration (integral, sigma) histories of integrals with the given
    event counts: integrals to rtol, sigma to rtol * kappa (or, where one of
    them is zero, within the rounding floor of a zero sigma)."""
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape
    for item in range(want.shape[0]):
        for k in range(want.shape[1]):
            (r, s), (wr, ws) = got[item, k], want[item, k]
            assert abs(r - wr) <= rtol * abs(wr), (item, k, r, wr)
            if ws == 0.0 or s == 0.0:
                assert max(s, ws) <= sigma_noise_floor(wr, events[item]), (item, k, s, ws)
            else:
                assert_sigma(s, ws, kappa_vegas(wr, ws, events[item]), rtol)
