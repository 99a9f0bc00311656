"""The reference's TensorFlow graph for one VegasFlow iteration, restated op
for op with multi-threaded torch CPU tensors: test infrastructure only (the
timed CPU baseline of the benchmark; never part of the product path).

Followed ops (src/vegasflow/):
* monte_carlo.py:264-270   uniforms in [TECH_CUT, 1 - TECH_CUT], w *= xjac
* vflow.py:113-126         xn = BINS * (1 - rnds^T) (a transposed copy)
* vflow.py:33-83           int cast, two batched gathers, x and the Jacobian
                           product, transposes back
* vflow.py:389-430         tmp = w f, tmp2 = tmp^2, the two reductions
* vflow.py:370-387,
  utils.py:40-43           the dense one-hot histogram: equal -> transpose ->
                           where -> reduce_sum, per dimension
* monte_carlo.py:438-459   events in steps of at most MAX_EVENTS_LIMIT = 1e6,
                           the step tuples summed on the host
* vflow.py:432-442         sigma, then the per-dimension refinement
                           (vflow.py:129-211, the serial 49-cut loop)
The uniforms come from torch's CPU generator (the reference's come from TF's
Philox): draws are not compared here, only time is measured.
"""
import math

import numpy as np
import torch

from .vegas import refine_grid

BINS = 50
TECH_CUT = 1e-8
MAX_EVENTS_LIMIT = 1_000_000


def gaussian(x, a=0.1, centre=0.5):
    """The benchmark integrand (simgauss-style, without the no-op term)."""
    d = x.shape[1]
    pref = (1.0 / (a * math.sqrt(math.pi))) ** d
    return pref * torch.exp(-torch.sum(torch.square(x - centre), dim=1) / a ** 2)


def _digest(xn, divisions):
    """vflow.py:33-83 on (d, n) xn."""
    ind = xn.to(torch.int32)
    ind64 = ind.to(torch.int64)
    x_ini = torch.gather(divisions, 1, ind64)
    x_fin = torch.gather(divisions, 1, ind64 + 1)
    xdelta = x_fin - x_ini
    rem = xn - torch.floor(xn)
    x = x_ini + xdelta * rem
    w = torch.prod(xdelta * BINS, dim=0)
    return x.t().contiguous(), w, ind.t().contiguous()


def _consume(tmp2, ind_col):
    """utils.py:40-43: dense one-hot accumulation into BINS bins."""
    all_bins = torch.arange(BINS, dtype=torch.int32)
    eq = torch.eq(ind_col, all_bins).t()
    res_tmp = torch.where(eq, tmp2, torch.zeros((), dtype=tmp2.dtype))
    return torch.sum(res_tmp, dim=1)


def run_step(divisions, n, n_total, integrand, gen, rnds=None):
    """One step of ``n`` events: returns (res, res2, arr_res2 (d, BINS)).
    ``rnds`` (n, d) replaces the generator's uniforms when given."""
    d = divisions.shape[0]
    if rnds is None:
        rnds = torch.rand((n, d), dtype=torch.float64, generator=gen) * (1 - 2 * TECH_CUT) + TECH_CUT
    xn = BINS * (1.0 - rnds.t())
    x, w, ind = _digest(xn, divisions)
    w = w * (1.0 / n_total)
    f = integrand(x)
    tmp = w * f
    tmp2 = torch.square(tmp)
    res = torch.sum(tmp)
    res2 = torch.sum(tmp2)
    arr = torch.stack([_consume(tmp2, ind[:, j:j + 1]) for j in range(d)])
    return res, res2, arr


def iteration(divisions, n_events, integrand, gen, train=True, max_steps=None):
    """One training iteration in steps of at most 1e6 events. Returns
    (res, sigma, events run); with ``max_steps`` only the first steps run
    (a bounded sample of the iteration's work, for timing)."""
    steps = []
    done = 0
    while done < n_events and (max_steps is None or len(steps) < max_steps):
        n = min(MAX_EVENTS_LIMIT, n_events - done)
        steps.append(run_step(divisions, n, n_events, integrand, gen))
        done += n
    res = sum(float(s[0]) for s in steps)
    res2 = sum(float(s[1]) for s in steps)
    arr = sum(s[2] for s in steps)
    sigma = math.sqrt(max((n_events * res2 - res ** 2) / (n_events - 1), 0.0))
    if train:
        new = refine_grid(arr.numpy(), divisions.numpy())
        divisions.copy_(torch.from_numpy(np.ascontiguousarray(new)))
    return res, sigma, done
