"""VegasFlowPlus: VEGAS+ adaptive importance + stratified sampling
(src/vegasflow/vflowplus.py, Lepage arXiv:2009.05112).

The stratification (n_strat, hypercube order, events per cube) follows the
reference exactly, including its float32 arithmetic for n_strat
(vflowplus.py:113-123). Per iteration with a built-in integrand:
``mv_vplus_fused`` (draws in the events' hypercubes, grid map, integrand,
per-cube segmented sums, bin histograms) then ``mv_vplus_finalize`` (result
and sigma over the cubes, the adaptive reallocation of events with
beta = 0.75, the new cube offsets, grid refinement), with no host round trip.
"""
import itertools
import logging

import numpy as np
import torch

from . import kernels
from .configflow import BETA, BINS_MAX, MAX_NEVAL_HCUBE
from .monte_carlo import MonteCarloFlow, native_hook, sampler, wrapper
from .vflow import VegasFlow

logger = logging.getLogger(__name__)


def stratification(n_dim, n_events, adaptive):
    """n_strat and the events per hypercube (vflowplus.py:105-137).

    ``tf.math.pow`` on Python floats computes in float32; so does this.
    Returns (adaptive, neval_eff, n_strat, n_cubes, min_neval).
    """
    if n_dim > 13 and adaptive:
        adaptive = False
        logger.warning("Disabling adaptive mode from VegasFlowPlus, too many dimensions!")
    neval_eff = int(n_events / 2) if adaptive else n_events
    f32 = np.float32
    n_strat = np.floor(np.power(f32(neval_eff / 2), f32(1 / n_dim)))
    if np.power(f32(n_strat), f32(n_dim)) > f32(MAX_NEVAL_HCUBE):
        n_strat = np.floor(np.power(f32(1e4), f32(1 / n_dim)))
    n_strat = int(np.int32(n_strat))
    n_cubes = n_strat ** n_dim
    min_neval = max(int(neval_eff // n_cubes), 2)
    return adaptive, neval_eff, n_strat, n_cubes, min_neval


class VegasFlowPlus(VegasFlow):
    """VEGAS+ (vflowplus.py:83-242): the VEGAS grid plus a stratification
    of the unit hypercube into n_strat^n_dim cubes, each sampled with its own
    number of events; with ``adaptive`` those numbers follow each cube's
    variance after every iteration."""

    def __init__(self, n_dim, n_events, train=True, adaptive=False, events_limit=None, **kwargs):
        # vflowplus.py:89-99: an iteration is held in one step
        if events_limit is None:
            logger.info("Events per device limit set to %d", n_events)
            events_limit = n_events
        elif events_limit < n_events:
            logger.warning(
                "VegasFlowPlus needs to hold all events in memory at once, "
                "setting the `events_limit` to be equal to `n_events=%d`",
                n_events,
            )
            events_limit = n_events
        super().__init__(n_dim, n_events, train, events_limit=events_limit, **kwargs)
        self._init_calls = n_events
        self._adaptive, _, self._n_strat, n_cubes, self.min_neval_hcube = stratification(
            n_dim, self.n_events, adaptive
        )
        cubes = list(itertools.product(range(self._n_strat), repeat=int(n_dim)))
        if len(cubes) != self._n_strat ** n_dim:
            raise ValueError("Hypercubes are not equal to n_strat^n_dim")
        self._hypercubes = torch.tensor(cubes, dtype=torch.int32, device=self.device).reshape(-1, n_dim)
        self._n_cubes = n_cubes
        self._n_ev = torch.full((n_cubes,), self.min_neval_hcube, dtype=torch.int32, device=self.device)
        self._offsets = torch.arange(n_cubes + 1, dtype=torch.int32, device=self.device) * self.min_neval_hcube
        self._n_events = int(n_cubes * self.min_neval_hcube)
        self._modified_jac = 1.0 / n_cubes
        self._ress = torch.zeros(n_cubes, dtype=torch.float64, device=self.device)
        self._ress2 = torch.zeros(n_cubes, dtype=torch.float64, device=self.device)
        self._last_out = None
        self._pending_events = []
        self.events_history = []  # events of each iteration run so far
        if self._adaptive:
            logger.warning("Variable number of events requires function signatures all across")

    # ----------------------------------------------------------- properties
    @property
    def xjac(self):
        return torch.tensor([self._modified_jac], dtype=self._dtype, device=self.device)

    def _xjac_value(self):
        return self._modified_jac

    @property
    def n_ev(self):
        """Events per hypercube (device int32 tensor)."""
        return self._n_ev

    @n_ev.setter
    def n_ev(self, val):
        v = torch.as_tensor(val, dtype=torch.int32, device=self.device).reshape(-1)
        if v.numel() != self._n_cubes:
            raise ValueError(f"n_ev needs {self._n_cubes} entries")
        self._n_ev = v.clone()
        self._offsets = torch.zeros(self._n_cubes + 1, dtype=torch.int32, device=self.device)
        self._offsets[1:] = torch.cumsum(self._n_ev, 0, dtype=torch.int32)

    def _events_bound(self):
        """Upper bound of the events of the next iteration: after a
        reallocation sum(max(min, x_c)) <= n_cubes * min + init_calls / 2."""
        if not self._adaptive:
            return self._n_events
        return max(self._n_events, self._n_cubes * self.min_neval_hcube + self._init_calls // 2 + 1)

    def _strata(self):
        return kernels.make_strata(self._n_strat, self._n_cubes, self._n_ev, self._offsets)

    _UNFUSED_HOOKS = VegasFlow._UNFUSED_HOOKS + ("redistribute_samples",)

    def _can_fuse(self):
        return self._hooks_native(VegasFlowPlus, self._UNFUSED_HOOKS)

    def _replay_ok(self):
        return self._hooks_native(VegasFlowPlus, ("_run_iteration", "refine_grid"))

    def redistribute_samples(self, arr_var):
        """vflowplus.py:153-163 on the device. Negative variances (rounding)
        are clamped to zero before the power; the reference yields NaN."""
        arr_var = torch.as_tensor(arr_var, dtype=torch.float64, device=self.device)
        damped = torch.pow(torch.clamp(arr_var, min=0.0), BETA / 2)
        total = damped.sum()
        if not float(total) > 0.0:
            return
        new_n_ev = torch.clamp(damped * self._init_calls / 2 / total, min=float(self.min_neval_hcube))
        self.n_ev = new_n_ev.to(torch.int32)
        self._n_events = int(self._n_ev.sum().item())

    def _current_events(self):
        return int(self._offsets[self._n_cubes].item())

    # ------------------------------------------------------------ sampling
    @native_hook
    def _digest_random_generation(self, rnds, n_ev=None):
        """Points in the hypercubes (vflowplus.py:37-80, 165-171): x, w,
        ind (uint8), segm (cube of each event).

        ``rnds`` an event count (a whole iteration): drawn and mapped in one
        kernel, w including xjac = 1/n_cubes and the limits. ``rnds`` a tensor
        of uniforms (the reference's contract): those uniforms placed in the
        cubes and mapped, w = prod(dx * 50) / n_ev without xjac or limits."""
        self._require_device()
        if n_ev is not None and n_ev is not self._n_ev:
            self.n_ev = n_ev
        total = self._current_events()
        if torch.is_tensor(rnds):
            u = rnds.to(self.device).to(torch.float64).contiguous()
            if tuple(u.shape) != (total, self.n_dim):
                raise ValueError(f"uniforms must have shape ({total}, {self.n_dim}), got {tuple(u.shape)}")
            step = kernels.make_step(self.n_dim, total, precision="f64", train=True, xjac=1.0, inject=u)
            x = torch.empty((total, self.n_dim), dtype=torch.float64, device=self.device)
            w = torch.empty((total,), dtype=torch.float64, device=self.device)
            ind = torch.empty((total, self.n_dim), dtype=torch.uint8, device=self.device)
            seg = torch.empty((total,), dtype=torch.int32, device=self.device)
            kernels.vplus_sample(step, self._strata(), self.divisions, x, w, ind, seg)
            return x.to(self._dtype), w.to(self._dtype), ind, seg
        n_events = int(rnds)
        if n_events != total:
            raise ValueError(f"VegasFlowPlus draws whole iterations ({total} events), got {n_events}")
        step = self._step(total, xjac=self._modified_jac)
        x = torch.empty((total, self.n_dim), dtype=self._dtype, device=self.device)
        w = torch.empty((total,), dtype=self._dtype, device=self.device)
        ind = torch.empty((total, self.n_dim), dtype=torch.uint8, device=self.device)
        seg = torch.empty((total,), dtype=torch.int32, device=self.device)
        kernels.vplus_sample(step, self._strata(), self.divisions, x, w, ind, seg)
        return x, w, ind, seg

    def generate_random_array(self, n_events, *args):
        """vflowplus.py:173-185: whole iterations of draws, cut to n_events."""
        rnds = []
        wgts = []
        for _ in range(n_events // self.n_events + 1):
            r, w = MonteCarloFlow.generate_random_array(self, self.n_events, self._n_ev)
            rnds.append(r)
            wgts.append(w)
        final_r = torch.cat(rnds, dim=0)[:n_events]
        final_w = torch.cat(wgts, dim=0)[:n_events] * self.n_events / n_events
        return final_r, final_w

    # ---------------------------------------------------------- iteration
    def _run_event(self, integrand, ncalls=None, n_ev=None):
        """vflowplus.py:187-220: ress, arr_var, arr_res2."""
        n = self._current_events()
        x, xjac, ind, segm = self._generate_random_array(n, n_ev)
        f = integrand(x, weight=xjac)
        if f.dtype != self._dtype:
            f = f.to(self._dtype)
        f = f.contiguous()
        xjac = torch.as_tensor(xjac, device=self.device).to(self._dtype).reshape(-1).contiguous()
        ind = ind.to(torch.uint8).contiguous()
        segm = segm.to(torch.int32).contiguous()
        custom_hist = self.train and not self._hooks_native(VegasFlowPlus, ("_importance_sampling_array_filling",))
        ress = torch.zeros(self._n_cubes, dtype=torch.float64, device=self.device)
        ress2 = torch.zeros(self._n_cubes, dtype=torch.float64, device=self.device)
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, 1, self.train))
        hist_in_kernel = self.train and not custom_hist
        step = self._step(n, train=hist_in_kernel, offset=0, xjac=self._modified_jac)
        ng = kernels.vplus_digest(step, self._strata(), f, xjac, ind if hist_in_kernel else None,
                                  segm, ress, ress2, ws)
        arr_res2 = []
        if custom_hist:
            # a subclass's histogram hook receives (w f)^2 and the bins
            # (vflowplus.py:218); the digest wrote per-cube sums only
            arr_res2 = self._importance_sampling_array_filling(torch.square(xjac * f), ind)
        elif self.train:
            arr_res2 = torch.empty((self.n_dim, BINS_MAX), dtype=torch.float64, device=self.device)
            kernels.vegas_finalize(self.n_dim, 1, True, False, n, ws, ng, hist_out=arr_res2, raw=True,
                                   status=self._status_word())
        fn_ev = self._n_ev.to(torch.float64)
        arr_var = ress2 * fn_ev - torch.square(ress)
        return ress, arr_var, arr_res2

    def run_event(self, tensorize_events=None, **kwargs):
        """One step holds the whole iteration (vflowplus.py:244-247)."""
        if not self.event:
            raise RuntimeError("Compile must be ran before running any iterations")
        return self.device_run(self.n_events, **kwargs)

    def _iteration_content(self):
        """vflowplus.py:222-242"""
        self._pending_events.append(self._current_events())
        self._next_draw()
        ress, arr_var, arr_res2 = self.run_event(n_ev=self._n_ev)
        sigmas2 = torch.clamp(arr_var, min=0.0)
        res = torch.sum(ress)
        sigma2 = torch.sum(sigmas2 / (self._n_ev.to(torch.float64) - 1.0))
        sigma = torch.sqrt(sigma2)
        if self._adaptive:
            self.redistribute_samples(arr_var)
        if self.train:
            self.refine_grid(arr_res2)
        return res, sigma

    def _graph_key(self):
        xmin, xdelta, _ = self._limits()
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, 1, self.train))
        return ("vplus", self._compile_serial, self.train, self._adaptive, self._events_bound(), self.precision,
                self._seed, self._n_strat, self._n_cubes, self.divisions.data_ptr(), self._n_ev.data_ptr(),
                self._offsets.data_ptr(), self._ress.data_ptr(), self._ress2.data_ptr(), ws.rows.data_ptr(),
                None if xmin is None else xmin.data_ptr(), None if xdelta is None else xdelta.data_ptr())

    def _graph_iteration(self, counter):
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, 1, self.train))
        step = self._step(self._events_bound(), train=self.train, offset=0, xjac=self._modified_jac,
                          draw_dev=counter.draw)
        strata = self._strata()
        ng = kernels.vplus_fused(step, strata, self._fused, self.divisions, self._ress, self._ress2, ws)
        return kernels.vplus_finalize(self.n_dim, strata, self._n_ev, self._offsets, self._ress, self._ress2,
                                      self._adaptive, self.min_neval_hcube, self._init_calls, BETA, ws, ng,
                                      self.train, self.train, self.divisions, counter=counter,
                                      status=self._status_word())

    def _replayed_events(self, events):
        # events of each replayed iteration, and the next iteration's count
        # from the last finalize
        self._pending_events.extend(int(e) for e in events)
        self._last_out = self._replay_last

    def _fused_iteration(self, inject=None):
        self._next_draw()
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, 1, self.train))
        step = self._step(self._events_bound(), train=self.train, offset=0, xjac=self._modified_jac,
                          inject=inject)
        strata = self._strata()
        ng = kernels.vplus_fused(step, strata, self._fused, self.divisions, self._ress, self._ress2, ws)
        own_refine = self._hooks_native(VegasFlowPlus, ("refine_grid",))
        hist = None
        if self.train and not own_refine:
            hist = torch.empty((self.n_dim, BINS_MAX), dtype=torch.float64, device=self.device)
        out = kernels.vplus_finalize(self.n_dim, strata, self._n_ev, self._offsets, self._ress, self._ress2,
                                     self._adaptive, self.min_neval_hcube, self._init_calls, BETA, ws, ng,
                                     self.train, self.train and own_refine, self.divisions, hist_out=hist,
                                     status=self._status_word())
        if hist is not None:
            # a subclass's refine_grid receives the histogram (vflowplus.py:239-240)
            self.refine_grid(hist)
        self._pending_events.append(self._last_out[2] if self._last_out is not None else self._n_events)
        self._last_out = out
        return out[0], out[1]

    def _device_extras(self, replayed):
        # the event counts still on the device, then the next iteration's
        last = self._replay_last if replayed else self._last_out
        pend = [e.reshape(()) for e in self._pending_events if torch.is_tensor(e)]
        self._extras_layout = (len(pend), last is not None)
        return pend + ([last[2].reshape(())] if last is not None else [])

    def _take_extras(self, values):
        n_pend, has_last = self._extras_layout
        vals = iter(values[:n_pend].tolist())
        for e in self._pending_events:
            self.events_history.append(int(next(vals)) if torch.is_tensor(e) else int(e))
        self._pending_events = []
        if has_last:
            self._n_events = int(values[n_pend])
        self._last_out = None


def vegasflowplus_wrapper(integrand, n_dim, n_iter, total_n_events, **kwargs):
    """Convenience wrapper (vflowplus.py:250-265)."""
    return wrapper(VegasFlowPlus, integrand, n_dim, n_iter, total_n_events, **kwargs)


def vegasflowplus_sampler(*args, **kwargs):
    """Trained grid as a sampler (vflowplus.py:268-282)."""
    return sampler(VegasFlowPlus, *args, **kwargs)
