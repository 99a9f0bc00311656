"""PlainFlow: plain Monte Carlo (src/vegasflow/plain.py)."""
import torch

from . import kernels
from .monte_carlo import MonteCarloFlow, sampler, wrapper


class PlainFlow(MonteCarloFlow):
    """Uniform sampling of the unit hypercube (or the limits): the integral
    is the mean of f w, with no grid to train (plain.py:11-43)."""

    _CAN_RUN_VECTORIAL = True

    def _can_fuse(self):
        return self._hooks_native(PlainFlow, ("_run_event", "_digest_random_generation", "_generate_random_array"))

    def _replay_ok(self):
        return self._hooks_native(PlainFlow, ("_run_iteration",))

    def _run_event(self, integrand, ncalls=None):
        """plain.py:18-35: res, raw res2 of one step."""
        n_events = self.n_events if ncalls is None else ncalls
        rnds, xjac, *_ = self._generate_random_array(n_events)
        f = integrand(rnds, weight=xjac)
        if f.dtype != self._dtype:
            f = f.to(self._dtype)
        f = f.contiguous()
        xjac = torch.as_tensor(xjac, device=self.device).to(self._dtype).reshape(-1).contiguous()
        n_out = 1 if f.dim() == 1 else f.shape[1]
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, n_out, False))
        step = self._step(n_events, train=False, offset=0)
        ng = kernels.vegas_digest(step, f, xjac, None, ws)
        out = kernels.vegas_finalize(self.n_dim, n_out, False, False, n_events, ws, ng, raw=True,
                                     status=self._status_word())
        if n_out == 1:
            return out[0], out[1]
        return out[:n_out], out[n_out:]

    def _run_iteration(self):
        """plain.py:37-43"""
        self._require_device()
        if self._fused is not None:
            return self._fused_iteration()
        self._next_draw()
        res, raw_res2 = self.run_event()
        n = float(self.n_events)
        res2 = raw_res2 * n
        err_tmp2 = (res2 - torch.square(res)) / (n - 1.0)
        sigma = torch.sqrt(torch.clamp(err_tmp2, min=0.0))
        return res, sigma

    def _graph_key(self):
        xmin, xdelta, _ = self._limits()
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, 1, False))
        return ("plain", self._compile_serial, self.n_events, self.precision, self._seed, ws.rows.data_ptr(),
                None if xmin is None else xmin.data_ptr(), None if xdelta is None else xdelta.data_ptr())

    def _graph_iteration(self, counter):
        n = self.n_events
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, 1, False))
        step = self._step(n, train=False, offset=0, draw_dev=counter.draw)
        ng = kernels.plain_fused(step, self._fused, ws)
        kernels.vegas_finalize(self.n_dim, 1, False, False, n, ws, ng, counter=counter,
                               status=self._status_word())

    def _fused_iteration(self, inject=None):
        self._next_draw()
        n = self.n_events
        ws = self._workspace_for(kernels.partial_stride(self.n_dim, 1, False))
        step = self._step(n, train=False, offset=0, inject=inject)
        ng = kernels.plain_fused(step, self._fused, ws)
        out = kernels.vegas_finalize(self.n_dim, 1, False, False, n, ws, ng, status=self._status_word())
        return out[0], out[1]


def plain_wrapper(*args, **kwargs):
    """``wrapper`` with PlainFlow: (integral, sigma) of n_iter iterations."""
    return wrapper(PlainFlow, *args, **kwargs)


def plain_sampler(*args, **kwargs):
    """``sampler`` with PlainFlow (uniform points, weight 1/n)."""
    return sampler(PlainFlow, *args, **kwargs)
