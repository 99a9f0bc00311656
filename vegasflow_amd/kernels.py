"""Torch-tensor front end of libmivegas.so.

One function per C entry point of include/mivegas.h. Each checks device,
dtype and contiguity of its tensors, fills the ABI structs and enqueues the
kernels on torch's current HIP stream; nothing here synchronises. The
integrator classes are built on these functions, and the parity tests call
them directly (e.g. with injected draws).
"""
import ctypes
import gc

import torch

from . import _abi

BINS = _abi.BINS


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_raw_device = getattr(torch._C, "_cuda_getDevice", None)


def _stream():
    """torch's current HIP stream of the current device, as a raw handle.
    torch.cuda.current_stream() spends ~10 us per call on device-index
    bookkeeping; the two C calls below give the same handle directly."""
    if _raw_stream is not None and _raw_device is not None:
        return ctypes.c_void_p(_raw_stream(_raw_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# A recording (see ``recording``) collects the C calls made while it is
# active, with every tensor and workspace they point into, so that the calls
# can be issued again with the same arguments.
_recording = None


class Recording:
    """C calls of the library made while active: ``calls`` holds (function,
    arguments, name); ``keep`` the tensors and workspaces whose device memory
    the arguments address (held so that the addresses stay valid)."""

    def __init__(self):
        self.calls = []
        self.keep = []


class recording:
    """Context manager: record the library calls made inside it."""

    def __enter__(self):
        global _recording
        if _recording is not None:
            raise RuntimeError("recordings do not nest")
        _recording = Recording()
        return _recording

    def __exit__(self, *exc):
        global _recording
        _recording = None
        return False


def replay_calls(calls):
    """Issue recorded calls again, in order, on the streams they named."""
    for fn, args, name in calls:
        _abi.check(fn(*args), name)


def _call(name, *args):
    rc = _abi.call(name, *args)
    if _recording is not None:
        _recording.calls.append((getattr(_abi.load(), name), args, name))
    return rc


def _ptr(t):
    if t is None:
        return None
    if _recording is not None:
        _recording.keep.append(t)
    return ctypes.c_void_p(t.data_ptr())


def _need_device(t, name, dtype=None):
    if t is None:
        return
    if not t.is_cuda:
        raise ValueError(f"{name} must be a HIP device tensor")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


_gpu_seen = False


def require_gpu():
    """Raise unless a ROCm GPU is available (checked until one is seen; the
    check costs microseconds and runs once per iteration)."""
    global _gpu_seen
    if _gpu_seen:
        return
    if not torch.cuda.is_available():
        raise RuntimeError(
            "vegasflow_amd runs its kernels on a ROCm GPU and none is available "
            "(torch.cuda.is_available() is False)"
        )
    _gpu_seen = True


def precision_code(precision):
    if precision in ("f64", _abi.F64):
        return _abi.F64
    if precision in ("mixed", _abi.MIXED):
        return _abi.MIXED
    raise ValueError(f"unknown precision {precision!r}")


def float_dtype(precision):
    return torch.float64 if precision_code(precision) == _abi.F64 else torch.float32


class Workspace:
    """Group rows of partial sums (caller-owned memory of the ABI): zero
    when allocated; producers add into them and the finalize calls consume
    and re-zero them."""

    def __init__(self, device, stride):
        self.device = device
        self.stride = int(stride)
        self.rows = torch.zeros(_abi.ROWS * self.stride, dtype=torch.float64, device=device)
        self._struct = _abi.Partials(self.rows.data_ptr(), self.stride)

    def partials(self):
        if _recording is not None:
            _recording.keep.append(self)
        return ctypes.byref(self._struct)


_WS = {}


def workspace(device, stride):
    """A cached workspace with at least ``stride`` doubles per row (shared
    by all stream-ordered calls on ``device``)."""
    key = (str(device),)
    ws = _WS.get(key)
    if ws is None or ws.stride < stride:
        ws = Workspace(device, max(stride, 64))
        _WS[key] = ws
    return ws


def partial_stride(n_dim, n_out=1, train=True):
    return 2 * n_out + (n_dim * BINS if train else 0)


def producer_rows(n_events):
    return _abi.load().mv_producer_rows(int(n_events))


def make_step(n_dim, n_events, *, precision="f64", train=True, event_offset=0,
              seed=0, draw_id=0, xjac=1.0, xmin=None, xdelta=None, xdeltajac=1.0, inject=None,
              draw_dev=None):
    _need_device(xmin, "xmin", torch.float64)
    _need_device(draw_dev, "draw_dev", torch.int32)
    _need_device(xdelta, "xdelta", torch.float64)
    if inject is not None:
        want = torch.float64 if precision_code(precision) == _abi.F64 else torch.int32
        _need_device(inject, "inject", want if want == torch.float64 else None)
        if want != torch.float64 and inject.dtype not in (torch.int32, torch.uint32):
            raise ValueError("mixed-precision inject takes the 32-bit words as int32/uint32")
        if tuple(inject.shape) != (int(n_events), int(n_dim)):
            raise ValueError(f"inject must have shape ({n_events}, {n_dim})")
    s = _abi.Step()
    s.n_dim = int(n_dim)
    s.precision = precision_code(precision)
    s.train = int(bool(train))
    s.n_events = int(n_events)
    s.event_offset = int(event_offset)
    s.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    s.draw_id = int(draw_id) & 0xFFFFFFFF
    s.xjac = float(xjac)
    s.xdeltajac = float(xdeltajac)
    s.xmin = None if xmin is None else xmin.data_ptr()
    s.xdelta = None if xdelta is None else xdelta.data_ptr()
    s.inject = None if inject is None else inject.data_ptr()
    s.draw_dev = None if draw_dev is None else draw_dev.data_ptr()
    # the struct holds raw addresses: keep the tensors alive with it, so that
    # temporaries passed here are not freed (and their memory handed to a
    # later allocation) before the launch that uses the struct
    s._tensors = (xmin, xdelta, inject, draw_dev)
    return s


_capture_streams = {}


def capture_graph(body):
    """Capture ``body()`` (launches on the current stream) as a HIP graph on
    a side stream ordered after the current one; returns (graph, what body
    returned). Python's cyclic garbage collector is off while the capture
    runs: a collection there can destroy an unreachable graph or release
    device memory, operations a stream capture does not allow. (Unlike
    torch.cuda.graph, no full collection and cache flush precede the
    capture: they cost milliseconds, more than a short run's replays save.)"""
    dev = torch.cuda.current_device()
    side = _capture_streams.get(dev)
    if side is None:
        side = _capture_streams[dev] = torch.cuda.Stream(dev)
    cur = torch.cuda.current_stream()
    side.wait_stream(cur)
    graph = torch.cuda.CUDAGraph()
    enabled = gc.isenabled()
    gc.disable()
    try:
        with torch.cuda.stream(side):
            graph.capture_begin()
            try:
                out = body()
            finally:
                graph.capture_end()
    finally:
        if enabled:
            gc.enable()
    cur.wait_stream(side)
    return graph, out


class IterationCounter:
    """Device iteration counter for replayed iterations: ``draw`` (int32 x 2)
    holds {draw id of the next iteration, draw id of history slot 0};
    steps read the draw id from it and the finalize kernels store {res,
    sigma, events} of each iteration in ``history`` (float64, 3 * capacity)
    and advance it."""

    def __init__(self, device, capacity):
        self.draw = torch.zeros(2, dtype=torch.int32, device=device)
        self.capacity = int(capacity)
        self.history = torch.zeros(3 * self.capacity, dtype=torch.float64, device=device)

    def start(self, draw_id):
        """Next iteration draws ``draw_id`` and lands in history slot 0 (a
        fill kernel on the current stream: no host-to-device copy)."""
        self.draw.fill_(int(draw_id))

    def results(self, n):
        """(n, 3) device view of the first n iterations' {res, sigma, events}."""
        return self.history[: 3 * n].view(n, 3)


# ---------------------------------------------------------------------------
# VegasFlow / PlainFlow
# ---------------------------------------------------------------------------
def vegas_fused(step, fn_struct, grid, ws):
    """mv_vegas_fused; returns the number of group rows of ``ws`` touched."""
    _need_device(grid, "grid", torch.float64)
    return _call("mv_vegas_fused", ctypes.byref(step), ctypes.byref(fn_struct), _ptr(grid),
                     ws.partials(), _stream())


def vegas_iteration_static(step, fn_struct, grid, refine, out, hist_out=None, status=None):
    """mv_vegas_iteration_static: a whole iteration with the library's static
    partial rows (no workspace); ``out`` (2,) float64 receives {res, sigma}."""
    _need_device(grid, "grid", torch.float64)
    _need_device(out, "out", torch.float64)
    _need_device(hist_out, "hist_out", torch.float64)
    if out.numel() < 2:
        raise ValueError("out needs 2 entries")
    if hist_out is not None and hist_out.numel() < step.n_dim * BINS:
        raise ValueError("hist_out needs n_dim * 50 entries")
    _need_device(status, "status", torch.int32)
    _call("mv_vegas_iteration_static", ctypes.byref(step), ctypes.byref(fn_struct), _ptr(grid),
              int(bool(refine)), _ptr(out), _ptr(hist_out), _ptr(status), _stream())
    return out


def plain_fused(step, fn_struct, ws):
    return _call("mv_plain_fused", ctypes.byref(step), ctypes.byref(fn_struct), ws.partials(),
                     _stream())


def vegas_sample(step, grid, x, w, idx=None):
    _need_device(grid, "grid", torch.float64)
    ft = float_dtype(step.precision)
    _need_device(x, "x", ft)
    _need_device(w, "w", ft)
    _need_device(idx, "idx", torch.uint8)
    n, d = step.n_events, step.n_dim
    if x.numel() < n * d or w.numel() < n or (idx is not None and idx.numel() < n * d):
        raise ValueError("sample buffers are too small")
    _call("mv_vegas_sample", ctypes.byref(step), _ptr(grid), _ptr(x), _ptr(w), _ptr(idx), _stream())


def plain_sample(step, x):
    _need_device(x, "x", float_dtype(step.precision))
    if x.numel() < step.n_events * step.n_dim:
        raise ValueError("sample buffer is too small")
    _call("mv_plain_sample", ctypes.byref(step), _ptr(x), _stream())


def vegas_digest(step, f, w, idx, ws, main_dim=0):
    """mv_vegas_digest for f of shape (n,) or (n, k); w of shape (n,) or (1,)."""
    ft = float_dtype(step.precision)
    _need_device(f, "f", ft)
    _need_device(w, "w", ft)
    _need_device(idx, "idx", torch.uint8)
    n = step.n_events
    if f.dim() == 1:
        n_out, f_stride = 1, 1
    else:
        n_out, f_stride = f.shape[1], f.stride(0)
    if f.shape[0] != n:
        raise ValueError(f"integrand output has {f.shape[0]} rows, expected {n}")
    w_stride = 0 if w.numel() == 1 else 1
    if w_stride and w.numel() != n:
        raise ValueError("w must have one entry per event or a single entry")
    return _call("mv_vegas_digest", ctypes.byref(step), _ptr(f), f_stride, n_out, int(main_dim),
                     _ptr(w), w_stride, _ptr(idx), ws.partials(), _stream())


def vegas_finalize(n_dim, n_out, train, refine, n_events, ws, n_rows, grid=None, hist_out=None,
                   raw=False, counter=None, status=None):
    """Reduce group rows; returns a (2 * n_out,) float64 tensor [res, sigma]
    (or [res, sum tmp^2] with raw). ``counter``: an IterationCounter the
    kernel records the result in and advances. ``status``: an int32 device
    word the kernel ORs MV_STATUS_NONFINITE into for non-finite sums."""
    out = torch.empty(2 * n_out, dtype=torch.float64, device=ws.device)
    _need_device(grid, "grid", torch.float64)
    _need_device(hist_out, "hist_out", torch.float64)
    fz = _abi.Finalize()
    fz.n_dim = int(n_dim)
    fz.n_out = int(n_out)
    fz.train = int(bool(train))
    fz.refine = int(bool(refine))
    fz.raw = int(bool(raw))
    fz.n_events = float(n_events)
    if counter is not None:
        fz.draw_dev = counter.draw.data_ptr()
        fz.history = counter.history.data_ptr()
        fz.history_len = counter.capacity
    if status is not None:
        _need_device(status, "status", torch.int32)
        fz.status = status.data_ptr()
    if _recording is not None:  # raw addresses in the struct
        _recording.keep.extend((counter, status))
    _call("mv_vegas_finalize", ctypes.byref(fz), _ptr(ws.rows), ws.stride, int(n_rows),
              _ptr(grid), _ptr(out), _ptr(hist_out), _stream())
    return out


def refine_grid(hist, grid):
    """mv_refine_grid: hist (n_dim, 50) float64, grid (n_dim, 51) in place."""
    _need_device(grid, "grid", torch.float64)
    if hist.dtype != torch.float64:
        hist = hist.to(torch.float64)
    hist = hist.contiguous()
    _need_device(hist, "hist", torch.float64)
    if grid.dim() != 2 or grid.shape[1] != BINS + 1 or hist.shape != (grid.shape[0], BINS):
        raise ValueError("refine_grid expects hist (n_dim, 50) and grid (n_dim, 51)")
    _call("mv_refine_grid", grid.shape[0], _ptr(hist), BINS, _ptr(grid), _stream())


def consume_array_into_indices(input_arr, indices, result_size):
    """mv_consume_array_into_indices (utils.py:10-44) on device tensors."""
    inp = input_arr.to(torch.float64).contiguous().reshape(-1)
    ind = indices.to(torch.int32).contiguous().reshape(-1)
    _need_device(inp, "input_arr", torch.float64)
    _need_device(ind, "indices", torch.int32)
    if inp.numel() != ind.numel():
        raise ValueError("input_arr and indices must have the same number of entries")
    out = torch.empty(int(result_size), dtype=torch.float64, device=inp.device)
    _call("mv_consume_array_into_indices", _ptr(inp), _ptr(ind), inp.numel(), int(result_size),
              _ptr(out), _stream())
    return out


# ---------------------------------------------------------------------------
# VegasFlowPlus
# ---------------------------------------------------------------------------
def make_strata(n_strat, n_cubes, n_ev, offsets):
    _need_device(n_ev, "n_ev", torch.int32)
    _need_device(offsets, "offsets", torch.int32)
    st = _abi.Strata()
    st.n_strat = int(n_strat)
    st.n_cubes = int(n_cubes)
    st.n_ev = n_ev.data_ptr()
    st.offsets = offsets.data_ptr()
    st._tensors = (n_ev, offsets)  # see make_step
    return st


def vplus_fused(step, strata, fn_struct, grid, ress, ress2, ws):
    _need_device(grid, "grid", torch.float64)
    _need_device(ress, "ress", torch.float64)
    _need_device(ress2, "ress2", torch.float64)
    return _call("mv_vplus_fused", ctypes.byref(step), ctypes.byref(strata), ctypes.byref(fn_struct),
                     _ptr(grid), _ptr(ress), _ptr(ress2), ws.partials(), _stream())


def vplus_sample(step, strata, grid, x, w, idx, seg):
    ft = float_dtype(step.precision)
    _need_device(grid, "grid", torch.float64)
    _need_device(x, "x", ft)
    _need_device(w, "w", ft)
    _need_device(idx, "idx", torch.uint8)
    _need_device(seg, "seg", torch.int32)
    _call("mv_vplus_sample", ctypes.byref(step), ctypes.byref(strata), _ptr(grid), _ptr(x), _ptr(w),
              _ptr(idx), _ptr(seg), _stream())


def vplus_digest(step, strata, f, w, idx, seg, ress, ress2, ws):
    ft = float_dtype(step.precision)
    _need_device(f, "f", ft)
    _need_device(w, "w", ft)
    _need_device(idx, "idx", torch.uint8)
    _need_device(seg, "seg", torch.int32)
    if f.dim() != 1:
        raise NotImplementedError("VegasFlowPlus does not support vectorial integrands")
    return _call("mv_vplus_digest", ctypes.byref(step), ctypes.byref(strata), _ptr(f), _ptr(w),
                     _ptr(idx), _ptr(seg), _ptr(ress), _ptr(ress2), ws.partials(), _stream())


def vplus_finalize(n_dim, strata, n_ev, offsets, ress, ress2, adaptive, min_neval, init_calls, beta,
                   ws, n_rows, train, refine, grid, arr_var_out=None, hist_out=None, counter=None,
                   status=None):
    """Returns a (3,) float64 tensor: res, sigma, events of the next iteration.
    ``counter``: an IterationCounter, as in vegas_finalize."""
    _need_device(status, "status", torch.int32)
    if _recording is not None:  # raw addresses below
        _recording.keep.append(counter)
    out = torch.empty(int(_abi.load().mv_vplus_out_len(int(strata.n_cubes))), dtype=torch.float64,
                      device=ws.device)
    _call("mv_vplus_finalize", int(n_dim), ctypes.byref(strata), _ptr(n_ev), _ptr(offsets), _ptr(ress),
              _ptr(ress2), int(bool(adaptive)), int(min_neval), float(init_calls), float(beta),
              _ptr(ws.rows), ws.stride, int(n_rows), int(bool(train)), int(bool(refine)),
              _ptr(grid), _ptr(out), _ptr(arr_var_out), _ptr(hist_out),
              None if counter is None else counter.draw.data_ptr(),
              None if counter is None else counter.history.data_ptr(),
              0 if counter is None else counter.capacity, _ptr(status), _stream())
    return out[:3]


# ---------------------------------------------------------------------------
# batched integrals
# ---------------------------------------------------------------------------
def batch_fused(items_dev, n_items, tiles_dev, n_tiles, precision, family, max_dim, draws, grids, rows,
                blocks):
    """mv_batch_fused; every argument is already on the device (no
    allocation here, so the call can be captured in a HIP graph)."""
    _call("mv_batch_fused", _ptr(items_dev), int(n_items), _ptr(tiles_dev), int(n_tiles),
              precision_code(precision), int(family), int(max_dim), _ptr(draws), _ptr(grids), _ptr(rows),
              int(blocks), _stream())


def batch_fused_pairs(items_dev, n_items, tiles_dev, n_tiles, family, max_dim, draws, grids, rows, blocks,
                      precision="mixed"):
    """mv_batch_fused_pairs (mixed) or mv_batch_fused_pairs_f64: pair
    histograms, 1024-thread workgroups; arguments as batch_fused."""
    name = "mv_batch_fused_pairs" if precision_code(precision) == _abi.MIXED else "mv_batch_fused_pairs_f64"
    _call(name, _ptr(items_dev), int(n_items), _ptr(tiles_dev), int(n_tiles),
              int(family), int(max_dim), _ptr(draws), _ptr(grids), _ptr(rows), int(blocks), _stream())


def batch_pairs_blocks(family, max_dim, precision="mixed"):
    """Resident 1024-thread workgroups of the pair batch kernel of this
    precision on the current device for this family and max_dim (0: the pair
    kernel does not apply)."""
    lib = _abi.load()
    fn = lib.mv_batch_pairs_blocks if precision_code(precision) == _abi.MIXED else lib.mv_batch_pairs_blocks_f64
    return int(fn(int(family), int(max_dim)))


def batch_finalize(items_dev, jobs, n_jobs, grids, rows, draws, stats, history=None, max_iter=0):
    _call("mv_batch_finalize", _ptr(items_dev), _ptr(jobs), int(n_jobs), _ptr(grids), _ptr(rows),
              _ptr(draws), _ptr(stats), _ptr(history), int(max_iter), _stream())


def batch_vplus_fused(items_dev, tiles_dev, n_tiles, precision, family, min_dim, max_dim, draws, grids, n_ev,
                      offsets, ress, ress2, rows, blocks):
    """mv_batch_vplus_fused; every argument is already on the device."""
    _call("mv_batch_vplus_fused", _ptr(items_dev), _ptr(tiles_dev), int(n_tiles), precision_code(precision),
          int(family), int(min_dim), int(max_dim), _ptr(draws), _ptr(grids), _ptr(n_ev), _ptr(offsets), _ptr(ress),
          _ptr(ress2), _ptr(rows), int(blocks), _stream())


def batch_vplus_finalize(items_dev, chunks, n_chunks, dims, n_dims, beta, grids, n_ev, offsets, ress, ress2,
                         rows, outs, draws, stats):
    """mv_batch_vplus_finalize (three launches); device arguments."""
    _call("mv_batch_vplus_finalize", _ptr(items_dev), _ptr(chunks), int(n_chunks), _ptr(dims), int(n_dims),
          float(beta), _ptr(grids), _ptr(n_ev), _ptr(offsets), _ptr(ress), _ptr(ress2), _ptr(rows),
          _ptr(outs), _ptr(draws), _ptr(stats), _stream())
