"""Continuous-batching integration server on one GPU.

Requests (a built-in integrand, n_dim, events per iteration, relative-error
target, iteration cap, seed) are admitted into slots of a device slab whose
size follows from a memory budget (by default from the device memory free
at start), optionally also within a cap on the events of one tick. Every
tick advances all live slots by one iteration with the batched kernels
(``mv_batch_fused`` per family present, then ``mv_batch_finalize``):
requests that arrived while others run join at the next tick, and their
work shares the same launches. After each tick the
host reads the per-slot iteration counters and running weighted averages,
retires the requests that reached their target (after ``min_iter``
iterations) or their cap, and admits waiting requests into the freed slots.

Failure handling: ``cancel`` withdraws a waiting or running request; a
request with a ``timeout`` (seconds from its arrival) is retired at the
first tick past it, or dropped unadmitted; a request whose integral is not
finite (its integrand gave NaN or Inf) is retired at once. Each ends with
``status`` "cancelled", "timeout" or "nonfinite" and its last result;
converged requests end "converged", those at their iteration cap "max_iter".

Two algorithms share the slots: VegasFlow requests (``algorithm="vegas"``)
run through ``mv_batch_fused`` / ``mv_batch_fused_pairs`` and
``mv_batch_finalize``; VegasFlowPlus requests (``algorithm="vegas+"``, the
stratified sampling with adaptive event allocation of vflowplus.py) through
``mv_batch_vplus_fused`` and ``mv_batch_vplus_finalize``, with their
hypercube allocations and per-cube sums resident in the slot.

A request's draws depend only on its seed, its iteration and its event
index, so its result is the same alone and in the mix up to summation order
(``run_alone`` reproduces it with VegasFlow or VegasFlowPlus).
"""
import collections
import ctypes
import dataclasses
import time

import numpy as np
import torch

from . import _abi, kernels
from .configflow import BETA, BINS_MAX, DTYPE, MAX_NEVAL_HCUBE, default_device, precision_of
from .integrands import BuiltinIntegrand

MAX_DIM = _abi.FUSED_MAX_DIM
GRID_STRIDE = MAX_DIM * (BINS_MAX + 1)
ROW_STRIDE = 2 + MAX_DIM * BINS_MAX
ITEM_BYTES = ctypes.sizeof(_abi.BatchItem)
VP_ITEM_BYTES = ctypes.sizeof(_abi.BatchVplusItem)
TILE_BYTES = ctypes.sizeof(_abi.BatchTile)
# VEGAS+ slot arrays: the stratification never makes more than
# MAX_NEVAL_HCUBE hypercubes (vflowplus.py:115-123)
MAX_CUBES = MAX_NEVAL_HCUBE
CUBE_CHUNK = 256  # cubes per workgroup of the VEGAS+ iteration end
OUT_LEN = 3 + 4 * ((MAX_CUBES + CUBE_CHUNK - 1) // CUBE_CHUNK)  # mv_vplus_out_len(MAX_CUBES)
ALGORITHMS = ("vegas", "vegas+")


@dataclasses.dataclass
class Request:
    integrand: BuiltinIntegrand
    n_dim: int
    n_events: int
    target: float = 1e-3
    max_iter: int = 20
    seed: int = 0
    arrival: float = 0.0
    rid: int = -1
    timeout: float = None  # seconds after arrival (None: no limit)
    algorithm: str = "vegas"  # "vegas" (VegasFlow) or "vegas+" (VegasFlowPlus)
    adaptive: bool = True  # VegasFlowPlus: reallocate events across hypercubes
    # filled by the server
    admitted: float = None
    done: float = None
    result: float = None
    error: float = None
    iterations: int = 0
    slot: int = -1
    status: str = "waiting"  # running, converged, max_iter, cancelled, timeout, nonfinite

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def slot_bytes():
    """Device bytes one slot holds: both item kinds, grid, row, counters and
    the VEGAS+ arrays (events and offsets per hypercube, per-cube sums, the
    iteration end's results and scratch), so that either algorithm fits."""
    vplus = 4 * MAX_CUBES + 4 * (MAX_CUBES + 1) + 16 * MAX_CUBES + 8 * OUT_LEN
    return ITEM_BYTES + VP_ITEM_BYTES + 8 * (GRID_STRIDE + ROW_STRIDE) + 4 + 8 * 4 + vplus


def free_memory_budget(device, share=16, cap=1 << 30):
    """A slab budget from the device memory free now: 1/share of it, at most
    ``cap`` bytes (1 GiB holds 4198 slots)."""
    free, _ = torch.cuda.mem_get_info(device)
    return min(int(free) // int(share), int(cap))


class Server:
    """Request loop over a slab of ``capacity`` slots.

    Admission takes the oldest waiting request while (1) a slot is free: the
    slab holds ``memory_budget`` bytes (None: ``free_memory_budget``, from
    the device memory free when the server starts), ``slot_bytes()`` per
    request whatever its event count, since the fused kernels keep no
    per-event memory; and (2) with ``max_tick_events`` set, the events of
    one tick (one iteration of every live request) stay within it, which
    bounds the time of a tick and so the latency that each admitted request
    adds to the others. A request larger than the cap is admitted alone.
    Admission is first come, first served: a request that does not fit
    holds back the ones behind it."""

    def __init__(self, memory_budget=None, dtype=None, min_iter=3, events_per_tile=16384,
                 blocks_per_cu=4, pair_histograms=None, pair_events_per_tile=65536,
                 max_tick_events=None):
        kernels.require_gpu()
        self.device = default_device()
        if memory_budget is None:
            memory_budget = free_memory_budget(self.device)
        self.memory_budget = int(memory_budget)
        self.max_tick_events = None if max_tick_events is None else int(max_tick_events)
        self.live_events = 0  # events of one tick over the live requests
        self.precision = precision_of(DTYPE if dtype is None else dtype)
        # training goes through the pair-histogram batch kernel
        # (mv_batch_fused_pairs / _f64) wherever it takes the family's max
        # n_dim: by default in mixed precision only (on the request mix of
        # the serving benchmark the FP64 pair kernel raised the median
        # latency, 4.4 -> 5.2 ms, for 3 % off the 99th percentile)
        self.pairs = self.precision == "mixed" if pair_histograms is None else bool(pair_histograms)
        self.pair_events_per_tile = int(pair_events_per_tile)
        self.capacity = int(memory_budget // slot_bytes())
        if self.capacity < 1:
            raise ValueError("memory budget below one slot")
        self.min_iter = int(min_iter)
        self.events_per_tile = int(events_per_tile)
        self.cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.blocks = int(blocks_per_cu) * self.cus
        S = self.capacity
        self.items_host = (_abi.BatchItem * S)()
        self.items = torch.zeros(S * ITEM_BYTES, dtype=torch.uint8, device=self.device)
        self.grids = torch.zeros(S * GRID_STRIDE, dtype=torch.float64, device=self.device)
        self.rows = torch.zeros(S * ROW_STRIDE, dtype=torch.float64, device=self.device)
        self.draws = torch.zeros(S, dtype=torch.int32, device=self.device)
        self.stats = torch.zeros(4 * S, dtype=torch.float64, device=self.device)
        # VegasFlowPlus slots: item, events and offsets per hypercube, per-cube
        # sums, iteration-end results (slot s owns entries s * MAX_CUBES ..)
        self.vp_items_host = (_abi.BatchVplusItem * S)()
        self.vp_items = torch.zeros(S * VP_ITEM_BYTES, dtype=torch.uint8, device=self.device)
        self.n_ev = torch.zeros(S * MAX_CUBES, dtype=torch.int32, device=self.device)
        self.offsets = torch.zeros(S * (MAX_CUBES + 1), dtype=torch.int32, device=self.device)
        self.ress = torch.zeros(S * MAX_CUBES, dtype=torch.float64, device=self.device)
        self.ress2 = torch.zeros(S * MAX_CUBES, dtype=torch.float64, device=self.device)
        self.outs = torch.zeros(S * OUT_LEN, dtype=torch.float64, device=self.device)
        self._linspace = torch.linspace(0, 1, BINS_MAX + 1, dtype=torch.float64, device=self.device)
        # pinned host copies of the per-slot state a tick reads back
        self._h_stats = torch.empty(4 * S, dtype=torch.float64, pin_memory=True)
        self._h_draws = torch.empty(S, dtype=torch.int32, pin_memory=True)
        self._h_next = torch.empty(S, dtype=torch.float64, pin_memory=True)
        self.free = list(range(S - 1, -1, -1))
        self.live = {}  # slot -> Request
        self.queue = collections.deque()
        self.finished = []
        self.ticks = 0
        self.events_done = 0
        self._tile_cache = {}
        self._launch = None  # device tiles / jobs / live slots, rebuilt when the live set changes
        self.tick_seconds = []  # wall time of every tick

    # ---------------------------------------------------------- admission
    def submit(self, req):
        if not isinstance(req.integrand, BuiltinIntegrand) or not req.integrand.fused_ok(req.n_dim) \
                or req.n_dim > MAX_DIM:
            raise ValueError(f"requests need a built-in integrand with n_dim <= {MAX_DIM}")
        if req.n_events < 2 or req.n_events >= 1 << 32:
            raise ValueError("a request needs 2 .. 2^32 - 1 events per iteration")
        if req.algorithm not in ALGORITHMS:
            raise ValueError(f"unknown algorithm {req.algorithm!r}; one of {ALGORITHMS}")
        if req.algorithm == "vegas+" and req.n_events >= 1 << 31:
            raise ValueError("a VegasFlowPlus request needs fewer than 2^31 events per iteration")
        self.queue.append(req)

    def _admit(self, req, now):
        s = self.free.pop()
        if req.algorithm == "vegas+":
            self._admit_vplus(req, s)
        it = self.items_host[s]
        it.n_dim = req.n_dim
        it.train = 1
        it.n_events = req.n_events
        it.seed = req.seed & 0xFFFFFFFFFFFFFFFF
        it.grid_offset = s * GRID_STRIDE
        it.row_offset = s * ROW_STRIDE
        it.fn = req.integrand.to_struct(req.n_dim)
        raw = torch.frombuffer(bytearray(bytes(it)), dtype=torch.uint8)
        self.items[s * ITEM_BYTES:(s + 1) * ITEM_BYTES].copy_(raw)
        g = self.grids[s * GRID_STRIDE:s * GRID_STRIDE + req.n_dim * (BINS_MAX + 1)]
        g.view(req.n_dim, BINS_MAX + 1).copy_(self._linspace.expand(req.n_dim, BINS_MAX + 1))
        self.draws[s] = 0
        self.stats[4 * s:4 * s + 4] = 0.0
        req.slot = s
        req.admitted = now
        req.status = "running"
        self.live[s] = req
        self._launch = None

    def _admit_vplus(self, req, s):
        """The VegasFlowPlus state of slot ``s``: stratification, an equal
        first allocation (vflowplus.py:105-139) and zeroed per-cube sums."""
        from .vflowplus import stratification

        adaptive, _, n_strat, n_cubes, min_neval = stratification(req.n_dim, req.n_events, req.adaptive)
        it = self.vp_items_host[s]
        it.n_dim = req.n_dim
        it.train = 1
        it.n_strat = n_strat
        it.n_cubes = n_cubes
        it.adaptive = int(adaptive)
        it.min_neval = min_neval
        it.init_calls = float(req.n_events)
        it.seed = req.seed & 0xFFFFFFFFFFFFFFFF
        it.grid_offset = s * GRID_STRIDE
        it.row_offset = s * ROW_STRIDE
        it.cube_offset = s * MAX_CUBES
        it.off_offset = s * (MAX_CUBES + 1)
        it.out_offset = s * OUT_LEN
        it.fn = req.integrand.to_struct(req.n_dim)
        raw = torch.frombuffer(bytearray(bytes(it)), dtype=torch.uint8)
        self.vp_items[s * VP_ITEM_BYTES:(s + 1) * VP_ITEM_BYTES].copy_(raw)
        c0 = s * MAX_CUBES
        self.n_ev[c0:c0 + n_cubes].fill_(min_neval)
        o0 = s * (MAX_CUBES + 1)
        self.offsets[o0:o0 + n_cubes + 1].copy_(
            torch.arange(n_cubes + 1, dtype=torch.int32, device=self.device) * min_neval)
        self.ress[c0:c0 + n_cubes].zero_()
        self.ress2[c0:c0 + n_cubes].zero_()
        first = n_cubes * min_neval
        # tiles cover the most events an iteration can have: after a
        # reallocation sum(max(min, x_c)) <= n_cubes * min + N / 2
        req._vp = {"n_cubes": n_cubes, "bound": first + (req.n_events // 2 + 1 if adaptive else 0),
                   "events": first}

    def _retire(self, req, now, status):
        """End ``req`` (running or waiting) with ``status``; a slot it held
        is free for the next admission (the last tick's finalize consumed
        its row, so the slot is clean)."""
        req.done = now
        req.status = status
        if req.slot in self.live and self.live[req.slot] is req:
            del self.live[req.slot]
            self.free.append(req.slot)
            self.live_events -= self._tick_events(req)
            self._launch = None
        self.finished.append(req)

    def cancel(self, req, now=None):
        """Withdraw a waiting or running request (between ticks); returns
        False if it had already finished."""
        now = time.perf_counter() if now is None else now
        if req.status == "waiting":
            at = next((i for i, q in enumerate(self.queue) if q is req), None)
            if at is None:
                return False
            del self.queue[at]
        elif not (req.status == "running" and self.live.get(req.slot) is req):
            return False
        self._retire(req, now, "cancelled")
        return True

    def _tiles_of(self, s, req, ept):
        n = req._vp["bound"] if req.algorithm == "vegas+" else req.n_events
        key = (s, req.algorithm, n, ept)
        t = self._tile_cache.get(key)
        if t is None:
            firsts = np.arange(0, n, ept, dtype=np.int64)
            counts = np.minimum(ept, n - firsts).astype(np.int32)
            t = np.zeros(len(firsts), dtype=[("item", "<i4"), ("count", "<i4"), ("first", "<i8")])
            t["item"] = s
            t["count"] = counts
            t["first"] = firsts
            self._tile_cache[key] = t
        return t

    # --------------------------------------------------------------- tick
    def _build_launch(self):
        """Device tiles per integrand family, finalize jobs and the live slot
        list for the current live set (reused by every tick until a request
        is admitted or retired)."""
        by_family = collections.defaultdict(list)
        vp_by_family = collections.defaultdict(list)
        jobs, vp_chunks, vp_dims = [], [], []
        slots = sorted(self.live)
        for s in slots:
            req = self.live[s]
            if req.algorithm == "vegas+":
                # items whose dimensions pair (2..14) share a pair-table launch
                vp_by_family[(req.integrand.family, 2 <= req.n_dim <= 14)].append((s, req))
                nb = (req._vp["n_cubes"] + CUBE_CHUNK - 1) // CUBE_CHUNK
                vp_chunks.extend((s, b) for b in range(nb))
                vp_dims.extend((s, j) for j in range(req.n_dim))
                continue
            by_family[req.integrand.family].append((s, req))
            jobs.append((s, -1))
            jobs.extend((s, j) for j in range(req.n_dim))
        # tiles of at most events_per_tile (pair kernel: pair_events_per_tile)
        # events, smaller (powers of two) when the live requests are too few to
        # give every workgroup two tiles, so that a light load still spreads
        # over the GPU
        total = sum(self.live[s].n_events for s in slots if self.live[s].algorithm == "vegas")
        vp_total = sum(self.live[s]._vp["events"] for s in slots if self.live[s].algorithm == "vegas+")
        if len(self._tile_cache) > 4 * self.capacity:
            self._tile_cache.clear()
        fams = []
        for fam, members in sorted(by_family.items()):
            max_dim = max(r.n_dim for _, r in members)
            pb = kernels.batch_pairs_blocks(fam, max_dim, self.precision) if self.pairs else 0
            if pb > 0:
                ept, blocks, floor = self.pair_events_per_tile, pb, 4096
            else:
                ept, blocks, floor = self.events_per_tile, self.blocks, 1024
            while ept > floor and total < 2 * blocks * ept:
                ept //= 2
            tiles = np.concatenate([self._tiles_of(s, r, ept) for s, r in members])
            tiles_dev = torch.from_numpy(tiles.view(np.uint8)).to(self.device)
            fams.append((fam, tiles_dev, len(tiles), max_dim, sum(r.n_events for _, r in members), pb))
        vp_fams = []
        for (fam, paired), members in sorted(vp_by_family.items()):
            min_dim = min(r.n_dim for _, r in members)
            max_dim = max(r.n_dim for _, r in members)
            if paired:  # 1024-thread workgroups, about one per CU
                ept, blocks, floor = self.pair_events_per_tile, self.cus, 4096
            else:
                ept, blocks, floor = self.events_per_tile, self.blocks, 1024
            while ept > floor and vp_total < 2 * blocks * ept:
                ept //= 2
            tiles = np.concatenate([self._tiles_of(s, r, ept) for s, r in members])
            tiles_dev = torch.from_numpy(tiles.view(np.uint8)).to(self.device)
            vp_fams.append((fam, tiles_dev, len(tiles), min_dim, max_dim))

        def dev_pairs(pairs):
            return torch.tensor(np.array(pairs, dtype=np.int32).reshape(-1), device=self.device)

        jobs_dev = dev_pairs(jobs) if jobs else None
        vp = (vp_fams, dev_pairs(vp_chunks), len(vp_chunks), dev_pairs(vp_dims), len(vp_dims)) if vp_chunks else None
        self._launch = (fams, jobs_dev, len(jobs), vp, slots)

    def tick(self):
        """One iteration for every live slot; returns the requests retired."""
        if not self.live:
            return []
        t_start = time.perf_counter()
        if self._launch is None:
            self._build_launch()
        fams, jobs_dev, n_jobs, vp, slots = self._launch
        for fam, tiles_dev, n_tiles, max_dim, events, pb in fams:
            if pb > 0:
                kernels.batch_fused_pairs(self.items, self.capacity, tiles_dev, n_tiles, fam, max_dim,
                                          self.draws, self.grids, self.rows, pb, self.precision)
            else:
                kernels.batch_fused(self.items, self.capacity, tiles_dev, n_tiles, self.precision, fam,
                                    max_dim, self.draws, self.grids, self.rows, self.blocks)
            self.events_done += events
        if n_jobs:
            kernels.batch_finalize(self.items, jobs_dev, n_jobs, self.grids, self.rows, self.draws, self.stats)
        if vp is not None:
            vp_fams, chunks_dev, n_chunks, dims_dev, n_dims = vp
            for fam, tiles_dev, n_tiles, min_dim, max_dim in vp_fams:
                kernels.batch_vplus_fused(self.vp_items, tiles_dev, n_tiles, self.precision, fam, min_dim, max_dim,
                                          self.draws, self.grids, self.n_ev, self.offsets, self.ress, self.ress2,
                                          self.rows, self.blocks)
            kernels.batch_vplus_finalize(self.vp_items, chunks_dev, n_chunks, dims_dev, n_dims, BETA, self.grids,
                                         self.n_ev, self.offsets, self.ress, self.ress2, self.rows, self.outs,
                                         self.draws, self.stats)
        # counters and running sums of the slots up to the highest live one
        # (slots are handed out lowest first), and the VEGAS+ slots' events of
        # the next iteration: asynchronous copies into pinned memory, then one
        # synchronisation
        n = slots[-1] + 1
        self._h_stats[:4 * n].copy_(self.stats[:4 * n], non_blocking=True)
        self._h_draws[:n].copy_(self.draws[:n], non_blocking=True)
        if vp is not None:
            self._h_next[:n].copy_(self.outs.view(-1, OUT_LEN)[:n, 2], non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        stats = self._h_stats[:4 * n].numpy().reshape(n, 4)
        draws = self._h_draws[:n].numpy()
        nexts = self._h_next[:n].numpy()
        now = time.perf_counter()
        self.ticks += 1
        self.tick_seconds.append(now - t_start)
        retired = []
        for s in slots:
            req = self.live[s]
            st = stats[s]
            k = int(draws[s])
            if req.algorithm == "vegas+":
                self.events_done += req._vp["events"]
                req._vp["events"] = int(nexts[s])
            with np.errstate(divide="ignore", invalid="ignore"):
                est = st[3] / st[2]
                err = np.sqrt(1.0 / st[2])
            req.iterations = k
            req.result, req.error = float(est), float(err)
            converged = k >= self.min_iter and np.isfinite(err) and err <= req.target * abs(est)
            if not (np.isfinite(st[0]) and np.isfinite(st[1])):
                status = "nonfinite"  # this iteration's integral: NaN or Inf
            elif converged:
                status = "converged"
            elif k >= req.max_iter:
                status = "max_iter"
            elif req.timeout is not None and now - req.arrival >= req.timeout:
                status = "timeout"
            else:
                continue
            self._retire(req, now, status)
            retired.append(req)
        return retired

    @staticmethod
    def _tick_events(req):
        """Events that ``req`` adds to a tick: its events per iteration (a
        VegasFlowPlus request's reallocation keeps its total within a few
        events per hypercube of that)."""
        return req.n_events

    def admit_waiting(self, now):
        while self.queue and self.free and self.queue[0].arrival <= now:
            req = self.queue[0]
            if req.timeout is not None and now - req.arrival >= req.timeout:
                self.queue.popleft()
                self._retire(req, now, "timeout")  # expired while waiting
                continue
            cost = self._tick_events(req)
            if self.max_tick_events is not None and self.live and \
                    self.live_events + cost > self.max_tick_events:
                break
            self.queue.popleft()
            self._admit(req, now)
            self.live_events += cost

    # -------------------------------------------------------------- loops
    def warm_up(self, requests):
        """Load the code objects the stream will use (one tiny iteration per
        integrand family and dimension present) before serving, so that no
        request waits for a first-launch load. Leaves the server empty."""
        seen = {}
        for r in requests:
            seen.setdefault((r.integrand.family, r.n_dim, r.algorithm), r)
        for (_, d, alg), r in seen.items():
            self.submit(Request(r.integrand, d, 64, target=1.0, max_iter=1, seed=1, algorithm=alg))
        self.admit_waiting(float("inf"))
        while self.live:
            self.tick()
        torch.cuda.synchronize()
        self.finished = []
        self.ticks = 0
        self.events_done = 0
        self.tick_seconds = []

    def run_stream(self, requests, clock=time.perf_counter, warm_up=True):
        """Serve ``requests`` (sorted by ``arrival``, seconds from the
        start) in real time; returns the finished requests."""
        if warm_up:
            self.warm_up(requests)
        t0 = clock()
        for r in requests:
            r.arrival = t0 + r.arrival
            self.submit(r)
        while self.queue or self.live:
            now = clock()
            self.admit_waiting(now)
            if self.live:
                self.tick()
            elif self.queue:
                wait = self.queue[0].arrival - clock()
                if wait > 0:
                    time.sleep(min(wait, 0.001))
        return self.finished


def poisson_requests(rate=200.0, duration=60.0, seed=5, events_gauss=(1_000_000, 20_000_000),
                     events_bw=(1_000_000, 20_000_000), targets=(1e-3, 1e-4), max_iter=20):
    """The serving workload: Poisson arrivals at ``rate`` per second for
    ``duration`` seconds; each request is the two-Gaussian mixture in d = 4
    or the Breit-Wigner x power law in d = 6, with a relative-error target
    drawn from ``targets`` (events per iteration: the first entry of the
    events pair for the looser target, the second for the tighter one). The
    mixture runs as VegasFlowPlus with adaptive allocation (the stratified
    configuration), the Breit-Wigner as VegasFlow."""
    from .integrands import BreitWignerPower, GaussianMix

    rng = np.random.default_rng(seed)
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        target = float(targets[int(rng.integers(len(targets)))])
        tight = target == min(targets)
        if rng.random() < 0.5:
            fn, d, n, alg = GaussianMix(), 4, events_gauss[1 if tight else 0], "vegas+"
        else:
            fn, d, n, alg = BreitWignerPower(), 6, events_bw[1 if tight else 0], "vegas"
        out.append(Request(fn, d, n, target=target, max_iter=max_iter, seed=10_000 + rid, arrival=t, rid=rid,
                           algorithm=alg))
        rid += 1
    return out


def run_alone(req, dtype=None, n_iter=None):
    """The request on its own through VegasFlow (or VegasFlowPlus): the same
    draws, iterations ``n_iter`` (default: what it ran in the server)."""
    from .vflow import VegasFlow
    from .vflowplus import VegasFlowPlus

    dtype = DTYPE if dtype is None else dtype
    if req.algorithm == "vegas+":
        inst = VegasFlowPlus(req.n_dim, req.n_events, adaptive=req.adaptive, seed=req.seed, verbose=False,
                             dtype=dtype)
    else:
        inst = VegasFlow(req.n_dim, req.n_events, seed=req.seed, verbose=False, dtype=dtype)
    inst.compile(req.integrand)
    return inst.run_integration(n_iter or req.iterations)


def latency_report(finished):
    lat = np.array([r.latency for r in finished if r.latency is not None])
    if lat.size == 0:
        return {}
    classes = collections.defaultdict(list)
    for r in finished:
        classes[f"{type(r.integrand).__name__}/d{r.n_dim}/{r.target:g}"].append(r)
    by_class = {}
    for key, rs in sorted(classes.items()):
        cl = np.array([r.latency for r in rs if r.latency is not None])
        by_class[key] = {
            "requests": len(rs),
            "events_per_iteration": int(np.median([r.n_events for r in rs])),
            "p50_ms": float(np.percentile(cl, 50) * 1e3) if cl.size else None,
            "p99_ms": float(np.percentile(cl, 99) * 1e3) if cl.size else None,
            "mean_iterations": float(np.mean([r.iterations for r in rs])),
            "met_target": float(np.mean([_met(r) for r in rs])),
        }
    waits = np.array([r.admitted - r.arrival for r in finished if r.admitted is not None])
    return {
        "requests": int(lat.size),
        "queue_wait_ms_p99": float(np.percentile(waits, 99) * 1e3) if waits.size else None,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "max_ms": float(lat.max() * 1e3),
        "mean_iterations": float(np.mean([r.iterations for r in finished])),
        "met_target": float(np.mean([_met(r) for r in finished])),
        "status": dict(collections.Counter(r.status for r in finished)),
        "by_class": by_class,
    }


def _met(req):
    return req.result is not None and req.error is not None and bool(req.error <= req.target * abs(req.result))
