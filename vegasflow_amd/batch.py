"""Batched VEGAS: many independent integrals advanced together on one GPU.

Each item is a VegasFlow integral (built-in integrand, n_dim <= 16, its own
event count and Philox seed). One iteration round is one persistent launch
of ``mv_batch_fused`` (or the pair-histogram kernel, ``mv_batch_fused_pairs``
or ``mv_batch_fused_pairs_f64``) per integrand family present (workgroups
stride over tiles of consecutive events of one item) plus one
``mv_batch_finalize``
(sums, sigma, running weighted average, history, grid refinement), with all
state resident on the device. Item i draws exactly what
``VegasFlow(n_dim_i, n_events_i, seed=seed_i)`` would, so its results equal a
lone run up to summation order.

Rounds can be captured once as a HIP graph (torch.cuda.CUDAGraph) and
replayed: the per-item iteration counters live on the device, so nothing in a
round depends on the host. The tile size, workgroup count and kernel come
from the autotuner (vegasflow_amd.autotune) unless given.
"""
import hashlib
import logging

import numpy as np
import torch

from . import _abi, autotune, kernels
from .configflow import BINS_MAX, DTYPE, default_device, precision_of
from .integrands import BuiltinIntegrand

logger = logging.getLogger(__name__)


def _bucket(n):
    return int(np.ceil(np.log2(max(int(n), 1))))


class BatchVegas:
    """Independent VegasFlow integrals on one device.

    Parameters
    ----------
        specs: list of (integrand, n_dim, n_events, seed)
        dtype: torch.float64 (FP64 drop-in) or torch.float32 (mixed precision)
        train: refine every item's grid after each iteration
        max_iter: iterations kept in the per-item history
        events_per_tile, blocks_per_cu: launch shape; autotuned when None
        pair_histograms: train through pair tables (mv_batch_fused_pairs,
            or mv_batch_fused_pairs_f64 in FP64; 1024-thread workgroups, as
            many as are resident; families whose max_dim it does not take
            fall back to mv_batch_fused with blocks_per_cu); autotuned when
            None
    """

    def __init__(self, specs, dtype=None, train=True, max_iter=64, events_per_tile=None, blocks_per_cu=None,
                 pair_histograms=None):
        kernels.require_gpu()
        self.device = default_device()
        self.precision = precision_of(DTYPE if dtype is None else dtype)
        self.train = bool(train)
        self.max_iter = int(max_iter)
        self.specs = []
        for spec in specs:
            fn, d, n, seed = spec
            if not isinstance(fn, BuiltinIntegrand) or not fn.fused_ok(int(d)) or int(d) > _abi.FUSED_MAX_DIM:
                raise ValueError(f"batch items need a built-in integrand with n_dim <= {_abi.FUSED_MAX_DIM}")
            if int(n) < 2 or int(n) >= 1 << 32:
                raise ValueError("every item needs 2 .. 2^32 - 1 events per iteration")
            self.specs.append((fn, int(d), int(n), int(seed)))
        self.n_items = len(self.specs)
        if self.n_items == 0:
            raise ValueError("empty batch")

        items = (_abi.BatchItem * self.n_items)()
        grid_off = row_off = 0
        grids = []
        jobs = []
        for i, (fn, d, n, seed) in enumerate(self.specs):
            it = items[i]
            it.n_dim = d
            it.train = int(self.train)
            it.n_events = n
            it.seed = seed & 0xFFFFFFFFFFFFFFFF
            it.grid_offset = grid_off
            it.row_offset = row_off
            it.fn = fn.to_struct(d)
            grids.append(np.tile(np.linspace(0, 1, BINS_MAX + 1), (d, 1)))
            grid_off += d * (BINS_MAX + 1)
            row_off += 2 + d * BINS_MAX
            jobs.append((i, -1))
            if self.train:
                jobs.extend((i, j) for j in range(d))
        self._items_host = items
        raw = np.frombuffer(bytes(items), dtype=np.uint8).copy()
        self.items = torch.from_numpy(raw).to(self.device)
        self.grids = torch.tensor(np.concatenate([g.reshape(-1) for g in grids]), dtype=torch.float64,
                                  device=self.device)
        self._grid_offsets = [items[i].grid_offset for i in range(self.n_items)]
        self.rows = torch.zeros(row_off, dtype=torch.float64, device=self.device)
        self.draws = torch.zeros(self.n_items, dtype=torch.int32, device=self.device)
        self.stats = torch.zeros(4 * self.n_items, dtype=torch.float64, device=self.device)
        self.history_buf = torch.zeros(2 * self.n_items * self.max_iter, dtype=torch.float64,
                                       device=self.device)
        self.jobs = torch.tensor(np.array(jobs, dtype=np.int32).reshape(-1), device=self.device)
        self.n_jobs = len(jobs)
        self._cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self._graph = None
        self._warmed = False

        if events_per_tile is None or (blocks_per_cu is None and not pair_histograms):
            ept, bpc, pairs = self._tune(pair_histograms)
            events_per_tile = events_per_tile or ept
            blocks_per_cu = blocks_per_cu or bpc
            if pair_histograms is None:
                pair_histograms = bool(pairs)
        self.events_per_tile = int(events_per_tile)
        self.pairs = bool(pair_histograms)
        self.blocks = int(blocks_per_cu or autotune.BLOCKS_PER_CU[-1]) * self._cus
        self._launches = self._tile_launches(self.events_per_tile)
        self._pair_blocks = self._pair_launch_blocks(self._launches) if self.pairs else {}

    # ------------------------------------------------------------ tiling
    def _tile_launches(self, events_per_tile):
        """One launch per family present: (family, tiles tensor, n_tiles, max_dim)."""
        by_family = {}
        for i, (fn, d, n, _) in enumerate(self.specs):
            by_family.setdefault(fn.family, []).append(i)
        launches = []
        for fam, members in sorted(by_family.items()):
            tiles = []
            for i in members:
                n = self.specs[i][2]
                for first in range(0, n, events_per_tile):
                    tiles.append((i, min(events_per_tile, n - first), first))
            # heaviest tiles (most dimensions, full tiles) first
            tiles.sort(key=lambda t: (-self.specs[t[0]][1] * t[1], t[0], t[2]))
            arr = (_abi.BatchTile * len(tiles))()
            for k, (i, c, f) in enumerate(tiles):
                arr[k].item, arr[k].count, arr[k].first = i, c, f
            dev_tiles = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(self.device)
            max_dim = max(self.specs[i][1] for i in members)
            launches.append((fam, dev_tiles, len(tiles), max_dim))
        return launches

    def _pair_launch_blocks(self, launches):
        """Resident pair-kernel workgroups per family launch (0: not taken)."""
        return {fam: kernels.batch_pairs_blocks(fam, maxd, self.precision) for fam, _, _, maxd in launches}

    def _signature(self):
        h = hashlib.sha1()
        pairs = sorted((fn.family, d, _bucket(n)) for fn, d, n, _ in self.specs)
        h.update(repr(pairs).encode())
        return f"batch|{self.precision}|{int(self.train)}|cus={self._cus}|{h.hexdigest()[:16]}"

    def _launch_round(self, launches, blocks, pair_blocks, rows):
        for fam, tiles, nt, maxd in launches:
            pb = pair_blocks.get(fam, 0)
            if pb > 0:
                kernels.batch_fused_pairs(self.items, self.n_items, tiles, nt, fam, maxd, self.draws,
                                          self.grids, rows, pb, self.precision)
            else:
                kernels.batch_fused(self.items, self.n_items, tiles, nt, self.precision, fam, maxd,
                                    self.draws, self.grids, rows, blocks)

    def _tune(self, pair_histograms=None):
        """Fastest (events per tile, workgroups per CU, pairs) for this batch.
        pair_histograms None: both kernels compete; True / False: only the
        pair / per-wave candidates."""
        scratch = torch.zeros_like(self.rows)

        def measure(cand):
            ept, bpc, pairs = cand
            launches = self._tile_launches(ept)
            pair_blocks = self._pair_launch_blocks(launches) if pairs else {}
            bpc = bpc or autotune.BLOCKS_PER_CU[-1]

            def launch():
                self._launch_round(launches, bpc * self._cus, pair_blocks, scratch)

            ms = autotune.time_ms(launch)
            scratch.zero_()
            return ms

        cands = []
        if not pair_histograms:
            cands += [(e, b, 0) for e in autotune.EVENTS_PER_TILE for b in autotune.BLOCKS_PER_CU]
        if pair_histograms is not False:
            cands += [(e, 0, 1) for e in autotune.PAIR_EVENTS_PER_TILE]
        choice = autotune.tune(self._signature() + f"|pairs={pair_histograms}", cands, measure)
        logger.info("batch launch shape: %d events per tile, %d workgroups per CU, pairs %d", *choice)
        return choice

    # ------------------------------------------------------------ rounds
    def _round(self):
        self._launch_round(self._launches, self.blocks, self._pair_blocks, self.rows)
        kernels.batch_finalize(self.items, self.jobs, self.n_jobs, self.grids, self.rows, self.draws,
                               self.stats, self.history_buf, self.max_iter)

    def _warm_up(self):
        """Launch both kernels once on scratch copies so that code objects
        are loaded before a graph capture; the batch state is untouched."""
        if self._warmed:
            return
        rows, draws = self.rows.clone(), self.draws.clone()
        grids, stats, hist = self.grids.clone(), self.stats.clone(), self.history_buf.clone()
        for fam, tiles, nt, maxd in self._launches:
            if self._pair_blocks.get(fam, 0) > 0:
                kernels.batch_fused_pairs(self.items, self.n_items, tiles, nt, fam, maxd, draws, grids, rows,
                                          self._pair_blocks[fam], self.precision)
            else:
                kernels.batch_fused(self.items, self.n_items, tiles, nt, self.precision, fam, maxd, draws,
                                    grids, rows, self.blocks)
        kernels.batch_finalize(self.items, self.jobs, self.n_jobs, grids, rows, draws, stats, hist,
                               self.max_iter)
        torch.cuda.synchronize()
        self._warmed = True

    def capture(self):
        """Capture one round as a HIP graph (done once)."""
        if self._graph is None:
            self._warm_up()
            self._graph, _ = kernels.capture_graph(self._round)
        return self._graph

    def step(self, graph=False):
        """Advance every item by one iteration."""
        if graph:
            self.capture().replay()
        else:
            self._round()

    def run(self, n_iter, graph=True):
        """Advance every item by ``n_iter`` iterations; returns results()."""
        if graph:
            self.capture()
        for _ in range(int(n_iter)):
            self.step(graph)
        return self.results()

    # ----------------------------------------------------------- results
    def results(self):
        """Per-item weighted average over all iterations so far and its
        error (monte_carlo.py:713-732): two float64 arrays."""
        st = self.stats.view(self.n_items, 4).cpu().numpy()
        with np.errstate(divide="ignore", invalid="ignore"):
            return st[:, 3] / st[:, 2], np.sqrt(1.0 / st[:, 2])

    def last(self):
        """Per-item (integral, sigma) of the latest iteration."""
        st = self.stats.view(self.n_items, 4).cpu().numpy()
        return st[:, 0], st[:, 1]

    def history(self):
        """(n_items, iterations, 2) array of per-iteration (integral, sigma)."""
        k = int(self.draws.max().item())
        h = self.history_buf.view(self.n_items, self.max_iter, 2).cpu().numpy()
        return h[:, : min(k, self.max_iter)]

    def chi2_dof(self):
        """Per-item chi^2 per degree of freedom of the recorded iterations
        around their weighted average (FP64, host; NaN below 2 iterations)."""
        h = self.history()
        if h.shape[1] < 2:
            return np.full(self.n_items, np.nan)
        with np.errstate(divide="ignore", invalid="ignore"):
            w = 1.0 / np.square(h[:, :, 1])
            avg = np.sum(w * h[:, :, 0], axis=1) / np.sum(w, axis=1)
            return np.sum(w * np.square(h[:, :, 0] - avg[:, None]), axis=1) / (h.shape[1] - 1)

    def iterations(self):
        return self.draws.cpu().numpy()

    def grid(self, i):
        d = self.specs[i][1]
        off = self._grid_offsets[i]
        return self.grids[off : off + d * (BINS_MAX + 1)].view(d, BINS_MAX + 1)

    def events_per_round(self):
        return sum(n for _, _, n, _ in self.specs)


def random_batch(n_items=256, seed=23, d_range=(2, 12), n_range=(1e4, 1e6), a_range=(0.05, 0.3),
                 c_range=(0.2, 0.8)):
    """Batch of Gaussian integrals: d ~ U{2..12}, events ~ logU[1e4, 1e6],
    width a ~ U[0.05, 0.3], centres ~ U[0.2, 0.8]^d; item i uses Philox seed
    ``seed * 1000 + i``. Returns a list of specs for BatchVegas."""
    from .integrands import Gaussian

    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_items):
        d = int(rng.integers(d_range[0], d_range[1] + 1))
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        a = float(rng.uniform(*a_range))
        c = [float(v) for v in rng.uniform(*c_range, size=d)]
        specs.append((Gaussian(a, c), d, n, seed * 1000 + i))
    return specs
