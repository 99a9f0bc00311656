// batch_vplus.hip - the producer of batched VegasFlowPlus integrals
// (mv_batch_vplus_fused): many VEGAS+ integrals advance one iteration in one
// persistent launch (vflowplus.py:37-80 + 187-220 per item). The iteration
// end of the items is in vplus.hip (mv_batch_vplus_finalize).
//
// Workgroups stride over tiles of consecutive events of one item. Per tile:
// the item's map table (P, Q, W = dx * 50) in LDS, histograms zeroed, then
// the tile's 64-event trips split evenly over the waves, each wave walking
// its run with the single-integral event body (vplus_run: hypercube lookup
// over the item's offsets, draws, map, integrand, per-cube segmented sums
// into the item's ress / ress2). The histogram goes into the item's row with
// atomics. A tile past the item's events of this iteration (tiles cover an
// upper bound of the adaptive event count) does nothing. Launches whose
// items share n_dim run a kernel compiled for it (the registers of one event
// body); mixed dimensions choose the body per tile.
#include <algorithm>
#include <array>
#include <type_traits>
#include <utility>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "rows.hpp"
#include "vplus.hpp"
#include "vplus_fused.hpp"

namespace mv {

// One wave's run of a tile for the tile's n_dim D: per-wave histograms
// (PAIRS false, hist + wave * 50 D) or the workgroup's pair tables (PAIRS,
// 2 <= D <= 14).
template <bool MIXED, int FAM, bool TRAIN, bool PAIRS, int D>
__device__ void vplus_tile_body(const mv_batch_vplus_item& it, const StepK& s, uint32_t draw,
                                const StrataK& st, int64_t first, int64_t end, const void* tab_,
                                const void* tw_, double* hist, double* __restrict__ ress,
                                double* __restrict__ ress2, double& acc, double& acc2) {
  using T = typename std::conditional<MIXED, float, double>::type;
  using T2 = typename std::conditional<MIXED, float2, double2>::type;
  constexpr bool P = PAIRS && D >= kPairMinDim && D <= kPairMaxDim;
  if (PAIRS && !P) return;  // a pair launch holds only items it can pair
  Integrand<T, FAM> F(it.fn);
  const double fstrat = (double)st.n_strat;
  const float rn = 1.0f / (float)st.n_strat;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* wh = P ? hist : hist + wave * (D * kBins);
  vplus_run<FAM, TRAIN, MIXED, D, false, P>(F, s, draw, st, first, end, fstrat, 1.0 / fstrat, rn,
                                            static_cast<const T2*>(tab_), static_cast<const T*>(tw_), s.xjac,
                                            wh, ress, ress2, acc, acc2);
}

template <bool MIXED, int FAM, bool TRAIN, bool PAIRS, int... Ds>
__device__ void vplus_tile_dispatch(std::integer_sequence<int, Ds...>, int d, const mv_batch_vplus_item& it,
                                    const StepK& s, uint32_t draw, const StrataK& st, int64_t first,
                                    int64_t end, const void* tab, const void* tw, double* hist,
                                    double* ress, double* ress2, double& acc, double& acc2) {
  // one call for the tile's n_dim (d is workgroup-uniform)
  ((d == Ds + 1 ? vplus_tile_body<MIXED, FAM, TRAIN, PAIRS, Ds + 1>(it, s, draw, st, first, end, tab, tw, hist,
                                                                    ress, ress2, acc, acc2)
                : void()),
   ...);
}

// PAIRS: BS = kPairBlock threads and one set of pair tables per workgroup
// (items with 2 <= n_dim <= 14 only): a wave's events lie in one or two
// hypercubes, whose few bins per dimension would take most of a per-wave
// histogram's adds; the pair tables spread them over up to 2500 cells.
// DSPEC > 0: every tiled item has n_dim DSPEC (one event body, the
// registers of that n_dim only); 0: the body is chosen per tile.
template <bool MIXED, int FAM, bool PAIRS, int BS, int DSPEC>
__global__ __launch_bounds__(BS) void batch_vplus_fused(const mv_batch_vplus_item* __restrict__ items,
                                                        const mv_batch_tile* __restrict__ tiles,
                                                        int n_tiles, const int32_t* __restrict__ draws,
                                                        const double* __restrict__ grids,
                                                        const int32_t* __restrict__ n_ev,
                                                        const int32_t* __restrict__ offsets,
                                                        double* __restrict__ ress,
                                                        double* __restrict__ ress2,
                                                        double* __restrict__ rows, int nbmax) {
  using T = typename std::conditional<MIXED, float, double>::type;
  using T2 = typename std::conditional<MIXED, float2, double2>::type;
  constexpr int W = BS / 64;
  extern __shared__ double smem[];
  // LDS: table (nbmax T2), W column (nbmax T), histograms (per-wave, or the
  // pair tables)
  T2* tab = reinterpret_cast<T2*>(smem);
  T* tw = reinterpret_cast<T*>(smem + (MIXED ? nbmax : 2 * nbmax));
  double* hist = smem + (MIXED ? nbmax + (nbmax + 1) / 2 : 3 * nbmax);
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const mv_batch_tile tile = tiles[t];
    const mv_batch_vplus_item& it = items[tile.item];
    const StrataK st{it.n_strat, it.n_cubes, n_ev + it.cube_offset, offsets + it.off_offset};
    const int64_t total = st.off[st.n_cubes];
    const int64_t first = tile.first;
    const int64_t end = min(first + (int64_t)tile.count, total);
    if (first >= end) continue;  // workgroup-uniform: no barrier skipped by part of it
    const int d = it.n_dim;
    const int nb = d * kBins;
    __syncthreads();  // the previous tile is done with the LDS
    const double* grid = grids + it.grid_offset;
    for (int i = threadIdx.x; i < nb; i += BS) {
      const int j = i / kBins, b = i - j * kBins;
      const double lo = grid[j * (kBins + 1) + b];
      const double dx = grid[j * (kBins + 1) + b + 1] - lo;  // vflow.py:73
      double P = lo, Q = dx;
      Integrand<T, FAM>::fold(it.fn, j, P, Q);
      T2 e;
      e.x = (T)P;
      e.y = (T)Q;
      tab[i] = e;
      tw[i] = (T)(dx * kFBins);  // vflow.py:78
    }
    const int hist_len = PAIRS ? pair_cells(d) : W * nb;
    if (it.train)
      for (int i = threadIdx.x; i < hist_len; i += BS) hist[i] = 0.0;
    __syncthreads();
    StepK s{};
    s.n_dim = d;
    s.groups = (d + 3) / 4;
    s.n_events = end - first;
    s.key.k0 = (uint32_t)(it.seed & 0xffffffffu);
    s.key.k1 = (uint32_t)(it.seed >> 32);
    s.xjac = 1.0 / (double)it.n_cubes;  // vflowplus.py:139
    s.xdeltajac = 1.0;
    const uint32_t draw = (uint32_t)draws[tile.item];
    // this wave's run: the tile's trips split evenly over the waves
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t trips = (end - first + 63) >> 6;
    const int64_t a = first + (wave * trips / W) * 64;
    const int64_t b = min(first + ((wave + 1) * trips / W) * 64, end);
    double acc = 0.0, acc2 = 0.0;
    double* cr = ress + it.cube_offset;
    double* cr2 = ress2 + it.cube_offset;
    if constexpr (DSPEC > 0) {
      if (it.train)
        vplus_tile_body<MIXED, FAM, true, PAIRS, DSPEC>(it, s, draw, st, a, b, tab, tw, hist, cr, cr2, acc, acc2);
      else
        vplus_tile_body<MIXED, FAM, false, PAIRS, DSPEC>(it, s, draw, st, a, b, tab, tw, hist, cr, cr2, acc, acc2);
    } else {
      constexpr auto dims = std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{};
      if (it.train)
        vplus_tile_dispatch<MIXED, FAM, true, PAIRS>(dims, d, it, s, draw, st, a, b, tab, tw, hist, cr, cr2, acc,
                                                     acc2);
      else
        vplus_tile_dispatch<MIXED, FAM, false, PAIRS>(dims, d, it, s, draw, st, a, b, tab, tw, hist, cr, cr2, acc,
                                                      acc2);
    }
    // histogram bins into the item's row (its sum columns stay unused: the
    // VEGAS+ sums are per cube)
    if (it.train) {
      if (PAIRS)
        finish_pairs<BS, false>(acc, acc2, hist, d, rows + it.row_offset);
      else
        finish_row<true, false>(acc, acc2, hist, nb, rows + it.row_offset);
    }
  }
}

using BatchVplusKernel = void (*)(const mv_batch_vplus_item*, const mv_batch_tile*, int, const int32_t*,
                                  const double*, const int32_t*, const int32_t*, double*, double*, double*,
                                  int);

template <bool MIXED, int FAM, bool PAIRS, int... Ds>
constexpr std::array<BatchVplusKernel, sizeof...(Ds)> batch_vplus_table(std::integer_sequence<int, Ds...>) {
  return {&batch_vplus_fused<MIXED, FAM, PAIRS, PAIRS ? kPairBlock : kBlock, Ds + 1>...};
}

// The kernel for (family, histogram form): compiled for n_dim ``dspec``
// (1..MV_FUSED_MAX_DIM), or choosing the body per tile (dspec 0).
template <bool MIXED, int FAM, bool PAIRS>
BatchVplusKernel batch_vplus_family(int dspec) {
  static constexpr auto table =
      batch_vplus_table<MIXED, FAM, PAIRS>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{});
  if (dspec >= 1 && dspec <= MV_FUSED_MAX_DIM) return table[dspec - 1];
  return &batch_vplus_fused<MIXED, FAM, PAIRS, PAIRS ? kPairBlock : kBlock, 0>;
}

template <bool MIXED, bool PAIRS>
BatchVplusKernel batch_vplus_kernel(int family, int dspec) {
  switch (family) {
    case MV_FN_CONSTANT: return batch_vplus_family<MIXED, MV_FN_CONSTANT, PAIRS>(dspec);
    case MV_FN_GAUSSIAN: return batch_vplus_family<MIXED, MV_FN_GAUSSIAN, PAIRS>(dspec);
    case MV_FN_PRODUCT: return batch_vplus_family<MIXED, MV_FN_PRODUCT, PAIRS>(dspec);
    case MV_FN_GAUSSMIX: return batch_vplus_family<MIXED, MV_FN_GAUSSMIX, PAIRS>(dspec);
    case MV_FN_BWPOWER: return batch_vplus_family<MIXED, MV_FN_BWPOWER, PAIRS>(dspec);
    default: return nullptr;
  }
}

// The producer of one precision (mv_batch_vplus_fused): picks the histogram
// form and the kernel, sizes the launch. Returns 0 or a negative status.
template <bool MIXED>
int launch_batch_vplus_precision(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int n_tiles,
                                 int family, int min_dim, int max_dim, const int32_t* draws, const double* grids,
                                 const int32_t* n_ev, const int32_t* offsets, double* ress, double* ress2,
                                 double* rows, int blocks, hipStream_t stream) {
  // pair tables whenever every tiled item can pair its dimensions and they
  // fit in a workgroup's LDS
  const int nbmax = max_dim * kBins;
  const size_t table = MIXED ? (size_t)nbmax + (nbmax + 1) / 2 : 3 * (size_t)nbmax;
  const bool pairs = min_dim >= kPairMinDim && max_dim <= kPairMaxDim &&
                     (table + (size_t)pair_cells(max_dim)) * sizeof(double) <= max_lds_per_block();
  const int dspec = min_dim == max_dim ? max_dim : 0;
  BatchVplusKernel k = pairs ? batch_vplus_kernel<MIXED, true>(family, dspec)
                             : batch_vplus_kernel<MIXED, false>(family, dspec);
  if (!k) return fail(MV_ERR_UNSUPPORTED, "mv_batch_vplus_fused: unknown family %d", family);
  const size_t hist = pairs ? (size_t)pair_cells(max_dim) : (size_t)kWaves * nbmax;
  const size_t lds = (table + hist) * sizeof(double);
  const int bs = pairs ? kPairBlock : kBlock;
  // at most the workgroups resident at once (a persistent launch)
  const int res = resident_blocks(reinterpret_cast<const void*>(k), lds, bs);
  if (res < 1) return fail(MV_ERR_UNSUPPORTED, "mv_batch_vplus_fused: a workgroup does not fit (n_dim %d)", max_dim);
  blocks = std::min(std::min(blocks, res), n_tiles);
  hipLaunchKernelGGL(k, dim3(blocks), dim3(bs), lds, stream, items, tiles, n_tiles, draws, grids, n_ev, offsets,
                     ress, ress2, rows, nbmax);
  return check_launch("mv_batch_vplus_fused");
}

int launch_batch_vplus_mixed(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int n_tiles, int family,
                             int min_dim, int max_dim, const int32_t* draws, const double* grids,
                             const int32_t* n_ev, const int32_t* offsets, double* ress, double* ress2, double* rows,
                             int blocks, hipStream_t stream) {
  return launch_batch_vplus_precision<true>(items, tiles, n_tiles, family, min_dim, max_dim, draws, grids, n_ev,
                                            offsets, ress, ress2, rows, blocks, stream);
}

int launch_batch_vplus_f64(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int n_tiles, int family,
                           int min_dim, int max_dim, const int32_t* draws, const double* grids,
                           const int32_t* n_ev, const int32_t* offsets, double* ress, double* ress2, double* rows,
                           int blocks, hipStream_t stream) {
  return launch_batch_vplus_precision<false>(items, tiles, n_tiles, family, min_dim, max_dim, draws, grids, n_ev,
                                             offsets, ress, ress2, rows, blocks, stream);
}

}  // namespace mv
