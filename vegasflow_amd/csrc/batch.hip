// batch.hip - many independent VegasFlow integrals in one persistent launch.
//
// The events of every item of an iteration round are cut into tiles of
// consecutive events of one item; ``blocks`` workgroups stride over the tile
// list. A tile loads its item's map table into LDS, runs the same per-event
// body as the single-integral fused kernels (fused_event.hpp) and adds its
// sums and histogram into the item's row with FP64 atomics. The finalize
// kernel gives one wave per (item, sums) and per (item, dimension) job.
// Grids, rows, per-item draw counters, running weighted averages and the
// per-iteration history all stay on the device, so a round can be replayed
// from a captured HIP graph: nothing in it depends on host values.
#include <algorithm>
#include <type_traits>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "refine.hpp"
#include "rows.hpp"

namespace mv {

template <bool MIXED, int FAM, bool TRAIN>
__device__ __forceinline__ void batch_events(const mv_batch_item& it, const mv_batch_tile& tile,
                                             uint32_t draw, const void* tab, const void* tw,
                                             const JacGroups& cs, double* hist, int nb, double* row) {
  using T = typename std::conditional<MIXED, float, double>::type;
  Integrand<T, FAM> F(it.fn);
  const Key key{(uint32_t)(it.seed & 0xffffffffu), (uint32_t)(it.seed >> 32)};
  const int d = it.n_dim;
  const int groups = (d + 3) / 4;
  const double wscale = 1.0 / (double)it.n_events;  // xjac, monte_carlo.py:224-227
  double* wh = hist + (threadIdx.x >> 6) * nb;
  double acc = 0.0, acc2 = 0.0;
  // a tile never crosses a multiple of 2^32 events (mv_batch_fused), so the
  // high counter word is uniform over the tile
  const uint32_t lo0 = (uint32_t)((uint64_t)tile.first & 0xffffffffu);
  const uint32_t hi = (uint32_t)((uint64_t)tile.first >> 32);
  for (int k = threadIdx.x; k < tile.count; k += kBlock) {
    double tmp;
    if constexpr (MIXED)
      tmp = event_mixed<FAM, TRAIN>(F, lo0 + (uint32_t)k, hi, draw, key, groups, d,
                                    static_cast<const float2*>(tab), cs, wh, nullptr, wscale);
    else
      tmp = event_f64<FAM, TRAIN>(F, lo0 + (uint32_t)k, hi, draw, key, groups, d,
                                  static_cast<const double2*>(tab), static_cast<const double*>(tw), wh,
                                  nullptr, wscale);
    acc += tmp;
    acc2 += tmp * tmp;
  }
  finish_row<TRAIN>(acc, acc2, hist, nb, row);
}

template <bool MIXED, int FAM>
__device__ __forceinline__ void batch_tile(const mv_batch_item& it, const mv_batch_tile& tile,
                                           uint32_t draw, const double* __restrict__ grids,
                                           double* __restrict__ rows, double* smem, int nbmax) {
  using T = typename std::conditional<MIXED, float, double>::type;
  using T2 = typename std::conditional<MIXED, float2, double2>::type;
  const int d = it.n_dim;
  const int nb = d * kBins;
  T2* tab = reinterpret_cast<T2*>(smem);
  // the mixed table has no W column (fused_event.hpp)
  T* tw = MIXED ? nullptr : reinterpret_cast<T*>(smem + 2 * nbmax);
  double* hist = smem + (MIXED ? nbmax : 3 * nbmax);
  __syncthreads();  // the previous tile is done with the LDS
  build_table<T, FAM>(it.fn, grids + it.grid_offset, d, nullptr, nullptr, tab, tw);
  if (it.train)
    for (int i = threadIdx.x; i < kWaves * nb; i += kBlock) hist[i] = 0.0;
  const JacGroups cs = MIXED ? jacobian_groups<FAM>(it.fn, d, nullptr) : JacGroups{};
  __syncthreads();
  double* row = rows + it.row_offset;
  if (it.train)
    batch_events<MIXED, FAM, true>(it, tile, draw, tab, tw, cs, hist, nb, row);
  else
    batch_events<MIXED, FAM, false>(it, tile, draw, tab, tw, cs, hist, nb, row);
}

// FAM >= 0: every tile belongs to an item of that family (smaller register
// footprint); FAM < 0: dispatch on each tile's item.
template <bool MIXED, int FAM>
__global__ __launch_bounds__(kBlock) void batch_fused(const mv_batch_item* __restrict__ items,
                                                      const mv_batch_tile* __restrict__ tiles,
                                                      int n_tiles, int nbmax,
                                                      const int32_t* __restrict__ draws,
                                                      const double* __restrict__ grids,
                                                      double* __restrict__ rows) {
  extern __shared__ double smem[];
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const mv_batch_tile tile = tiles[t];
    const mv_batch_item& it = items[tile.item];
    const uint32_t draw = (uint32_t)draws[tile.item];
    if constexpr (FAM >= 0) {
      batch_tile<MIXED, FAM>(it, tile, draw, grids, rows, smem, nbmax);
    } else {
      switch (it.fn.family) {
        case MV_FN_CONSTANT: batch_tile<MIXED, MV_FN_CONSTANT>(it, tile, draw, grids, rows, smem, nbmax); break;
        case MV_FN_GAUSSIAN: batch_tile<MIXED, MV_FN_GAUSSIAN>(it, tile, draw, grids, rows, smem, nbmax); break;
        case MV_FN_PRODUCT: batch_tile<MIXED, MV_FN_PRODUCT>(it, tile, draw, grids, rows, smem, nbmax); break;
        case MV_FN_GAUSSMIX: batch_tile<MIXED, MV_FN_GAUSSMIX>(it, tile, draw, grids, rows, smem, nbmax); break;
        default: batch_tile<MIXED, MV_FN_BWPOWER>(it, tile, draw, grids, rows, smem, nbmax); break;
      }
    }
  }
}

// Tiles with pair histograms (fused_event.hpp), BS threads. LDS = the map
// table (mixed: nbmax float2; FP64: nbmax (P, Q) pairs and nbmax W), then
// pair_cells(max_dim) doubles. Training items add one LDS atomic per pair of
// dimensions; frozen items run the plain event body of their precision.
template <int FAM, int BS, bool MIXED, int D = 0>
__device__ __forceinline__ void batch_tile_pairs(const mv_batch_item& it, const mv_batch_tile& tile,
                                                 uint32_t draw, const double* __restrict__ grids,
                                                 double* __restrict__ rows, double* smem, int nbmax) {
  const int d = D ? D : it.n_dim;  // D > 0: compiled for that n_dim (straight-line event body)
  // 16-byte aligned: nbmax * 8 is a multiple of 400
  double* H = smem + (MIXED ? nbmax : 3 * nbmax);
  __syncthreads();  // the previous tile is done with the LDS
  const Key key{(uint32_t)(it.seed & 0xffffffffu), (uint32_t)(it.seed >> 32)};
  const double wscale = 1.0 / (double)it.n_events;  // xjac, monte_carlo.py:224-227
  // a tile never crosses a multiple of 2^32 events: uniform high counter word
  const uint32_t lo0 = (uint32_t)((uint64_t)tile.first & 0xffffffffu);
  const uint32_t hi = (uint32_t)((uint64_t)tile.first >> 32);
  const int groups = (d + 3) / 4;
  double acc = 0.0, acc2 = 0.0;
  if constexpr (MIXED) {
    float2* tab = reinterpret_cast<float2*>(smem);
    build_table<float, FAM>(it.fn, grids + it.grid_offset, d, nullptr, nullptr, tab, nullptr);
    if (it.train) {
      const int half = pair_cells(d) / 2;
      for (int i = threadIdx.x; i < half; i += BS) reinterpret_cast<double2*>(H)[i] = make_double2(0.0, 0.0);
    }
    const JacGroups cs = jacobian_groups<FAM>(it.fn, d, nullptr);
    __syncthreads();
    Integrand<float, FAM> F(it.fn);
    if (it.train) {
      for (int k = threadIdx.x; k < tile.count; k += BS) {
        const double tmp = event_mixed_pairs<FAM, false>(F, lo0 + (uint32_t)k, hi, draw, key, d, tab,
                                                         cs, H, nullptr, wscale);
        acc += tmp;
        acc2 += tmp * tmp;
      }
    } else {
      for (int k = threadIdx.x; k < tile.count; k += BS) {
        const double tmp = event_mixed<FAM, false>(F, lo0 + (uint32_t)k, hi, draw, key, groups, d, tab,
                                                   cs, nullptr, nullptr, wscale);
        acc += tmp;
        acc2 += tmp * tmp;
      }
    }
  } else {
    double2* tab = reinterpret_cast<double2*>(smem);
    double* tw = smem + 2 * nbmax;
    build_table<double, FAM>(it.fn, grids + it.grid_offset, d, nullptr, nullptr, tab, tw);
    if (it.train)
      for (int i = threadIdx.x; i < pair_cells(d); i += BS) H[i] = 0.0;
    __syncthreads();
    Integrand<double, FAM> F(it.fn);
    if (it.train) {
      for (int k = threadIdx.x; k < tile.count; k += BS) {
        const double tmp = event_f64<FAM, true, true, 0>(F, lo0 + (uint32_t)k, hi, draw, key, groups, d,
                                                          tab, tw, H, nullptr, wscale);
        acc += tmp;
        acc2 += tmp * tmp;
      }
    } else {
      for (int k = threadIdx.x; k < tile.count; k += BS) {
        const double tmp = event_f64<FAM, false, false, 0>(F, lo0 + (uint32_t)k, hi, draw, key, groups, d,
                                                            tab, tw, nullptr, nullptr, wscale);
        acc += tmp;
        acc2 += tmp * tmp;
      }
    }
  }
  finish_pairs<BS, true, MIXED>(acc, acc2, H, it.train ? d : 0, rows + it.row_offset);
}

template <int FAM, int BS, bool MIXED>
__global__ __launch_bounds__(BS) void batch_fused_pairs(const mv_batch_item* __restrict__ items,
                                                        const mv_batch_tile* __restrict__ tiles,
                                                        int n_tiles, int nbmax,
                                                        const int32_t* __restrict__ draws,
                                                        const double* __restrict__ grids,
                                                        double* __restrict__ rows) {
  extern __shared__ double smem[];
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const mv_batch_tile tile = tiles[t];
    const mv_batch_item& it = items[tile.item];
    const uint32_t draw = (uint32_t)draws[tile.item];
    if constexpr (FAM >= 0) {
      // one event body per n_dim 1..kPairMaxDim, chosen per tile
#define MV_BTP(D) case D: batch_tile_pairs<FAM, BS, MIXED, D>(it, tile, draw, grids, rows, smem, nbmax); break
      switch (it.n_dim) {
        MV_BTP(1); MV_BTP(2); MV_BTP(3); MV_BTP(4); MV_BTP(5); MV_BTP(6); MV_BTP(7);
        MV_BTP(8); MV_BTP(9); MV_BTP(10); MV_BTP(11); MV_BTP(12); MV_BTP(13); MV_BTP(14);
        default: batch_tile_pairs<FAM, BS, MIXED>(it, tile, draw, grids, rows, smem, nbmax); break;
      }
#undef MV_BTP
      static_assert(kPairMaxDim == 14, "the switch above covers n_dim 1..14");
    } else {
      switch (it.fn.family) {
        case MV_FN_CONSTANT: batch_tile_pairs<MV_FN_CONSTANT, BS, MIXED>(it, tile, draw, grids, rows, smem, nbmax); break;
        case MV_FN_GAUSSIAN: batch_tile_pairs<MV_FN_GAUSSIAN, BS, MIXED>(it, tile, draw, grids, rows, smem, nbmax); break;
        case MV_FN_PRODUCT: batch_tile_pairs<MV_FN_PRODUCT, BS, MIXED>(it, tile, draw, grids, rows, smem, nbmax); break;
        case MV_FN_GAUSSMIX: batch_tile_pairs<MV_FN_GAUSSMIX, BS, MIXED>(it, tile, draw, grids, rows, smem, nbmax); break;
        default: batch_tile_pairs<MV_FN_BWPOWER, BS, MIXED>(it, tile, draw, grids, rows, smem, nbmax); break;
      }
    }
  }
}

__global__ __launch_bounds__(64) void batch_finalize_kernel(const mv_batch_item* __restrict__ items,
                                                            const int32_t* __restrict__ jobs,
                                                            double* __restrict__ grids,
                                                            double* __restrict__ rows,
                                                            int32_t* __restrict__ draws,
                                                            double* __restrict__ stats,
                                                            double* __restrict__ history,
                                                            int max_iter) {
  __shared__ double hb[kBins];
  const int item = jobs[2 * blockIdx.x], dim = jobs[2 * blockIdx.x + 1];
  const mv_batch_item& it = items[item];
  double* row = rows + it.row_offset;
  const int t = threadIdx.x;
  if (dim < 0) {
    if (t == 0) {
      const double N = (double)it.n_events;
      const double res = row[0], res2 = row[1];
      row[0] = 0.0;
      row[1] = 0.0;
      const double err = (N * res2 - res * res) / (N - 1.0);  // vflow.py:437
      const double sig = sqrt_clamped(err);                   // vflow.py:438
      double* st = stats + 4 * (int64_t)item;
      const double wgt = 1.0 / (sig * sig);                   // monte_carlo.py:718
      st[0] = res;
      st[1] = sig;
      st[2] += wgt;                                           // monte_carlo.py:720
      st[3] += res * wgt;                                     // monte_carlo.py:719
      const int k = draws[item];
      if (history && k < max_iter) {
        history[2 * ((int64_t)item * max_iter + k)] = res;
        history[2 * ((int64_t)item * max_iter + k) + 1] = sig;
      }
      draws[item] = k + 1;
    }
    return;
  }
  const int64_t c = 2 + (int64_t)dim * kBins + t;
  if (t < kBins) {
    hb[t] = row[c];
    row[c] = 0.0;
  }
  __syncthreads();
  if (it.train && wave_hist_finite(hb)) refine_wave(hb, grids + it.grid_offset + (int64_t)dim * (kBins + 1));
}

}  // namespace mv

using namespace mv;

extern "C" int mv_batch_fused(const mv_batch_item* items, int32_t n_items, const mv_batch_tile* tiles,
                              int32_t n_tiles, int32_t precision, int32_t family, int32_t max_dim,
                              const int32_t* draws, const double* grids, double* rows, int32_t blocks,
                              void* stream) {
  if (n_items < 0 || n_tiles < 0) return fail(MV_ERR_ARG, "mv_batch_fused: negative counts");
  if (n_tiles == 0) return MV_OK;
  if (!items || !tiles || !draws || !grids || !rows)
    return fail(MV_ERR_ARG, "mv_batch_fused: missing buffer");
  if (max_dim < 1 || max_dim > MV_FUSED_MAX_DIM)
    return fail(MV_ERR_UNSUPPORTED, "mv_batch_fused: max_dim must be 1..%d", MV_FUSED_MAX_DIM);
  if (precision != MV_F64 && precision != MV_MIXED)
    return fail(MV_ERR_ARG, "mv_batch_fused: unknown precision %d", precision);
  if (blocks < 1) return fail(MV_ERR_ARG, "mv_batch_fused: blocks < 1");
  const int nbmax = max_dim * kBins;
  hipStream_t s = as_stream(stream);
  const bool mixed = precision == MV_MIXED;
  const size_t lds = mixed ? (size_t)(nbmax + kWaves * nbmax) * sizeof(double)
                           : (size_t)(3 * nbmax + kWaves * nbmax) * sizeof(double);
#define MV_BF(MX, F)                                                                                \
  hipLaunchKernelGGL((batch_fused<MX, F>), dim3(blocks), dim3(kBlock), lds, s, items, tiles, n_tiles, \
                     nbmax, draws, grids, rows)
#define MV_BF_ALL(MX)                                             \
  switch (family) {                                               \
    case MV_FN_CONSTANT: MV_BF(MX, MV_FN_CONSTANT); break;        \
    case MV_FN_GAUSSIAN: MV_BF(MX, MV_FN_GAUSSIAN); break;        \
    case MV_FN_PRODUCT: MV_BF(MX, MV_FN_PRODUCT); break;          \
    case MV_FN_GAUSSMIX: MV_BF(MX, MV_FN_GAUSSMIX); break;        \
    case MV_FN_BWPOWER: MV_BF(MX, MV_FN_BWPOWER); break;          \
    default: MV_BF(MX, -1); break;                                \
  }
  if (mixed) { MV_BF_ALL(true) } else { MV_BF_ALL(false) }
#undef MV_BF_ALL
#undef MV_BF
  return check_launch("mv_batch_fused");
}

namespace {
constexpr int kBatchPairBlock = 1024;

size_t batch_pairs_lds(int max_dim, bool mixed) {
  if (max_dim < 1 || max_dim > kPairMaxDim) return 0;
  const size_t table = (size_t)max_dim * kBins * (mixed ? 1 : 3);
  const size_t lds = (table + pair_cells(max_dim)) * sizeof(double);
  return lds <= max_lds_per_block() ? lds : 0;
}

template <int F, bool MIXED>
const void* batch_pairs_kernel() {
  return reinterpret_cast<const void*>(batch_fused_pairs<F, kBatchPairBlock, MIXED>);
}

template <bool MIXED>
int32_t pairs_blocks(int32_t family, int32_t max_dim) {
  const size_t lds = batch_pairs_lds(max_dim, MIXED);
  if (!lds) return 0;
  const void* k = nullptr;
  switch (family) {
    case MV_FN_CONSTANT: k = batch_pairs_kernel<MV_FN_CONSTANT, MIXED>(); break;
    case MV_FN_GAUSSIAN: k = batch_pairs_kernel<MV_FN_GAUSSIAN, MIXED>(); break;
    case MV_FN_PRODUCT: k = batch_pairs_kernel<MV_FN_PRODUCT, MIXED>(); break;
    case MV_FN_GAUSSMIX: k = batch_pairs_kernel<MV_FN_GAUSSMIX, MIXED>(); break;
    case MV_FN_BWPOWER: k = batch_pairs_kernel<MV_FN_BWPOWER, MIXED>(); break;
    default: k = batch_pairs_kernel<-1, MIXED>(); break;
  }
  return resident_blocks(k, lds, kBatchPairBlock);
}

template <bool MIXED>
int fused_pairs(const mv_batch_item* items, int32_t n_items, const mv_batch_tile* tiles, int32_t n_tiles,
                int32_t family, int32_t max_dim, const int32_t* draws, const double* grids, double* rows,
                int32_t blocks, void* stream, const char* what) {
  if (n_items < 0 || n_tiles < 0) return fail(MV_ERR_ARG, "%s: negative counts", what);
  if (n_tiles == 0) return MV_OK;
  if (!items || !tiles || !draws || !grids || !rows) return fail(MV_ERR_ARG, "%s: missing buffer", what);
  const size_t lds = batch_pairs_lds(max_dim, MIXED);
  if (!lds) return fail(MV_ERR_UNSUPPORTED, "%s: max_dim must be 1..%d and fit the LDS", what, kPairMaxDim);
  if (blocks < 1) return fail(MV_ERR_ARG, "%s: blocks < 1", what);
  const int nbmax = max_dim * kBins;
  hipStream_t s = as_stream(stream);
#define MV_BP(F)                                                                                     \
  hipLaunchKernelGGL((batch_fused_pairs<F, kBatchPairBlock, MIXED>), dim3(blocks), dim3(kBatchPairBlock), \
                     lds, s, items, tiles, n_tiles, nbmax, draws, grids, rows)
  switch (family) {
    case MV_FN_CONSTANT: MV_BP(MV_FN_CONSTANT); break;
    case MV_FN_GAUSSIAN: MV_BP(MV_FN_GAUSSIAN); break;
    case MV_FN_PRODUCT: MV_BP(MV_FN_PRODUCT); break;
    case MV_FN_GAUSSMIX: MV_BP(MV_FN_GAUSSMIX); break;
    case MV_FN_BWPOWER: MV_BP(MV_FN_BWPOWER); break;
    default: MV_BP(-1); break;
  }
#undef MV_BP
  return check_launch(what);
}
}  // namespace

extern "C" int32_t mv_batch_pairs_blocks(int32_t family, int32_t max_dim) {
  return pairs_blocks<true>(family, max_dim);
}

extern "C" int32_t mv_batch_pairs_blocks_f64(int32_t family, int32_t max_dim) {
  return pairs_blocks<false>(family, max_dim);
}

extern "C" int mv_batch_fused_pairs(const mv_batch_item* items, int32_t n_items,
                                    const mv_batch_tile* tiles, int32_t n_tiles, int32_t family,
                                    int32_t max_dim, const int32_t* draws, const double* grids,
                                    double* rows, int32_t blocks, void* stream) {
  return fused_pairs<true>(items, n_items, tiles, n_tiles, family, max_dim, draws, grids, rows, blocks, stream,
                           "mv_batch_fused_pairs");
}

extern "C" int mv_batch_fused_pairs_f64(const mv_batch_item* items, int32_t n_items,
                                        const mv_batch_tile* tiles, int32_t n_tiles, int32_t family,
                                        int32_t max_dim, const int32_t* draws, const double* grids,
                                        double* rows, int32_t blocks, void* stream) {
  return fused_pairs<false>(items, n_items, tiles, n_tiles, family, max_dim, draws, grids, rows, blocks, stream,
                            "mv_batch_fused_pairs_f64");
}

extern "C" int mv_batch_finalize(const mv_batch_item* items, const int32_t* jobs, int32_t n_jobs,
                                 double* grids, double* rows, int32_t* draws, double* stats,
                                 double* history, int32_t max_iter, void* stream) {
  if (n_jobs < 0) return fail(MV_ERR_ARG, "mv_batch_finalize: negative n_jobs");
  if (n_jobs == 0) return MV_OK;
  if (!items || !jobs || !grids || !rows || !draws || !stats)
    return fail(MV_ERR_ARG, "mv_batch_finalize: missing buffer");
  if (history && max_iter < 1) return fail(MV_ERR_ARG, "mv_batch_finalize: max_iter < 1");
  hipLaunchKernelGGL(batch_finalize_kernel, dim3(n_jobs), dim3(64), 0, as_stream(stream), items, jobs,
                     grids, rows, draws, stats, history, max_iter);
  return check_launch("mv_batch_finalize");
}
