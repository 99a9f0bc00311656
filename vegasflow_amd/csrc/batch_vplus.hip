// batch_vplus.hip - the producer of batched VegasFlowPlus integrals
// (mv_batch_vplus_fused): many VEGAS+ integrals advance one iteration in one
// persistent launch (vflowplus.py:37-80 + 187-220 per item). The iteration
// end of the items is in vplus.hip (mv_batch_vplus_finalize).
//
// Workgroups stride over tiles of consecutive events of one item. Per tile:
// the item's map table (P, Q, W = dx * 50) in LDS, per-wave histograms, then
// the tile's 64-event trips split evenly over the four waves, each wave
// walking its run with the single-integral event body (vplus_run:
// hypercube lookup over the item's offsets, draws, map, integrand, per-cube
// segmented sums into the item's ress / ress2), compiled per n_dim and
// selected per tile. The histogram goes into the item's row with atomics.
// A tile past the item's events of this iteration (tiles cover an upper
// bound of the adaptive event count) does nothing.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "rows.hpp"
#include "vplus.hpp"
#include "vplus_fused.hpp"

namespace mv {

namespace {

template <bool MIXED, int FAM, bool TRAIN, int D>
__device__ void vplus_tile_body(const mv_batch_vplus_item& it, const StepK& s, uint32_t draw,
                                const StrataK& st, int64_t first, int64_t end, const void* tab_,
                                const void* tw_, double* hist, double* __restrict__ ress,
                                double* __restrict__ ress2, double& acc, double& acc2) {
  using T = typename std::conditional<MIXED, float, double>::type;
  using T2 = typename std::conditional<MIXED, float2, double2>::type;
  Integrand<T, FAM> F(it.fn);
  const double fstrat = (double)st.n_strat;
  const float rn = 1.0f / (float)st.n_strat;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* wh = hist + wave * (D * kBins);
  vplus_run<FAM, TRAIN, MIXED, D, false, false>(F, s, draw, st, first, end, fstrat, 1.0 / fstrat, rn,
                                         static_cast<const T2*>(tab_), static_cast<const T*>(tw_), s.xjac,
                                         wh, ress, ress2, acc, acc2);
}

template <bool MIXED, int FAM, bool TRAIN, int... Ds>
__device__ void vplus_tile_dispatch(std::integer_sequence<int, Ds...>, int d, const mv_batch_vplus_item& it,
                                    const StepK& s, uint32_t draw, const StrataK& st, int64_t first,
                                    int64_t end, const void* tab, const void* tw, double* hist,
                                    double* ress, double* ress2, double& acc, double& acc2) {
  // one call for the tile's n_dim (d is workgroup-uniform)
  ((d == Ds + 1 ? vplus_tile_body<MIXED, FAM, TRAIN, Ds + 1>(it, s, draw, st, first, end, tab, tw, hist, ress,
                                                             ress2, acc, acc2)
                : void()),
   ...);
}

template <bool MIXED, int FAM>
__global__ __launch_bounds__(kBlock) void batch_vplus_fused(const mv_batch_vplus_item* __restrict__ items,
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
  extern __shared__ double smem[];
  // LDS: table (nbmax T2), W column (nbmax T), per-wave histograms
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
    for (int i = threadIdx.x; i < nb; i += kBlock) {
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
    if (it.train)
      for (int i = threadIdx.x; i < kWaves * nb; i += kBlock) hist[i] = 0.0;
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
    const int64_t a = first + (wave * trips / kWaves) * 64;
    const int64_t b = min(first + ((wave + 1) * trips / kWaves) * 64, end);
    double acc = 0.0, acc2 = 0.0;
    double* cr = ress + it.cube_offset;
    double* cr2 = ress2 + it.cube_offset;
    if (it.train)
      vplus_tile_dispatch<MIXED, FAM, true>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{}, d, it, s, draw,
                                            st, a, b, tab, tw, hist, cr, cr2, acc, acc2);
    else
      vplus_tile_dispatch<MIXED, FAM, false>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{}, d, it, s, draw,
                                             st, a, b, tab, tw, hist, cr, cr2, acc, acc2);
    // histogram bins into the item's row (its sum columns stay unused: the
    // VEGAS+ sums are per cube)
    if (it.train)
      finish_row<true, false>(acc, acc2, hist, nb, rows + it.row_offset);
  }
}

using BatchVplusKernel = void (*)(const mv_batch_vplus_item*, const mv_batch_tile*, int, const int32_t*,
                                  const double*, const int32_t*, const int32_t*, double*, double*, double*,
                                  int);

template <bool MIXED>
BatchVplusKernel batch_vplus_kernel(int family) {
  switch (family) {
    case MV_FN_CONSTANT: return batch_vplus_fused<MIXED, MV_FN_CONSTANT>;
    case MV_FN_GAUSSIAN: return batch_vplus_fused<MIXED, MV_FN_GAUSSIAN>;
    case MV_FN_PRODUCT: return batch_vplus_fused<MIXED, MV_FN_PRODUCT>;
    case MV_FN_GAUSSMIX: return batch_vplus_fused<MIXED, MV_FN_GAUSSMIX>;
    case MV_FN_BWPOWER: return batch_vplus_fused<MIXED, MV_FN_BWPOWER>;
    default: return nullptr;
  }
}

}  // namespace

}  // namespace mv

using namespace mv;

extern "C" int mv_batch_vplus_fused(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int32_t n_tiles,
                                    int32_t precision, int32_t family, int32_t max_dim, const int32_t* draws,
                                    const double* grids, const int32_t* n_ev, const int32_t* offsets,
                                    double* ress, double* ress2, double* rows, int32_t blocks, void* stream) {
  if (n_tiles < 0 || blocks < 1) return fail(MV_ERR_ARG, "mv_batch_vplus_fused: bad n_tiles / blocks");
  if (n_tiles == 0) return MV_OK;
  if (!items || !tiles || !draws || !grids || !n_ev || !offsets || !ress || !ress2 || !rows)
    return fail(MV_ERR_ARG, "mv_batch_vplus_fused: missing buffer");
  if (precision != MV_F64 && precision != MV_MIXED)
    return fail(MV_ERR_ARG, "mv_batch_vplus_fused: unknown precision %d", precision);
  if (max_dim < 1 || max_dim > MV_FUSED_MAX_DIM)
    return fail(MV_ERR_UNSUPPORTED, "mv_batch_vplus_fused: max_dim %d outside 1..%d", max_dim, MV_FUSED_MAX_DIM);
  const bool mixed = precision == MV_MIXED;
  BatchVplusKernel k = mixed ? batch_vplus_kernel<true>(family) : batch_vplus_kernel<false>(family);
  if (!k) return fail(MV_ERR_UNSUPPORTED, "mv_batch_vplus_fused: unknown family %d", family);
  const int nbmax = max_dim * kBins;
  const size_t table = mixed ? (size_t)nbmax + (nbmax + 1) / 2 : 3 * (size_t)nbmax;
  const size_t lds = (table + (size_t)kWaves * nbmax) * sizeof(double);
  hipLaunchKernelGGL(k, dim3(std::min(blocks, n_tiles)), dim3(kBlock), lds, as_stream(stream), items, tiles,
                     n_tiles, draws, grids, n_ev, offsets, ress, ress2, rows, nbmax);
  return check_launch("mv_batch_vplus_fused");
}
