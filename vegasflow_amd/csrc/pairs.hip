// pairs.hip - the mixed-precision training producer of VegasFlow with pair
// histograms (vflow.py:87-126 + 33-83 + 389-430 in FP32 map and integrand,
// FP64 sums and histograms), compiled once per (integrand family, n_dim).
//
// With n_dim a compile-time constant the event body is straight-line code:
// the per-group and per-dimension tests of a runtime n_dim, and the branches
// around them, disappear (measured 68.5 -> 56.3 us at d = 8, 1e7 events).
// Histograms are pair tables (fused_event.hpp): dimensions 2p and 2p + 1
// share one 50 x 50 table, one LDS add per pair of dimensions. LDS (static):
// the map table (n_dim * 50 float2), then pair_cells(n_dim) doubles, about
// 83 KB at d = 8, so one kPairBlock-thread workgroup per CU.
#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "rows.hpp"

namespace mv {

template <int FAM, int D>
__global__ __launch_bounds__(kPairBlock) void vegas_fused_pairs(StepK s, mv_integrand fn,
                                                                const double* __restrict__ grid,
                                                                RowsK R) {
  const uint32_t draw = draw_of(s);
  // Static LDS (its size is a compile-time function of D): its addresses are
  // constants that the LDS instructions take as offsets, where the base of
  // dynamic LDS would cost an add per table address.
  constexpr int nb = D * kBins;
  __shared__ __align__(16) double smem[nb + pair_cells(D)];
  float2* tab = reinterpret_cast<float2*>(smem);
  double* H = smem + nb;  // 16-byte aligned: nb * 8 is a multiple of 400
  build_table<float, FAM>(fn, grid, D, s.xmin, s.xdelta, tab, nullptr);
  constexpr int half = pair_cells(D) / 2;
  for (int i = threadIdx.x; i < half; i += kPairBlock)
    reinterpret_cast<double2*>(H)[i] = make_double2(0.0, 0.0);
  const JacGroups cs = jacobian_groups<FAM>(fn, D, s.xdelta);
  __syncthreads();
  Integrand<float, FAM> F(fn);
  const uint32_t* inj = static_cast<const uint32_t*>(s.inject);
  const double wscale = s.xmin ? s.xjac * s.xdeltajac : s.xjac;  // monte_carlo.py:270-274
  double acc = 0.0, acc2 = 0.0;
  const int64_t step = (int64_t)gridDim.x * kPairBlock;
  const int64_t i0 = spread_first_event();
  if (inj) {
    for (int64_t i = i0; i < s.n_events; i += step) {
      const double tmp = event_mixed_pairs<FAM, true>(F, 0u, 0u, 0u, s.key, D, tab, cs, H,
                                                      inj + i * D, wscale);
      acc += tmp;
      acc2 += tmp * tmp;
    }
  } else {
    // The wave's trips in which all 64 lanes have an event run without a
    // per-lane test, counted on the scalar unit; at most one trip with a
    // per-lane test follows. The Philox counter word advances by the step
    // (mod 2^32) as the 64-bit event index would.
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t w0 = ((int64_t)wave * gridDim.x + blockIdx.x) * 64;  // the wave's first event (i0 - lane)
    const int64_t full = s.n_events - 64 - w0 >= 0 ? (s.n_events - 64 - w0) / step + 1 : 0;
    uint32_t c = s.ev_lo + (uint32_t)i0;
    for (int64_t t = 0; t < full; ++t) {
      const double tmp = event_mixed_pairs<FAM, false>(F, c, s.ev_hi, draw, s.key, D, tab, cs, H, nullptr,
                                                       wscale);
      acc += tmp;
      acc2 += tmp * tmp;
      c += (uint32_t)step;
    }
    if (i0 + full * step < s.n_events) {
      const double tmp = event_mixed_pairs<FAM, false>(F, c, s.ev_hi, draw, s.key, D, tab, cs, H, nullptr,
                                                       wscale);
      acc += tmp;
      acc2 += tmp * tmp;
    }
  }
  finish_pairs<kPairBlock, true, true>(acc, acc2, H, D, group_row(R));
}

namespace {

template <int FAM, int D>
int launch_d(const StepK& k, const mv_integrand& fn, const double* grid, const RowsK& R, hipStream_t st) {
  return launch_producer<kPairBlock>(vegas_fused_pairs<FAM, D>, k.n_events, 0, st, k, fn, grid, R);
}

template <int FAM>
int launch_fam(const StepK& k, const mv_integrand& fn, const double* grid, const RowsK& R, hipStream_t st) {
  switch (k.n_dim) {
    case 2: return launch_d<FAM, 2>(k, fn, grid, R, st);
    case 3: return launch_d<FAM, 3>(k, fn, grid, R, st);
    case 4: return launch_d<FAM, 4>(k, fn, grid, R, st);
    case 5: return launch_d<FAM, 5>(k, fn, grid, R, st);
    case 6: return launch_d<FAM, 6>(k, fn, grid, R, st);
    case 7: return launch_d<FAM, 7>(k, fn, grid, R, st);
    case 8: return launch_d<FAM, 8>(k, fn, grid, R, st);
    case 9: return launch_d<FAM, 9>(k, fn, grid, R, st);
    case 10: return launch_d<FAM, 10>(k, fn, grid, R, st);
    case 11: return launch_d<FAM, 11>(k, fn, grid, R, st);
    case 12: return launch_d<FAM, 12>(k, fn, grid, R, st);
    case 13: return launch_d<FAM, 13>(k, fn, grid, R, st);
    case 14: return launch_d<FAM, 14>(k, fn, grid, R, st);
    default: return fail(MV_ERR_UNSUPPORTED, "pair kernel: n_dim %d outside %d..%d", k.n_dim,
                         kPairMinDim, kPairMaxDim);
  }
}
static_assert(kPairMinDim == 2 && kPairMaxDim == 14, "launch_fam covers n_dim 2..14");

}  // namespace

int launch_vegas_fused_pairs(int family, const StepK& k, const mv_integrand& fn, const double* grid,
                             const RowsK& R, hipStream_t st) {
  switch (family) {
    case MV_FN_CONSTANT: return launch_fam<MV_FN_CONSTANT>(k, fn, grid, R, st);
    case MV_FN_GAUSSIAN: return launch_fam<MV_FN_GAUSSIAN>(k, fn, grid, R, st);
    case MV_FN_PRODUCT: return launch_fam<MV_FN_PRODUCT>(k, fn, grid, R, st);
    case MV_FN_GAUSSMIX: return launch_fam<MV_FN_GAUSSMIX>(k, fn, grid, R, st);
    case MV_FN_BWPOWER: return launch_fam<MV_FN_BWPOWER>(k, fn, grid, R, st);
    default: return fail(MV_ERR_UNSUPPORTED, "pair kernel: unknown family %d", family);
  }
}

}  // namespace mv
