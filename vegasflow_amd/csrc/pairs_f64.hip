// pairs_f64.hip - the FP64 training producer of VegasFlow with pair
// histograms (vflow.py:87-126 + 33-83 + 389-430 in the reference's FP64
// operation order), compiled per (integrand family, n_dim, injected draws).
// Dimensions 2p and 2p + 1 share one 50 x 50 table of FP64 cells in LDS, so
// an event makes one LDS add per pair of dimensions, spread over 2500 cells.
// LDS: the (P, Q) and W tables (3 doubles per bin), then pair_cells(n_dim)
// doubles, about 90 KB at d = 8: one kPairBlock-thread workgroup per CU.
#include <array>
#include <utility>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "rows.hpp"

namespace mv {

template <int FAM, int D, bool INJ>
__global__ __launch_bounds__(kPairBlock) void vegas_fused_pairs_f64(StepK s, mv_integrand fn,
                                                                    const double* __restrict__ grid,
                                                                    RowsK R) {
  const uint32_t draw = draw_of(s);
  extern __shared__ double smem[];
  constexpr int nb = D * kBins;
  double2* tab = reinterpret_cast<double2*>(smem);
  double* tw = smem + 2 * nb;
  double* H = tw + nb;
  build_table<double, FAM>(fn, grid, D, s.xmin, s.xdelta, tab, tw);
  for (int i = threadIdx.x; i < pair_cells(D); i += kPairBlock) H[i] = 0.0;
  __syncthreads();
  Integrand<double, FAM> F(fn);
  const double* inj = static_cast<const double*>(s.inject);
  const double wscale = s.xmin ? s.xjac * s.xdeltajac : s.xjac;  // monte_carlo.py:270-274
  double acc = 0.0, acc2 = 0.0;
  const int64_t step = (int64_t)gridDim.x * kPairBlock;
  for (int64_t i = spread_first_event(); i < s.n_events; i += step) {
    const double tmp = event_f64<FAM, true, true, INJ ? 1 : 0>(
        F, s.ev_lo + (uint32_t)i, s.ev_hi, draw, s.key, (D + 3) / 4, D, tab, tw, H,
        INJ ? inj + i * D : nullptr, wscale);
    acc += tmp;
    acc2 += tmp * tmp;
  }
  finish_pairs<kPairBlock>(acc, acc2, H, D, group_row(R));
}

namespace {

using Kernel = void (*)(StepK, mv_integrand, const double*, RowsK);

template <int FAM, bool INJ, int... Ds>
constexpr std::array<Kernel, sizeof...(Ds)> kernels_of(std::integer_sequence<int, Ds...>) {
  return {&vegas_fused_pairs_f64<FAM, Ds + kPairMinDim, INJ>...};
}

template <int FAM>
int launch_fam(const StepK& k, const mv_integrand& fn, const double* grid, const RowsK& R, size_t lds,
               hipStream_t st) {
  constexpr int n = kPairMaxDim - kPairMinDim + 1;
  static constexpr auto table = kernels_of<FAM, false>(std::make_integer_sequence<int, n>{});
  static constexpr auto table_inj = kernels_of<FAM, true>(std::make_integer_sequence<int, n>{});
  if (k.n_dim < kPairMinDim || k.n_dim > kPairMaxDim)
    return fail(MV_ERR_UNSUPPORTED, "FP64 pair kernel: n_dim %d outside %d..%d", k.n_dim, kPairMinDim,
                kPairMaxDim);
  const Kernel kern = (k.inject ? table_inj : table)[k.n_dim - kPairMinDim];
  return launch_producer<kPairBlock>(kern, k.n_events, lds, st, k, fn, grid, R);
}

}  // namespace

size_t fused_pairs_f64_lds(int d) { return (size_t)(3 * d * kBins + pair_cells(d)) * sizeof(double); }

int launch_vegas_fused_pairs_f64(int family, const StepK& k, const mv_integrand& fn, const double* grid,
                                 const RowsK& R, size_t lds, hipStream_t st) {
  switch (family) {
    case MV_FN_CONSTANT: return launch_fam<MV_FN_CONSTANT>(k, fn, grid, R, lds, st);
    case MV_FN_GAUSSIAN: return launch_fam<MV_FN_GAUSSIAN>(k, fn, grid, R, lds, st);
    case MV_FN_PRODUCT: return launch_fam<MV_FN_PRODUCT>(k, fn, grid, R, lds, st);
    case MV_FN_GAUSSMIX: return launch_fam<MV_FN_GAUSSMIX>(k, fn, grid, R, lds, st);
    case MV_FN_BWPOWER: return launch_fam<MV_FN_BWPOWER>(k, fn, grid, R, lds, st);
    default: return fail(MV_ERR_UNSUPPORTED, "FP64 pair kernel: unknown family %d", family);
  }
}

}  // namespace mv
