// fused_mixed.hip - the mixed-precision fused VegasFlow producer with
// per-wave histograms (frozen grids at every n_dim; training where the pair
// kernel of pairs.hip does not apply), compiled per (integrand family,
// train, n_dim) so that the event body is straight-line code.
//
// Per event: one Philox word per dimension (mixed_mirrored), FP32 map and
// integrand from the mirrored (P + 2Q, -Q) table in LDS, FP64 sums; when training,
// (w f)^2 into the wave's private LDS histogram. Eight waves per SIMD
// (amdgpu_waves_per_eu): more resident waves overlap the histogram atomics
// with the Philox VALU work.
#include <array>
#include <utility>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "rows.hpp"

namespace mv {

template <int FAM, bool TRAIN, int D, bool INJ>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void vegas_fused_mixed(
    StepK s, mv_integrand fn, const double* __restrict__ grid, RowsK R) {
  const uint32_t draw = draw_of(s);
  extern __shared__ double smem[];
  constexpr int nb = D * kBins;
  float2* tab = reinterpret_cast<float2*>(smem);  // nb float2
  double* hist = smem + nb;
  build_table<float, FAM>(fn, grid, D, s.xmin, s.xdelta, tab, nullptr);
  if (TRAIN)
    for (int i = threadIdx.x; i < kWaves * nb; i += kBlock) hist[i] = 0.0;
  const JacGroups cs = jacobian_groups<FAM>(fn, D, s.xdelta);
  __syncthreads();
  double* wh = hist + (threadIdx.x >> 6) * nb;
  Integrand<float, FAM> F(fn);
  const uint32_t* inj = static_cast<const uint32_t*>(s.inject);
  const double wscale = s.xmin ? s.xjac * s.xdeltajac : s.xjac;  // monte_carlo.py:270-274
  double acc = 0.0, acc2 = 0.0;
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t i = spread_first_event(); i < s.n_events; i += step) {
    const double tmp = event_mixed<FAM, TRAIN, false, INJ ? 1 : 0>(
        F, s.ev_lo + (uint32_t)i, s.ev_hi, draw, s.key, (D + 3) / 4, D, tab, cs, wh,
        INJ ? inj + i * D : nullptr, wscale);
    acc += tmp;
    acc2 += tmp * tmp;
  }
  finish_block<TRAIN>(acc, acc2, hist, nb, R);
}

namespace {

using Kernel = void (*)(StepK, mv_integrand, const double*, RowsK);

template <int FAM, bool TRAIN, bool INJ, int... Ds>
constexpr std::array<Kernel, sizeof...(Ds)> kernels_of(std::integer_sequence<int, Ds...>) {
  return {&vegas_fused_mixed<FAM, TRAIN, Ds + 1, INJ>...};
}

template <int FAM, bool TRAIN>
int launch_fam(const StepK& k, const mv_integrand& fn, const double* grid, const RowsK& R,
               hipStream_t st) {
  // draws from Philox, or injected (s.inject, parity tests)
  static constexpr auto table =
      kernels_of<FAM, TRAIN, false>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{});
  static constexpr auto table_inj =
      kernels_of<FAM, TRAIN, true>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{});
  if (k.n_dim < 1 || k.n_dim > MV_FUSED_MAX_DIM)
    return fail(MV_ERR_UNSUPPORTED, "fused mixed kernel: n_dim %d outside 1..%d", k.n_dim,
                MV_FUSED_MAX_DIM);
  const int nb = k.n_dim * kBins;
  const size_t lds = (size_t)(nb + (TRAIN ? kWaves * nb : 0)) * sizeof(double);
  const Kernel kern = (k.inject ? table_inj : table)[k.n_dim - 1];
  return launch_producer(kern, k.n_events, lds, st, k, fn, grid, R);
}

}  // namespace

int launch_vegas_fused_mixed(int family, bool train, const StepK& k, const mv_integrand& fn,
                             const double* grid, const RowsK& R, hipStream_t st) {
#define MV_MIX_FAM(F) \
  return train ? launch_fam<F, true>(k, fn, grid, R, st) : launch_fam<F, false>(k, fn, grid, R, st)
  switch (family) {
    case MV_FN_CONSTANT: MV_MIX_FAM(MV_FN_CONSTANT);
    case MV_FN_GAUSSIAN: MV_MIX_FAM(MV_FN_GAUSSIAN);
    case MV_FN_PRODUCT: MV_MIX_FAM(MV_FN_PRODUCT);
    case MV_FN_GAUSSMIX: MV_MIX_FAM(MV_FN_GAUSSMIX);
    case MV_FN_BWPOWER: MV_MIX_FAM(MV_FN_BWPOWER);
    default: return fail(MV_ERR_UNSUPPORTED, "fused mixed kernel: unknown family %d", family);
  }
#undef MV_MIX_FAM
}

}  // namespace mv
