// plain_fused.hip - the fused PlainFlow producer (plain.py:18-35,
// monte_carlo.py:249-275): uniform points from Philox, the built-in
// integrand, w f sums. Compiled per (integrand family, precision, n_dim,
// injected draws) for n_dim 1..MV_FUSED_MAX_DIM, where the event body is
// straight-line code; a runtime-n_dim instance takes larger n_dim (families
// that allow it).
#include <array>
#include <type_traits>
#include <utility>

#include "abi_util.hpp"
#include "device.hpp"
#include "rows.hpp"

namespace mv {

// The events of this thread (plain.py:18-35): uniform points, with the
// integration limits (LIM, monte_carlo.py:271-274) or without, the
// integrand, and the sums of w f and (w f)^2. INJ: injected draws
// (s.inject, parity tests) instead of Philox. Both are compile-time choices,
// so that the event body has no test per dimension.
template <int FAM, bool MIXED, int D, bool INJ, bool LIM, typename T>
__device__ __forceinline__ void plain_events(const StepK& s, uint32_t draw, Integrand<T, FAM>& F, double wt,
                                             double& acc, double& acc2) {
  const int d = D ? D : s.n_dim;
  const int groups = D ? (D + 3) / 4 : s.groups;
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t i = spread_first_event(); i < s.n_events; i += step) {
    const uint32_t c0 = s.ev_lo + (uint32_t)i, c1 = s.ev_hi;  // launch-uniform high word
    F.begin();
#pragma unroll
    for (int g = 0; g < groups; ++g) {
      uint4 h = make_uint4(0, 0, 0, 0), l = make_uint4(0, 0, 0, 0);
      if (!INJ) {
        h = philox(c0, c1, draw, (uint32_t)g, s.key);
        if (!MIXED) l = philox(c0, c1, draw, (uint32_t)(groups + g), s.key);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = 4 * g + q;
        if (j < d) {
          T x;
          if (MIXED) {
            const uint32_t hw = INJ ? static_cast<const uint32_t*>(s.inject)[i * d + j] : word(h, q);
            float xf = fmaf((float)(hw >> 9), 0x1p-23f, 0x1p-24f);
            if (LIM) xf = fmaf(xf, (float)s.xdelta[j], (float)s.xmin[j]);
            x = (T)xf;
          } else {
            const double u = INJ ? static_cast<const double*>(s.inject)[i * d + j]
                                 : draw_f64(word(h, q), word(l, q));
            x = (T)(LIM ? s.xmin[j] + u * s.xdelta[j] : u);
          }
          F.add(j, x);
        }
      }
    }
    const double tmp = (double)F.value() * wt;  // plain.py:27
    acc += tmp;
    acc2 += tmp * tmp;
  }
}

template <int FAM, bool MIXED, int D, bool INJ>
__global__ __launch_bounds__(kBlock) void plain_fused(StepK s, mv_integrand fn, RowsK R) {
  const uint32_t draw = draw_of(s);
  using T = typename std::conditional<MIXED, float, double>::type;
  Integrand<T, FAM> F(fn);
  const bool limits = s.xmin != nullptr;
  const double wt = limits ? s.xjac * s.xdeltajac : s.xjac;  // monte_carlo.py:270-274
  double acc = 0.0, acc2 = 0.0;
  if (limits)
    plain_events<FAM, MIXED, D, INJ, true>(s, draw, F, wt, acc, acc2);
  else
    plain_events<FAM, MIXED, D, INJ, false>(s, draw, F, wt, acc, acc2);
  finish_block<false>(acc, acc2, nullptr, 0, R);
}

namespace {

using Kernel = void (*)(StepK, mv_integrand, RowsK);

template <int FAM, bool MIXED, bool INJ, int... Ds>
constexpr std::array<Kernel, sizeof...(Ds)> kernels_of(std::integer_sequence<int, Ds...>) {
  return {&plain_fused<FAM, MIXED, Ds + 1, INJ>...};
}

template <int FAM, bool MIXED>
int launch_fam(const StepK& k, const mv_integrand& fn, const RowsK& R, hipStream_t st) {
  static constexpr auto table =
      kernels_of<FAM, MIXED, false>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{});
  static constexpr auto table_inj =
      kernels_of<FAM, MIXED, true>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{});
  const bool inj = k.inject != nullptr;
  const Kernel kern = k.n_dim >= 1 && k.n_dim <= MV_FUSED_MAX_DIM
                          ? (inj ? table_inj : table)[k.n_dim - 1]
                          : (inj ? &plain_fused<FAM, MIXED, 0, true> : &plain_fused<FAM, MIXED, 0, false>);
  return launch_producer(kern, k.n_events, 0, st, k, fn, R);
}

}  // namespace

int launch_plain_fused_dims(int family, bool mixed, const StepK& k, const mv_integrand& fn,
                            const RowsK& R, hipStream_t st) {
#define MV_PLAIN_FAM(F) \
  return mixed ? launch_fam<F, true>(k, fn, R, st) : launch_fam<F, false>(k, fn, R, st)
  switch (family) {
    case MV_FN_CONSTANT: MV_PLAIN_FAM(MV_FN_CONSTANT);
    case MV_FN_GAUSSIAN: MV_PLAIN_FAM(MV_FN_GAUSSIAN);
    case MV_FN_PRODUCT: MV_PLAIN_FAM(MV_FN_PRODUCT);
    case MV_FN_GAUSSMIX: MV_PLAIN_FAM(MV_FN_GAUSSMIX);
    case MV_FN_BWPOWER: MV_PLAIN_FAM(MV_FN_BWPOWER);
    default: return fail(MV_ERR_UNSUPPORTED, "fused plain kernel: unknown family %d", family);
  }
#undef MV_PLAIN_FAM
}

}  // namespace mv
