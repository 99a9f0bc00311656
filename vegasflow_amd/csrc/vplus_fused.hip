// vplus_fused.hip - the fused VEGAS+ producer (mv_vplus_fused): the kernel
// over one integral's events and its launchers, compiled per (integrand
// family, train, precision, n_dim, injected draws) for n_dim
// 1..MV_FUSED_MAX_DIM, so that the event body (vplus_fused.hpp) is
// straight-line code.
#include "vplus_fused.hpp"

namespace mv {

template <int FAM, bool TRAIN, bool MIXED, int D, bool INJ>
__global__ __launch_bounds__((vplus_block<TRAIN, D>())) void vplus_fused(StepK s, StrataK st, mv_integrand fn,
                                                      const double* __restrict__ grid,
                                                      double* __restrict__ ress,
                                                      double* __restrict__ ress2, RowsK R) {
  constexpr bool PAIRS = vplus_pairs<TRAIN, D>();
  constexpr int BS = vplus_block<TRAIN, D>();
  constexpr int W = BS / 64;
  const uint32_t draw = draw_of(s);
  using T = typename std::conditional<MIXED, float, double>::type;
  using T2 = typename std::conditional<MIXED, float2, double2>::type;
  extern __shared__ double smem[];
  constexpr int nb = D * kBins;
  T2* tab = reinterpret_cast<T2*>(smem);
  T* tw = reinterpret_cast<T*>(smem + (MIXED ? nb : 2 * nb));
  double* hist = smem + (MIXED ? nb + (nb + 1) / 2 : 3 * nb);
  const bool limits = s.xmin != nullptr;
  for (int i = threadIdx.x; i < nb; i += BS) {
    const int j = i / kBins, b = i - j * kBins;
    const double lo = grid[j * (kBins + 1) + b];
    const double dx = grid[j * (kBins + 1) + b + 1] - lo;
    double P = lo, Q = dx;
    if (limits) {
      P = s.xmin[j] + lo * s.xdelta[j];
      Q = dx * s.xdelta[j];
    }
    Integrand<T, FAM>::fold(fn, j, P, Q);
    T2 e;
    e.x = (T)P;
    e.y = (T)Q;
    tab[i] = e;
    tw[i] = (T)(dx * kFBins);
  }
  constexpr int hist_len = PAIRS ? pair_cells(D) : (TRAIN ? kWaves * nb : 0);
  for (int i = threadIdx.x; i < hist_len; i += BS) hist[i] = 0.0;
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* wh = PAIRS ? hist : hist + wave * nb;

  Integrand<T, FAM> F(fn);
  const double wscale = limits ? s.xjac * s.xdeltajac : s.xjac;
  const double fstrat = (double)st.n_strat;
  const float rn = 1.0f / (float)st.n_strat;
  const int64_t total = st.off[st.n_cubes];
  // this wave's run: trips gw * trips / nwaves up to (gw + 1) * trips / nwaves
  // (exclusive) of 64 events each, the trips split evenly over the waves
  const int64_t trips = (total + 63) >> 6;
  const int64_t nwaves = (int64_t)gridDim.x * W;
  const int64_t gw = (int64_t)blockIdx.x * W + wave;
  const int64_t first = (gw * trips / nwaves) * 64;
  const int64_t end = min(((gw + 1) * trips / nwaves) * 64, total);
  double acc = 0.0, acc2 = 0.0;
  vplus_run<FAM, TRAIN, MIXED, D, INJ>(F, s, draw, st, first, end, fstrat, 1.0 / fstrat, rn, tab, tw, wscale,
                                       wh, ress, ress2, acc, acc2);
  if (PAIRS)
    finish_pairs<BS, false>(acc, acc2, hist, D, group_row(R));
  else
    finish_block<TRAIN, false>(acc, acc2, hist, nb, R);
}

using VplusKernel = void (*)(StepK, StrataK, mv_integrand, const double*, double*, double*, RowsK);

template <int FAM, bool TRAIN, bool MIXED, bool INJ, int... Ds>
constexpr std::array<VplusKernel, sizeof...(Ds)> vplus_kernels_of(std::integer_sequence<int, Ds...>) {
  return {&vplus_fused<FAM, TRAIN, MIXED, Ds + 1, INJ>...};
}

template <int FAM, bool TRAIN, bool MIXED, bool INJ>
int launch_vplus_fam(const StepK& k, const StrataK& st, const mv_integrand& fn, const double* grid,
                     double* ress, double* ress2, const RowsK& R, hipStream_t s) {
  static constexpr auto table =
      vplus_kernels_of<FAM, TRAIN, MIXED, INJ>(std::make_integer_sequence<int, MV_FUSED_MAX_DIM>{});
  if (k.n_dim < 1 || k.n_dim > MV_FUSED_MAX_DIM)
    return fail(MV_ERR_UNSUPPORTED, "fused VEGAS+ kernel: n_dim %d outside 1..%d", k.n_dim,
                MV_FUSED_MAX_DIM);
  const int nb = k.n_dim * kBins;
  const bool pairs = TRAIN && k.n_dim >= kPairMinDim && k.n_dim <= kPairMaxDim;
  const size_t hist = pairs ? (size_t)pair_cells(k.n_dim) : (TRAIN ? (size_t)kWaves * nb : 0);
  const size_t lds = ((MIXED ? nb + (nb + 1) / 2 : 3 * nb) + hist) * sizeof(double);
  if (pairs)
    return launch_producer<kPairBlock>(table[k.n_dim - 1], k.n_events, lds, s, k, st, fn, grid, ress, ress2, R);
  return launch_producer(table[k.n_dim - 1], k.n_events, lds, s, k, st, fn, grid, ress, ress2, R);
}

template <int FAM, bool MIXED>
int launch_vplus_family(bool train, const StepK& k, const StrataK& st, const mv_integrand& fn,
                        const double* grid, double* ress, double* ress2, const RowsK& R, hipStream_t s) {
  const bool inj = k.inject != nullptr;
  if (train)
    return inj ? launch_vplus_fam<FAM, true, MIXED, true>(k, st, fn, grid, ress, ress2, R, s)
               : launch_vplus_fam<FAM, true, MIXED, false>(k, st, fn, grid, ress, ress2, R, s);
  return inj ? launch_vplus_fam<FAM, false, MIXED, true>(k, st, fn, grid, ress, ress2, R, s)
             : launch_vplus_fam<FAM, false, MIXED, false>(k, st, fn, grid, ress, ress2, R, s);
}

// The launcher of one precision: the kernel for (family, train, n_dim,
// injected draws). Returns the workgroups launched or a negative error.
template <bool MIXED>
int launch_vplus_precision(int family, bool train, const StepK& k, const StrataK& st,
                           const mv_integrand& fn, const double* grid, double* ress, double* ress2,
                           const RowsK& R, hipStream_t s) {
#define MV_VPF(F) return launch_vplus_family<F, MIXED>(train, k, st, fn, grid, ress, ress2, R, s)
  switch (family) {
    case MV_FN_CONSTANT: MV_VPF(MV_FN_CONSTANT);
    case MV_FN_GAUSSIAN: MV_VPF(MV_FN_GAUSSIAN);
    case MV_FN_PRODUCT: MV_VPF(MV_FN_PRODUCT);
    case MV_FN_GAUSSMIX: MV_VPF(MV_FN_GAUSSMIX);
    case MV_FN_BWPOWER: MV_VPF(MV_FN_BWPOWER);
    default: return fail(MV_ERR_UNSUPPORTED, "fused VEGAS+ kernel: unknown family %d", family);
  }
#undef MV_VPF
}

int launch_vplus_fused_mixed(int family, bool train, const StepK& k, const StrataK& st,
                             const mv_integrand& fn, const double* grid, double* ress, double* ress2,
                             const RowsK& R, hipStream_t s) {
  return launch_vplus_precision<true>(family, train, k, st, fn, grid, ress, ress2, R, s);
}

int launch_vplus_fused_f64(int family, bool train, const StepK& k, const StrataK& st,
                           const mv_integrand& fn, const double* grid, double* ress, double* ress2,
                           const RowsK& R, hipStream_t s) {
  return launch_vplus_precision<false>(family, train, k, st, fn, grid, ress, ress2, R, s);
}

}  // namespace mv
