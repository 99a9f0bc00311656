// vplus_fused.hpp - the fused VEGAS+ producer (vflowplus.py:37-80 + 187-220,
// monte_carlo.py:249-275), compiled per (integrand family, train, precision,
// n_dim, injected draws) for n_dim 1..MV_FUSED_MAX_DIM, so that the event
// body is straight-line code; vplus_fused_mixed.hip and vplus_fused_f64.hip
// instantiate one precision each.
//
// FP64: xn = ((h + u) * 50) / n_strat with three roundings (vflowplus.py:72),
// the division by n_strat as div_by_strat (vplus.hpp), correctly rounded.
// Mixed: the same index arithmetic in FP64 from u = hi * 2^-32, FP32 map and
// integrand, FP64 sums.
//
// Every wave owns one contiguous run of the iteration's events, taken 64 at
// a time. The run's first cube comes from one binary search over the
// offsets; each later trip starts in the cube where the previous one ended.
// A trip inside one cube or across one cube boundary needs no per-lane
// lookup, and its sums stay in the lanes until the wave leaves the cube (one
// wave reduction and one FP64 global atomic pair per cube and wave); a trip
// across more cubes finds each lane's cube by a lifting over the next 64
// offsets and adds its sums by a segmented scan, one atomic pair per (trip,
// cube) segment.
#pragma once

#include <array>
#include <type_traits>
#include <utility>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "rows.hpp"
#include "vplus.hpp"

namespace mv {

// Training at kPairMinDim <= n_dim <= kPairMaxDim keeps pair histograms
// (fused_event.hpp) in kPairBlock-thread workgroups: the events of a wave
// lie in one or two hypercubes, so per-dimension histograms would take most
// adds of a wave on the few bins a cube spans; the pair tables spread them
// over up to 2500 cells and need one add per pair of dimensions.
template <bool TRAIN, int D>
constexpr bool vplus_pairs() {
  return TRAIN && D >= kPairMinDim && D <= kPairMaxDim;
}
template <bool TRAIN, int D>
constexpr int vplus_block() {
  return vplus_pairs<TRAIN, D>() ? kPairBlock : kBlock;
}

namespace {

// per-cube sums of one trip (vflowplus.py:213-214)
__device__ __forceinline__ void trip_cube_sums(int cube, double v, double v2, double* __restrict__ ress,
                                               double* __restrict__ ress2) {
  const int first = __builtin_amdgcn_readlane(cube, 0);
  const int last = __builtin_amdgcn_readlane(cube, 63);
  if (first == last) {
    v = wave_sum(v);
    v2 = wave_sum(v2);
    if ((threadIdx.x & 63) == 0) {
      unsafeAtomicAdd(&ress[first], v);
      unsafeAtomicAdd(&ress2[first], v2);
    }
  } else {
    segment_add(cube, v, v2, ress, ress2);
  }
}

// One event of cube ``cube`` (nev events): returns tmp = w f / nev and, when
// training, adds tmp^2 to the histograms wh (the workgroup's pair tables if
// PAIRS, else the wave's per-dimension histogram). INJ: the draws come from
// s.inject (event i) instead of Philox.
template <int FAM, bool TRAIN, bool MIXED, int D, bool INJ, bool PAIRS = vplus_pairs<TRAIN, D>(),
          typename T, typename T2>
__device__ __forceinline__ double vplus_event(Integrand<T, FAM>& F, const StepK& s, uint32_t draw, int cube,
                                              int nev, int64_t i, int n_strat, double fstrat, double ystrat,
                                              float rn,
                                              const T2* tab, const T* tw, double wscale, double* wh) {
  constexpr int groups = (D + 3) / 4;
  int h[D];
  cube_digits_d<D>(cube, n_strat, rn, h);
  const uint64_t ge = (uint64_t)(s.event_offset + i);
  const uint32_t c0w = (uint32_t)ge, c1w = (uint32_t)(ge >> 32);
  F.begin();
  T w = (T)1;
  uint32_t pk[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int g = 0; g < groups; ++g) {
    uint4 hw4 = make_uint4(0, 0, 0, 0), lw4 = make_uint4(0, 0, 0, 0);
    if (!INJ) {
      hw4 = philox(c0w, c1w, draw, (uint32_t)g, s.key);
      if (!MIXED) lw4 = philox(c0w, c1w, draw, (uint32_t)(groups + g), s.key);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = 4 * g + q;
      if (j < D) {
        double u;
        if (MIXED) {
          const uint32_t hw = INJ ? static_cast<const uint32_t*>(s.inject)[i * D + j] : word(hw4, q);
          u = (double)hw * 0x1p-32;
        } else {
          u = INJ ? static_cast<const double*>(s.inject)[i * D + j] : draw_f64(word(hw4, q), word(lw4, q));
        }
        const double xn = div_by_strat(((double)h[j] + u) * kFBins, fstrat, ystrat);  // vflowplus.py:72
        const int ind = clamp_bin((int)xn);
        const T frac = (T)(xn - floor(xn));
        const T2 t = tab[j * kBins + ind];
        F.add_folded(j, MIXED ? (T)fmaf((float)t.y, (float)frac, (float)t.x)
                              : (T)fma((double)t.y, (double)frac, (double)t.x));
        w = w * tw[j * kBins + ind];
        if (TRAIN) pk[g] |= (uint32_t)ind << (8 * q);
      }
    }
  }
  // vflowplus.py:77, monte_carlo.py:270-274, vflowplus.py:209
  const double wev = (double)w / (double)nev;
  const double tmp = (wev * wscale) * (double)F.value();
  if (TRAIN && PAIRS)
    pair_adds(wh, pk, D, tmp * tmp);
  else if (TRAIN)
    hist_adds(wh, pk, D, tmp * tmp);
  return tmp;
}

// Per-cube sums that the lanes of a wave hold for hypercube ``cube`` (-1:
// none): one wave reduction and one atomic pair, then cleared.
__device__ __forceinline__ void flush_held_cube(int cube, double& v, double& v2, double* __restrict__ ress,
                                                double* __restrict__ ress2) {
  if (cube < 0) return;  // wave-uniform
  const double a = wave_sum(v), a2 = wave_sum(v2);
  if ((threadIdx.x & 63) == 0) {
    unsafeAtomicAdd(&ress[cube], a);
    unsafeAtomicAdd(&ress2[cube], a2);
  }
  v = v2 = 0.0;
}

// The wave's run of events first .. end - 1, 64 per trip, into acc, acc2
// (per-lane sums) and the per-cube sums ress, ress2 (vflowplus.py:213-214).
// A trip that lies inside one hypercube, or across the boundary of two,
// takes cubes and n_ev from scalar reads and keeps the current cube's sums
// in the lanes until the wave moves on to another cube (one reduction per
// cube and wave); a trip across more cubes finds each lane's cube with
// window_cube and adds its per-cube sums at once (trip_cube_sums).
template <int FAM, bool TRAIN, bool MIXED, int D, bool INJ, bool PAIRS = vplus_pairs<TRAIN, D>(),
          typename T, typename T2>
__device__ __forceinline__ void vplus_run(Integrand<T, FAM>& F, const StepK& s, uint32_t draw,
                                          const StrataK& st, int64_t first, int64_t end, double fstrat,
                                          double ystrat, float rn, const T2* tab, const T* tw,
                                          double wscale, double* wh, double* __restrict__ ress,
                                          double* __restrict__ ress2,
                                          double& acc, double& acc2) {
  const int lane = threadIdx.x & 63;
  // c0: the cube of the trip's first event (wave-uniform); c0_end: its end
  int c0 = first < end ? __builtin_amdgcn_readfirstlane(find_cube(st.off, st.n_cubes, (int)first)) : 0;
  int64_t c0_end = first < end ? (int64_t)st.off[c0 + 1] : 0;
  int held = -1;  // the cube whose sums the lanes hold in hv, hv2
  double hv = 0.0, hv2 = 0.0;
  for (int64_t base = first; base < end; base += 64) {
    while (c0_end <= base) c0_end = (int64_t)st.off[++c0 + 1];  // cubes are never empty
    const int64_t i = base + lane;
    const bool live = i < end;
    // the trip lies in cube c0 (inside), in c0 and c0 + 1 (two: lanes from
    // event c0_end on are in c0 + 1), or across more cubes
    const int64_t last = min(base + 63, end - 1);
    const bool inside = last < c0_end;
    int64_t c1_end = 0;
    if (!inside) c1_end = (int64_t)st.off[c0 + 2];  // in bounds: cube c0 + 1 holds event `last`
    const bool two = !inside && last < c1_end;
    const bool upper = i >= c0_end;  // lane in c0 + 1 (two)
    int cube = c0, nev;
    if (inside) {
      nev = st.n_ev[c0];
    } else if (two) {
      const int n0 = st.n_ev[c0], n1 = st.n_ev[c0 + 1];
      cube = upper ? c0 + 1 : c0;
      nev = upper ? n1 : n0;
    } else {
      cube = window_cube(st.off, st.n_ev, st.n_cubes, c0, (int)(live ? i : end - 1), nev);
    }
    double tmp = 0.0, tmp2 = 0.0;
    if (live) {
      tmp = vplus_event<FAM, TRAIN, MIXED, D, INJ, PAIRS>(F, s, draw, cube, nev, i, st.n_strat, fstrat, ystrat,
                                                          rn, tab, tw, wscale, wh);
      tmp2 = tmp * tmp;
      acc += tmp;
      acc2 += tmp2;
    }
    if (inside || two) {
      if (held != c0) {
        flush_held_cube(held, hv, hv2, ress, ress2);
        held = c0;
      }
      hv += upper && two ? 0.0 : tmp;
      hv2 += upper && two ? 0.0 : tmp2;
      if (two) {  // c0 ends in this trip: its sums go out, c0 + 1 is held next
        flush_held_cube(held, hv, hv2, ress, ress2);
        held = ++c0;
        c0_end = c1_end;
        hv = upper ? tmp : 0.0;
        hv2 = upper ? tmp2 : 0.0;
      }
    } else {
      trip_cube_sums(cube, tmp, tmp2, ress, ress2);
      // the next trip starts in the cube of this trip's last event or later
      c0 = __builtin_amdgcn_readlane(cube, 63);
      c0_end = (int64_t)st.off[c0 + 1];
    }
  }
  flush_held_cube(held, hv, hv2, ress, ress2);
}

}  // namespace

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

}  // namespace mv
