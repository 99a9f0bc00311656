// vplus_fused.hpp - the event body and the per-wave run loop of the fused
// VEGAS+ producers (vflowplus.py:37-80 + 187-220, monte_carlo.py:249-275),
// shared by the single-integral kernel (vplus_fused.hip) and the batched one
// (batch_vplus.hip).
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

}  // namespace mv
