// fused_event.hpp - the per-event body of the fused VegasFlow kernels and
// their LDS tables, shared by the single-integral and the batched kernels.
//
// Per (dimension, bin) the LDS table holds the map frac -> y = P + Q frac of
// the integrand coordinate (Integrand::fold, integration limits folded in)
// and, in FP64, the Jacobian factor W = dx * 50 (vflow.py:73, 78). The FP32
// table of the mixed path is indexed by the mirrored bin k = 49 - bin and
// takes 1 + mfrac, mfrac = 1 - frac (mixed_mirrored), so entry k holds
// (P + 2Q, -Q): y = (P + 2Q) - Q (1 + mfrac) = P + Q frac. It has no W
// column: W = Q * c_j with the per-dimension constant c_j = 50 / (xdelta_j *
// fold scale_j) (jacobian_scale), so an event multiplies its slopes Q_j
// and, after each group of four dimensions, the product of that group's c_j
// (jacobian_groups), instead of one more LDS read per dimension.
#pragma once

#include <type_traits>

#include "device.hpp"

namespace mv {

__device__ __forceinline__ int clamp_bin(int b) { return b < 0 ? 0 : (b > kBins - 1 ? kBins - 1 : b); }

template <typename T, int FAM>
__device__ __forceinline__ void build_table(const mv_integrand& fn, const double* __restrict__ grid,
                                            int d, const double* xmin, const double* xdelta,
                                            typename std::conditional<sizeof(T) == 8, double2, float2>::type* tab,
                                            T* tw) {
  const int nb = d * kBins;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const int j = i / kBins, b = i - j * kBins;
    const double lo = grid[j * (kBins + 1) + b];
    const double dx = grid[j * (kBins + 1) + b + 1] - lo;  // vflow.py:73
    double P = lo, Q = dx;
    if (xmin) {  // monte_carlo.py:273
      P = xmin[j] + lo * xdelta[j];
      Q = dx * xdelta[j];
    }
    Integrand<T, FAM>::fold(fn, j, P, Q);
    if constexpr (sizeof(T) == 4) {  // mirrored: entry 49 - b
      tab[i + kBins - 1 - 2 * b].x = (T)(P + 2.0 * Q);
      tab[i + kBins - 1 - 2 * b].y = (T)(-Q);
    } else {
      tab[i].x = (T)P;
      tab[i].y = (T)Q;
    }
    if (tw) tw[i] = (T)(dx * kFBins);  // vflow.py:78
  }
}

// c_j of the mixed path's Jacobian: W = dx * 50 = Q * c_j. A dimension of
// zero width (xmin == xmax, allowed by monte_carlo.py:171) has Q = 0 and
// takes c_j = 0, so its events weigh 0 as in FP64 (w * prod(xdelta) = 0).
template <int FAM>
__device__ __forceinline__ double jacobian_scale(const mv_integrand& fn, int j, const double* xdelta) {
  const double xd = xdelta ? xdelta[j] : 1.0;
  if (xd == 0.0) return 0.0;
  return kFBins / (xd * Integrand<float, FAM>::fold_scale(fn, j));
}

// Products of c_j per group of four dimensions (g[k] for dimensions 4k ..
// 4k + 3), each formed in FP64 and rounded once. An event multiplies its
// slopes Q_j and, after each group, that group's product, so the running
// FP32 product never strays from the Jacobian prod (dx_j * 50) of the
// dimensions done by more than one group's factor. (One product of all the
// c_j overflows FP32 with narrow limits: 500^16 on [0, 0.1]^16, against a
// Jacobian of order 1.) Exact in range while every |log2 c_j| < 30.
struct JacGroups {
  float g[4];
};

template <int FAM>
__device__ __forceinline__ JacGroups jacobian_groups(const mv_integrand& fn, int d, const double* xdelta) {
  JacGroups c;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double p = 1.0;
    for (int j = 4 * k; j < 4 * k + 4 && j < d; ++j) p *= jacobian_scale<FAM>(fn, j, xdelta);
    c.g[k] = (float)p;
  }
  return c;
}

// Histogram adds of one event: tmp2 into bin pk[j] of every dimension j
// (byte j & 3 of pk[j >> 2]). Spreading one wave instruction's adds over
// several dimensions (lane-rotated order) was measured slower on gfx950: the
// extra VALU selects cost more than the LDS conflicts they save.
__device__ __forceinline__ void hist_adds(double* wh, const uint32_t (&pk)[4], int d, double tmp2) {
#pragma unroll
  for (int j = 0; j < MV_FUSED_MAX_DIM; ++j)
    if (j < d) lds_add(&wh[j * kBins + ((pk[j >> 2] >> (8 * (j & 3))) & 0xffu)], tmp2);
}

// Pair histograms: dimensions 2p and 2p + 1 share one 50 x 50 table of
// cells, so an event makes one LDS add per pair instead of one per
// dimension, and the adds of a wave spread over 2500 cells instead of 50,
// which removes most same-address conflicts of the LDS atomics. The
// per-dimension histograms are the row and column sums (pair_marginal_half).
// LDS: ceil(d / 2) * 2500 doubles, so only for d <= kPairMaxDim.
constexpr int kPairBins = kBins * kBins;
constexpr int kPairMinDim = 2;
constexpr int kPairMaxDim = 14;
constexpr int kPairBlock = 1024;  // threads of a pair-histogram workgroup

__host__ __device__ constexpr int pair_cells(int d) { return ((d + 1) >> 1) * kPairBins; }

// cell (b_2p, b_2p+1) of pair p gets tmp2; b_2p+1 = 0 (its byte of pk is
// never set) for the unpaired last dimension of an odd d
__device__ __forceinline__ void pair_adds(double* H, const uint32_t (&pk)[4], int d, double tmp2) {
#pragma unroll
  for (int p = 0; p < kPairMaxDim / 2; ++p) {
    if (2 * p < d) {
      const uint32_t w = pk[p >> 1] >> (16 * (p & 1));
      lds_add(&H[p * kPairBins + (w & 0xffu) * kBins + ((w >> 8) & 0xffu)], tmp2);
    }
  }
}

// Bin b of dimension j = m / 50 (b = m % 50) from the pair histograms is the
// sum of row b of pair j / 2 for an even j, of column b for an odd j. This is
// half h of that sum (h = 0: cells 0..24, h = 1: cells 25..49) in order; the
// bin is half 0 + half 1.
__device__ __forceinline__ double pair_marginal_half(const double* H, int m, int h) {
  const int j = m / kBins, b = m - j * kBins;
  const int stride = (j & 1) ? kBins : 1;
  const double* c = H + (j >> 1) * kPairBins + ((j & 1) ? b : b * kBins) + h * (kBins / 2) * stride;
  double a = 0.0;
#pragma unroll 5
  for (int k = 0; k < kBins / 2; ++k) a += c[k * stride];
  return a;
}

template <bool PAIRS>
__device__ __forceinline__ void event_hist(double* wh, const uint32_t (&pk)[4], int d, double tmp2) {
  if (PAIRS)
    pair_adds(wh, pk, d, tmp2);
  else
    hist_adds(wh, pk, d, tmp2);
}

// One event, FP64 (vflow.py:87-126 + 33-83 + 389-430). Returns tmp = w f
// and, when training, adds tmp^2 to the histogram ``wh`` (the wave's
// per-dimension histograms, or the workgroup's pair histograms if PAIRS).
// INJ: injected draws taken when inj != nullptr (-1, tested per event), never
// (0) or always (1); a compile-time 0 or 1 leaves no test in the event body.
template <int FAM, bool TRAIN, bool PAIRS = false, int INJ = -1>
__device__ __forceinline__ double event_f64(Integrand<double, FAM>& F, uint32_t c0, uint32_t c1,
                                            uint32_t draw, Key key, int groups, int d,
                                            const double2* tab, const double* tw, double* wh,
                                            const double* inj, double wscale) {
  const bool injected = INJ < 0 ? inj != nullptr : INJ > 0;
  F.begin();
  double w = 1.0;
  uint32_t pk[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    if (g < groups) {
      uint4 h = make_uint4(0, 0, 0, 0), l = make_uint4(0, 0, 0, 0);
      if (!injected) {
        h = philox(c0, c1, draw, (uint32_t)g, key);
        l = philox(c0, c1, draw, (uint32_t)(groups + g), key);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = 4 * g + q;
        if (j < d) {
          const double u = injected ? inj[j] : draw_f64(word(h, q), word(l, q));
          const double xn = kFBins * (1.0 - u);  // vflow.py:117
          const int ind = clamp_bin((int)xn);    // vflow.py:67
          const double2 t = tab[j * kBins + ind];
          const double frac = xn - floor(xn);    // vflow.py:75
          F.add_folded(j, fma(t.y, frac, t.x));  // vflow.py:76
          w = w * tw[j * kBins + ind];           // vflow.py:78
          if (TRAIN) pk[g] |= (uint32_t)ind << (8 * q);
        }
      }
    }
  }
  const double tmp = (w * wscale) * F.value();  // vflow.py:416
  if (TRAIN) event_hist<PAIRS>(wh, pk, d, tmp * tmp);  // vflow.py:417, 370-387
  return tmp;
}

// One event, mixed precision: FP32 map and integrand from the high word of
// each draw (mixed_mirrored), FP64 result.
template <int FAM, bool TRAIN, bool PAIRS = false, int INJ = -1>
__device__ __forceinline__ double event_mixed(Integrand<float, FAM>& F, uint32_t c0, uint32_t c1,
                                              uint32_t draw, Key key, int groups, int d,
                                              const float2* tab, const JacGroups& cs, double* wh,
                                              const uint32_t* inj, double wscale) {
  const bool injected = INJ < 0 ? inj != nullptr : INJ > 0;
  F.begin();
  float w = 1.0f;
  uint32_t pk[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    if (g < groups) {
      uint4 h = make_uint4(0, 0, 0, 0);
      if (!injected) h = philox(c0, c1, draw, (uint32_t)g, key);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = 4 * g + q;
        if (j < d) {
          float f1;
          const int k = mixed_mirrored(injected ? inj[j] : word(h, q), f1);
          const float2 t = tab[j * kBins + k];
          F.add_folded(j, fmaf(t.y, f1, t.x));
          w = w * -t.y;
          if (TRAIN) pk[g] |= (uint32_t)(kBins - 1 - k) << (8 * q);
        }
      }
      w = w * cs.g[g];
    }
  }
  // w * f in FP32; the 1/N (and limits) factor in FP64 so that its rounding
  // is not a systematic bias
  const double tmp = (double)(w * F.value()) * wscale;
  if (TRAIN) event_hist<PAIRS>(wh, pk, d, tmp * tmp);
  return tmp;
}

// One dimension of a mixed event: position of draw word hw in dimension j,
// the integrand coordinate into F, the map slope into w; returns the
// mirrored bin 49 - bin.
template <int FAM>
__device__ __forceinline__ uint32_t mixed_dim(Integrand<float, FAM>& F, int j, uint32_t hw,
                                              const float2* tab, float& w) {
  float f1;
  const int k = mixed_mirrored(hw, f1);
  const float2 t = tab[j * kBins + k];
  F.add_folded(j, fmaf(t.y, f1, t.x));
  w = w * -t.y;
  return (uint32_t)k;
}

template <int FAM>
__device__ __forceinline__ void pair_dim(Integrand<float, FAM>& F, int j, uint32_t hw,
                                         const float2* tab, float& w,
                                         uint32_t (&cell)[kPairMaxDim / 2]) {
  const uint32_t b = mixed_dim<FAM>(F, j, hw, tab, w);
  if (j & 1)
    cell[j >> 1] += b * (uint32_t)sizeof(double);
  else
    cell[j >> 1] = b * (uint32_t)(kBins * sizeof(double));
}

// One training event of the pair-histogram mixed kernel (event_mixed with
// PAIRS, restructured for fewer instructions): the byte offset of the event's
// cell in each pair table is built while the dimensions are mapped (no bin
// packing), groups of four dimensions that are all present skip the
// per-dimension checks, and INJ (injected words instead of Philox) is a
// template argument rather than a test per dimension. The pair tables are
// mirrored like the map table: bins (b, b') of a pair add into cell
// (49 - b) * 50 + 49 - b' (finish_pairs with MIRROR). Adds tmp^2 to the pair
// tables at H and returns tmp.
template <int FAM, bool INJ>
__device__ __forceinline__ double event_mixed_pairs(Integrand<float, FAM>& F, uint32_t c0, uint32_t c1,
                                                    uint32_t draw, Key key, int d, const float2* tab,
                                                    const JacGroups& cs, double* H, const uint32_t* inj,
                                                    double wscale) {
  F.begin();
  float w = 1.0f;
  uint32_t cell[kPairMaxDim / 2];
#pragma unroll
  for (int p = 0; p < kPairMaxDim / 2; ++p) cell[p] = 0u;
#pragma unroll
  for (int g = 0; g < (kPairMaxDim + 3) / 4; ++g) {
    if (4 * g < d) {
      const bool full = 4 * g + 4 <= d;
      uint4 h;
      if (INJ) {
        h.x = inj[4 * g];
        h.y = full || 4 * g + 1 < d ? inj[4 * g + 1] : 0u;
        h.z = full || 4 * g + 2 < d ? inj[4 * g + 2] : 0u;
        h.w = full ? inj[4 * g + 3] : 0u;
      } else {
        h = philox(c0, c1, draw, (uint32_t)g, key);
      }
      if (full) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (4 * g + q < kPairMaxDim) pair_dim<FAM>(F, 4 * g + q, word(h, q), tab, w, cell);
      } else {
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (4 * g + q < kPairMaxDim && 4 * g + q < d) pair_dim<FAM>(F, 4 * g + q, word(h, q), tab, w, cell);
      }
      w = w * cs.g[g];
    }
  }
  const double tmp = (double)(w * F.value()) * wscale;
  const double tmp2 = tmp * tmp;
  char* Hb = reinterpret_cast<char*>(H);
#pragma unroll
  for (int p = 0; p < kPairMaxDim / 2; ++p)
    if (2 * p < d) lds_add(reinterpret_cast<double*>(Hb + p * kPairBins * sizeof(double) + cell[p]), tmp2);
  return tmp;
}

}  // namespace mv
