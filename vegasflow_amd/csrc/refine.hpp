// refine.hpp - grid refinement and the finalize building blocks shared by
// the VegasFlow and VegasFlowPlus kernels.
#pragma once

#include "device.hpp"

namespace mv {

// refine_grid_per_dimension (vflow.py:129-211) on one wave.
// Lanes 0..49 own a bin, lanes 0..50 an edge; the 49-cut rebin loop runs on
// wave-uniform values with readlane broadcasts, in the reference's order.
// One addition: the inner loop stops at the last bin (n <= 49); the
// reference has no guard there.
__device__ inline void refine_wave(const double* h /* LDS, 50 */, double* __restrict__ div /* 51 */) {
  const int lane = threadIdx.x & 63;
  const double edge = lane <= kBins ? div[lane] : 0.0;
  double s = 0.0;
  if (lane < kBins) {
    const double r = h[lane];
    const double rp = lane + 1 < kBins ? h[lane + 1] : 0.0;
    const double rm = lane > 0 ? h[lane - 1] : 0.0;
    const double tmp = (r + rp) + rm;                          // vflow.py:158
    const double m = (lane == 0 || lane == kBins - 1) ? 2.0 : 3.0;
    s = fmax(tmp / m, 1e-30);                                  // vflow.py:159
  }
  const double sum_t = bcast_lane(wave_sum(s), 0);            // vflow.py:162
  double wei = 0.0;
  if (lane < kBins) {
    const double aux = (1.0 - s / sum_t) / (log(sum_t) - log(s));  // vflow.py:163-164
    wei = pow(aux, 1.5);                                           // vflow.py:165 (ALPHA)
  }
  const double ave = bcast_lane(wave_sum(wei), 0) / (double)kBins;  // vflow.py:166

  double mine = lane == 0 ? 0.0 : 1.0;  // edges 0 and 50 (vflow.py:195, 208)
  double bw = 0.0, cur = 0.0, prev = 0.0;
  int n = -1;
  for (int k = 1; k < kBins; ++k) {  // vflow.py:201-207
    while (bw < ave && n < kBins - 1) {
      ++n;
      bw += bcast_lane(wei, n);
      prev = cur;
      cur = bcast_lane(edge, n + 1);
    }
    bw -= ave;
    // n < 0 only if the weights are not finite (NaN integrand values)
    const double delta = (cur - prev) * bw / bcast_lane(wei, n < 0 ? 0 : n);
    if (lane == k) mine = cur - delta;
  }
  if (lane <= kBins) div[lane] = mine;
}

struct FzK {
  int n_dim, n_out, train, refine, raw;
  double n_events;
};

// Sum column c of the group rows in order and zero it for the next producer.
__device__ __forceinline__ double take_column(double* __restrict__ P, int64_t stride, int nrows,
                                              int64_t c) {
  double a = P[c];
  P[c] = 0.0;
  for (int r = 1; r < nrows; ++r) {
    a += P[(int64_t)r * stride + c];
    P[(int64_t)r * stride + c] = 0.0;
  }
  return a;
}

// Sums and sigma from the group rows (vflow.py:432-442): thread v < 2 n_out
// takes column v.
__device__ inline void finalize_sums_block(const FzK& z, double* __restrict__ P, int64_t stride,
                                           int nrows, double* __restrict__ out) {
  __shared__ double red[2 * MV_MAX_OUT];
  const int t = threadIdx.x;
  if (t < 2 * z.n_out) red[t] = take_column(P, stride, nrows, t);
  __syncthreads();
  if (t < z.n_out) {
    const double N = z.n_events;
    const double res = red[t];
    const double res2 = red[z.n_out + t];
    out[t] = res;
    if (z.raw) {
      out[z.n_out + t] = res2;
    } else {
      const double err = (N * res2 - res * res) / (N - 1.0);  // vflow.py:437
      out[z.n_out + t] = sqrt(fmax(err, 0.0));                // vflow.py:438
    }
  }
}

// Histogram of dimension j from the group rows (fixed order, columns zeroed
// afterwards), exported to hist_out and/or used to refine grid row j
// (vflow.py:349-362).
__device__ inline void hist_refine_block(const FzK& z, double* __restrict__ P, int64_t stride,
                                         int nrows, int j, double* __restrict__ grid,
                                         double* __restrict__ hist_out) {
  __shared__ double hb[kBins];
  const int t = threadIdx.x;
  if (t < kBins) {
    const double a = take_column(P, stride, nrows, 2 * z.n_out + j * kBins + t);
    hb[t] = a;
    if (hist_out) hist_out[j * kBins + t] = a;
  }
  __syncthreads();
  if (z.refine && t < 64) refine_wave(hb, grid + j * (kBins + 1));
}

}  // namespace mv
