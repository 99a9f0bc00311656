// refine.hpp - grid refinement and the finalize building blocks shared by
// the VegasFlow and VegasFlowPlus kernels.
#pragma once

#include "device.hpp"

namespace mv {

// refine_grid_per_dimension (vflow.py:129-211) on one wave.
// Lanes 0..49 own a bin, lanes 0..50 an edge. The reference's 49 serial cuts
// (vflow.py:194-207) are computed independently, lane k making cut k: the
// serial loop stops in the first bin n whose running weight reaches k * ave,
// with bw = cumsum(wei)[n] - k * ave left over, so lane k finds n, the count
// of prefix sums below k * ave (a binary search over the non-decreasing
// prefix sums, held one per lane), and interpolates as the reference does,
// prev = 0.0 (not div[0]) for n = 0 included. Only the rounding of bw differs
// from the serial accumulation (a few ulp of sum(wei)). n is capped at the
// last bin; the reference has no guard there. The wave runs alone on its CU,
// so its instruction latency is the kernel's time: the weight
// (1 - r)/(log(sum) - log(s))^1.5 is taken as aux * sqrt(aux) with one log of
// sum / s, equal to the reference's pow and two logs to a few ulp.
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
    const double aux = (1.0 - s / sum_t) / log(sum_t / s);  // vflow.py:163-164
    wei = aux * sqrt(aux);                                   // vflow.py:165, ALPHA = 1.5
  }
  const double ave = bcast_lane(wave_sum(wei), 0) / (double)kBins;  // vflow.py:166

  // inclusive prefix sums of the weights (lanes >= 50 hold 0)
  const double c = wave_inclusive_scan(wei);
  const double target = (double)lane * ave;  // cut k = lane
  // n = #{b < 50 : c_b < target}: c is non-decreasing over lanes 0..63 (lanes
  // >= 50 repeat the total), so binary lifting over lanes; every lane runs
  // every shuffle
  int n = 0;
#pragma unroll
  for (int stp = 32; stp >= 1; stp >>= 1) {
    const double v = __shfl(c, n + stp - 1, 64);  // c_(n + stp - 1), index <= 62
    if (n + stp <= kBins && v < target) n += stp;
  }
  if (n > kBins - 1) n = kBins - 1;
  // every shuffle runs on the whole wave (a lane-gather from an inactive
  // lane reads 0); the n == 0 case is selected afterwards
  const double cn = __shfl(c, n, 64);
  const double wn = __shfl(wei, n, 64);
  const double en = __shfl(edge, n, 64);
  const double cur = __shfl(edge, n + 1, 64);      // vflow.py:193
  const double bw = cn - target;                   // vflow.py:205
  const double prev = n == 0 ? 0.0 : en;           // vflow.py:197-200
  const double delta = (cur - prev) * bw / wn;     // vflow.py:206
  double mine = lane == 0 ? 0.0 : 1.0;  // edges 0 and 50 (vflow.py:195, 208)
  if (lane >= 1 && lane < kBins) mine = cur - delta;             // vflow.py:207
  if (lane <= kBins) div[lane] = mine;
}

// Iteration counter advanced by the finalize kernels (mv_finalize): one
// thread stores {res, sigma, events} in history slot draw[0] - draw[1] when
// that is below len, then moves draw[0] to the next draw id.
struct CounterK {
  uint32_t* draw;
  double* history;
  int len;
};

__device__ __forceinline__ void advance_counter(const CounterK& c, double res, double sig,
                                                double events) {
  if (!c.draw) return;
  const uint32_t k = c.draw[0] - c.draw[1];
  if (c.history && k < (uint32_t)c.len) {
    double* h = c.history + 3 * (int64_t)k;
    h[0] = res;
    h[1] = sig;
    h[2] = events;
  }
  c.draw[0] = c.draw[0] + 1u;
}

struct FzK {
  int n_dim, n_out, train, refine, raw;
  double n_events;
  CounterK counter;  // zero: no counter
  uint32_t* status;  // nullable: MV_STATUS_* bits
};

// Non-finite iteration sums (an integrand that returned NaN or Inf at some
// event: both survive the additions and squares of the sums) set
// MV_STATUS_NONFINITE in the caller's status word.
__device__ __forceinline__ void flag_nonfinite(uint32_t* status, double a, double b) {
  if (status && !(isfinite(a) && isfinite(b))) atomicOr(status, MV_STATUS_NONFINITE);
}

// sqrt(max(err, 0)) with the reference's NaN behaviour (vflow.py:438:
// tf.maximum passes a NaN through; fmax would turn it into 0).
__device__ __forceinline__ double sqrt_clamped(double err) { return sqrt(err < 0.0 ? 0.0 : err); }

// The 50 bins of a histogram are all finite (whole wave votes; lanes >= 50
// vote finite). Refinement is skipped otherwise, so that an iteration with
// non-finite sums leaves the grid as it was.
__device__ __forceinline__ bool wave_hist_finite(const double* hb) {
  const int lane = threadIdx.x & 63;
  const bool bad = lane < kBins && !isfinite(hb[lane]);
  return !__any(bad);
}

// Sum column c of the (at most MV_ROWS) group rows in order and zero it for
// the next producer. All loads are issued before the first store so that
// they are in flight together instead of one memory latency per row.
__device__ __forceinline__ double take_column(double* __restrict__ P, int64_t stride, int nrows,
                                              int64_t c) {
  double v[MV_ROWS];
#pragma unroll
  for (int r = 0; r < MV_ROWS; ++r) v[r] = r < nrows ? P[(int64_t)r * stride + c] : 0.0;
  double a = v[0];
#pragma unroll
  for (int r = 1; r < MV_ROWS; ++r)
    if (r < nrows) a += v[r];
#pragma unroll
  for (int r = 0; r < MV_ROWS; ++r)
    if (r < nrows) P[(int64_t)r * stride + c] = 0.0;
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
    double second = res2;
    if (!z.raw) {
      const double err = (N * res2 - res * res) / (N - 1.0);  // vflow.py:437
      second = sqrt_clamped(err);                             // vflow.py:438
    }
    out[z.n_out + t] = second;
    flag_nonfinite(z.status, res, res2);
    if (t == 0) advance_counter(z.counter, res, second, N);
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
  if (z.refine && t < 64 && wave_hist_finite(hb)) refine_wave(hb, grid + j * (kBins + 1));
}

}  // namespace mv
