// rows.hpp - launch parameters shared by the producers, and how their
// partial sums reach the finalize kernels.
//
// Every workgroup of a producer reduces its events to one row
// [sum_k, sum_k^2, histogram] in registers and LDS, then adds that row into
// group row (blockIdx.x mod MV_ROWS) with FP64 global atomics at agent scope.
// Atomics are coherent across the XCDs without any L2 write-back or
// invalidate, which a fence-based "last workgroup reduces" scheme would pay
// once per workgroup. The finalize kernels read the MV_ROWS group rows in a
// fixed order and zero them for the next producer, so the rows are zero
// whenever no iteration is in flight. The order in which workgroups land in a
// group row is not fixed: the sums are reproducible to rounding, not bitwise.
#pragma once

#include "abi_util.hpp"
#include "device.hpp"

namespace mv {

constexpr int kRows = MV_ROWS;

struct StepK {
  int n_dim;
  int groups;  // Philox blocks of high words: ceil(n_dim / 4)
  int64_t n_events;
  int64_t event_offset;
  Key key;
  uint32_t draw_id;
  double xjac;
  double xdeltajac;
  const double* xmin;
  const double* xdelta;
  const void* inject;
};

inline StepK make_stepk(const mv_step& s) {
  StepK k;
  k.n_dim = s.n_dim;
  k.groups = (s.n_dim + 3) / 4;
  k.n_events = s.n_events;
  k.event_offset = s.event_offset;
  k.key.k0 = (uint32_t)(s.seed & 0xffffffffu);
  k.key.k1 = (uint32_t)(s.seed >> 32);
  k.draw_id = s.draw_id;
  k.xjac = s.xjac;
  k.xdeltajac = s.xdeltajac;
  k.xmin = s.xmin;
  k.xdelta = s.xdelta;
  k.inject = s.inject;
  return k;
}

struct RowsK {
  double* rows;
  int64_t stride;
};

inline int producer_blocks(int64_t n_events) {
  int64_t b = (n_events + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > MV_MAX_BLOCKS) b = MV_MAX_BLOCKS;
  return (int)b;
}

inline int rows_used(int blocks) { return blocks < kRows ? blocks : kRows; }

inline int make_rows(const mv_partials* p, int64_t need_stride, const char* what, RowsK* out) {
  if (!p || !p->rows) return fail(MV_ERR_ARG, "%s: partials workspace missing", what);
  if (p->stride < need_stride)
    return fail(MV_ERR_WORKSPACE, "%s: row stride %lld < %lld", what, (long long)p->stride,
                (long long)need_stride);
  out->rows = p->rows;
  out->stride = p->stride;
  return MV_OK;
}

// this workgroup's group row
__device__ __forceinline__ double* group_row(const RowsK& R) {
  return R.rows + (int64_t)(blockIdx.x % kRows) * R.stride;
}

__device__ __forceinline__ void row_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// Workgroup end of the fused producers: reduce the per-thread sums over the
// workgroup (fixed order) and the kWaves private LDS histograms, and add the
// resulting row [sum, sum2, hist(hist_len)] into this workgroup's group row.
// SUMS = false leaves columns 0 and 1 alone (VEGAS+ sums per hypercube).
template <bool TRAIN, bool SUMS = true>
__device__ __forceinline__ void finish_row(double acc, double acc2, const double* hist, int hist_len,
                                           double* row) {
  __shared__ double red[2 * kWaves];
  if (SUMS) {
    acc = wave_sum(acc);
    acc2 = wave_sum(acc2);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[wave] = acc;
      red[kWaves + wave] = acc2;
    }
  }
  __syncthreads();
  if (SUMS && threadIdx.x < 2) {
    const double* r = red + threadIdx.x * kWaves;
    double a = r[0];
    for (int k = 1; k < kWaves; ++k) a += r[k];
    row_add(&row[threadIdx.x], a);
  }
  if (TRAIN) {
    for (int i = threadIdx.x; i < hist_len; i += kBlock) {
      double h = hist[i];
      for (int k = 1; k < kWaves; ++k) h += hist[k * hist_len + i];
      if (h != 0.0) row_add(&row[2 + i], h);
    }
  }
}

template <bool TRAIN, bool SUMS = true>
__device__ __forceinline__ void finish_block(double acc, double acc2, const double* hist,
                                             int hist_len, const RowsK& R) {
  finish_row<TRAIN, SUMS>(acc, acc2, hist, hist_len, group_row(R));
}

}  // namespace mv
