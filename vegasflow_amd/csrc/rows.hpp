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
#include "fused_event.hpp"

namespace mv {

constexpr int kRows = MV_ROWS;

struct StepK {
  int n_dim;
  int groups;  // Philox blocks of high words: ceil(n_dim / 4)
  int64_t n_events;
  int64_t event_offset;
  // event_offset split into the two low Philox counter words; for launches
  // cut by for_each_counter_range, event i of the launch has counter words
  // (ev_lo + i, ev_hi) with no carry, so the high word is launch-uniform
  uint32_t ev_lo, ev_hi;
  Key key;
  uint32_t draw_id;
  const uint32_t* draw_ptr;  // nullable: device word holding the draw id
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
  k.ev_lo = (uint32_t)((uint64_t)s.event_offset & 0xffffffffu);
  k.ev_hi = (uint32_t)((uint64_t)s.event_offset >> 32);
  k.key.k0 = (uint32_t)(s.seed & 0xffffffffu);
  k.key.k1 = (uint32_t)(s.seed >> 32);
  k.draw_id = s.draw_id;
  k.draw_ptr = s.draw_dev;
  k.xjac = s.xjac;
  k.xdeltajac = s.xdeltajac;
  k.xmin = s.xmin;
  k.xdelta = s.xdelta;
  k.inject = s.inject;
  return k;
}

// The draw id of a launch: the device word draw_ptr points to (a counter
// that finalize kernels advance, so that captured graphs can be replayed),
// else the host value. Read once per kernel, before the event loop.
__device__ __forceinline__ uint32_t draw_of(const StepK& s) {
  return s.draw_ptr ? *s.draw_ptr : s.draw_id;
}

// Run ``launch(StepK, first)`` over pieces of the step's events that do not
// cross a multiple of 2^32 in the event counter (a single piece unless the
// range crosses one), advancing event_offset and the injected draws; ``first``
// is the index of the piece's first event within the step (for outputs).
// Returns the largest value ``launch`` returned (its workgroup count), or
// the first negative status it returned (no later piece is launched).
template <typename F>
int for_each_counter_range(const mv_step& step, size_t inject_elem_bytes, F&& launch) {
  mv_step piece = step;
  int most = 0;
  int64_t left = step.n_events;
  do {
    const uint64_t lo = (uint64_t)piece.event_offset & 0xffffffffull;
    const int64_t room = (int64_t)((1ull << 32) - lo);
    piece.n_events = left < room ? left : room;
    const int b = launch(make_stepk(piece), step.n_events - left);
    if (b < 0) return b;
    most = b > most ? b : most;
    left -= piece.n_events;
    if (piece.inject)
      piece.inject = static_cast<const char*>(piece.inject) +
                     (size_t)piece.n_events * (size_t)piece.n_dim * inject_elem_bytes;
    piece.event_offset += piece.n_events;
  } while (left > 0);
  return most;
}

struct RowsK {
  double* rows;
  int64_t stride;
};

inline int producer_blocks(int64_t n_events, int block = kBlock) {
  int64_t b = (n_events + block - 1) / block;
  if (b < 1) b = 1;
  if (b > MV_MAX_BLOCKS) b = MV_MAX_BLOCKS;
  return (int)b;
}

inline int rows_used(int blocks) { return blocks < kRows ? blocks : kRows; }

// Workgroups of ``kernel`` (``block`` threads, ``lds`` dynamic LDS bytes)
// that are resident on the current device at once: occupancy per CU x CUs,
// cached per (device, kernel, block, lds). 0 if the runtime cannot tell or
// the workgroup does not fit.
int resident_blocks(const void* kernel, size_t lds, int block = kBlock);

// Dynamic LDS bytes one workgroup may use on the current device (cached).
size_t max_lds_per_block();

// Launch a grid-stride producer over ``n`` events with producer_blocks(n)
// workgroups capped at one device-wide round of resident workgroups: every
// workgroup owns an equal share of the events, so a grid larger than what
// fits at once would leave a second, mostly idle round at the end.
// Returns the number of workgroups launched.
template <int BS = kBlock, typename... KArgs, typename... Args>
int launch_producer(void (*kernel)(KArgs...), int64_t n, size_t lds, hipStream_t st, Args... args) {
  int blocks = producer_blocks(n, BS);
  const int fit = resident_blocks(reinterpret_cast<const void*>(kernel), lds, BS);
  if (fit > 0 && blocks > fit) blocks = fit;
  kernel<<<dim3(blocks), dim3(BS), lds, st>>>(args...);
  return blocks;
}

inline int make_rows(const mv_partials* p, int64_t need_stride, const char* what, RowsK* out) {
  if (!p || !p->rows) return fail(MV_ERR_ARG, "%s: partials workspace missing", what);
  if (p->stride < need_stride)
    return fail(MV_ERR_WORKSPACE, "%s: row stride %lld < %lld", what, (long long)p->stride,
                (long long)need_stride);
  out->rows = p->rows;
  out->stride = p->stride;
  return MV_OK;
}

// First event of this thread in a grid-stride loop whose stride is all the
// launch's threads (blockDim.x a multiple of 64): the lanes of a wave take 64
// consecutive events, and wave w of workgroup b takes wave slot
// w * gridDim.x + b, so consecutive slots lie in different workgroups. The
// last, partial round of the loop then falls on a few waves of every
// workgroup (at most one more round per SIMD) instead of on every wave of the
// first workgroups, whose CUs would finish one round after all the others.
__device__ __forceinline__ int64_t spread_first_event() {
  return ((int64_t)(threadIdx.x >> 6) * gridDim.x + blockIdx.x) * 64 + (threadIdx.x & 63);
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

// Workgroup end of the pair-histogram producers (BS threads): sums as in
// finish_row, histogram bins from the pair tables (pair_marginal_half). SUMS =
// false leaves columns 0 and 1 alone (VEGAS+ sums per hypercube). MIRROR:
// the tables hold bin b of each dimension at 49 - b (the mixed kernels).
template <int BS, bool SUMS = true, bool MIRROR = false>
__device__ __forceinline__ void finish_pairs(double acc, double acc2, const double* H, int d,
                                             double* row) {
  constexpr int W = BS / 64;
  __shared__ double red[2 * W];
  if (SUMS) {
    acc = wave_sum(acc);
    acc2 = wave_sum(acc2);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[wave] = acc;
      red[W + wave] = acc2;
    }
  }
  __syncthreads();  // also: every histogram add has landed
  if (SUMS && threadIdx.x < 2) {
    const double* r = red + threadIdx.x * W;
    double a = r[0];
    for (int k = 1; k < W; ++k) a += r[k];
    row_add(&row[threadIdx.x], a);
  }
  // two threads per bin (adjacent lanes), each summing half of its cells
  for (int t = threadIdx.x; t < 2 * d * kBins; t += BS) {
    const int m = t >> 1, half = t & 1;
    // bin b of dimension j: table index j * 50 + (49 - b) when mirrored
    const double part = pair_marginal_half(H, MIRROR ? (m / kBins) * (2 * kBins) + kBins - 1 - m : m, half);
    const double h = part + __shfl_xor(part, 1, 64);  // half 0 + half 1 in both lanes
    if (half == 0 && h != 0.0) row_add(&row[2 + m], h);
  }
}

// The fused producers with per-wave histograms (fused.hip), compiled per
// (family, train, n_dim) for 1 <= n_dim <= MV_FUSED_MAX_DIM: FP64 and mixed
// precision. Return the workgroups launched.
int launch_vegas_fused_f64(int family, bool train, const StepK& k, const mv_integrand& fn,
                           const double* grid, const RowsK& R, hipStream_t st);
int launch_vegas_fused_mixed(int family, bool train, const StepK& k, const mv_integrand& fn,
                             const double* grid, const RowsK& R, hipStream_t st);

// The fused PlainFlow producer (plain_fused.hip), compiled per (family,
// precision, n_dim) for 1 <= n_dim <= MV_FUSED_MAX_DIM.
int launch_plain_fused_dims(int family, bool mixed, const StepK& k, const mv_integrand& fn,
                            const RowsK& R, hipStream_t st);

// The mixed-precision training producer with pair histograms (pairs.hip):
// kPairBlock-thread workgroups, n_dim * 50 float2 table entries and
// pair_cells(n_dim) doubles of static LDS, one kernel per (family, n_dim)
// for kPairMinDim <= n_dim <= kPairMaxDim. Returns the workgroups launched.
int launch_vegas_fused_pairs(int family, const StepK& k, const mv_integrand& fn, const double* grid,
                             const RowsK& R, hipStream_t st);

// The FP64 training producer with pair histograms (pairs_f64.hip), as
// launch_vegas_fused_pairs; ``lds`` = fused_pairs_f64_lds(n_dim).
size_t fused_pairs_f64_lds(int d);
int launch_vegas_fused_pairs_f64(int family, const StepK& k, const mv_integrand& fn, const double* grid,
                                 const RowsK& R, size_t lds, hipStream_t st);

template <bool TRAIN, bool SUMS = true>
__device__ __forceinline__ void finish_block(double acc, double acc2, const double* hist,
                                             int hist_len, const RowsK& R) {
  finish_row<TRAIN, SUMS>(acc, acc2, hist, hist_len, group_row(R));
}

}  // namespace mv
