// rows.hpp - launch parameters shared by the producers and the two-level
// fixed-order reduction of their partial rows.
//
// Every workgroup of a producer writes one block row. The last workgroup to
// finish within each group of kGroup consecutive workgroups (it learns so
// from a ticket counter it then resets to zero) adds the group's block rows,
// in increasing workgroup order, into one group row. The finalize kernels
// then reduce at most MV_MAX_GROUPS group rows. No spin waits: a workgroup
// that is not last simply exits.
#pragma once

#include "abi_util.hpp"
#include "device.hpp"

namespace mv {

constexpr int kGroup = MV_GROUP;

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
  double* blk;
  double* grp;
  int32_t* tickets;
  int64_t stride;
  int accumulate;
};

inline int producer_blocks(int64_t n_events) {
  int64_t b = (n_events + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > MV_MAX_BLOCKS) b = MV_MAX_BLOCKS;
  return (int)b;
}

inline int groups_of(int blocks) { return (blocks + kGroup - 1) / kGroup; }

inline int make_rows(const mv_partials* p, int64_t need_stride, int accumulate, const char* what,
                     RowsK* out) {
  if (!p || !p->block_rows || !p->group_rows || !p->tickets)
    return fail(MV_ERR_ARG, "%s: incomplete partials workspace", what);
  if (p->stride < need_stride)
    return fail(MV_ERR_WORKSPACE, "%s: row stride %lld < %lld", what, (long long)p->stride,
                (long long)need_stride);
  out->blk = p->block_rows;
  out->grp = p->group_rows;
  out->tickets = p->tickets;
  out->stride = p->stride;
  out->accumulate = accumulate != 0;
  return MV_OK;
}

// Fold columns [c0, c1) of this workgroup's group; ``slot`` selects an
// independent ticket set (grid.y slices of the digest kernel).
__device__ inline void group_fold_cols(const RowsK& R, int c0, int c1, int slot) {
  __shared__ int last;
  __threadfence();  // publish this workgroup's block row device-wide
  __syncthreads();
  const int g = blockIdx.x / kGroup;
  const int r0 = g * kGroup;
  const int gs = min(kGroup, (int)gridDim.x - r0);
  int32_t* tk = R.tickets + (int64_t)slot * MV_MAX_GROUPS + g;
  if (threadIdx.x == 0) last = (atomicAdd(tk, 1) == gs - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();  // acquire the other block rows of the group
  for (int c = c0 + (int)threadIdx.x; c < c1; c += (int)blockDim.x) {
    double s = R.blk[(int64_t)r0 * R.stride + c];
    for (int r = 1; r < gs; ++r) s += R.blk[(int64_t)(r0 + r) * R.stride + c];
    double* o = R.grp + (int64_t)g * R.stride + c;
    *o = R.accumulate ? *o + s : s;
  }
  if (threadIdx.x == 0) *tk = 0;
}

__device__ inline void group_fold(const RowsK& R, int ncols) { group_fold_cols(R, 0, ncols, 0); }

}  // namespace mv
