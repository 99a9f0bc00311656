// vplus.hpp - hypercube bookkeeping shared by the VEGAS+ kernels
// (vflowplus.py:37-80, 187-220): cube lookup from the event offsets, cube
// coordinates, per-cube segmented sums, and the fused producer's launcher.
//
// Events of an iteration are laid out cube after cube: cube c owns events
// offsets[c] .. offsets[c+1]-1 (offsets = exclusive prefix sum of n_ev,
// vflowplus.py:67).
#pragma once

#include "abi_util.hpp"
#include "device.hpp"
#include "rows.hpp"

namespace mv {

struct StrataK {
  int n_strat;
  int n_cubes;
  const int32_t* n_ev;
  const int32_t* off;
};

// largest c with off[c] <= e (off strictly increasing, off[0] = 0)
__device__ __forceinline__ int find_cube(const int32_t* __restrict__ off, int n_cubes, int e) {
  int lo = 0, hi = n_cubes;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Cube and n_ev of event e (lane-varying) for a wave whose events all lie in
// cubes c0 .. c0 + 63 (c0 wave-uniform, off[c0] <= every e): a lane-parallel
// binary lifting over the 64 offsets from c0, held one per lane and read with
// shuffles that every lane executes. 64 consecutive events starting in cube
// c0 lie in that window when cubes hold at least one event; a lane whose
// event reaches the window's last cube (possible only with empty cubes past
// it) searches the whole array instead.
__device__ __forceinline__ int window_cube(const int32_t* __restrict__ off,
                                          const int32_t* __restrict__ n_ev, int n_cubes, int c0, int e,
                                          int& nev) {
  const int lane = threadIdx.x & 63;
  const int lim = min(n_cubes - c0, 64);
  const int ok = lane < lim ? off[c0 + lane] : 0x7fffffff;
  const int nk = lane < lim ? n_ev[c0 + lane] : 0;
  int lo = 0;
#pragma unroll
  for (int stp = 32; stp >= 1; stp >>= 1) {
    const int probe = lo + stp;  // <= 63
    const int v = __shfl(ok, probe, 64);
    if (probe < lim && v <= e) lo = probe;
  }
  nev = __shfl(nk, lo, 64);
  int c = c0 + lo;
  if (lo == 63) {
    c = find_cube(off, n_cubes, e);
    nev = n_ev[c];
  }
  return c;
}

// q = c / n, r = c % n for 0 <= c < 2^24 through an FP32 reciprocal and
// one correction step
__device__ __forceinline__ void divmod_small(int c, int n, float rn, int& q, int& r) {
  q = (int)((float)c * rn);
  r = c - q * n;
  if (r < 0) {
    --q;
    r += n;
  } else if (r >= n) {
    ++q;
    r -= n;
  }
}

// t / n rounded correctly, for a positive integer n and 0 <= t < 2^31, from
// y = 1 / n rounded correctly (one division per launch): q = t y is within
// 1.5 ulp; one FMA residual correction makes it faithful, and a second,
// from a faithful quotient and a correctly rounded reciprocal, gives the
// correctly rounded quotient (Markstein's theorem). Five FP64 operations
// instead of a full division per dimension.
__device__ __forceinline__ double div_by_strat(double t, double n, double y) {
  double q = t * y;
  double r = fma(-q, n, t);
  q = fma(r, y, q);
  r = fma(-q, n, t);
  return fma(r, y, q);
}

// hypercube coordinates for a compile-time n_dim D, last dimension fastest
// (vflowplus.py:126-128)
template <int D>
__device__ __forceinline__ void cube_digits_d(int cube, int n_strat, float rn, int (&h)[D]) {
  int c = cube;
#pragma unroll
  for (int j = D - 1; j >= 0; --j) {
    int q, r;
    divmod_small(c, n_strat, rn, q, r);
    h[j] = r;
    c = q;
  }
}

// Segmented suffix sum over the wave's (sorted) cube ids; the first lane of
// each segment adds the segment total to ress / ress2 (vflowplus.py:213-214).
__device__ __forceinline__ void segment_add(int cube, double v, double v2, double* __restrict__ ress,
                                            double* __restrict__ ress2) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double ov = __shfl_down(v, off, 64);
    const double ov2 = __shfl_down(v2, off, 64);
    const int oc = __shfl_down(cube, off, 64);
    if (lane + off < 64 && oc == cube) {
      v += ov;
      v2 += ov2;
    }
  }
  const int pc = __shfl_up(cube, 1, 64);
  const bool head = lane == 0 || pc != cube;
  if (head && cube >= 0) {
    unsafeAtomicAdd(&ress[cube], v);
    unsafeAtomicAdd(&ress2[cube], v2);
  }
}




// The fused VEGAS+ producer (vplus_fused.hip), compiled per (integrand
// family, train, n_dim, injected draws) for 1 <= n_dim <= MV_FUSED_MAX_DIM,
// one launcher per precision. Return the workgroups launched or a negative
// error.
int launch_vplus_fused_mixed(int family, bool train, const StepK& k, const StrataK& st,
                             const mv_integrand& fn, const double* grid, double* ress, double* ress2,
                             const RowsK& R, hipStream_t s);
int launch_vplus_fused_f64(int family, bool train, const StepK& k, const StrataK& st,
                           const mv_integrand& fn, const double* grid, double* ress, double* ress2,
                           const RowsK& R, hipStream_t s);

// The batched VEGAS+ producer (batch_vplus.hip) of one precision: 0 or a
// negative status.
int launch_batch_vplus_mixed(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int n_tiles, int family,
                             int min_dim, int max_dim, const int32_t* draws, const double* grids,
                             const int32_t* n_ev, const int32_t* offsets, double* ress, double* ress2, double* rows,
                             int blocks, hipStream_t stream);
int launch_batch_vplus_f64(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int n_tiles, int family,
                           int min_dim, int max_dim, const int32_t* draws, const double* grids,
                           const int32_t* n_ev, const int32_t* offsets, double* ress, double* ress2, double* rows,
                           int blocks, hipStream_t stream);

}  // namespace mv
