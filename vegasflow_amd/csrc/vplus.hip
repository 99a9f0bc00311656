// vplus.hip - VegasFlowPlus (VEGAS+, importance + adaptive stratified
// sampling, vflowplus.py) kernels for gfx950 and their C ABI.
//
// Events of an iteration are laid out cube after cube: cube c owns events
// offsets[c] .. offsets[c+1]-1 (offsets = exclusive prefix sum of n_ev,
// vflowplus.py:67). A wave handles 64 consecutive events, so the cube ids in
// a wave are sorted; per-cube sums are a segmented wave scan followed by one
// FP64 global atomic per (wave, cube) segment.
#include <algorithm>
#include <type_traits>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "refine.hpp"
#include "rows.hpp"
#include "vplus.hpp"

namespace mv {

// ---------------------------------------------------------------------------
// unfused VEGAS+: sample (vflowplus.py:37-80) and digest (209-218)
// ---------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void vplus_sample(StepK s, StrataK st,
                                                       const double* __restrict__ grid,
                                                       void* __restrict__ xo_, void* __restrict__ wo_,
                                                       uint8_t* __restrict__ io,
                                                       int32_t* __restrict__ so) {
  const uint32_t draw = draw_of(s);
  const int d = s.n_dim;
  const bool limits = s.xmin != nullptr;
  const double fstrat = (double)st.n_strat;
  const float rn = 1.0f / (float)st.n_strat;
  const int total = st.off[st.n_cubes];
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += step) {
    const int cube = find_cube(st.off, st.n_cubes, (int)i);
    const uint64_t ge = (uint64_t)(s.event_offset + i);
    const uint32_t c0 = (uint32_t)ge, c1 = (uint32_t)(ge >> 32);
    // digits from the last dimension backwards
    int rest = cube;
    int hcache[4];
    double w = 1.0;
    float wf = 1.0f;
    for (int g = s.groups - 1; g >= 0; --g) {
      for (int q = 3; q >= 0; --q) {
        const int j = 4 * g + q;
        hcache[q] = 0;
        if (j < d) {
          int qq, r;
          divmod_small(rest, st.n_strat, rn, qq, r);
          hcache[q] = r;
          rest = qq;
        }
      }
      uint4 hw4 = make_uint4(0, 0, 0, 0), lw4 = make_uint4(0, 0, 0, 0);
      if (!s.inject) {
        hw4 = philox(c0, c1, draw, (uint32_t)g, s.key);
        if (!MIXED) lw4 = philox(c0, c1, draw, (uint32_t)(s.groups + g), s.key);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = 4 * g + q;
        if (j < d) {
          double u;
          if (MIXED) {
            const uint32_t hw = s.inject ? static_cast<const uint32_t*>(s.inject)[i * d + j]
                                         : word(hw4, q);
            u = (double)hw * 0x1p-32;
          } else {
            u = s.inject ? static_cast<const double*>(s.inject)[i * d + j]
                         : draw_f64(word(hw4, q), word(lw4, q));
          }
          const double xn = (((double)hcache[q] + u) * kFBins) / fstrat;
          const int ind = clamp_bin((int)xn);
          const double lo = grid[j * (kBins + 1) + ind];
          const double dx = grid[j * (kBins + 1) + ind + 1] - lo;
          const double frac = xn - floor(xn);
          if (MIXED) {
            float x = fmaf((float)dx, (float)frac, (float)lo);
            wf = wf * (float)(dx * kFBins);
            if (limits) x = fmaf(x, (float)s.xdelta[j], (float)s.xmin[j]);
            static_cast<float*>(xo_)[i * d + j] = x;
          } else {
            double x = lo + dx * frac;  // vflow.py:76
            w = w * (dx * kFBins);      // vflow.py:78
            if (limits) x = s.xmin[j] + x * s.xdelta[j];
            static_cast<double*>(xo_)[i * d + j] = x;
          }
          if (io) io[i * d + j] = (uint8_t)ind;
        }
      }
    }
    if (MIXED) {
      static_cast<float*>(wo_)[i] =
          (float)(((double)wf / (double)st.n_ev[cube]) * (limits ? s.xjac * s.xdeltajac : s.xjac));
    } else {
      double wt = (w / (double)st.n_ev[cube]) * s.xjac;  // vflowplus.py:77, monte_carlo.py:270
      if (limits) wt = wt * s.xdeltajac;
      static_cast<double*>(wo_)[i] = wt;
    }
    if (so) so[i] = cube;
  }
}

template <typename FT, bool TRAIN>
__global__ __launch_bounds__(kBlock) void vplus_digest(StepK s, StrataK st, const FT* __restrict__ f,
                                                       const FT* __restrict__ w,
                                                       const uint8_t* __restrict__ idx,
                                                       const int32_t* __restrict__ seg,
                                                       double* __restrict__ ress,
                                                       double* __restrict__ ress2, RowsK R) {
  extern __shared__ double hist[];
  const int d = s.n_dim;
  const int nb = d * kBins;
  if (TRAIN)
    for (int i = threadIdx.x; i < kWaves * nb; i += kBlock) hist[i] = 0.0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* wh = hist + wave * nb;
  const int total = st.off[st.n_cubes];
  double acc = 0.0, acc2 = 0.0;
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t base = (int64_t)blockIdx.x * kBlock + wave * 64; base < total; base += step) {
    const int64_t i = base + lane;
    int cube = -1;
    double tmp = 0.0, tmp2 = 0.0;
    if (i < total) {
      cube = seg[i];
      tmp = (double)w[i] * (double)f[i];  // vflowplus.py:209
      tmp2 = tmp * tmp;                   // vflowplus.py:210
      acc += tmp;
      acc2 += tmp2;
      if (TRAIN)
        for (int j = 0; j < d; ++j) lds_add(&wh[j * kBins + clamp_bin(idx[i * d + j])], tmp2);
    }
    segment_add(cube, tmp, tmp2, ress, ress2);
  }
  finish_block<TRAIN, false>(acc, acc2, hist, nb, R);
}

// ---------------------------------------------------------------------------
// iteration end over the cubes (vflowplus.py:222-242 + 153-163), one
// workgroup of 1024 threads; thread t owns a contiguous chunk of cubes.
// ---------------------------------------------------------------------------
// VEGAS+ iteration end over the cubes, in three launches of kBlock-thread
// workgroups with one thread per cube (coalesced, spread over the device):
// vplus_cube_stats, vplus_realloc, vplus_offsets. ``out`` carries the
// results and, after them, the per-workgroup scratch (3 + 4 nb doubles,
// nb = ceil(n_cubes / kBlock)): out[3 + 3 b + k] partial sums, out[3 + 3 nb
// + b] event counts per workgroup.

// Sum of v over the kBlock threads, wave totals added in order.
__device__ __forceinline__ double block_sum(double v, double* part) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  double a = part[0];
  for (int k = 1; k < kWaves; ++k) a += part[k];
  __syncthreads();
  return a;
}

// Exclusive prefix sum of v over the kBlock threads; *total = the sum.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* total) {
  __shared__ int64_t wtot[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t inc = wave_inclusive_scan(v);
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  int64_t before = 0, all = 0;
  for (int k = 0; k < kWaves; ++k) {
    before += k < w ? wtot[k] : 0;
    all += wtot[k];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// Part 1 (vflowplus.py:216-233, 157): per cube the variance
// ress2 * n_ev - ress^2 (to arr_var_out) and, when adaptive, the damped
// variance max(var, 0)^(beta/2), left in ress2 for part 2; per workgroup the
// partial sums of ress, max(var, 0) / (n_ev - 1) and the damped variances.
// The bodies of the three kernels take the workgroup's chunk ``b`` of cubes
// explicitly, so that the batched iteration end (batch_vplus.hip) runs them
// for many integrals in one launch each.
__device__ inline void cube_stats_body(int b, int n_cubes, const int32_t* __restrict__ n_ev,
                                       const double* __restrict__ ress, double* __restrict__ ress2,
                                       int adaptive, double beta, double* __restrict__ out,
                                       double* __restrict__ arr_var_out) {
  __shared__ double part[kWaves];
  const int c = b * kBlock + threadIdx.x;
  double rs = 0.0, s2 = 0.0, dl = 0.0;
  if (c < n_cubes) {
    const double fn = (double)n_ev[c];
    const double r = ress[c];
    const double var = ress2[c] * fn - r * r;  // vflowplus.py:216-217
    if (arr_var_out) arr_var_out[c] = var;
    rs = r;                                    // vflowplus.py:231
    s2 = fmax(var, 0.0) / (fn - 1.0);          // vflowplus.py:230, 232
    if (adaptive) {
      dl = pow(fmax(var, 0.0), 0.5 * beta);    // vflowplus.py:157, clamped at zero first
      ress2[c] = dl;
    }
  }
  rs = block_sum(rs, part);
  s2 = block_sum(s2, part);
  dl = block_sum(dl, part);
  if (threadIdx.x == 0) {
    double* o = out + 3 + 3 * (int64_t)b;
    o[0] = rs;
    o[1] = s2;
    o[2] = dl;
  }
}

__global__ __launch_bounds__(kBlock) void vplus_cube_stats(int n_cubes, const int32_t* __restrict__ n_ev,
                                                           const double* __restrict__ ress,
                                                           double* __restrict__ ress2, int adaptive,
                                                           double beta, double* __restrict__ out,
                                                           double* __restrict__ arr_var_out) {
  cube_stats_body(blockIdx.x, n_cubes, n_ev, ress, ress2, adaptive, beta, out, arr_var_out);
}

// The partial sums of part 1 over all workgroups, in order (every workgroup
// of parts 2 and 3 recomputes them; there are few).
__device__ __forceinline__ void cube_totals(const double* __restrict__ out, int nb, double* tot) {
  __shared__ double buf[3][kBlock];
  double acc = 0.0;
  for (int b0 = 0; b0 < nb; b0 += kBlock) {
    if (b0 + (int)threadIdx.x < nb)
      for (int k = 0; k < 3; ++k) buf[k][threadIdx.x] = out[3 + 3 * (int64_t)(b0 + threadIdx.x) + k];
    __syncthreads();
    if (threadIdx.x < 3) {
      const int m = min(kBlock, nb - b0);
      for (int i = 0; i < m; ++i) acc = (b0 + i == 0) ? buf[threadIdx.x][i] : acc + buf[threadIdx.x][i];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) tot[threadIdx.x] = acc;
  __syncthreads();
}

// Part 2 (vflowplus.py:158-162): n_ev = int(max(min_neval, damped *
// init_calls / 2 / sum damped)) when adaptive (and the sum is positive), the
// events per workgroup, and out[0..1] = {res, sigma} (vflowplus.py:231-233).
// Chunk 0's thread 0 writes out[0..1] and returns true (the caller may then
// read them in that thread).
__device__ inline bool realloc_body(int b, int n_cubes, int32_t* __restrict__ n_ev,
                                    const int32_t* __restrict__ offsets, const double* __restrict__ ress2,
                                    int adaptive, int min_neval, double init_calls, double* __restrict__ out,
                                    CounterK counter, uint32_t* status) {
  __shared__ double tot[3];
  __shared__ double part[kWaves];
  const int nb = (n_cubes + kBlock - 1) / kBlock;
  cube_totals(out, nb, tot);
  const double dsum = tot[2];
  const int c = b * kBlock + threadIdx.x;
  double ne = 0.0;
  if (c < n_cubes) {
    int v = n_ev[c];
    if (adaptive && dsum > 0.0) {
      v = (int32_t)fmax((double)min_neval, ress2[c] * init_calls / 2.0 / dsum);
      n_ev[c] = v;
    }
    ne = (double)v;
  }
  ne = block_sum(ne, part);  // an integer < 2^53: exact
  if (threadIdx.x == 0) {
    out[3 + 3 * (int64_t)nb + b] = ne;
    if (b == 0) {
      out[0] = tot[0];
      out[1] = sqrt(tot[1]);  // vflowplus.py:233
      flag_nonfinite(status, out[0], tot[1]);
      // this iteration's events: offsets are rewritten only by vplus_offsets
      advance_counter(counter, out[0], out[1], (double)offsets[n_cubes]);
      return true;
    }
  }
  return false;
}

__global__ __launch_bounds__(kBlock) void vplus_realloc(int n_cubes, int32_t* __restrict__ n_ev,
                                                        const int32_t* __restrict__ offsets,
                                                        const double* __restrict__ ress2, int adaptive,
                                                        int min_neval, double init_calls,
                                                        double* __restrict__ out, CounterK counter,
                                                        uint32_t* status) {
  realloc_body(blockIdx.x, n_cubes, n_ev, offsets, ress2, adaptive, min_neval, init_calls, out, counter,
               status);
}

// Part 3: offsets of the next iteration, the exclusive prefix sum of n_ev
// (vflowplus.py:67), out[2] = events of the next iteration, ress and ress2
// zeroed. Workgroups beyond the cubes' (with train): the histogram rows of one
// dimension each and the grid refinement (vflow.py:349-362).
__device__ inline void offsets_body(int b, int n_cubes, const int32_t* __restrict__ n_ev,
                                    int32_t* __restrict__ off, double* __restrict__ ress,
                                    double* __restrict__ ress2, double* __restrict__ out) {
  const int nb = (n_cubes + kBlock - 1) / kBlock;
  __shared__ double part[kWaves];
  const double* bt = out + 3 + 3 * (int64_t)nb;
  // events of the workgroups before this one (integers: order-free, exact)
  double before = 0.0;
  for (int k = threadIdx.x; k < b; k += kBlock) before += bt[k];
  before = block_sum(before, part);
  const int c = b * kBlock + threadIdx.x;
  const int64_t v = c < n_cubes ? n_ev[c] : 0;
  int64_t total = 0;
  const int64_t ex = block_exclusive_scan(v, &total) + (int64_t)before;
  if (c < n_cubes) {
    off[c] = (int32_t)ex;
    ress[c] = 0.0;
    ress2[c] = 0.0;
  }
  if (b == nb - 1 && threadIdx.x == 0) {
    off[n_cubes] = (int32_t)((int64_t)before + total);
    out[2] = before + (double)total;
  }
}

__global__ __launch_bounds__(kBlock) void vplus_offsets(int n_cubes, const int32_t* __restrict__ n_ev,
                                                        int32_t* __restrict__ off, double* __restrict__ ress,
                                                        double* __restrict__ ress2, double* __restrict__ out,
                                                        FzK z, double* __restrict__ P, int64_t stride,
                                                        int nrows, double* __restrict__ grid,
                                                        double* __restrict__ hist_out) {
  const int nb = (n_cubes + kBlock - 1) / kBlock;
  if ((int)blockIdx.x >= nb) {
    hist_refine_block(z, P, stride, nrows, (int)blockIdx.x - nb, grid, hist_out);
    return;
  }
  offsets_body(blockIdx.x, n_cubes, n_ev, off, ress, ress2, out);
}

// ---------------------------------------------------------------------------
// batched VegasFlowPlus iteration end (mv_batch_vplus_finalize): the three
// kernels above for many items, one workgroup per {item, chunk of 256 cubes}
// job, and in the third launch one more per {item, dimension} job for the
// histogram columns and the refinement.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void batch_vplus_stats(const mv_batch_vplus_item* __restrict__ items,
                                                            const int32_t* __restrict__ chunks, double beta,
                                                            const int32_t* __restrict__ n_ev,
                                                            const double* __restrict__ ress,
                                                            double* __restrict__ ress2,
                                                            double* __restrict__ outs) {
  const int item = chunks[2 * blockIdx.x], b = chunks[2 * blockIdx.x + 1];
  const mv_batch_vplus_item& it = items[item];
  cube_stats_body(b, it.n_cubes, n_ev + it.cube_offset, ress + it.cube_offset, ress2 + it.cube_offset,
                  it.adaptive, beta, outs + it.out_offset, nullptr);
}

__global__ __launch_bounds__(kBlock) void batch_vplus_realloc(const mv_batch_vplus_item* __restrict__ items,
                                                              const int32_t* __restrict__ chunks,
                                                              int32_t* __restrict__ n_ev,
                                                              const int32_t* __restrict__ offsets,
                                                              const double* __restrict__ ress2,
                                                              double* __restrict__ outs,
                                                              int32_t* __restrict__ draws,
                                                              double* __restrict__ stats) {
  const int item = chunks[2 * blockIdx.x], b = chunks[2 * blockIdx.x + 1];
  const mv_batch_vplus_item& it = items[item];
  double* out = outs + it.out_offset;
  if (realloc_body(b, it.n_cubes, n_ev + it.cube_offset, offsets + it.off_offset, ress2 + it.cube_offset,
                   it.adaptive, it.min_neval, it.init_calls, out, CounterK{}, nullptr)) {
    // the item's running weighted average (monte_carlo.py:713-732) and its
    // iteration counter
    const double res = out[0], sig = out[1];
    double* st = stats + 4 * (int64_t)item;
    const double wgt = 1.0 / (sig * sig);  // monte_carlo.py:718
    st[0] = res;
    st[1] = sig;
    st[2] += wgt;        // monte_carlo.py:720
    st[3] += res * wgt;  // monte_carlo.py:719
    draws[item] = draws[item] + 1;
  }
}

__global__ __launch_bounds__(kBlock) void batch_vplus_offsets(const mv_batch_vplus_item* __restrict__ items,
                                                              const int32_t* __restrict__ chunks, int n_chunks,
                                                              const int32_t* __restrict__ dims,
                                                              const int32_t* __restrict__ n_ev,
                                                              int32_t* __restrict__ offsets,
                                                              double* __restrict__ ress,
                                                              double* __restrict__ ress2,
                                                              double* __restrict__ outs,
                                                              double* __restrict__ grids,
                                                              double* __restrict__ rows) {
  if ((int)blockIdx.x < n_chunks) {
    const int item = chunks[2 * blockIdx.x], b = chunks[2 * blockIdx.x + 1];
    const mv_batch_vplus_item& it = items[item];
    offsets_body(b, it.n_cubes, n_ev + it.cube_offset, offsets + it.off_offset, ress + it.cube_offset,
                 ress2 + it.cube_offset, outs + it.out_offset);
    return;
  }
  const int job = (int)blockIdx.x - n_chunks;
  const int item = dims[2 * job], dim = dims[2 * job + 1];
  const mv_batch_vplus_item& it = items[item];
  __shared__ double hb[kBins];
  double* row = rows + it.row_offset + 2 + (int64_t)dim * kBins;
  const int t = threadIdx.x;
  if (t < kBins) {
    hb[t] = row[t];
    row[t] = 0.0;
  }
  __syncthreads();
  if (it.train && t < 64 && wave_hist_finite(hb))
    refine_wave(hb, grids + it.grid_offset + (int64_t)dim * (kBins + 1));  // vflowplus.py:239-240
}

}  // namespace mv

using namespace mv;

namespace {

int validate_strata(const mv_strata* st, int n_dim, const char* what) {
  if (!st || !st->n_ev || !st->offsets) return fail(MV_ERR_ARG, "%s: strata incomplete", what);
  if (st->n_strat < 1 || st->n_cubes < 1) return fail(MV_ERR_ARG, "%s: bad stratification", what);
  if (st->n_cubes >= (1 << 24)) return fail(MV_ERR_UNSUPPORTED, "%s: too many hypercubes", what);
  int64_t p = 1;
  for (int j = 0; j < n_dim && p <= st->n_cubes; ++j) p *= st->n_strat;
  if (p != st->n_cubes) return fail(MV_ERR_ARG, "%s: n_cubes != n_strat^n_dim", what);
  return MV_OK;
}

StrataK make_strata(const mv_strata& s) { return StrataK{s.n_strat, s.n_cubes, s.n_ev, s.offsets}; }


}  // namespace

extern "C" int mv_vplus_fused(const mv_step* step, const mv_strata* strata, const mv_integrand* fn,
                              const double* grid, double* ress, double* ress2,
                              const mv_partials* parts, void* stream) {
  int rc = validate_step(step, "mv_vplus_fused", MV_FUSED_MAX_DIM);
  if (rc) return rc;
  if ((rc = validate_fn(fn, "mv_vplus_fused"))) return rc;
  if ((rc = validate_strata(strata, step->n_dim, "mv_vplus_fused"))) return rc;
  if (!grid || !ress || !ress2) return fail(MV_ERR_ARG, "mv_vplus_fused: grid, ress, ress2 required");
  if (step->n_events >= (int64_t)1 << 31) return fail(MV_ERR_UNSUPPORTED, "mv_vplus_fused: too many events");
  RowsK R;
  if ((rc = make_rows(parts, mv_partial_stride(step->n_dim, 1, step->train),
                      "mv_vplus_fused", &R)))
    return rc;
  const StepK k = make_stepk(*step);
  const StrataK st = make_strata(*strata);
  const bool mixed = step->precision == MV_MIXED, train = step->train != 0;
  hipStream_t s = as_stream(stream);
  const int blocks = mixed ? launch_vplus_fused_mixed(fn->family, train, k, st, *fn, grid, ress, ress2, R, s)
                           : launch_vplus_fused_f64(fn->family, train, k, st, *fn, grid, ress, ress2, R, s);
  if (blocks < 0) return blocks;
  if ((rc = check_launch("mv_vplus_fused"))) return rc;
  return rows_used(blocks);
}

extern "C" int mv_vplus_sample(const mv_step* step, const mv_strata* strata, const double* grid,
                               void* x, void* w, uint8_t* idx, int32_t* seg, void* stream) {
  int rc = validate_step(step, "mv_vplus_sample", 0);
  if (rc) return rc;
  if ((rc = validate_strata(strata, step->n_dim, "mv_vplus_sample"))) return rc;
  if (!grid || !x || !w) return fail(MV_ERR_ARG, "mv_vplus_sample: grid, x and w are required");
  if (step->n_events >= (int64_t)1 << 31) return fail(MV_ERR_UNSUPPORTED, "mv_vplus_sample: too many events");
  const StepK k = make_stepk(*step);
  const StrataK st = make_strata(*strata);
  int64_t b = (step->n_events + kBlock - 1) / kBlock;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(b, 8192));
  if (step->precision == MV_MIXED)
    hipLaunchKernelGGL(vplus_sample<true>, dim3(blocks), dim3(kBlock), 0, as_stream(stream), k, st, grid, x, w, idx, seg);
  else
    hipLaunchKernelGGL(vplus_sample<false>, dim3(blocks), dim3(kBlock), 0, as_stream(stream), k, st, grid, x, w, idx, seg);
  return check_launch("mv_vplus_sample");
}

extern "C" int mv_vplus_digest(const mv_step* step, const mv_strata* strata, const void* f,
                               const void* w, const uint8_t* idx, const int32_t* seg, double* ress,
                               double* ress2, const mv_partials* parts, void* stream) {
  int rc = validate_step(step, "mv_vplus_digest", 0);
  if (rc) return rc;
  if ((rc = validate_strata(strata, step->n_dim, "mv_vplus_digest"))) return rc;
  if (!f || !w || !seg || !ress || !ress2) return fail(MV_ERR_ARG, "mv_vplus_digest: missing buffer");
  if (step->train && !idx) return fail(MV_ERR_ARG, "mv_vplus_digest: idx required when training");
  if (step->train && step->n_dim > 40)
    return fail(MV_ERR_UNSUPPORTED, "mv_vplus_digest: n_dim > 40 with training");
  RowsK R;
  if ((rc = make_rows(parts, mv_partial_stride(step->n_dim, 1, step->train),
                      "mv_vplus_digest", &R)))
    return rc;
  const int blocks = producer_blocks(step->n_events);
  const StepK k = make_stepk(*step);
  const StrataK st = make_strata(*strata);
  const size_t lds = step->train ? (size_t)kWaves * step->n_dim * kBins * sizeof(double) : 0;
  hipStream_t s = as_stream(stream);
#define MV_VPD(FT, TR) \
  hipLaunchKernelGGL((vplus_digest<FT, TR>), dim3(blocks), dim3(kBlock), lds, s, k, st, \
                     static_cast<const FT*>(f), static_cast<const FT*>(w), idx, seg, ress, ress2, R)
  if (step->precision == MV_MIXED) { if (step->train) MV_VPD(float, true); else MV_VPD(float, false); }
  else { if (step->train) MV_VPD(double, true); else MV_VPD(double, false); }
#undef MV_VPD
  if ((rc = check_launch("mv_vplus_digest"))) return rc;
  return rows_used(blocks);
}

extern "C" int64_t mv_vplus_out_len(int32_t n_cubes) {
  return 3 + 4 * (int64_t)((n_cubes + kBlock - 1) / kBlock);
}

extern "C" int mv_vplus_finalize(int32_t n_dim, const mv_strata* strata, int32_t* n_ev, int32_t* offsets,
                                 double* ress, double* ress2, int32_t adaptive, int32_t min_neval,
                                 double init_calls, double beta, double* rows, int64_t row_stride,
                                 int32_t n_rows, int32_t train, int32_t refine, double* grid,
                                 double* out, double* arr_var_out, double* hist_out,
                                 uint32_t* draw_dev, double* history, int32_t history_len,
                                 uint32_t* status, void* stream) {
  int rc = validate_strata(strata, n_dim, "mv_vplus_finalize");
  if (rc) return rc;
  if (history && history_len < 0) return fail(MV_ERR_ARG, "mv_vplus_finalize: history_len < 0");
  if (!n_ev || !offsets || !ress || !ress2 || !out)
    return fail(MV_ERR_ARG, "mv_vplus_finalize: missing buffer");
  if (adaptive && min_neval < 1) return fail(MV_ERR_ARG, "mv_vplus_finalize: min_neval < 1");
  if (train) {  // histogram columns are always consumed
    if (!rows || n_rows < 1 || n_rows > MV_ROWS)
      return fail(MV_ERR_ARG, "mv_vplus_finalize: histogram rows missing or n_rows %d not in 1..%d", n_rows,
                  MV_ROWS);
    if (row_stride < mv_partial_stride(n_dim, 1, 1))
      return fail(MV_ERR_ARG, "mv_vplus_finalize: row_stride too small");
    if (refine && !grid) return fail(MV_ERR_ARG, "mv_vplus_finalize: grid required to refine");
  }
  const int n_cubes = strata->n_cubes;
  const int nb = (n_cubes + kBlock - 1) / kBlock;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(vplus_cube_stats, dim3(nb), dim3(kBlock), 0, s, n_cubes, n_ev, ress, ress2, adaptive,
                     beta, out, arr_var_out);
  if ((rc = check_launch("mv_vplus_finalize"))) return rc;
  hipLaunchKernelGGL(vplus_realloc, dim3(nb), dim3(kBlock), 0, s, n_cubes, n_ev, offsets, ress2, adaptive,
                     min_neval, init_calls, out, CounterK{draw_dev, history, history_len}, status);
  if ((rc = check_launch("mv_vplus_finalize"))) return rc;
  const FzK z{n_dim, 1, 1, refine, 0, 0.0, CounterK{}, nullptr};
  hipLaunchKernelGGL(vplus_offsets, dim3(nb + (train ? n_dim : 0)), dim3(kBlock), 0, s, n_cubes, n_ev,
                     offsets, ress, ress2, out, z, rows, row_stride, n_rows, grid, hist_out);
  return check_launch("mv_vplus_finalize");
}

extern "C" int mv_batch_vplus_finalize(const mv_batch_vplus_item* items, const int32_t* chunks, int32_t n_chunks,
                                       const int32_t* dims, int32_t n_dims, double beta, double* grids,
                                       int32_t* n_ev, int32_t* offsets, double* ress, double* ress2, double* rows,
                                       double* outs, int32_t* draws, double* stats, void* stream) {
  if (n_chunks < 0 || n_dims < 0) return fail(MV_ERR_ARG, "mv_batch_vplus_finalize: negative job count");
  if (n_chunks == 0 && n_dims == 0) return MV_OK;
  if (!items || (n_chunks && !chunks) || (n_dims && !dims) || !grids || !n_ev || !offsets || !ress || !ress2 ||
      !rows || !outs || !draws || !stats)
    return fail(MV_ERR_ARG, "mv_batch_vplus_finalize: missing buffer");
  hipStream_t s = as_stream(stream);
  int rc;
  if (n_chunks) {
    hipLaunchKernelGGL(batch_vplus_stats, dim3(n_chunks), dim3(kBlock), 0, s, items, chunks, beta, n_ev, ress,
                       ress2, outs);
    if ((rc = check_launch("mv_batch_vplus_finalize"))) return rc;
    hipLaunchKernelGGL(batch_vplus_realloc, dim3(n_chunks), dim3(kBlock), 0, s, items, chunks, n_ev, offsets,
                       ress2, outs, draws, stats);
    if ((rc = check_launch("mv_batch_vplus_finalize"))) return rc;
  }
  hipLaunchKernelGGL(batch_vplus_offsets, dim3(n_chunks + n_dims), dim3(kBlock), 0, s, items, chunks, n_chunks,
                     dims, n_ev, offsets, ress, ress2, outs, grids, rows);
  return check_launch("mv_batch_vplus_finalize");
}

extern "C" int mv_batch_vplus_fused(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int32_t n_tiles,
                                    int32_t precision, int32_t family, int32_t min_dim, int32_t max_dim,
                                    const int32_t* draws, const double* grids, const int32_t* n_ev,
                                    const int32_t* offsets, double* ress, double* ress2, double* rows,
                                    int32_t blocks, void* stream) {
  if (n_tiles < 0 || blocks < 1) return fail(MV_ERR_ARG, "mv_batch_vplus_fused: bad n_tiles / blocks");
  if (n_tiles == 0) return MV_OK;
  if (!items || !tiles || !draws || !grids || !n_ev || !offsets || !ress || !ress2 || !rows)
    return fail(MV_ERR_ARG, "mv_batch_vplus_fused: missing buffer");
  if (precision != MV_F64 && precision != MV_MIXED)
    return fail(MV_ERR_ARG, "mv_batch_vplus_fused: unknown precision %d", precision);
  if (min_dim < 1 || max_dim < min_dim || max_dim > MV_FUSED_MAX_DIM)
    return fail(MV_ERR_UNSUPPORTED, "mv_batch_vplus_fused: n_dim range %d..%d outside 1..%d", min_dim, max_dim,
                MV_FUSED_MAX_DIM);
  hipStream_t s = as_stream(stream);
  if (precision == MV_MIXED)
    return launch_batch_vplus_mixed(items, tiles, n_tiles, family, min_dim, max_dim, draws, grids, n_ev, offsets,
                                    ress, ress2, rows, blocks, s);
  return launch_batch_vplus_f64(items, tiles, n_tiles, family, min_dim, max_dim, draws, grids, n_ev, offsets, ress,
                                ress2, rows, blocks, s);
}
