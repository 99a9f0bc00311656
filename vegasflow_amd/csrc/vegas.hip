// vegas.hip - VegasFlow and PlainFlow kernels for gfx950 and their C ABI.
//
// Kernels
//   vegas_fused_f64 / vegas_fused_mixed   draws + grid map + built-in
//       integrand + sums + per-wave LDS bin histograms; no per-event HBM
//       traffic. Each workgroup adds one row of partial sums into a group row.
//   vegas_sample_{f64,mixed} / vegas_digest   the unfused path around a
//       torch-callable integrand (x, w, idx through HBM).
//   finalize_kernel   fixed-order reduction of group rows (zeroed after), sigma, and the
//       per-dimension refinement (one workgroup per dimension).
//   plain_*   PlainFlow variants.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "abi_util.hpp"
#include "device.hpp"
#include "fused_event.hpp"
#include "refine.hpp"
#include "rows.hpp"

namespace mv {

std::string& last_error_slot() {
  static thread_local std::string s;
  return s;
}

int resident_blocks(const void* kernel, size_t lds, int block) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, int, size_t>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  const auto key = std::make_tuple(dev, kernel, block, lds);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  const int r = cus * per_cu;
  std::lock_guard<std::mutex> g(mu);
  cache[key] = r;
  return r;
}

size_t max_lds_per_block() {
  static std::mutex mu;
  static std::map<int, size_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
    (void)hipGetLastError();
    v = 0;
  }
  cache[dev] = (size_t)v;
  return (size_t)v;
}

// ---------------------------------------------------------------------------
// unfused path: sample (any n_dim)
// With STAGED, each block writes its 256 events' rows of x (and bins) to LDS
// first and then stores the block's contiguous part of the row-major (n, d)
// outputs with consecutive threads on consecutive elements (full cache lines
// instead of one element per thread every d elements). STAGED needs
// stage_bytes(d) of LDS; beyond that the kernels store directly.
// ---------------------------------------------------------------------------
constexpr size_t kStageMax = 64 * 1024;

inline size_t stage_bytes(int d, size_t elem, bool bins) {
  return (size_t)kBlock * d * elem + (bins ? (((size_t)kBlock * d + 3) & ~(size_t)3) : 0);
}

// Copy ``bytes`` bytes from the LDS stage to dst: 16 bytes per thread and
// store where both sides are 16-byte aligned, then the rest byte-wise by
// element. ``st`` is 16-byte aligned (x rows at the stage start, bins after
// kBlock * d elements of 4 or 8 bytes).
__device__ __forceinline__ void stage_copy(const uint8_t* st, uint8_t* __restrict__ dst, int bytes,
                                           int elem) {
  int done = 0;
  if (((uintptr_t)dst & 15u) == 0 && ((uintptr_t)st & 15u) == 0) {
    const int vecs = bytes >> 4;
    const uint4* s16 = reinterpret_cast<const uint4*>(st);
    uint4* d16 = reinterpret_cast<uint4*>(dst);
    for (int k = threadIdx.x; k < vecs; k += kBlock) d16[k] = s16[k];
    done = vecs << 4;
  }
  if (elem == 8) {
    const double* s8 = reinterpret_cast<const double*>(st);
    double* d8 = reinterpret_cast<double*>(dst);
    for (int k = (done >> 3) + threadIdx.x; k < (bytes >> 3); k += kBlock) d8[k] = s8[k];
  } else if (elem == 4) {
    const float* s4 = reinterpret_cast<const float*>(st);
    float* d4 = reinterpret_cast<float*>(dst);
    for (int k = (done >> 2) + threadIdx.x; k < (bytes >> 2); k += kBlock) d4[k] = s4[k];
  } else {
    for (int k = done + threadIdx.x; k < bytes; k += kBlock) dst[k] = st[k];
  }
}

template <typename T>
__device__ __forceinline__ void stage_store(const T* st, T* __restrict__ dst, int elems) {
  stage_copy(reinterpret_cast<const uint8_t*>(st), reinterpret_cast<uint8_t*>(dst),
             elems * (int)sizeof(T), (int)sizeof(T));
}

__device__ __forceinline__ void stage_store_bytes(const uint8_t* st, uint8_t* __restrict__ dst,
                                                  int elems) {
  stage_copy(st, dst, elems, 1);
}

// vflow.py:87-126 + 33-83, monte_carlo.py:249-275. FP64: the reference's
// operation order (x = lo + dx * frac with two roundings). Mixed: the
// positions of mixed_mirrored, map in FP32.
template <bool MIXED, bool STAGED>
__global__ __launch_bounds__(kBlock) void vegas_sample(StepK s, const double* __restrict__ grid,
                                                       void* __restrict__ xo_, void* __restrict__ wo_,
                                                       uint8_t* __restrict__ io) {
  const uint32_t draw = draw_of(s);
  using T = typename std::conditional<MIXED, float, double>::type;
  extern __shared__ double stage_mem[];
  T* xo = static_cast<T*>(xo_);
  T* wo = static_cast<T*>(wo_);
  const int d = s.n_dim;
  T* xs = reinterpret_cast<T*>(stage_mem);
  uint8_t* bs = reinterpret_cast<uint8_t*>(xs + (size_t)kBlock * d);
  const bool limits = s.xmin != nullptr;
  const int64_t n = s.n_events;
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += step) {  // block-uniform
    const int64_t i = base + threadIdx.x;
    if (i < n) {
      const uint32_t c0 = s.ev_lo + (uint32_t)i, c1 = s.ev_hi;  // launch-uniform high word
      T w = (T)1;
      for (int g = 0; g < s.groups; ++g) {
        uint4 h = make_uint4(0, 0, 0, 0), l = make_uint4(0, 0, 0, 0);
        if (!s.inject) {
          h = philox(c0, c1, draw, (uint32_t)g, s.key);
          if (!MIXED) l = philox(c0, c1, draw, (uint32_t)(s.groups + g), s.key);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = 4 * g + q;
          if (j < d) {
            int ind;
            T x;
            if constexpr (MIXED) {
              const uint32_t* inj = static_cast<const uint32_t*>(s.inject);
              float f1;
              ind = kBins - 1 - mixed_mirrored(inj ? inj[i * d + j] : word(h, q), f1);
              const double lo = grid[j * (kBins + 1) + ind];
              const double dx = grid[j * (kBins + 1) + ind + 1] - lo;
              x = fmaf((float)dx, 2.0f - f1, (float)lo);  // frac = 2 - f1 = 1 - mfrac is exact
              w = w * (float)(dx * kFBins);
              if (limits) x = fmaf(x, (float)s.xdelta[j], (float)s.xmin[j]);
            } else {
              const double* inj = static_cast<const double*>(s.inject);
              const double u = inj ? inj[i * d + j] : draw_f64(word(h, q), word(l, q));
              const double xn = kFBins * (1.0 - u);  // vflow.py:117
              ind = clamp_bin((int)xn);               // vflow.py:67
              const double lo = grid[j * (kBins + 1) + ind];
              const double dx = grid[j * (kBins + 1) + ind + 1] - lo;  // vflow.py:73
              const double frac = xn - floor(xn);     // vflow.py:75
              x = lo + dx * frac;                     // two roundings, as vflow.py:76
              w = w * (dx * kFBins);                  // vflow.py:78
              if (limits) x = s.xmin[j] + x * s.xdelta[j];  // monte_carlo.py:273
            }
            if (STAGED) {
              xs[threadIdx.x * d + j] = x;
              if (io) bs[threadIdx.x * d + j] = (uint8_t)ind;
            } else {
              xo[i * d + j] = x;
              if (io) io[i * d + j] = (uint8_t)ind;
            }
          }
        }
      }
      if constexpr (MIXED) {
        wo[i] = w * (float)(limits ? s.xjac * s.xdeltajac : s.xjac);
      } else {
        double wt = w * s.xjac;  // monte_carlo.py:270
        if (limits) wt = wt * s.xdeltajac;
        wo[i] = wt;
      }
    }
    if (STAGED) {
      __syncthreads();
      const int elems = (int)(n - base < kBlock ? n - base : kBlock) * d;
      stage_store(xs, xo + base * d, elems);
      if (io) stage_store_bytes(bs, io + base * d, elems);
      __syncthreads();
    }
  }
}

// monte_carlo.py:264-274 (PlainFlow draws): x = u, or xmin + u * xdelta.
template <bool MIXED, bool STAGED>
__global__ __launch_bounds__(kBlock) void plain_sample(StepK s, void* __restrict__ xo_) {
  const uint32_t draw = draw_of(s);
  using T = typename std::conditional<MIXED, float, double>::type;
  extern __shared__ double stage_mem[];
  T* xo = static_cast<T*>(xo_);
  T* xs = reinterpret_cast<T*>(stage_mem);
  const int d = s.n_dim;
  const bool limits = s.xmin != nullptr;
  const int64_t n = s.n_events;
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += step) {  // block-uniform
    const int64_t i = base + threadIdx.x;
    if (i < n) {
      const uint32_t c0 = s.ev_lo + (uint32_t)i, c1 = s.ev_hi;  // launch-uniform high word
      for (int g = 0; g < s.groups; ++g) {
        uint4 h = make_uint4(0, 0, 0, 0), l = make_uint4(0, 0, 0, 0);
        if (!s.inject) {
          h = philox(c0, c1, draw, (uint32_t)g, s.key);
          if (!MIXED) l = philox(c0, c1, draw, (uint32_t)(s.groups + g), s.key);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = 4 * g + q;
          if (j < d) {
            T x;
            if constexpr (MIXED) {
              const uint32_t hw = s.inject ? static_cast<const uint32_t*>(s.inject)[i * d + j]
                                           : word(h, q);
              float xf = fmaf((float)(hw >> 9), 0x1p-23f, 0x1p-24f);
              if (limits) xf = fmaf(xf, (float)s.xdelta[j], (float)s.xmin[j]);
              x = xf;
            } else {
              const double u = s.inject ? static_cast<const double*>(s.inject)[i * d + j]
                                        : draw_f64(word(h, q), word(l, q));
              x = limits ? s.xmin[j] + u * s.xdelta[j] : u;
            }
            if (STAGED)
              xs[threadIdx.x * d + j] = x;
            else
              xo[i * d + j] = x;
          }
        }
      }
    }
    if (STAGED) {
      __syncthreads();
      stage_store(xs, xo + base * d, (int)(n - base < kBlock ? n - base : kBlock) * d);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// unfused path: digest (vflow.py:414-430, plain.py:26-35)
// grid.y splits the dimensions into slices of kDigestDims so that any n_dim
// fits the per-wave LDS histograms; slice 0 also accumulates the sums.
// ---------------------------------------------------------------------------
constexpr int kDigestDims = MV_DIGEST_DIMS;

template <typename FT, bool VEC, bool TRAIN>
__global__ __launch_bounds__(kBlock) void vegas_digest(StepK s, const FT* __restrict__ f,
                                                       int64_t fstride, int nout, int main_dim,
                                                       const FT* __restrict__ w, int64_t wstride,
                                                       const uint8_t* __restrict__ idx, RowsK R) {
  __shared__ double hist[kWaves * kDigestDims * kBins];
  __shared__ double red[kWaves * 2 * MV_MAX_OUT];
  const int d = s.n_dim;
  const int j0 = blockIdx.y * kDigestDims;
  const int nd = min(kDigestDims, d - j0);
  const int nb = nd * kBins;
  if (TRAIN) {
    for (int i = threadIdx.x; i < kWaves * nb; i += kBlock) hist[i] = 0.0;
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6;
  double* wh = hist + wave * nb;
  const bool sums = blockIdx.y == 0;
  constexpr int KM = VEC ? MV_MAX_OUT : 1;
  double acc[KM], acc2[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) acc[k] = acc2[k] = 0.0;
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < s.n_events; i += step) {
    const double wi = (double)w[i * wstride];
    double t2main;
    if (VEC) {
      t2main = 0.0;
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        if (k < nout) {
          const double tmp = wi * (double)f[i * fstride + k];  // vflow.py:416
          const double tmp2 = tmp * tmp;
          acc[k] += tmp;
          acc2[k] += tmp2;
          if (k == main_dim) t2main = tmp2;  // vflow.py:425-426
        }
      }
    } else {
      const double tmp = wi * (double)f[i * fstride];
      t2main = tmp * tmp;
      acc[0] += tmp;
      acc2[0] += t2main;
    }
    if (TRAIN) {
      const uint8_t* ir = idx + i * d + j0;
      for (int j = 0; j < nd; ++j) lds_add(&wh[j * kBins + clamp_bin(ir[j])], t2main);
    }
  }
  const int no = VEC ? nout : 1;
  if (sums) {
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      if (k < no) {
        const double a = wave_sum(acc[k]);
        const double b = wave_sum(acc2[k]);
        if ((threadIdx.x & 63) == 0) {
          red[(2 * k) * kWaves + wave] = a;
          red[(2 * k + 1) * kWaves + wave] = b;
        }
      }
    }
  }
  __syncthreads();
  // every y-slice adds its own histogram columns; slice 0 owns the sums
  double* row = group_row(R);
  if (sums && threadIdx.x < 2 * no) {
    const int v = threadIdx.x;  // 2k: sum of output k, 2k+1: its sum of squares
    double a = red[v * kWaves];
    for (int q = 1; q < kWaves; ++q) a += red[v * kWaves + q];
    row_add(&row[(v & 1) ? no + (v >> 1) : (v >> 1)], a);
  }
  if (TRAIN) {
    for (int i = threadIdx.x; i < nb; i += kBlock) {
      double h = hist[i];
      for (int q = 1; q < kWaves; ++q) h += hist[q * nb + i];
      if (h != 0.0) row_add(&row[2 * no + j0 * kBins + i], h);
    }
  }
}

// Training digest of a scalar integrand with pair histograms (fused_event.hpp)
// for 2 <= n_dim <= kPairMaxDim: one LDS add per pair of dimensions into a
// workgroup's pair tables (LDS: pair_cells(d) doubles), so the kernel is
// bound by its HBM reads (f, w, bins) rather than by LDS atomics. Rows of
// bins whose length is a multiple of 4 are read four bins per load. Bins
// are clamped to 0..49 as in vegas_digest.
constexpr int kDigestU = 4;  // events per trip of the pair digest's word path

template <typename FT, int BS>
__global__ __launch_bounds__(BS) void vegas_digest_pairs(StepK s, const FT* __restrict__ f,
                                                         int64_t fstride, const FT* __restrict__ w,
                                                         int64_t wstride, const uint8_t* __restrict__ idx,
                                                         RowsK R) {
  extern __shared__ double H[];
  const int d = s.n_dim;
  const int half = pair_cells(d) / 2;
  for (int i = threadIdx.x; i < half; i += BS) reinterpret_cast<double2*>(H)[i] = make_double2(0.0, 0.0);
  __syncthreads();
  const int np = (d + 1) >> 1;
  const bool words = (d & 3) == 0 && ((uintptr_t)idx & 3u) == 0;
  const uint32_t top = kBins - 1;
  double acc = 0.0, acc2 = 0.0;
  const int64_t step = (int64_t)gridDim.x * BS;
  const int64_t i0 = spread_first_event();
  if (words) {
    // kDigestU events per trip, all their loads issued before the first is
    // used: with one workgroup per CU the loop is otherwise bound by one
    // memory latency per event. Sums are added in event order.
    constexpr int U = kDigestU;
    constexpr int MAXW = kPairMaxDim / 4;
    const int nw = d >> 2;
    for (int64_t i = i0; i < s.n_events; i += U * step) {
      double tmp[U];
      uint32_t v[U][MAXW];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t e = i + u * step;
        ok[u] = e < s.n_events;
        tmp[u] = 0.0;
#pragma unroll
        for (int g = 0; g < MAXW; ++g) v[u][g] = 0u;
        if (ok[u]) {
          tmp[u] = (double)w[e * wstride] * (double)f[e * fstride];  // vflow.py:416
          const uint32_t* ir4 = reinterpret_cast<const uint32_t*>(idx + e * d);  // 4-aligned
#pragma unroll
          for (int g = 0; g < MAXW; ++g)
            if (g < nw) v[u][g] = ir4[g];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        const double tmp2 = tmp[u] * tmp[u];
        acc += tmp[u];
        acc2 += tmp2;
#pragma unroll
        for (int g = 0; g < MAXW; ++g) {
          if (g < nw) {
            const uint32_t x = v[u][g];
            const uint32_t b0 = min(x & 0xffu, top), b1 = min((x >> 8) & 0xffu, top);
            const uint32_t b2 = min((x >> 16) & 0xffu, top), b3 = min(x >> 24, top);
            lds_add(&H[(2 * g) * kPairBins + b0 * kBins + b1], tmp2);
            lds_add(&H[(2 * g + 1) * kPairBins + b2 * kBins + b3], tmp2);
          }
        }
      }
    }
  } else {
    for (int64_t i = i0; i < s.n_events; i += step) {
      const double tmp = (double)w[i * wstride] * (double)f[i * fstride];  // vflow.py:416
      const double tmp2 = tmp * tmp;
      acc += tmp;
      acc2 += tmp2;
      const uint8_t* ir = idx + i * d;
      for (int p = 0; p < np; ++p) {
        const uint32_t b0 = min((uint32_t)ir[2 * p], top);
        const uint32_t b1 = 2 * p + 1 < d ? min((uint32_t)ir[2 * p + 1], top) : 0u;
        lds_add(&H[p * kPairBins + b0 * kBins + b1], tmp2);
      }
    }
  }
  finish_pairs<BS>(acc, acc2, H, d, group_row(R));
}
// ---------------------------------------------------------------------------
// finalize: block 0 sums and sigma, block 1 + j dimension j (refine.hpp)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void finalize_kernel(FzK z, double* __restrict__ P,
                                                          int64_t stride, int nrows,
                                                          double* __restrict__ grid,
                                                          double* __restrict__ out,
                                                          double* __restrict__ hist_out) {
  if (blockIdx.x == 0)
    finalize_sums_block(z, P, stride, nrows, out);
  else
    hist_refine_block(z, P, stride, nrows, (int)blockIdx.x - 1, grid, hist_out);
}

__global__ __launch_bounds__(64) void refine_kernel(const double* __restrict__ hist, int64_t hstride,
                                                    double* __restrict__ grid) {
  __shared__ double hb[kBins];
  const int j = blockIdx.x, t = threadIdx.x;
  if (t < kBins) hb[t] = hist[j * hstride + t];
  __syncthreads();
  if (wave_hist_finite(hb)) refine_wave(hb, grid + j * (kBins + 1));
}

// ---------------------------------------------------------------------------
// consume_array_into_indices (utils.py:10-44)
// ---------------------------------------------------------------------------
__global__ void fill_kernel(double* p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

constexpr int kConsumeLds = 4096;

__global__ __launch_bounds__(kBlock) void consume_kernel(const double* __restrict__ in,
                                                         const int32_t* __restrict__ ind,
                                                         int64_t n, int size,
                                                         double* __restrict__ out) {
  __shared__ double h[kConsumeLds];
  const bool lds = size <= kConsumeLds;
  if (lds) {
    for (int i = threadIdx.x; i < size; i += kBlock) h[i] = 0.0;
    __syncthreads();
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int b = ind[i];
    if (b >= 0 && b < size) {
      if (lds) lds_add(&h[b], in[i]);
      else unsafeAtomicAdd(&out[b], in[i]);
    }
  }
  if (lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < size; i += kBlock)
      if (h[i] != 0.0) unsafeAtomicAdd(&out[i], h[i]);
  }
}

}  // namespace mv

using namespace mv;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int mv_abi_version(void) { return MV_ABI_VERSION; }

extern "C" const char* mv_last_error(void) { return last_error_slot().c_str(); }

extern "C" int64_t mv_struct_size(int32_t which) {
  switch (which) {
    case 0: return (int64_t)sizeof(mv_integrand);
    case 1: return (int64_t)sizeof(mv_step);
    case 2: return (int64_t)sizeof(mv_partials);
    case 3: return (int64_t)sizeof(mv_strata);
    case 4: return (int64_t)sizeof(mv_finalize);
    case 5: return (int64_t)sizeof(mv_batch_item);
    case 6: return (int64_t)sizeof(mv_batch_tile);
    case 7: return (int64_t)sizeof(mv_batch_vplus_item);
    default: return -1;
  }
}

extern "C" int64_t mv_partial_stride(int32_t n_dim, int32_t n_out, int32_t train) {
  return 2 * (int64_t)n_out + (train ? (int64_t)n_dim * kBins : 0);
}

extern "C" int32_t mv_producer_blocks(int64_t n_events) { return producer_blocks(n_events); }

extern "C" int32_t mv_producer_rows(int64_t n_events) {
  return rows_used(producer_blocks(n_events));
}

namespace {

// LDS bytes of vegas_fused_pairs' table and cells (static in the kernel),
// 0 where it does not apply (FP64, no training, n_dim outside
// kPairMinDim..kPairMaxDim, or more LDS than a workgroup may have)
size_t pairs_lds(int d, bool mixed, bool train) {
  if (!mixed || !train || d < kPairMinDim || d > kPairMaxDim) return 0;
  const size_t lds = (size_t)(d * kBins + pair_cells(d)) * sizeof(double);
  return lds <= max_lds_per_block() ? lds : 0;
}

template <int FAM>
int launch_vegas_fused(const StepK& k, const mv_integrand& fn, bool mixed, bool train,
                       const double* grid, const RowsK& R, hipStream_t st) {
  if (pairs_lds(k.n_dim, mixed, train) != 0) return launch_vegas_fused_pairs(FAM, k, fn, grid, R, st);
  if (!mixed && train && k.n_dim >= kPairMinDim && k.n_dim <= kPairMaxDim &&
      fused_pairs_f64_lds(k.n_dim) <= max_lds_per_block())
    return launch_vegas_fused_pairs_f64(FAM, k, fn, grid, R, fused_pairs_f64_lds(k.n_dim), st);
  return mixed ? launch_vegas_fused_mixed(FAM, train, k, fn, grid, R, st)
               : launch_vegas_fused_f64(FAM, train, k, fn, grid, R, st);
}

template <int FAM>
int launch_plain_fused(const StepK& k, const mv_integrand& fn, bool mixed, const RowsK& R,
                       hipStream_t st) {
  return launch_plain_fused_dims(FAM, mixed, k, fn, R, st);
}

int sample_blocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}

}  // namespace

extern "C" int mv_vegas_fused(const mv_step* step, const mv_integrand* fn, const double* grid,
                              const mv_partials* parts, void* stream) {
  int rc = validate_step(step, "mv_vegas_fused", MV_FUSED_MAX_DIM);
  if (rc) return rc;
  if ((rc = validate_fn(fn, "mv_vegas_fused"))) return rc;
  if (!grid) return fail(MV_ERR_ARG, "mv_vegas_fused: grid is required");
  RowsK R;
  if ((rc = make_rows(parts, mv_partial_stride(step->n_dim, 1, step->train),
                      "mv_vegas_fused", &R)))
    return rc;
  const bool mixed = step->precision == MV_MIXED, train = step->train != 0;
  hipStream_t st = as_stream(stream);
  const int blocks = for_each_counter_range(*step, mixed ? 4 : 8, [&](const StepK& k, int64_t) {
    switch (fn->family) {
      case MV_FN_CONSTANT: return launch_vegas_fused<MV_FN_CONSTANT>(k, *fn, mixed, train, grid, R, st);
      case MV_FN_GAUSSIAN: return launch_vegas_fused<MV_FN_GAUSSIAN>(k, *fn, mixed, train, grid, R, st);
      case MV_FN_PRODUCT: return launch_vegas_fused<MV_FN_PRODUCT>(k, *fn, mixed, train, grid, R, st);
      case MV_FN_GAUSSMIX: return launch_vegas_fused<MV_FN_GAUSSMIX>(k, *fn, mixed, train, grid, R, st);
      default: return launch_vegas_fused<MV_FN_BWPOWER>(k, *fn, mixed, train, grid, R, st);
    }
  });
  if (blocks < 0) return blocks;
  if ((rc = check_launch("mv_vegas_fused"))) return rc;
  return rows_used(blocks);
}

extern "C" int mv_plain_fused(const mv_step* step, const mv_integrand* fn, const mv_partials* parts,
                              void* stream) {
  int rc = validate_step(step, "mv_plain_fused", 0);
  if (rc) return rc;
  if ((rc = validate_fn(fn, "mv_plain_fused"))) return rc;
  if ((fn->family == MV_FN_GAUSSIAN || fn->family == MV_FN_GAUSSMIX) &&
      step->n_dim > MV_FUSED_MAX_DIM)
    return fail(MV_ERR_UNSUPPORTED, "mv_plain_fused: n_dim %d exceeds %d for this family",
                step->n_dim, MV_FUSED_MAX_DIM);
  RowsK R;
  if ((rc = make_rows(parts, mv_partial_stride(step->n_dim, 1, 0),
                      "mv_plain_fused", &R)))
    return rc;
  const bool mixed = step->precision == MV_MIXED;
  hipStream_t st = as_stream(stream);
  const int blocks = for_each_counter_range(*step, mixed ? 4 : 8, [&](const StepK& k, int64_t) {
    switch (fn->family) {
      case MV_FN_CONSTANT: return launch_plain_fused<MV_FN_CONSTANT>(k, *fn, mixed, R, st);
      case MV_FN_GAUSSIAN: return launch_plain_fused<MV_FN_GAUSSIAN>(k, *fn, mixed, R, st);
      case MV_FN_PRODUCT: return launch_plain_fused<MV_FN_PRODUCT>(k, *fn, mixed, R, st);
      case MV_FN_GAUSSMIX: return launch_plain_fused<MV_FN_GAUSSMIX>(k, *fn, mixed, R, st);
      default: return launch_plain_fused<MV_FN_BWPOWER>(k, *fn, mixed, R, st);
    }
  });
  if (blocks < 0) return blocks;
  if ((rc = check_launch("mv_plain_fused"))) return rc;
  return rows_used(blocks);
}

extern "C" int mv_vegas_sample(const mv_step* step, const double* grid, void* x, void* w,
                               uint8_t* idx, void* stream) {
  int rc = validate_step(step, "mv_vegas_sample", 0);
  if (rc) return rc;
  if (!grid || !x || !w) return fail(MV_ERR_ARG, "mv_vegas_sample: grid, x and w are required");
  if (step->n_events == 0) return MV_OK;
  hipStream_t st = as_stream(stream);
  const bool f64 = step->precision == MV_F64;
  const int64_t d = step->n_dim;
  const size_t elem = f64 ? 8 : 4;
  const size_t lds = stage_bytes(step->n_dim, elem, idx != nullptr);
  const bool staged = lds <= kStageMax;
  for_each_counter_range(*step, elem, [&](const StepK& k, int64_t e0) {
    const int blocks = sample_blocks(k.n_events);
    uint8_t* io = idx ? idx + e0 * d : nullptr;
    void* xp = static_cast<char*>(x) + e0 * d * (int64_t)elem;
    void* wp = static_cast<char*>(w) + e0 * (int64_t)elem;
    if (f64) {
      if (staged) hipLaunchKernelGGL((vegas_sample<false, true>), dim3(blocks), dim3(kBlock), lds, st, k, grid, xp, wp, io);
      else hipLaunchKernelGGL((vegas_sample<false, false>), dim3(blocks), dim3(kBlock), 0, st, k, grid, xp, wp, io);
    } else {
      if (staged) hipLaunchKernelGGL((vegas_sample<true, true>), dim3(blocks), dim3(kBlock), lds, st, k, grid, xp, wp, io);
      else hipLaunchKernelGGL((vegas_sample<true, false>), dim3(blocks), dim3(kBlock), 0, st, k, grid, xp, wp, io);
    }
    return blocks;
  });
  return check_launch("mv_vegas_sample");
}

extern "C" int mv_plain_sample(const mv_step* step, void* x, void* stream) {
  int rc = validate_step(step, "mv_plain_sample", 0);
  if (rc) return rc;
  if (!x) return fail(MV_ERR_ARG, "mv_plain_sample: x is required");
  if (step->n_events == 0) return MV_OK;
  hipStream_t st = as_stream(stream);
  const bool f64 = step->precision == MV_F64;
  const int64_t d = step->n_dim;
  const size_t elem = f64 ? 8 : 4;
  const size_t lds = stage_bytes(step->n_dim, elem, false);
  const bool staged = lds <= kStageMax;
  for_each_counter_range(*step, elem, [&](const StepK& k, int64_t e0) {
    const int blocks = sample_blocks(k.n_events);
    void* xp = static_cast<char*>(x) + e0 * d * (int64_t)elem;
    if (f64) {
      if (staged) hipLaunchKernelGGL((plain_sample<false, true>), dim3(blocks), dim3(kBlock), lds, st, k, xp);
      else hipLaunchKernelGGL((plain_sample<false, false>), dim3(blocks), dim3(kBlock), 0, st, k, xp);
    } else {
      if (staged) hipLaunchKernelGGL((plain_sample<true, true>), dim3(blocks), dim3(kBlock), lds, st, k, xp);
      else hipLaunchKernelGGL((plain_sample<true, false>), dim3(blocks), dim3(kBlock), 0, st, k, xp);
    }
    return blocks;
  });
  return check_launch("mv_plain_sample");
}

extern "C" int mv_vegas_digest(const mv_step* step, const void* f, int64_t f_stride, int32_t n_out,
                               int32_t main_dim, const void* w, int64_t w_stride,
                               const uint8_t* idx, const mv_partials* parts, void* stream) {
  int rc = validate_step(step, "mv_vegas_digest", 0);
  if (rc) return rc;
  if (!f || !w) return fail(MV_ERR_ARG, "mv_vegas_digest: f and w are required");
  if (n_out < 1 || n_out > MV_MAX_OUT)
    return fail(MV_ERR_UNSUPPORTED, "mv_vegas_digest: n_out must be 1..%d (got %d)", MV_MAX_OUT, n_out);
  if (main_dim < 0 || main_dim >= n_out)
    return fail(MV_ERR_ARG, "mv_vegas_digest: main_dim %d out of range", main_dim);
  if (step->train && !idx) return fail(MV_ERR_ARG, "mv_vegas_digest: idx required when training");
  if (f_stride < n_out) return fail(MV_ERR_ARG, "mv_vegas_digest: f_stride < n_out");
  if (w_stride < 0) return fail(MV_ERR_ARG, "mv_vegas_digest: negative w_stride");
  RowsK R;
  if ((rc = make_rows(parts, mv_partial_stride(step->n_dim, n_out, step->train),
                      "mv_vegas_digest", &R)))
    return rc;
  const StepK k = make_stepk(*step);
  hipStream_t st = as_stream(stream);
  const int d = step->n_dim;
  if (step->train && n_out == 1 && d >= kPairMinDim && d <= kPairMaxDim) {
    const size_t lds = (size_t)pair_cells(d) * sizeof(double);
    if (lds <= max_lds_per_block()) {
      const int b = step->precision == MV_F64
          ? launch_producer<kPairBlock>(vegas_digest_pairs<double, kPairBlock>, step->n_events, lds, st, k,
                                        static_cast<const double*>(f), f_stride,
                                        static_cast<const double*>(w), w_stride, idx, R)
          : launch_producer<kPairBlock>(vegas_digest_pairs<float, kPairBlock>, step->n_events, lds, st, k,
                                        static_cast<const float*>(f), f_stride,
                                        static_cast<const float*>(w), w_stride, idx, R);
      if ((rc = check_launch("mv_vegas_digest"))) return rc;
      return rows_used(b);
    }
  }
  const int blocks = producer_blocks(step->n_events);
  const int slices = step->train ? (step->n_dim + kDigestDims - 1) / kDigestDims : 1;
  if (slices > 65535) return fail(MV_ERR_UNSUPPORTED, "mv_vegas_digest: n_dim %d too large", step->n_dim);
  const dim3 g(blocks, slices);
  const bool vec = n_out > 1, train = step->train != 0;
#define MV_DIGEST(FT, V, TR)                                                                   \
  hipLaunchKernelGGL((vegas_digest<FT, V, TR>), g, dim3(kBlock), 0, st, k,                     \
                     static_cast<const FT*>(f), f_stride, n_out, main_dim,                     \
                     static_cast<const FT*>(w), w_stride, idx, R)
  if (step->precision == MV_F64) {
    if (vec) { if (train) MV_DIGEST(double, true, true); else MV_DIGEST(double, true, false); }
    else { if (train) MV_DIGEST(double, false, true); else MV_DIGEST(double, false, false); }
  } else {
    if (vec) { if (train) MV_DIGEST(float, true, true); else MV_DIGEST(float, true, false); }
    else { if (train) MV_DIGEST(float, false, true); else MV_DIGEST(float, false, false); }
  }
#undef MV_DIGEST
  if ((rc = check_launch("mv_vegas_digest"))) return rc;
  return rows_used(blocks);
}

extern "C" int mv_vegas_finalize(const mv_finalize* fz, double* rows, int64_t row_stride,
                                 int32_t n_rows, double* grid, double* out, double* hist_out,
                                 void* stream) {
  if (!fz) return fail(MV_ERR_ARG, "mv_vegas_finalize: fz is NULL");
  if (fz->n_dim < 1 || fz->n_out < 1 || fz->n_out > MV_MAX_OUT)
    return fail(MV_ERR_ARG, "mv_vegas_finalize: bad n_dim/n_out");
  if (!rows || !out || n_rows < 1 || n_rows > MV_ROWS)
    return fail(MV_ERR_ARG, "mv_vegas_finalize: rows/out missing or n_rows %d not in 1..%d", n_rows, MV_ROWS);
  if (row_stride < mv_partial_stride(fz->n_dim, fz->n_out, fz->train))
    return fail(MV_ERR_ARG, "mv_vegas_finalize: row_stride too small");
  if (fz->refine && (!fz->train || !grid))
    return fail(MV_ERR_ARG, "mv_vegas_finalize: refine needs train rows and a grid");
  if (hist_out && !fz->train) return fail(MV_ERR_ARG, "mv_vegas_finalize: no histogram without train");
  if (fz->history && fz->history_len < 0) return fail(MV_ERR_ARG, "mv_vegas_finalize: history_len < 0");
  FzK z{fz->n_dim, fz->n_out, fz->train, fz->refine, fz->raw, fz->n_events,
        CounterK{fz->draw_dev, fz->history, fz->history_len}, fz->status};
  const int blocks = 1 + (fz->train ? fz->n_dim : 0);  // histogram columns are always consumed
  hipLaunchKernelGGL(finalize_kernel, dim3(blocks), dim3(kBlock), 0, as_stream(stream), z, rows,
                     row_stride, n_rows, grid, out, hist_out);
  return check_launch("mv_vegas_finalize");
}

extern "C" int mv_refine_grid(int32_t n_dim, const double* hist, int64_t hist_stride, double* grid,
                              void* stream) {
  if (n_dim < 1 || !hist || !grid || hist_stride < kBins)
    return fail(MV_ERR_ARG, "mv_refine_grid: bad arguments");
  hipLaunchKernelGGL(refine_kernel, dim3(n_dim), dim3(64), 0, as_stream(stream), hist, hist_stride, grid);
  return check_launch("mv_refine_grid");
}

extern "C" int mv_consume_array_into_indices(const double* input, const int32_t* indices, int64_t n,
                                             int32_t result_size, double* out, void* stream) {
  if (result_size < 0 || n < 0) return fail(MV_ERR_ARG, "mv_consume_array_into_indices: bad sizes");
  if (result_size > 0 && !out) return fail(MV_ERR_ARG, "mv_consume_array_into_indices: NULL out");
  if (n > 0 && (!input || !indices)) return fail(MV_ERR_ARG, "mv_consume_array_into_indices: NULL input");
  hipStream_t st = as_stream(stream);
  if (result_size == 0) return MV_OK;
  hipLaunchKernelGGL(fill_kernel, dim3((result_size + 255) / 256), dim3(256), 0, st, out,
                     (int64_t)result_size, 0.0);
  if (n > 0) {
    const int blocks = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 1024);
    hipLaunchKernelGGL(consume_kernel, dim3(blocks), dim3(kBlock), 0, st, input, indices, n,
                       result_size, out);
  }
  return check_launch("mv_consume_array_into_indices");
}

// ---------------------------------------------------------------------------
// One iteration without any allocation (mv_vegas_iteration_static)
// ---------------------------------------------------------------------------
// The partial rows of mv_vegas_fused + mv_vegas_finalize live in a
// file-scope device array instead of a caller workspace: MV_ROWS rows of
// kStaticStride doubles (2 sums + MV_FUSED_MAX_DIM * MV_BINS histogram
// bins), 205 KB held by every GPU the library runs on until the process
// ends. Host code never passes the array's name to a kernel: that is a
// host-side stand-in, not a device address. The address of each GPU's copy
// comes from hipGetSymbolAddress, once per device with that device current,
// and is kept per device.
namespace {

constexpr int64_t kStaticStride = 2 + (int64_t)MV_FUSED_MAX_DIM * MV_BINS;
constexpr int64_t kStaticDoubles = (int64_t)MV_ROWS * kStaticStride;
constexpr int kStaticDevices = 64;

}  // namespace

static __device__ __align__(256) double scratch[kStaticDoubles];

namespace {

std::mutex g_static_mu;
double* g_static_rows[kStaticDevices] = {};

// This device's copy of ``scratch``. The first call on a device looks up
// its address and zeroes it on ``st`` (a fill kernel, ordered before the
// caller's work); every finalize zeroes the rows it consumes, so the rows
// are zero again between calls.
int static_rows(hipStream_t st, double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(MV_ERR_DEVICE, "mv_vegas_iteration_static: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kStaticDevices)
    return fail(MV_ERR_DEVICE, "mv_vegas_iteration_static: device %d beyond %d", dev, kStaticDevices);
  std::lock_guard<std::mutex> g(g_static_mu);
  if (!g_static_rows[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MV_ERR_DEVICE, "mv_vegas_iteration_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
    double* rows = static_cast<double*>(p);
    hipLaunchKernelGGL(fill_kernel, dim3((kStaticDoubles + kBlock - 1) / kBlock), dim3(kBlock), 0, st, rows,
                       kStaticDoubles, 0.0);
    int rc = check_launch("mv_vegas_iteration_static");
    if (rc) return rc;
    g_static_rows[dev] = rows;
  }
  *out = g_static_rows[dev];
  return MV_OK;
}

}  // namespace

extern "C" int mv_vegas_iteration_static(const mv_step* step, const mv_integrand* fn, double* grid,
                                         int32_t refine, double* out, double* hist_out, uint32_t* status,
                                         void* stream) {
  int rc = validate_step(step, "mv_vegas_iteration_static", MV_FUSED_MAX_DIM);
  if (rc) return rc;
  if ((rc = validate_fn(fn, "mv_vegas_iteration_static"))) return rc;
  if (step->n_events < 1) return fail(MV_ERR_ARG, "mv_vegas_iteration_static: n_events must be >= 1");
  if (!grid || !out) return fail(MV_ERR_ARG, "mv_vegas_iteration_static: grid and out are required");
  if ((refine || hist_out) && !step->train)
    return fail(MV_ERR_ARG, "mv_vegas_iteration_static: refine and hist_out need a training step");
  hipStream_t st = as_stream(stream);
  double* rows = nullptr;
  if ((rc = static_rows(st, &rows))) return rc;
  const mv_partials parts{rows, kStaticStride};
  const int n_rows = mv_vegas_fused(step, fn, grid, &parts, stream);
  if (n_rows < 0) return n_rows;
  mv_finalize fz{};
  fz.n_dim = step->n_dim;
  fz.n_out = 1;
  fz.train = step->train != 0;
  fz.refine = refine != 0;
  fz.n_events = (double)step->n_events;
  fz.status = status;
  return mv_vegas_finalize(&fz, rows, kStaticStride, n_rows, grid, out, hist_out, stream);
}
