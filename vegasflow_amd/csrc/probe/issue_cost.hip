// issue_cost.hip - measures the issue cost, in shader cycles per wave64
// instruction per SIMD, of the VALU instruction kinds that the fused kernels
// are made of (64-bit integer multiply-add and three-input XOR of Philox,
// FP32 and FP64 arithmetic, conversions, the FP32 exponential, shifts), and
// the cost per CU of their LDS instructions (map-table reads, pair-cell
// FP64 atomic adds) at the pair kernel's launch shape.
//
// Built into its own library (libmvprobe.so), not into libmivegas.so: it
// serves the measurement of the fused kernels' instruction bound, not the
// integration path. Each kernel runs 8 independent dependency chains of one
// instruction kind per thread over 4096 x 256 threads (16 waves per SIMD),
// so the time is set by issue throughput, not latency. The time comes from
// HIP events around the launch; the shader clock from wave 0's s_memtime
// against the 100 MHz s_memrealtime over its own run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../device.hpp"

namespace {

constexpr int kChains = 8;
constexpr int kBlocks = 4096;
constexpr int kThreads = 256;

struct Clock {
  unsigned long long t0, t1, r0, r1;
};

__device__ __forceinline__ void clock_begin(Clock* c) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->t0 = __builtin_amdgcn_s_memtime();
    c->r0 = __builtin_amdgcn_s_memrealtime();
  }
}

__device__ __forceinline__ void clock_end(Clock* c) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->t1 = __builtin_amdgcn_s_memtime();
    c->r1 = __builtin_amdgcn_s_memrealtime();
  }
}

// kernel NAME: kChains chains of element type T, each advanced ``iters`` times by STEP
#define MV_PROBE(NAME, T, INIT, STEP, FOLD)                                    \
  __global__ void NAME(uint32_t* out, Clock* clk, int iters, uint32_t K) {     \
    clock_begin(clk);                                                          \
    T a[kChains], b[kChains];                                                  \
    for (int c = 0; c < kChains; ++c) { INIT; }                                \
    for (int i = 0; i < iters; ++i) {                                          \
      _Pragma("unroll") for (int c = 0; c < kChains; ++c) { STEP; }            \
    }                                                                          \
    uint32_t s = 0;                                                            \
    for (int c = 0; c < kChains; ++c) { FOLD; }                                \
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;                            \
    clock_end(clk);                                                            \
  }

MV_PROBE(p_mad_u64, uint64_t,
         (a[c] = ((uint64_t)(blockIdx.x + c) << 32) | (threadIdx.x * 7 + c), b[c] = 0),
         a[c] = (uint64_t)(uint32_t)a[c] * K + a[c], s ^= (uint32_t)a[c] ^ (uint32_t)(a[c] >> 32))
MV_PROBE(p_xor3, uint32_t, (a[c] = threadIdx.x * 7 + c, b[c] = blockIdx.x + c),
         (a[c] = __builtin_amdgcn_bitop3_b32(a[c], b[c], i * K, 0x96),
          b[c] = __builtin_amdgcn_bitop3_b32(b[c], a[c], i + K, 0x96)),
         s ^= a[c] ^ b[c])
MV_PROBE(p_fma_f32, float, (a[c] = threadIdx.x * 1e-3f + c, b[c] = 0.999f),
         (a[c] = __builtin_fmaf(a[c], b[c], 1e-4f * K), a[c] = __builtin_fmaf(a[c], b[c], 2e-4f)),
         s ^= __float_as_uint(a[c]))
MV_PROBE(p_mul_f32, float, (a[c] = threadIdx.x * 1e-3f + c, b[c] = 0.999f + 1e-7f * K),
         (a[c] = a[c] * b[c], a[c] = a[c] * b[c]), s ^= __float_as_uint(a[c]))
MV_PROBE(p_fma_f64, double, (a[c] = threadIdx.x * 1e-3 + c, b[c] = 0.999),
         (a[c] = __builtin_fma(a[c], b[c], 1e-4 * K), a[c] = __builtin_fma(a[c], b[c], 2e-4)),
         s ^= (uint32_t)__double_as_longlong(a[c]))
MV_PROBE(p_add_f64, double, (a[c] = threadIdx.x * 1e-3 + c, b[c] = 1e-9 * K),
         (a[c] = a[c] + b[c], a[c] = a[c] + b[c]), s ^= (uint32_t)__double_as_longlong(a[c]))
MV_PROBE(p_cvt, float, (a[c] = threadIdx.x * 1e-3f + c, b[c] = 0),
         a[c] = (float)((double)a[c] * (1.0 + 1e-9 * K)), s ^= __float_as_uint(a[c]))
MV_PROBE(p_exp_f32, float, (a[c] = threadIdx.x * 1e-6f + c * 1e-3f, b[c] = 0),
         a[c] = __builtin_amdgcn_exp2f(a[c]) * -0.5f, s ^= __float_as_uint(a[c]))
MV_PROBE(p_shift_add, uint32_t, (a[c] = threadIdx.x * 7 + c, b[c] = blockIdx.x + c),
         (a[c] = (a[c] << 3) + b[c], b[c] = (b[c] << 5) + a[c]), s ^= a[c] ^ b[c])

// Philox4x32-10 blocks of the fused kernels, four independent per trip,
// folded with two three-input XORs each
__global__ void p_philox(uint32_t* out, Clock* clk, int iters, uint32_t K) {
  clock_begin(clk);
  uint32_t acc = 0;
  const uint32_t c0 = blockIdx.x * blockDim.x + threadIdx.x;
  const mv::Key key{K * 0x9E3779B9u, K ^ 0x1234567u};
  for (int i = 0; i < iters / 4; ++i) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint4 h = mv::philox(c0, (uint32_t)i, 7u, (uint32_t)c, key);
      acc = __builtin_amdgcn_bitop3_b32(acc, h.x, h.y, 0x96);
      acc = __builtin_amdgcn_bitop3_b32(acc, h.z, h.w, 0x96);
    }
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
  clock_end(clk);
}

// ---- LDS: one 1024-thread workgroup per CU, 83.2 KB of dynamic LDS laid
// out as the d = 8 pair kernel's (400 float2 map entries, then 4 tables of
// 50 x 50 FP64 cells), so that the CU's LDS serves 16 waves as it does there.
constexpr int kLdsBytes = 83200;
constexpr int kTabBytes = 3200;
constexpr int kCells = 2500;

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// MODE 0: 8 reads per trip of random map entries (entry 50 r + bin, r = 0..7)
// MODE 1: 8 reads of lane-consecutive entries (no bank conflict)
// MODE 2: 4 FP64 atomic adds per trip to random cells of the 4 pair tables
// MODE 3: 4 FP64 atomic adds to lane-consecutive cells (no bank conflict)
// MODE 4: the pair kernel's LDS work per event: modes 0 and 2 together
// Addresses change between trips by a wave-uniform XOR of bits 3..5 of the
// byte offset, which keeps each one inside its 64-byte block: reads below
// byte 3200 of the map table, adds below byte 80000 of the cells.
template <int MODE>
__global__ __launch_bounds__(1024) void p_lds(uint32_t* out, Clock* clk, int iters, uint32_t K) {
  extern __shared__ double smem[];
  float2* tab = reinterpret_cast<float2*>(smem);
  char* H = reinterpret_cast<char*>(smem) + kTabBytes;
  for (int i = threadIdx.x; i < kLdsBytes / 8; i += 1024) smem[i] = 0.0;
  __syncthreads();
  for (int i = threadIdx.x; i < kTabBytes / 8; i += 1024) tab[i] = make_float2((float)i, 0.5f * i);
  __syncthreads();
  clock_begin(clk);
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t h0 = mix32(K ^ (blockIdx.x * 1024 + threadIdx.x) * 0x9E3779B9u);
  uint32_t ra[8], ca[4];
#pragma unroll
  for (int r = 0; r < 8; ++r)
    ra[r] = (uint32_t)(r * 50 + (MODE == 1 ? lane % 50 : mix32(h0 + r) % 50u)) * 8u;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const uint32_t cell = MODE == 3 ? threadIdx.x % kCells
                                    : (mix32(h0 + 17 * p + 3) % 50u) * 50u + mix32(h0 + 17 * p + 5) % 50u;
    ca[p] = (uint32_t)(p * kCells + cell) * 8u;
  }
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const double v = 1.0 + lane;
  for (int it = 0; it < iters; ++it) {
    const uint32_t k = ((uint32_t)it & 7u) << 3;
    if (MODE == 0 || MODE == 1 || MODE == 4) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const float2 t = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(tab) + (ra[r] ^ k));
        acc[r] = __builtin_fmaf(t.x, t.y, acc[r]);
      }
    }
    if (MODE >= 2) {
#pragma unroll
      for (int p = 0; p < 4; ++p) mv::lds_add(reinterpret_cast<double*>(H + (ca[p] ^ k)), v);
    }
  }
  __syncthreads();
  float a = 0.f;
#pragma unroll
  for (int r = 0; r < 8; ++r) a += acc[r];
  out[blockIdx.x * 1024 + threadIdx.x] = __float_as_uint(a) ^ (uint32_t)__double2loint(reinterpret_cast<const double*>(H)[threadIdx.x]);
  clock_end(clk);
}

typedef void (*LdsProbe)(uint32_t*, Clock*, int, uint32_t);
const LdsProbe kLdsKinds[] = {p_lds<0>, p_lds<1>, p_lds<2>, p_lds<3>, p_lds<4>};
const double kLdsOps[] = {8, 8, 4, 4, 12};  // LDS wave instructions per trip

typedef void (*Probe)(uint32_t*, Clock*, int, uint32_t);

struct Kind {
  Probe fn;
  double per_step;  // wave instructions of the kind per chain step (Philox: blocks per trip / chains)
};

const Kind kKinds[] = {
    {p_mad_u64, 1}, {p_xor3, 2},    {p_fma_f32, 2}, {p_mul_f32, 2},     {p_fma_f64, 2},
    {p_add_f64, 2}, {p_cvt, 3},     {p_exp_f32, 2}, {p_shift_add, 2},  {p_philox, 1.0 / kChains},
};

}  // namespace

extern "C" {

// Number of probe kinds.
int mvp_kinds(void) { return (int)(sizeof(kKinds) / sizeof(kKinds[0])); }

// Runs probe ``kind`` with ``iters`` chain steps per thread (a warm-up launch
// first) and stores the launch time in ms, the shader clock in GHz, and the
// wave instructions of the kind per SIMD (4 SIMDs per CU; Philox: blocks
// per SIMD).
// Returns 0, or -1 for an unknown kind / a HIP error.
int mvp_issue_cost(int kind, int iters, double* ms, double* ghz, double* per_simd) {
  if (kind < 0 || kind >= mvp_kinds() || iters < 4) return -1;
  uint32_t* out = nullptr;
  Clock* clk = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1, dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    return -1;
  if (hipMalloc(&out, sizeof(uint32_t) * kBlocks * kThreads) != hipSuccess) return -1;
  if (hipMalloc(&clk, sizeof(Clock)) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  for (int rep = 0; rep < 2; ++rep) {
    if (hipEventRecord(e0, 0) != hipSuccess) goto done;
    hipLaunchKernelGGL(kKinds[kind].fn, dim3(kBlocks), dim3(kThreads), 0, 0, out, clk, iters, 3u);
    if (hipGetLastError() != hipSuccess) goto done;
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
  }
  {
    float t = 0.f;
    Clock h;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) goto done;
    if (hipMemcpy(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost) != hipSuccess) goto done;
    *ms = t;
    *ghz = (double)(h.t1 - h.t0) / (double)(h.r1 - h.r0) * 0.1;
    *per_simd = (double)kBlocks * kThreads / 64 * iters * kChains * kKinds[kind].per_step / (4.0 * cus);
    rc = 0;
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (clk) (void)hipFree(clk);
  (void)hipFree(out);
  return rc;
}

// Number of LDS probe kinds.
int mvp_lds_kinds(void) { return (int)(sizeof(kLdsKinds) / sizeof(kLdsKinds[0])); }

// Runs LDS probe ``kind`` (``iters`` trips per wave, one workgroup per CU;
// a warm-up launch first) and stores the launch time in ms, the shader
// clock in GHz, and the LDS wave instructions per CU.
// Returns 0, or -1 for an unknown kind / a HIP error.
int mvp_lds_cost(int kind, int iters, double* ms, double* ghz, double* per_cu) {
  if (kind < 0 || kind >= mvp_lds_kinds() || iters < 8) return -1;
  uint32_t* out = nullptr;
  Clock* clk = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1, dev = 0, cus = 0;
  const void* fn = reinterpret_cast<const void*>(kLdsKinds[kind]);
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    return -1;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes) != hipSuccess) return -1;
  if (hipMalloc(&out, sizeof(uint32_t) * cus * 1024) != hipSuccess) return -1;
  if (hipMalloc(&clk, sizeof(Clock)) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  for (int rep = 0; rep < 2; ++rep) {
    if (hipEventRecord(e0, 0) != hipSuccess) goto done;
    hipLaunchKernelGGL(kLdsKinds[kind], dim3(cus), dim3(1024), kLdsBytes, 0, out, clk, iters, 3u + rep);
    if (hipGetLastError() != hipSuccess) goto done;
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
  }
  {
    float t = 0.f;
    Clock h;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) goto done;
    if (hipMemcpy(&h, clk, sizeof(Clock), hipMemcpyDeviceToHost) != hipSuccess) goto done;
    *ms = t;
    *ghz = (double)(h.t1 - h.t0) / (double)(h.r1 - h.r0) * 0.1;
    *per_cu = 16.0 * iters * kLdsOps[kind];
    rc = 0;
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (clk) (void)hipFree(clk);
  (void)hipFree(out);
  return rc;
}

}  // extern "C"
