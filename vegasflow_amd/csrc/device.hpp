// device.hpp - gfx950 device helpers shared by the VEGAS kernels:
// Philox4x32-10, draw conversion, wave/block reductions, built-in
// integrand families evaluated one dimension at a time.
//
// Everything here is compiled with -ffp-contract=off: a * b + c is two
// rounded operations unless fma() is written, so the index arithmetic
// reproduces the reference's operation order exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mivegas.h"

namespace mv {

constexpr int kBins = MV_BINS;                     // configflow.py:13
constexpr double kTechCut = 1e-8;                  // configflow.py:16
constexpr double kSpan = (1.0 - kTechCut) - kTechCut;  // tf.random.uniform max - min
constexpr double kFBins = 50.0;                    // FBINS, vflow.py:30
constexpr int kBlock = 256;                        // threads per workgroup
constexpr int kWaves = kBlock / 64;
constexpr int kMaxRows = 2048;                     // partial rows per producer launch

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11). The key schedule depends only on the
// seed, which is wave-uniform, so it stays in scalar registers. Each round's
// two three-way XORs are one v_bitop3_b32 each (truth table 0x96 = a^b^c,
// new in gfx950) instead of two v_xor_b32.
// ---------------------------------------------------------------------------
struct Key {
  uint32_t k0, k1;
};

__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        Key key) {
  uint32_t k0 = key.k0, k1 = key.k1;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);  // hi1 ^ c1 ^ k0
    c1 = lo1;
    c2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);  // hi0 ^ c3 ^ k1
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return make_uint4(c0, c1, c2, c3);
}

__device__ __forceinline__ uint32_t word(const uint4& v, int q) {
  return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}

// FP64 draw in [TECH_CUT, 1 - TECH_CUT]. The 52 mantissa bits of a double
// in [1, 2) are the top 52 bits of the 64-bit word hi:lo, r = that - 1 in
// [0, 1) (exact), then r * span + TECH_CUT as tf.random.uniform orders it
// (monte_carlo.py:264-266).
__device__ __forceinline__ double draw_f64(uint32_t hi, uint32_t lo) {
  const uint32_t mlo = __builtin_amdgcn_alignbit(hi, lo, 12);  // (hi:lo) >> 12, low word
  const uint32_t mhi = 0x3FF00000u | (hi >> 12);
  const double r = __hiloint2double((int)mhi, (int)mlo) - 1.0;
  return r * kSpan + kTechCut;
}

// Mixed-precision position of a draw in one dimension, from the 32-bit word
// hw alone, in the mirrored coordinate m = 50 - xn (xn = 50 (1 - u),
// vflow.py:117, counted from the top of the unit interval), to 23 fraction
// bits: q = floor((K (hw + 1/2) + 5 2^32) / 2^32) with K = 50 2^23 - 10, the
// high word of one 64-bit multiply-add (v_mad_u64_u32). q runs from 5 to
// 50 2^23 - 6, so u = q / (50 2^23) stays inside [TECH_CUT, 1 - TECH_CUT]
// like the reference's draws (monte_carlo.py:264-266), and the point tracks
// the FP64 draw of the same word to about 2^-22 of a bin. Returns the
// mirrored bin k = q >> 23 (always 0..49; the bin is 49 - k) and sets
// one_plus_mfrac to 1 + mfrac, the float whose mantissa is the 23 fraction
// bits of q: the fraction within bin 49 - k is 1 - mfrac (mfrac = 0: the
// bin's upper edge).
__device__ __forceinline__ int mixed_mirrored(uint32_t hw, float& one_plus_mfrac) {
  constexpr uint64_t kMul = (50ull << 23) - 10ull;
  constexpr uint64_t kAdd = (5ull << 32) + (kMul >> 1);
  const uint32_t q = (uint32_t)(((uint64_t)hw * kMul + kAdd) >> 32);
  one_plus_mfrac = __uint_as_float((q & 0x7FFFFFu) | 0x3F800000u);
  return (int)(q >> 23);
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bcast_lane(double v, int lane) {
  // lane is wave-uniform
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// The 64-bit value of the lane CTRL selects, through DPP moves of both
// 32-bit halves; lanes outside ROW_MASK, or whose source lies outside their
// row (row_shr), read 0. EVERY: every lane has a source (quad_perm,
// row_ror with the full mask), so no old value is needed.
template <int CTRL, int ROW_MASK, bool EVERY = false>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const int l = (int)(uint32_t)v, h = (int)(uint32_t)(v >> 32);
  int lo, hi;
  if (EVERY) {
    lo = __builtin_amdgcn_mov_dpp(l, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(h, CTRL, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, l, CTRL, ROW_MASK, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, h, CTRL, ROW_MASK, 0xF, false);
  }
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

template <int CTRL, int ROW_MASK, bool EVERY = false>
__device__ __forceinline__ double dpp_f64(double v) {
  return __longlong_as_double((long long)dpp_u64<CTRL, ROW_MASK, EVERY>((uint64_t)__double_as_longlong(v)));
}

// Sum of v over the 64 lanes of a wave, returned in every lane; the whole
// wave must be active. Data moves by DPP within the VALU (no LDS round trip):
// lane pairs and quads (quad_perm), the 16 lanes of a row (row_ror 4, 8),
// then rows 0 + 1 and 2 + 3 (row_bcast 15 into rows 1 and 3) and all four
// (row_bcast 31 into rows 2 and 3); the total ends in lane 63. A fixed
// order of additions.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1, 0xF, true>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E, 0xF, true>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x124, 0xF, true>(v);  // row_ror 4
  v += dpp_f64<0x128, 0xF, true>(v);  // row_ror 8
  v += dpp_f64<0x142, 0xA>(v);        // row_bcast 15 into rows 1 and 3
  v += dpp_f64<0x143, 0xC>(v);        // row_bcast 31 into rows 2 and 3
  return bcast_lane(v, 63);
}

// Inclusive prefix sum over the lanes (lane l: lanes 0 .. l, added in a
// fixed order) by DPP, for double or int64_t; the whole wave must be active.
// Within each row of 16 lanes row_shr 1, 2, 4, 8 (lanes shifted in from
// outside the row add 0), then row 0's total into row 1 and row 2's into
// row 3 (row_bcast 15), then the total of rows 0 and 1 into rows 2 and 3
// (row_bcast 31).
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  static_assert(sizeof(T) == 8, "64-bit scans only");
  auto step = [](T x, uint64_t moved) {
    T y;
    __builtin_memcpy(&y, &moved, 8);
    return x + y;
  };
  auto bits = [](T x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
  };
  v = step(v, dpp_u64<0x111, 0xF>(bits(v)));  // row_shr 1
  v = step(v, dpp_u64<0x112, 0xF>(bits(v)));  // row_shr 2
  v = step(v, dpp_u64<0x114, 0xF>(bits(v)));  // row_shr 4
  v = step(v, dpp_u64<0x118, 0xF>(bits(v)));  // row_shr 8
  v = step(v, dpp_u64<0x142, 0xA>(bits(v)));  // row_bcast 15 into rows 1 and 3
  v = step(v, dpp_u64<0x143, 0xC>(bits(v)));  // row_bcast 31 into rows 2 and 3
  return v;
}

// LDS float add without return (ds_add_f64 / ds_add_f32 on gfx950).
__device__ __forceinline__ void lds_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------------------
// built-in integrands, streamed one dimension at a time so that no x array
// is ever held: begin(), add(j, x_j), value().
//
// The fused kernels keep, per (dimension, bin), the affine map frac -> y
// with y = P + Q * frac, where y is x_j (integration limits folded in) or,
// for the Gaussian, already (x_j - c_j) / a: fold() turns the x map into the
// y map once per workgroup and add_folded(j, y) consumes y.
// ---------------------------------------------------------------------------
template <typename T, int FAM>
struct Integrand;

template <typename T>
struct Integrand<T, MV_FN_CONSTANT> {
  T c;
  __device__ explicit Integrand(const mv_integrand& p) : c((T)p.comp[0][0]) {}
  __device__ static void fold(const mv_integrand&, int, double&, double&) {}
  __device__ static double fold_scale(const mv_integrand&, int) { return 1.0; }
  __device__ __forceinline__ void begin() {}
  __device__ __forceinline__ void add(int, T) {}
  __device__ __forceinline__ void add_folded(int, T) {}
  __device__ __forceinline__ T value() const { return c; }
};

template <typename T>
struct Integrand<T, MV_FN_PRODUCT> {
  T p;
  __device__ explicit Integrand(const mv_integrand&) : p(1) {}
  __device__ static void fold(const mv_integrand&, int, double&, double&) {}
  __device__ static double fold_scale(const mv_integrand&, int) { return 1.0; }
  __device__ __forceinline__ void begin() { p = (T)1; }
  __device__ __forceinline__ void add(int, T x) { p *= x; }
  __device__ __forceinline__ void add_folded(int j, T x) { add(j, x); }
  __device__ __forceinline__ T value() const { return p; }
};

template <typename T>
struct Integrand<T, MV_FN_GAUSSIAN> {
  const mv_integrand* prm;
  T inv_a, pref, coef;
  __device__ explicit Integrand(const mv_integrand& p)
      : prm(&p), inv_a((T)p.comp[0][1]), pref((T)p.comp[0][2]), coef(0) {}
  __device__ static void fold(const mv_integrand& p, int j, double& P, double& Q) {
    P = (P - p.centre[0][j]) * p.comp[0][1];
    Q = Q * p.comp[0][1];
  }
  // the factor fold() applies to Q
  __device__ static double fold_scale(const mv_integrand& p, int) { return p.comp[0][1]; }
  __device__ __forceinline__ void begin() { coef = (T)0; }
  __device__ __forceinline__ void add(int j, T x) {
    const T t = (x - (T)prm->centre[0][j]) * inv_a;
    coef = coef + t * t;
  }
  __device__ __forceinline__ void add_folded(int, T y) { coef = fma(y, y, coef); }
  __device__ __forceinline__ T value() const;
};
template <>
__device__ __forceinline__ double Integrand<double, MV_FN_GAUSSIAN>::value() const {
  return pref * exp(-coef);
}
template <>
__device__ __forceinline__ float Integrand<float, MV_FN_GAUSSIAN>::value() const {
  return pref * __builtin_amdgcn_exp2f(-coef * 1.4426950408889634f);
}

template <typename T>
struct Integrand<T, MV_FN_GAUSSMIX> {
  const mv_integrand* prm;
  int nc;
  T coef[MV_MAX_COMP];
  __device__ explicit Integrand(const mv_integrand& p) : prm(&p), nc(p.n_comp) {}
  __device__ static void fold(const mv_integrand&, int, double&, double&) {}
  __device__ static double fold_scale(const mv_integrand&, int) { return 1.0; }
  __device__ __forceinline__ void add_folded(int j, T x) { add(j, x); }
  __device__ __forceinline__ void begin() {
#pragma unroll
    for (int k = 0; k < MV_MAX_COMP; ++k) coef[k] = (T)0;
  }
  __device__ __forceinline__ void add(int j, T x) {
#pragma unroll
    for (int k = 0; k < MV_MAX_COMP; ++k) {
      if (k < nc) {
        const T t = (x - (T)prm->centre[k][j]) * (T)prm->comp[k][1];
        coef[k] = coef[k] + t * t;
      }
    }
  }
  __device__ __forceinline__ T value() const;
};
template <>
__device__ __forceinline__ double Integrand<double, MV_FN_GAUSSMIX>::value() const {
  double f = 0.0;
#pragma unroll
  for (int k = 0; k < MV_MAX_COMP; ++k)
    if (k < nc) f = f + prm->comp[k][0] * (prm->comp[k][2] * exp(-coef[k]));
  return f;
}
template <>
__device__ __forceinline__ float Integrand<float, MV_FN_GAUSSMIX>::value() const {
  float f = 0.0f;
#pragma unroll
  for (int k = 0; k < MV_MAX_COMP; ++k)
    if (k < nc)
      f = f + (float)prm->comp[k][0] *
                  ((float)prm->comp[k][2] * __builtin_amdgcn_exp2f(-coef[k] * 1.4426950408889634f));
  return f;
}

template <typename T>
struct Integrand<T, MV_FN_BWPOWER> {
  const mv_integrand* prm;
  T bw, pw;
  __device__ explicit Integrand(const mv_integrand& p) : prm(&p), bw(1), pw(1) {}
  __device__ static void fold(const mv_integrand&, int, double&, double&) {}
  __device__ static double fold_scale(const mv_integrand&, int) { return 1.0; }
  __device__ __forceinline__ void add_folded(int j, T x) { add(j, x); }
  __device__ __forceinline__ void begin() {
    bw = (T)1;
    pw = (T)1;
  }
  __device__ __forceinline__ void add(int j, T x);
  __device__ __forceinline__ T value() const { return bw * pw; }
};
template <>
__device__ __forceinline__ void Integrand<double, MV_FN_BWPOWER>::add(int j, double x) {
  if (j == 0) {
    const double d = x - prm->comp[0][0];
    const double g2 = prm->comp[0][1];
    bw = prm->comp[0][2] / (d * d + g2 * g2);
  } else if (j == 1) {
    pw = pow(x, prm->comp[0][3]);
  }
}
template <>
__device__ __forceinline__ void Integrand<float, MV_FN_BWPOWER>::add(int j, float x) {
  if (j == 0) {
    const float d = x - (float)prm->comp[0][0];
    const float g2 = (float)prm->comp[0][1];
    bw = (float)prm->comp[0][2] / (d * d + g2 * g2);
  } else if (j == 1) {
    pw = __builtin_amdgcn_exp2f((float)prm->comp[0][3] * __builtin_amdgcn_logf(x));
  }
}

}  // namespace mv
