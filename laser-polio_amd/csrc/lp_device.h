// Device-side building blocks of the daily step: counter-based random numbers,
// float32 elementary functions with a fixed operation order, fixed-point
// weights and probability thresholds.  Compiled with -ffp-contract=off so every
// a*b+c below is two rounded operations unless written as fmaf().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lp {

// Philox stream ids (low 8 bits of the 4th counter word).
enum : uint32_t { ST_PARALYSIS = 1, ST_RI = 2, ST_SIA = 3, ST_STRAIN = 4, ST_KM = 5,
                  ST_BIRTH = 6, ST_COUNT = 7, ST_PROBE = 8 };

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with Weyl key bumps.
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  return make_uint4(c0, c1, c2, c3);
}

// Agent-keyed draw: counter = (index lo, index hi, day, stream).
__device__ __forceinline__ uint4 draw(uint64_t seed, uint64_t idx, int32_t t, uint32_t stream) {
  return philox((uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)t, stream, seed);
}

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// Fixed-point selection / infectiousness weight: rint(v * 2^24), 0 for v <= 0 or NaN.
__device__ __forceinline__ int64_t weight_q24(float v) {
  if (!(v > 0.0f)) return 0;
  return (int64_t)__builtin_rintf(v * 16777216.0f);
}

// ceil(p * 2^32) clamped to [0, 2^32]: r32 < thr  <=>  r32 * 2^-32 < p.
__device__ __forceinline__ uint64_t thr_from_prob(float p) {
  if (!(p > 0.0f)) return 0;
  const float s = p * 4294967296.0f;
  if (s >= 4294967296.0f) return 4294967296ull;
  return (uint64_t)ceilf(s);
}

// The IEEE-754 binary64 product y = fl(x * n) (round to nearest, ties to even)
// of a non-negative double x, given by its bit pattern, and an integer n >= 0,
// computed with integer arithmetic only: returns floor(y) and sets *thr =
// ceil((y - floor(y)) * 2^32), the Bernoulli threshold of its fractional part.
// x * n is exact in 128 bits (53 + 32 bits); it is rounded to 53 significant
// bits as the floating-point multiply would round it.  Negative, infinite and
// NaN x give 0.  Values below 2^-32 (subnormal results included) have floor 0
// and threshold 1, which no rounding below that scale changes.
__device__ __forceinline__ int64_t f64_mul_floor(uint64_t xbits, uint32_t n, uint64_t* thr) {
  *thr = 0;
  const uint32_t ef = (uint32_t)(xbits >> 52) & 0x7ffu;
  uint64_t m = xbits & ((1ull << 52) - 1);
  if ((xbits >> 63) || ef == 0x7ffu || n == 0) return 0;
  int32_t e2;
  if (ef == 0) e2 = -1074;
  else { m |= 1ull << 52; e2 = (int32_t)ef - 1075; }
  if (m == 0) return 0;
  const unsigned __int128 P = (unsigned __int128)m * n;
  const uint64_t hi = (uint64_t)(P >> 64), lo = (uint64_t)P;
  const int L = hi ? 128 - __clzll(hi) : 64 - __clzll(lo);   // bit length of P
  uint64_t M;
  int32_t ex = e2;
  if (L > 53) {
    int s = L - 53;
    M = (uint64_t)(P >> s);
    const unsigned __int128 R = P & ((((unsigned __int128)1) << s) - 1);
    const unsigned __int128 half = ((unsigned __int128)1) << (s - 1);
    if (R > half || (R == half && (M & 1))) ++M;
    if (M == (1ull << 53)) { M >>= 1; ++s; }
    ex += s;
  } else {
    M = lo;
  }
  const int lm = 64 - __clzll(M);                               // bit length of M
  if (ex + lm - 1 < -32) { *thr = 1; return 0; }               // 0 < y < 2^-32
  if (ex >= 0) return ex >= 10 ? INT64_MAX : (int64_t)(M << ex);
  const int sh = -ex;
  if (sh >= 64) {                                               // floor 0, y = M 2^-sh, sh - 32 < 64 here
    const int k = sh - 32;
    *thr = (M >> k) + ((M & ((1ull << k) - 1)) != 0);
    return 0;
  }
  const uint64_t ip = M >> sh, fb = M & ((1ull << sh) - 1);
  if (sh <= 32) *thr = fb << (32 - sh);
  else {
    const int k = sh - 32;
    *thr = (fb >> k) + ((fb & ((1ull << k) - 1)) != 0);
  }
  return (int64_t)ip;
}

__device__ __forceinline__ float u24(uint32_t r) { return (float)(r >> 8) * 5.9604644775390625e-08f; }

// ---- float32 elementary functions --------------------------------------------
// exp/expm1: Cody & Waite argument reduction x = k ln2 + r (ln2 split in a
// 15-bit head and a tail so k*head is exact), Taylor polynomial of expm1 on
// |r| <= ln2/2 through r^8/8!, reconstruction by ldexp.  log: x = 2^e f,
// f in (sqrt(1/2), sqrt(2)], log f = 2 atanh(s), s = (f-1)/(f+1), odd series
// through s^13.  Every step is one IEEE-rounded operation in a fixed order.
constexpr float LN2_HI = 0.693145751953125f;
constexpr float LN2_LO = 1.42860676533018672e-06f;
constexpr float LOG2E = 1.44269502162933349609f;

__device__ __forceinline__ float em1_poly(float r) {
  float p = 2.48015873015873e-05f;
  p = p * r + 1.98412698412698e-04f;
  p = p * r + 1.38888888888889e-03f;
  p = p * r + 8.33333333333333e-03f;
  p = p * r + 4.16666666666667e-02f;
  p = p * r + 1.66666666666667e-01f;
  p = p * r + 0.5f;
  return r + (r * r) * p;
}

__device__ __forceinline__ float expm1f_lp(float x) {
  if (x < -17.5f) return -1.0f;
  if (x > 88.0f) return __builtin_inff();
  if (fabsf(x) <= 0.34657359f) return em1_poly(x);
  const float kf = __builtin_rintf(x * LOG2E);
  const float r = (x - kf * LN2_HI) - kf * LN2_LO;
  const float e = em1_poly(r);
  const float s = ldexpf(1.0f, (int)kf);
  return __builtin_fmaf(s, e, s - 1.0f);
}

__device__ __forceinline__ float expf_lp(float x) {
  if (x < -87.0f) return 0.0f;
  if (x > 88.0f) return __builtin_inff();
  const float kf = __builtin_rintf(x * LOG2E);
  const float r = (x - kf * LN2_HI) - kf * LN2_LO;
  return ldexpf(1.0f + em1_poly(r), (int)kf);
}

__device__ __forceinline__ float logf_lp(float y) {
  uint32_t b = __float_as_uint(y);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float f = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
  if (f > 1.41421356f) { f = f * 0.5f; e += 1; }
  const float s = (f - 1.0f) / (f + 1.0f);
  const float z = s * s;
  float p = 0.0769230769231f;
  p = p * z + 0.0909090909091f;
  p = p * z + 0.111111111111f;
  p = p * z + 0.142857142857f;
  p = p * z + 0.2f;
  p = p * z + 0.333333333333f;
  const float lf = (s + s) + (s + s) * (z * p);
  const float ef = (float)e;
  return ef * LN2_HI + (ef * LN2_LO + lf);
}

__device__ __forceinline__ float log1pf_lp(float x) {
  const float y = 1.0f + x;
  if (y == 1.0f) return x;
  const float c = (y - 1.0f) - x;
  return logf_lp(y) - c / y;
}

// ---- wavefront helpers (wave64) -----------------------------------------------
__device__ __forceinline__ int32_t wave_sum_i32(int32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) { const int32_t u = __shfl_xor(v, o, 64); v = u < v ? u : v; }
  return v;
}
// Inclusive prefix sum across the wave (Hillis-Steele on 64 lanes).
__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

}  // namespace lp
