// gs_math.h — portable single-precision exp/log used by every kernel whose
// results must be reproducible bit for bit on the host.
//
// The GPU's expf/logf (ocml) and glibc's differ by an ULP here and there.
// The trainer's discrete decisions (cull if opa*comp < A_MIN, the tile
// rectangle from the opacity-aware radius, skip if alpha < A_MIN, stop when
// T < 1e-4) turn such ULP differences into different sort keys or different
// pixels. These two functions use only IEEE-754 basic operations (+, *, /,
// fused multiply-add, round-to-nearest-even) and integer bit manipulation,
// so a C compiler on x86-64 and hipcc on gfx950 produce identical bits from
// the same source, provided both compile with -ffp-contract=off.
//
//   exp: Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, then the
//        degree-7 Taylor polynomial of e^r in Horner form (truncation error
//        r^8/8! < 5.3e-9 relative), then scaling by 2^k through the
//        exponent field. Max error about 1 ulp.
//   log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//        log m = 2 atanh(s) = 2(s + s^3/3 + ... + s^9/9)
//        (|s| <= 0.1716, truncation < 4e-10), plus e*ln2 split in two parts.
//        Max error about 2 ulp.
//
// Included from C (the FP32 restatement) and from HIP device code.
#ifndef GS_MATH_H
#define GS_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define GS_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <string.h>
#define GS_HD static inline
#endif

GS_HD uint32_t gs_f2u(float x) {
#if defined(__HIPCC__)
  return __builtin_bit_cast(uint32_t, x);
#else
  uint32_t u; memcpy(&u, &x, 4); return u;
#endif
}

GS_HD float gs_u2f(uint32_t u) {
#if defined(__HIPCC__)
  return __builtin_bit_cast(float, u);
#else
  float x; memcpy(&x, &u, 4); return x;
#endif
}

// e^x, correctly handling NaN, overflow, underflow and denormal results.
GS_HD float gs_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283172607421875f) return gs_u2f(0x7f800000u);  // +inf
  if (x < -103.972084045410156f) return 0.0f;
  const float kf = rintf(x * 1.44269502162933349609375f);      // round(x/ln2)
  // ln2 = LN2_HI + LN2_LO, LN2_HI has 15 significant bits so k*LN2_HI is exact
  float r = fmaf(kf, -0.693145751953125f, x);
  r = fmaf(kf, -1.428606765330187045e-06f, r);
  float p = 1.984126984126984e-04f;          // 1/7!
  p = fmaf(p, r, 1.388888888888889e-03f);    // 1/6!
  p = fmaf(p, r, 8.333333333333333e-03f);    // 1/5!
  p = fmaf(p, r, 4.166666666666667e-02f);    // 1/4!
  p = fmaf(p, r, 1.666666666666667e-01f);    // 1/3!
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  int k = (int)kf;
  if (k > 127) {             // only k == 128 reaches here: p*2^127*2
    return (p * gs_u2f(0x7f000000u)) * 2.0f;
  }
  if (k < -126) {            // denormal result: one rounding, in the last multiply
    return (p * gs_u2f((uint32_t)(k + 126 + 127) << 23)) * gs_u2f(0x00800000u);
  }
  return p * gs_u2f((uint32_t)(k + 127) << 23);
}

// gs_expf for x in [-87, 87] without its range branches: there k =
// round(x/ln2) lies in [-126, 126] and gs_expf takes its normal-result path,
// whose operations these are, so the bits agree. For the clamped arguments
// of the scales (exp of [-12, 6]) and the sigmoids (exp of [-9, 9]).
GS_HD float gs_expf_mid(float x) {
  const float kf = rintf(x * 1.44269502162933349609375f);
  float r = fmaf(kf, -0.693145751953125f, x);
  r = fmaf(kf, -1.428606765330187045e-06f, r);
  float p = 1.984126984126984e-04f;
  p = fmaf(p, r, 1.388888888888889e-03f);
  p = fmaf(p, r, 8.333333333333333e-03f);
  p = fmaf(p, r, 4.166666666666667e-02f);
  p = fmaf(p, r, 1.666666666666667e-01f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const int k = (int)kf;
  return p * gs_u2f((uint32_t)(k + 127) << 23);
}

// e^x for the rasterizer's inner loops, x in [-87, 0] (the caller clamps;
// a pixel takes a splat only at x >= -eCut >= -80). Branch-free: the range
// reduction of gs_expf (k = round(x/ln2) in [-126, 0], a normal result),
// then a degree-5 polynomial for e^r on [-ln2/2, ln2/2] with the constant and
// linear coefficients fixed at 1 and the other four fitted for minimal
// relative error (Lawson's iteratively reweighted least squares on a dense
// grid, rounded to f32). Two fewer FMAs than gs_expf's degree-7 Taylor
// polynomial; at most 2 ulp from the correctly rounded e^x on every float of
// [-87, 0] (gs_expf: at most 1), inside WGSL's 3 + 2|x| ulp for exp.
GS_HD float gs_expf_raster(float xc) {
  const float kf = rintf(xc * 1.44269502162933349609375f);
  float r = fmaf(kf, -0.693145751953125f, xc);
  r = fmaf(kf, -1.428606765330187045e-06f, r);
  float p = gs_u2f(0x3c0840d6u);             // 8.316239e-03
  p = fmaf(p, r, gs_u2f(0x3d2b9b09u));       // 4.1895900e-02
  p = fmaf(p, r, gs_u2f(0x3e2aabc8u));       // 1.6667092e-01
  p = fmaf(p, r, gs_u2f(0x3efffeeeu));       // 4.9999183e-01
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const int k = (int)kf;
#if defined(__HIP_DEVICE_COMPILE__)
  // p 2^k in one v_ldexp_f32 instead of building 2^k and multiplying. Both
  // are exact and equal: k = -126 only for x < -86.99, where r >= 0.337 and
  // p >= 1.4, and p >= 0.707 otherwise, so p 2^k is always normal.
  return __builtin_amdgcn_ldexpf(p, k);
#else
  return p * gs_u2f((uint32_t)(k + 127) << 23);
#endif
}

// natural log; NaN for x < 0, -inf at 0, +inf at +inf.
GS_HD float gs_logf(float x) {
  if (x != x) return x;
  if (x < 0.0f) return gs_u2f(0x7fc00000u);
  if (x == 0.0f) return gs_u2f(0xff800000u);
  uint32_t u = gs_f2u(x);
  if (u == 0x7f800000u) return x;
  int e = 0;
  if (u < 0x00800000u) {     // denormal input: scale by 2^23 first
    x = x * 8388608.0f;
    u = gs_f2u(x);
    e = -23;
  }
  e += (int)(u >> 23) - 127;
  float m = gs_u2f((u & 0x007fffffu) | 0x3f800000u);   // [1, 2)
  if (m > 1.41421353816986083984375f) { m = m * 0.5f; e += 1; }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float p = 1.111111111111111e-01f;          // 1/9
  p = fmaf(p, s2, 1.428571428571429e-01f);   // 1/7
  p = fmaf(p, s2, 2.0e-01f);                 // 1/5
  p = fmaf(p, s2, 3.333333333333333e-01f);   // 1/3
  const float s2x = s + s;
  const float lm = fmaf(s2x * s2, p, s2x);
  const float ef = (float)e;
  return fmaf(ef, 0.693145751953125f, fmaf(ef, 1.428606765330187045e-06f, lm));
}

// ---- bfloat16 storage (mixed precision) ----
// bf16 keeps the f32 sign, exponent and top 7 mantissa bits. Values are
// decoded exactly; arithmetic stays f32. Stores round stochastically: the
// 16 bits below the kept mantissa are added to a random 16-bit number before
// truncation, so a value rounds up with probability equal to its distance
// from the lower neighbour. Small Adam updates (lr 1e-4 on O(0.1) SH
// coefficients, the 0.1% per-step decay of v) then survive in expectation
// instead of vanishing under round-to-nearest. Inf and NaN are truncated (a
// NaN keeps a mantissa bit set).

// 32-bit integer hash (xor-shift-multiply avalanche): the random bits of
// stochastic rounding come from a counter of (Gaussian, step, stream, slot).
GS_HD uint32_t gs_hash32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Seed of one (Gaussian, step): two full hash rounds.
GS_HD uint32_t gs_sr_seed(uint64_t gaussian, uint32_t step) {
  return gs_hash32((uint32_t)gaussian * 0x9E3779B1u ^ gs_hash32((uint32_t)(gaussian >> 32) + step * 0xC2B2AE3Du));
}

// Random bits of that seed for the pair of values k = 2p, 2p + 1 (k = 64 *
// stream + slot): one multiply-xorshift round; value 2p takes the low 16
// bits and 2p + 1 the high 16 (gs_sr_bits). Each step is a bijection of the
// word, so for a uniform seed both halves are exactly uniform.
GS_HD uint32_t gs_sr_pair(uint32_t seed, uint32_t p) {
  uint32_t x = seed ^ (p * 0x9E3779B9u);
  x *= 0x85ebca6bu;
  return x ^ (x >> 16);
}

// 16 random bits for value k of that seed
GS_HD uint32_t gs_sr_bits(uint32_t seed, uint32_t k) {
  const uint32_t x = gs_sr_pair(seed, k >> 1);
  return (k & 1u) ? (x >> 16) : (x & 0xffffu);
}

GS_HD float gs_bf2f(uint32_t b) { return gs_u2f((b & 0xffffu) << 16); }

// round to nearest even (initial upload of f32 data); NaN stays NaN
GS_HD uint32_t gs_f2bf_rne(float x) {
  const uint32_t u = gs_f2u(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// Adds 16 random bits below the kept mantissa and truncates. An infinity
// has a zero mantissa, so it cannot carry and stays infinite; a NaN keeps
// its sign and exponent and a set mantissa bit.
GS_HD uint32_t gs_f2bf_sr(float x, uint32_t rnd) {
  const uint32_t u = gs_f2u(x);
  const uint32_t r = (u + (rnd & 0xffffu)) >> 16;
  return x != x ? (u >> 16) | 0x40u : r;
}

// ---- IEEE binary16 (fp16) storage ----
// The same roles as bf16 above with 10 mantissa bits and 5 exponent bits
// (normal range [2^-14, 65504], subnormal spacing 2^-24). Integer and exact
// f32 operations only, so host and device round identically.

// exact decode
GS_HD float gs_h2f(uint32_t h) {
  const uint32_t s = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0u) {
    const float v = (float)m * 5.9604644775390625e-08f;   // m * 2^-24, exact
    return s ? -v : v;
  }
  if (e == 31u) return gs_u2f(s | 0x7f800000u | (m << 13));
  return gs_u2f(s | ((e + 112u) << 23) | (m << 13));
}

// Round to nearest even. Normal results: rebias the exponent and round the
// 13 dropped mantissa bits as gs_f2bf_rne does; ties at 65520 and above go
// to infinity. Subnormal results: |x| * 2^24 (exact) rounded to an integer
// by the 2^23 shift (IEEE round to nearest even).
GS_HD uint32_t gs_f2h_rne(float x) {
  const uint32_t u = gs_f2u(x);
  const uint32_t s = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return s | 0x7e00u;
  if (a >= 0x477ff000u) return s | 0x7c00u;
  if (a >= 0x38800000u) {
    const uint32_t r = a - 0x38000000u;
    return s | ((r + 0x0fffu + ((r >> 13) & 1u)) >> 13);
  }
  const float v = gs_u2f(a) * 16777216.0f;
  const float q = (v + 8388608.0f) - 8388608.0f;
  return s | (uint32_t)q;
}

// Stochastic rounding with 16 random bits: a value rounds up with
// probability equal to its distance from the lower neighbour (to 2^-13 for
// normal results, 2^-16 for subnormal ones). Infinity and NaN are kept;
// results past 65504 may round to infinity.
GS_HD uint32_t gs_f2h_sr(float x, uint32_t rnd) {
  const uint32_t u = gs_f2u(x);
  const uint32_t s = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return s | 0x7e00u;
  if (a >= 0x47800000u) return s | 0x7c00u;
  if (a >= 0x38800000u) {
    const uint32_t h = (a - 0x38000000u + ((rnd >> 3) & 0x1fffu)) >> 13;
    return s | (h > 0x7c00u ? 0x7c00u : h);
  }
  const float v = gs_u2f(a) * 16777216.0f;   // < 1024, exact
  const uint32_t fl = (uint32_t)v;           // floor
  const uint32_t frac16 = (uint32_t)((v - (float)fl) * 65536.0f);
  return s | (fl + ((rnd & 0xffffu) < frac16 ? 1u : 0u));
}

// smallest fp16 at or above x (x >= 0 finite): a conservative bound
GS_HD uint32_t gs_f2h_up(float x) {
  const uint32_t h = gs_f2h_rne(x);
  return gs_h2f(h) < x ? h + 1u : h;
}

#if defined(__cplusplus)
// 16-bit storage formats: 1 bf16, 2 fp16
template <int kFmt>
GS_HD float gs_16_to_f(uint32_t b) {
  return kFmt == 2 ? gs_h2f(b & 0xffffu) : gs_bf2f(b);
}
template <int kFmt>
GS_HD uint32_t gs_f_to_16_rne(float x) {
  return kFmt == 2 ? gs_f2h_rne(x) : gs_f2bf_rne(x);
}
template <int kFmt>
GS_HD uint32_t gs_f_to_16_sr(float x, uint32_t rnd) {
  return kFmt == 2 ? gs_f2h_sr(x, rnd) : gs_f2bf_sr(x, rnd);
}
#endif

// 1 / (1 + e^-clamp(x, -9, 9)): the activation for colour and opacity logits.
GS_HD float gs_sigmoid_clamped(float x) {
  const float c = fminf(fmaxf(x, -9.0f), 9.0f);
  return 1.0f / (1.0f + gs_expf_mid(-c));
}

#endif  // GS_MATH_H
