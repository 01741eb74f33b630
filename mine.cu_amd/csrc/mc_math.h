// mc_math.h - IEEE-754 binary32 numerics shared by the gfx950 kernels and the
// host-side code that must reproduce them bit for bit.
//
// The reference computes cosf/sinf/tanf with NVIDIA fast-math intrinsics
// (src/kernels.cu:118-135, 331-352, 1539-1540, 1617-1620; setup.py:19
// `--use_fast_math`).  Those results are not reproducible off NVIDIA hardware,
// so this project defines the transcendental functions itself:
//
//   * argument reduction: Payne & Hanek, "Radian reduction for trigonometric
//     functions" (SIGNUM Newsletter 18(1), 1983) -- the binary32 input is an
//     exact integer mantissa times a power of two, which is multiplied by a
//     96-bit window of the bits of 2/pi with 32x32->64 integer products, so the
//     reduction is exact to 2^-64 of a quadrant for every finite float;
//   * evaluation: the Taylor series of sin and cos about 0, truncated after
//     r^19 / r^20, in binary64 on |r| <= pi/4 (truncation error < 1e-19);
//   * the binary64 result is rounded once to binary32.
//
// Only IEEE binary64 multiply/add, integer ops and one final rounding are
// used, all compiled with FP contraction disabled, so a host build (g++) and a
// device build (hipcc, gfx950) return identical bits.  The result is within
// 1 ulp of the exact value (almost always correctly rounded); tests bound it
// against the host libm.
#pragma once
#include <stdint.h>
#include <string.h>

#ifdef __clang__
#pragma clang fp contract(off)
#endif

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MC_HD __host__ __device__ __forceinline__
#else
#define MC_HD static inline
#endif

namespace mcm {

// Bits 1..256 of 2/pi after the binary point (2/pi = 0.a2f9836e 4e441529 ...),
// computed with exact integer arithmetic from Machin's formula.
MC_HD uint32_t two_over_pi_word(int i) {
  switch (i) {
    case 0: return 0xa2f9836eu;
    case 1: return 0x4e441529u;
    case 2: return 0xfc2757d1u;
    case 3: return 0xf534ddc0u;
    case 4: return 0xdb629599u;
    case 5: return 0x3c439041u;
    case 6: return 0xfe5163abu;
    default: return 0xdebbc561u;
  }
}

MC_HD uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
MC_HD float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

// Reduce |x| (finite, |x| >= ~pi/4) to r in [-pi/4, pi/4] and a quadrant q
// such that |x| = q*pi/2 + r (mod 2pi).
MC_HD double reduce_quadrant(uint32_t ax_bits, int* q_out) {
  const uint32_t m = (ax_bits & 0x7fffffu) | 0x800000u;     // 24-bit mantissa
  const int e = (int)((ax_bits >> 23) & 0xffu) - 150;        // |x| = m * 2^e
  // Window of 2/pi bits j0 .. j0+95 (1-based); bits before j0 only contribute
  // multiples of 4 quadrants and are dropped.
  const int j0 = (e >= 2) ? (e - 1) : 1;
  const int off = j0 - 1;
  const int wi = off >> 5, sh = off & 31;
  uint32_t w0, w1, w2;
  if (off == 0) {
    // |x| < 2^25 (every angle an agent turns to in practice): the window is
    // the first three words, the general case below with wi = sh = 0.
    w0 = 0xa2f9836eu; w1 = 0x4e441529u; w2 = 0xfc2757d1u;
  } else if (sh == 0) {
    w0 = two_over_pi_word(wi); w1 = two_over_pi_word(wi + 1); w2 = two_over_pi_word(wi + 2);
  } else {
    const uint32_t a0 = two_over_pi_word(wi), a1 = two_over_pi_word(wi + 1);
    const uint32_t a2 = two_over_pi_word(wi + 2), a3 = two_over_pi_word(wi + 3);
    w0 = (a0 << sh) | (a1 >> (32 - sh));
    w1 = (a1 << sh) | (a2 >> (32 - sh));
    w2 = (a2 << sh) | (a3 >> (32 - sh));
  }
  // P = m * (w0:w1:w2), a 120-bit integer held as hi:lo.
  const uint64_t p0 = (uint64_t)m * w0;
  const uint64_t p1 = (uint64_t)m * w1;
  const uint64_t p2 = (uint64_t)m * w2;
  const uint64_t lo = p2 + (p1 << 32);
  const uint64_t carry = (lo < p2) ? 1u : 0u;
  const uint64_t hi = p0 + (p1 >> 32) + carry;
  // |x| * 2/pi = P * 2^-S with S fractional bits, 64 < S <= 120.
  const int S = j0 + 95 - e;
  int q = (int)((hi >> (S - 64)) & 3u);
  const uint64_t F = (hi << (128 - S)) | (lo >> (S - 64));   // top 64 fraction bits
  int64_t Fs = (int64_t)F;
  if (F >> 63) q = (q + 1) & 3;                               // round to nearest quadrant
  *q_out = q;
  // pi/2 * 2^-64
  const double k = 0x1.921fb54442d18p-64;
  return (double)Fs * k;
}

// sin(r), cos(r) for |r| <= pi/4 + tiny, Taylor series in binary64.
MC_HD double sin_poly(double r) {
  const double r2 = r * r;
  double s = -0x1.2f49b46814157p-57;            // -1/19!
  s = s * r2 + 0x1.952c77030ad4ap-49;           // +1/17!
  s = s * r2 - 0x1.ae7f3e733b81fp-41;           // -1/15!
  s = s * r2 + 0x1.6124613a86d09p-33;           // +1/13!
  s = s * r2 - 0x1.ae64567f544e4p-26;           // -1/11!
  s = s * r2 + 0x1.71de3a556c734p-19;           // +1/9!
  s = s * r2 - 0x1.a01a01a01a01ap-13;           // -1/7!
  s = s * r2 + 0x1.1111111111111p-7;            // +1/5!
  s = s * r2 - 0x1.5555555555555p-3;            // -1/3!
  return r + (r * r2) * s;
}

MC_HD double cos_poly(double r) {
  const double r2 = r * r;
  double c = 0x1.e542ba4020225p-62;             // +1/20!
  c = c * r2 - 0x1.6827863b97d97p-53;           // -1/18!
  c = c * r2 + 0x1.ae7f3e733b81fp-45;           // +1/16!
  c = c * r2 - 0x1.93974a8c07c9dp-37;           // -1/14!
  c = c * r2 + 0x1.1eed8eff8d898p-29;           // +1/12!
  c = c * r2 - 0x1.27e4fb7789f5cp-22;           // -1/10!
  c = c * r2 + 0x1.a01a01a01a01ap-16;           // +1/8!
  c = c * r2 - 0x1.6c16c16c16c17p-10;           // -1/6!
  c = c * r2 + 0x1.5555555555555p-5;            // +1/4!
  c = c * r2 - 0.5;                             // -1/2!
  return 1.0 + r2 * c;
}

// sin and cos of a binary32 argument, computed in binary64 (not yet rounded).
MC_HD void sincos_d(float x, double* s_out, double* c_out) {
  const uint32_t xb = f2u(x);
  const uint32_t ab = xb & 0x7fffffffu;
  if (ab >= 0x7f800000u) {                       // inf or NaN -> NaN
    const double nan = (double)u2f(0x7fc00000u);
    *s_out = nan; *c_out = nan;
    return;
  }
  double r;
  int q;
  if (ab < 0x3f490fdau) {                        // |x| < pi/4 (0x3f490fdb rounds up)
    r = (double)u2f(ab);
    q = 0;
  } else {
    r = reduce_quadrant(ab, &q);
  }
  const double sp = sin_poly(r), cp = cos_poly(r);
  // Quadrant q: (s, c) = (sp, cp), (cp, -sp), (-sp, -cp), (-cp, sp), as
  // selects (one straight-line sequence on the GPU).
  double s = (q & 1) ? cp : sp;
  double c = (q & 1) ? sp : cp;
  if (q & 2) s = -s;
  if ((q + 1) & 2) c = -c;
  if (xb >> 31) s = -s;
  *s_out = s; *c_out = c;
}

}  // namespace mcm

// Public binary32 entry points.
MC_HD void mc_sincosf(float x, float* s, float* c) {
  double sd, cd;
  mcm::sincos_d(x, &sd, &cd);
  *s = (float)sd;
  *c = (float)cd;
}

MC_HD float mc_sinf(float x) { float s, c; mc_sincosf(x, &s, &c); return s; }
MC_HD float mc_cosf(float x) { float s, c; mc_sincosf(x, &s, &c); return c; }

// tan as the binary64 quotient of the binary64 sin and cos, rounded once.
MC_HD float mc_tanf(float x) {
  double sd, cd;
  mcm::sincos_d(x, &sd, &cd);
  return (float)(sd / cd);
}

// Half-angle tangent of the vertical field of view, with the degree->radian
// conversion done in binary32 in the reference's operation order
// (src/kernels.cu:347-349: fov * 3.14159265f / 180.0f, then * 0.5f).
// Evaluated once per launch on the host and passed to the kernels.
MC_HD float mc_half_fov_tan(float fov_degrees) {
  const float fov_rad = fov_degrees * 3.14159265f / 180.0f;
  return mc_tanf(fov_rad * 0.5f);
}

// The flat-colour render variants convert with a single constant instead
// (src/kernels.cu:536-537, 1036-1038: fov * 0.01745329f, then * 0.5f).
MC_HD float mc_half_fov_tan_flat(float fov_degrees) {
  const float fov_rad = fov_degrees * 0.01745329f;
  return mc_tanf(fov_rad * 0.5f);
}
