// Deterministic FP32 math shared by the host and the gfx950 device code.
//
// Every function here is built only from IEEE-754 single-precision +, -, *, /,
// fmaf, sqrtf, rintf and exact bit manipulation, with FMA contraction disabled
// at compile time (-ffp-contract=off).  Those operations are correctly rounded
// on both x86-64 (SSE/FMA) and CDNA4, so each function returns the same bits on
// the CPU and on the GPU.  That is what lets a CPU restatement of the renderer
// reproduce the GPU's rays, hits and colours exactly instead of "within a few
// ulp" (the reference relies on CUDA fast-math intrinsics, whose results are
// not reproducible off the GPU).
//
// Methods (all textbook, coefficients derived from the series themselves):
//   sin/cos : Cody-Waite reduction by pi/2 (three-part constant applied with
//             fmaf), Taylor polynomials to x^11 / x^12 on [-pi/4, pi/4].
//   log     : exponent/mantissa split, mantissa in [sqrt(1/2), sqrt(2)),
//             ln(m) = 2 atanh((m-1)/(m+1)) series to s^11.
//   exp2    : round-to-nearest integer split, Taylor series of 2^f for
//             |f| <= 1/2 to f^9, exact power-of-two scaling.
//   asin    : Maclaurin series on |x| <= 1/2, asin(x) = pi/2 - 2 asin(sqrt((1-x)/2))
//             above.
//   atan2   : octant reduction, atan(t) = pi/4 + atan((t-1)/(t+1)) above
//             tan(pi/8), alternating series to t^21 below.
// Accuracy is within a few ulp of the correctly rounded result on the ranges
// the renderer uses; the tests check that against the host libm.
#pragma once

#include <stdint.h>
#include <math.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define EYE_HD __host__ __device__ __forceinline__
#else
#define EYE_HD inline
#endif

namespace eye {

constexpr float kPiF = 3.14159265358979323846f;       // M_PIf of the reference (sutil/vec_math.h:43)
constexpr float kPiO2Hi = 1.57079637050628662109375f;  // float(pi/2)
constexpr float kPiO2Mid = -4.371138828673793e-8f;    // float(pi/2 - hi)
constexpr float kPiO2Lo = -1.7151245100058819e-15f;   // float(pi/2 - hi - mid)
constexpr float kTwoOverPi = 0.636619772367581343f;
constexpr float kLn2Hi = 0.693145751953125f;           // ln 2 with 16 trailing zero bits
constexpr float kLn2Lo = 1.42860682030941723e-6f;      // ln 2 - kLn2Hi
constexpr float kInvLn2 = 1.44269504088896341f;

EYE_HD uint32_t f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u; memcpy(&u, &f, 4); return u;
#endif
}
EYE_HD float u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float f; memcpy(&f, &u, 4); return f;
#endif
}

// sin and cos of x (radians), |x| < 2^14.
EYE_HD void det_sincos(float x, float* s_out, float* c_out) {
  const float kf = rintf(x * kTwoOverPi);
  float r = fmaf(-kf, kPiO2Hi, x);
  r = fmaf(-kf, kPiO2Mid, r);
  r = fmaf(-kf, kPiO2Lo, r);
  const int q = ((int)kf) & 3;
  const float r2 = r * r;
  // sin r = r + r^3 * P(r^2), Taylor coefficients (-1)^n / (2n+1)!
  float ps = -2.50521083854417188e-8f;            // -1/11!
  ps = fmaf(ps, r2, 2.75573192239858907e-6f);     //  1/9!
  ps = fmaf(ps, r2, -1.98412698412698413e-4f);    // -1/7!
  ps = fmaf(ps, r2, 8.33333333333333333e-3f);     //  1/5!
  ps = fmaf(ps, r2, -1.66666666666666667e-1f);    // -1/3!
  const float s = fmaf(ps * r2, r, r);
  // cos r = 1 + r^2 * Q(r^2), Taylor coefficients (-1)^n / (2n)!
  float pc = 2.08767569878680990e-9f;             //  1/12!
  pc = fmaf(pc, r2, -2.75573192239858907e-7f);    // -1/10!
  pc = fmaf(pc, r2, 2.48015873015873016e-5f);     //  1/8!
  pc = fmaf(pc, r2, -1.38888888888888889e-3f);    // -1/6!
  pc = fmaf(pc, r2, 4.16666666666666667e-2f);     //  1/4!
  pc = fmaf(pc, r2, -0.5f);                        // -1/2!
  const float c = fmaf(pc, r2, 1.0f);
  float so, co;
  switch (q) {
    case 0: so = s; co = c; break;
    case 1: so = c; co = -s; break;
    case 2: so = -s; co = -c; break;
    default: so = -c; co = s; break;
  }
  *s_out = so;
  *c_out = co;
}

EYE_HD float det_cos(float x) { float s, c; det_sincos(x, &s, &c); return c; }
EYE_HD float det_sin(float x) { float s, c; det_sincos(x, &s, &c); return s; }

// Natural logarithm for finite x > 0 (returns -inf for 0, NaN for x < 0).
EYE_HD float det_log(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return INFINITY;
  uint32_t b = f2u(x);
  int e = 0;
  if (b < 0x00800000u) {          // subnormal: scale by 2^23 (exact)
    x = x * 8388608.0f;
    b = f2u(x);
    e = -23;
  }
  e += (int)((b >> 23) & 0xffu) - 127;
  uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
  if (mb > 0x3fb504f3u) {          // m > sqrt(2): halve it (exact)
    mb -= 0x00800000u;
    e += 1;
  }
  const float m = u2f(mb);
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  float p = 9.09090909090909091e-2f;              // 1/11
  p = fmaf(p, s2, 1.11111111111111111e-1f);       // 1/9
  p = fmaf(p, s2, 1.42857142857142857e-1f);       // 1/7
  p = fmaf(p, s2, 2.0e-1f);                       // 1/5
  p = fmaf(p, s2, 3.33333333333333333e-1f);       // 1/3
  const float two_s = s + s;
  const float lnm = fmaf(two_s * s2, p, two_s);
  const float ef = (float)e;
  return fmaf(ef, kLn2Hi, fmaf(ef, kLn2Lo, lnm));
}

// 2^x.  Underflows to 0 below -160, overflows to +inf above 128.
EYE_HD float det_exp2(float x) {
  if (x != x) return x;
  if (x < -160.0f) return 0.0f;
  if (x > 128.0f) return INFINITY;
  const float kf = rintf(x);
  const float f = x - kf;          // exact, |f| <= 1/2
  // 2^f = sum (f ln2)^n / n!, coefficients ln2^n / n!
  float p = 1.01780860092396997e-7f;              // ln2^9/9!
  p = fmaf(p, f, 1.32154867901443094e-6f);        // ln2^8/8!
  p = fmaf(p, f, 1.52527338040598403e-5f);        // ln2^7/7!
  p = fmaf(p, f, 1.54035303933816099e-4f);        // ln2^6/6!
  p = fmaf(p, f, 1.33335581464284434e-3f);        // ln2^5/5!
  p = fmaf(p, f, 9.61812910762847716e-3f);        // ln2^4/4!
  p = fmaf(p, f, 5.55041086648215800e-2f);        // ln2^3/3!
  p = fmaf(p, f, 2.40226506959100712e-1f);        // ln2^2/2!
  p = fmaf(p, f, 6.93147180559945309e-1f);        // ln2
  p = fmaf(p, f, 1.0f);
  int k = (int)kf;
  // scale by 2^k with at most two exact power-of-two multiplies
  if (k < -126) {
    p = p * u2f(0x00800000u);      // 2^-126
    k += 126;
  } else if (k > 127) {
    p = p * u2f(0x7f000000u);      // 2^127
    k -= 127;
  }
  return p * u2f((uint32_t)(k + 127) << 23);
}

// x^y for x >= 0 (x <= 0 returns 0, x == 1 returns 1), via 2^(y log2 x).
EYE_HD float det_pow(float x, float y) {
  if (!(x > 0.0f)) return 0.0f;
  if (x == 1.0f) return 1.0f;
  const float l2 = det_log(x) * kInvLn2;
  return det_exp2(y * l2);
}

// asin on [-1, 1]; |x| > 1 gives NaN.
EYE_HD float det_asin_core(float x) {  // |x| <= 1/2
  const float x2 = x * x;
  // Maclaurin coefficients (2n)! / (4^n (n!)^2 (2n+1)), n = 10 .. 1
  float p = 8.39033580961681500e-3f;
  p = fmaf(p, x2, 9.76160952919440038e-3f);
  p = fmaf(p, x2, 1.15518008961397050e-2f);
  p = fmaf(p, x2, 1.39648437500000000e-2f);
  p = fmaf(p, x2, 1.73527644230769231e-2f);
  p = fmaf(p, x2, 2.23721590909090909e-2f);
  p = fmaf(p, x2, 3.03819444444444444e-2f);
  p = fmaf(p, x2, 4.46428571428571429e-2f);
  p = fmaf(p, x2, 7.50000000000000000e-2f);
  p = fmaf(p, x2, 1.66666666666666667e-1f);
  return fmaf(x * x2, p, x);
}
EYE_HD float det_asin(float x) {
  const float ax = fabsf(x);
  if (ax <= 0.5f) return det_asin_core(x);
  if (ax > 1.0f) return NAN;
  const float z = (1.0f - ax) * 0.5f;   // exact for ax >= 1/2
  const float r = sqrtf(z);
  const float a = det_asin_core(r);
  const float v = fmaf(-2.0f, a, kPiO2Hi) + kPiO2Mid;  // (pi/2 - 2a), pi/2 = hi + mid
  return x < 0.0f ? -v : v;
}

// atan(t) for 0 <= t <= 1.
EYE_HD float det_atan_unit(float t) {
  float off = 0.0f;
  if (t > 0.414213562373095049f) {     // tan(pi/8)
    t = (t - 1.0f) / (t + 1.0f);
    off = 0.785398163397448310f;       // pi/4
  }
  const float t2 = t * t;
  float p = 4.76190476190476190e-2f;            //  1/21
  p = fmaf(p, t2, -5.26315789473684211e-2f);    // -1/19
  p = fmaf(p, t2, 5.88235294117647059e-2f);     //  1/17
  p = fmaf(p, t2, -6.66666666666666667e-2f);    // -1/15
  p = fmaf(p, t2, 7.69230769230769231e-2f);     //  1/13
  p = fmaf(p, t2, -9.09090909090909091e-2f);    // -1/11
  p = fmaf(p, t2, 1.11111111111111111e-1f);     //  1/9
  p = fmaf(p, t2, -1.42857142857142857e-1f);    // -1/7
  p = fmaf(p, t2, 2.0e-1f);                     //  1/5
  p = fmaf(p, t2, -3.33333333333333333e-1f);    // -1/3
  return off + fmaf(t * t2, p, t);
}
EYE_HD float det_atan2(float y, float x) {
  if (x != x || y != y) return NAN;
  const float ax = fabsf(x), ay = fabsf(y);
  if (ax == 0.0f && ay == 0.0f) return 0.0f;
  float a;
  if (ay <= ax) a = det_atan_unit(ay / ax);
  else a = kPiO2Hi - det_atan_unit(ax / ay);
  if (x < 0.0f) a = kPiF - a;
  return y < 0.0f ? -a : a;
}

}  // namespace eye
