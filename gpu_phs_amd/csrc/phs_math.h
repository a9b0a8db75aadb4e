// Bounded-range elementary functions for the kernels: sin and cos of 2 pi x, tan, exp, log,
// atan, and the azimuth atan2(y, x) in turns.
//
// The maps need sin and cos of 2*pi*x for a random x in [0, 1] (the azimuth, phs_cpu.cpp:301-307),
// tan of an angle inside (-pi/2, pi/2) (the Breit-Wigner map, phs_cpu.cpp:25-28), and exp and
// log of moderate arguments (phs_cpu.cpp:29-59, 128-140).  The device library's routines accept
// every double and cost 40-190 VALU instructions each; worse, inside the kernels' branch loop
// the compiler hoists their polynomial coefficients out of the loop into vector registers, some
// sixty of them for the whole set, which halves the number of resident waves.  Here every
// coefficient is materialised in a scalar register pair right where it is used (two scalar
// moves, issued beside the vector work): GPHS_FMA and GPHS_LIN for the polynomial steps,
// GPHS_PIN for the other constants.
//
// Methods (all with explicit fused multiply-adds, so they do not depend on contraction):
//   * sin/cos: reduction to |theta| <= pi/4 -- exact for turns (t = 4x, r = t - rint(t));
//     for radians, Cody-Waite subtraction of n * pi/2 with pi/2 split into two 33-bit parts
//     and a 53-bit tail (W. J. Cody, W. Waite, Software Manual for the Elementary Functions,
//     1980), exact in its first two steps for |n| < 2^20 -- then the Taylor series cut after
//     theta^17/17! and theta^18/18! (first dropped terms at pi/4: 1.1e-19 and 3.3e-21);
//   * exp: x = k ln 2 + r with |r| <= ln2 / 2 (ln 2 as a double-double), the Taylor series of
//     e^r to r^13/13! (first dropped term 4.2e-18), then scaling by 2^k;
//   * log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m - 1)/(m + 1),
//     |s| <= 0.1716, series to s^21/21 (first dropped term 2.3e-17 relative), plus e ln 2.
//   * atan2 in turns: see atan2_turns; atan in radians: see atan_radians.
// The tests check them against 300-bit arithmetic: within 2 units in the last place (sin,
// cos, exp, log) and 3 (tan, atan2 in turns).  tan(-0) is +0.  The header has no HIP dependency, so the same
// code is compiled for the host in those tests.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GPHS_MATH_FN __host__ __device__ __forceinline__
#else
#define GPHS_MATH_FN inline
#endif

namespace gphs {
namespace math {

// 1/n! as the correctly rounded doubles of exact quotients (every n! here is below 2^53).
constexpr double kS3 = -1.0 / 6.0;
constexpr double kS5 = 1.0 / 120.0;
constexpr double kS7 = -1.0 / 5040.0;
constexpr double kS9 = 1.0 / 362880.0;
constexpr double kS11 = -1.0 / 39916800.0;
constexpr double kS13 = 1.0 / 6227020800.0;
constexpr double kS15 = -1.0 / 1307674368000.0;
constexpr double kS17 = 1.0 / 355687428096000.0;
constexpr double kC4 = 1.0 / 24.0;
constexpr double kC6 = -1.0 / 720.0;
constexpr double kC8 = 1.0 / 40320.0;
constexpr double kC10 = -1.0 / 3628800.0;
constexpr double kC12 = 1.0 / 479001600.0;
constexpr double kC14 = -1.0 / 87178291200.0;
constexpr double kC16 = 1.0 / 20922789888000.0;
constexpr double kC18 = -1.0 / 6402373705728000.0;

// pi/2 = kPio2Hi + kPio2Lo to 106 bits.
constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
constexpr double kPio2Lo = 0x1.1a62633145c07p-54;
// pi/2 = kCw1 + kCw2 + kCw3: kCw1 and kCw2 hold 33 significant bits, so n * kCw1 and n * kCw2
// are exact for |n| < 2^20.
constexpr double kCw1 = 0x1.921fb544p+0;
constexpr double kCw2 = 0x1.0b4611a6p-34;
constexpr double kCw3 = 0x1.3198a2e037073p-69;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
// ln 2 = kLn2Hi + kLn2Lo to 106 bits; log2(e); sqrt(1/2).
constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;
constexpr double kLog2e = 0x1.71547652b82fep+0;
constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;
// 1 / (2 pi) and atan(k / 4) / (2 pi), k = 1, 2, 3, correctly rounded.
constexpr double kInv2Pi = 0x1.45f306dc9c883p-3;
constexpr double kAtanQ1 = 0x1.3f670b6bdc73dp-5;
constexpr double kAtanQ2 = 0x1.2e4051d9df308p-4;
constexpr double kAtanQ3 = 0x1.a37f5c4c419efp-4;

// A constant materialised in a scalar register pair at its point of use (see the file comment):
// two scalar moves inside volatile asm, which the compiler neither hoists out of loops nor
// places before `after` has been computed.
template <uint64_t B>
GPHS_MATH_FN double pin_bits(double after) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t lo, hi;
  asm volatile("s_mov_b32 %0, %1" : "=s"(lo) : "i"((int32_t)(uint32_t)B), "v"(after));
  asm volatile("s_mov_b32 %0, %1" : "=s"(hi) : "i"((int32_t)(uint32_t)(B >> 32)), "v"(after));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
#else
  (void)after;
  return __builtin_bit_cast(double, B);
#endif
}
#define GPHS_PIN(c, after) ::gphs::math::pin_bits<__builtin_bit_cast(uint64_t, (double)(c))>(after)

// a * b + C and z * C1 + C0 (a product and a sum, rounded twice) for compile-time constants:
// each in one asm statement that forms the constants in the scalar pair s[96:97] right before
// the vector instruction that reads them, so that they are neither hoisted into vector
// registers nor kept live in scalar ones.  One statement per step matters: the compiler pads
// every inline asm statement with a wait state.  (Left to itself it also picks v_fmac_f64,
// whose addend must be the destination vector register, and copies a scalar constant into
// vector registers first.)
template <uint64_t C>
GPHS_MATH_FN double fma_k(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r;
  asm volatile("s_mov_b32 s96, %3\n\ts_mov_b32 s97, %4\n\tv_fma_f64 %0, %1, %2, s[96:97]"
               : "=v"(r)
               : "v"(a), "v"(b), "i"((int32_t)(uint32_t)C), "i"((int32_t)(uint32_t)(C >> 32))
               : "s96", "s97");
  return r;
#else
  return __builtin_fma(a, b, __builtin_bit_cast(double, C));
#endif
}
template <uint64_t C1, uint64_t C0>
GPHS_MATH_FN double lin_k(double z) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r;
  asm volatile("s_mov_b32 s96, %2\n\ts_mov_b32 s97, %3\n\tv_mul_f64 %0, %1, s[96:97]\n\t"
               "s_mov_b32 s96, %4\n\ts_mov_b32 s97, %5\n\tv_add_f64 %0, %0, s[96:97]"
               : "=&v"(r)
               : "v"(z), "i"((int32_t)(uint32_t)C1), "i"((int32_t)(uint32_t)(C1 >> 32)),
                 "i"((int32_t)(uint32_t)C0), "i"((int32_t)(uint32_t)(C0 >> 32))
               : "s96", "s97");
  return r;
#else
  return z * __builtin_bit_cast(double, C1) + __builtin_bit_cast(double, C0);
#endif
}
#define GPHS_BITS(c) __builtin_bit_cast(uint64_t, (double)(c))
#define GPHS_FMA(a, b, c) ::gphs::math::fma_k<GPHS_BITS(c)>((a), (b))
#define GPHS_LIN(z, c1, c0) ::gphs::math::lin_k<GPHS_BITS(c1), GPHS_BITS(c0)>(z)

// a / b where an ulp or so does not matter: the reciprocal estimate refined by two Newton
// steps (r <- r + r (1 - b r)), times a; within a few ulps of the quotient for normal b, and
// half the instructions of a correctly rounded division.
GPHS_MATH_FN double quot(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  return a * r;
#else
  return a * (1.0 / b);   // the reciprocal, then the product: rounded twice, as on the device
#endif
}

// 1 / b: the reciprocal estimate refined by two Newton steps, within an ulp or so for normal b.
GPHS_MATH_FN double recip(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  return __builtin_fma(r, e, r);
#else
  return 1.0 / b;
#endif
}

// sqrt(x): the reciprocal square root estimate refined by one Goldschmidt step and two Newton
// corrections -- the compiler's own sequence for a double sqrt without its rescaling of inputs
// below 2^-767, which the kernels' arguments (squared masses and momenta of physical size,
// or exactly 0) never reach; below 2^-1022 the result is not accurate.  sqrt(0) = 0,
// sqrt(+inf) = +inf, sqrt(x < 0) = NaN.
GPHS_MATH_FN double sqrt_phys(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return (x == 0 || x == __builtin_inf()) ? x : g;
#else
  return __builtin_sqrt(x);
#endif
}

// sin and cos of |t| <= pi/4.
GPHS_MATH_FN void sincos_reduced(double t, double& s, double& c) {
  const double t2 = t * t;
  // the first step as a product and a sum, so that both constants can stay scalar
  double ps = GPHS_LIN(t2, kS17, kS15);
  ps = GPHS_FMA(ps, t2, kS13);
  ps = GPHS_FMA(ps, t2, kS11);
  ps = GPHS_FMA(ps, t2, kS9);
  ps = GPHS_FMA(ps, t2, kS7);
  ps = GPHS_FMA(ps, t2, kS5);
  ps = GPHS_FMA(ps, t2, kS3);
  s = __builtin_fma(t * t2, ps, t);
  double pc = GPHS_LIN(t2, kC18, kC16);
  pc = GPHS_FMA(pc, t2, kC14);
  pc = GPHS_FMA(pc, t2, kC12);
  pc = GPHS_FMA(pc, t2, kC10);
  pc = GPHS_FMA(pc, t2, kC8);
  pc = GPHS_FMA(pc, t2, kC6);
  pc = GPHS_FMA(pc, t2, kC4);
  c = __builtin_fma(t2 * t2, pc, __builtin_fma(t2, -0.5, 1.0));
}

// s = sin(2 pi x), c = cos(2 pi x) for |x| < 2^28 (a full turn per unit of x).
GPHS_MATH_FN void sincos_turns(double x, double& s, double& c) {
  const double t = 4.0 * x;
  const double q = __builtin_rint(t);
  const double r = t - q;  // exact, |r| <= 1/2
  double s0, c0;
  sincos_reduced(__builtin_fma(r, GPHS_PIN(kPio2Hi, r), r * GPHS_PIN(kPio2Lo, r)), s0, c0);
  const int n = (int)q;  // quadrant; n & 3 is right for negative n too
  const double a = (n & 1) ? c0 : s0;
  const double b = (n & 1) ? s0 : c0;
  s = (n & 2) ? -a : a;
  c = ((n + 1) & 2) ? -b : b;
}

// tan(z) for |z| < 2^20 * pi/2; the kernels call it for |z| < pi/2.
GPHS_MATH_FN double tan_radians(double z) {
  const double q = __builtin_rint(z * GPHS_PIN(kTwoOverPi, z));
  const double r = ((z - q * GPHS_PIN(kCw1, q)) - q * GPHS_PIN(kCw2, q)) - q * GPHS_PIN(kCw3, q);
  double s0, c0;
  sincos_reduced(r, s0, c0);
  const bool odd = ((int)q) & 1;
  return (odd ? -c0 : s0) / (odd ? s0 : c0);
}

// e^x; overflows to +inf above ~709.8 and underflows to 0 below ~-745 like exp().
GPHS_MATH_FN double exp_bounded(double x) {
  const double xc = x > 709.8 ? 709.8 : (x < -745.2 ? -745.2 : x);   // keeps k in int range
  const double k = __builtin_rint(xc * GPHS_PIN(kLog2e, xc));
  const double r = __builtin_fma(-k, GPHS_PIN(kLn2Lo, k), __builtin_fma(-k, GPHS_PIN(kLn2Hi, k), xc));
  double p = GPHS_LIN(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
  p = GPHS_FMA(p, r, 1.0 / 39916800.0);
  p = GPHS_FMA(p, r, 1.0 / 3628800.0);
  p = GPHS_FMA(p, r, 1.0 / 362880.0);
  p = GPHS_FMA(p, r, 1.0 / 40320.0);
  p = GPHS_FMA(p, r, 1.0 / 5040.0);
  p = GPHS_FMA(p, r, 1.0 / 720.0);
  p = GPHS_FMA(p, r, 1.0 / 120.0);
  p = GPHS_FMA(p, r, 1.0 / 24.0);
  p = GPHS_FMA(p, r, 1.0 / 6.0);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r * r, r) + 1.0;   // 1 + r + r^2 (1/2 + r/6 + ...)
  // 2^k in two steps, so that 2^k itself need not be representable near the ends of the range
  const int ki = (int)k;
  const double y = __builtin_ldexp(__builtin_ldexp(p, ki / 2), ki - ki / 2);
  return x != x ? x : y;
}

// Natural logarithm; log(0) = -inf, log(x < 0) = NaN, log(+inf) = +inf, like log().
GPHS_MATH_FN double log_bounded(double x) {
  int e;
  double m = __builtin_frexp(x, &e);   // x = m 2^e, m in [1/2, 1)
  if (m < GPHS_PIN(kSqrtHalf, m)) { m = 2 * m; e -= 1; }
  const double f = m - 1.0;             // exact
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = GPHS_LIN(z, 1.0 / 21.0, 1.0 / 19.0);
  p = GPHS_FMA(p, z, 1.0 / 17.0);
  p = GPHS_FMA(p, z, 1.0 / 15.0);
  p = GPHS_FMA(p, z, 1.0 / 13.0);
  p = GPHS_FMA(p, z, 1.0 / 11.0);
  p = GPHS_FMA(p, z, 1.0 / 9.0);
  p = GPHS_FMA(p, z, 1.0 / 7.0);
  p = GPHS_FMA(p, z, 1.0 / 5.0);
  p = GPHS_FMA(p, z, 1.0 / 3.0);
  const double lm = __builtin_fma(2 * s * z, p, 2 * s);     // 2 atanh(s)
  const double de = (double)e;
  const double y = __builtin_fma(de, GPHS_PIN(kLn2Hi, lm), __builtin_fma(de, GPHS_PIN(kLn2Lo, lm), lm));
  if (!(x > 0)) return x == 0 ? -__builtin_inf() : __builtin_nan("");
  return x == __builtin_inf() ? x : y;
}

// atan(k / 4), k = 1 .. 4, as double-doubles (hi correctly rounded, lo the rest).
constexpr double kAtanR1 = 0x1.f5b75f92c80ddp-3, kAtanR1Lo = 0x1.8ab6e3cf7afbdp-57;
constexpr double kAtanR2 = 0x1.dac670561bb4fp-2, kAtanR2Lo = 0x1.a2b7f222f65e2p-56;
constexpr double kAtanR3 = 0x1.4978fa3269ee1p-1, kAtanR3Lo = 0x1.2419a87f2a458p-56;
constexpr double kAtanR4 = 0x1.921fb54442d18p-1, kAtanR4Lo = 0x1.1a62633145c07p-55;

// atan(x) in radians, for any x (the Breit-Wigner map's constant atan((m_max^2 - M^2) / (M Gamma))
// at each event's sqrt(s), phs_cpu.cpp:479-480), to about half an ulp, so that it agrees with a
// correctly rounded atan for nearly every argument (the constant enters differences a1 - a0 that
// can cancel).  With a = |x| for |x| <= 1 and 1 / |x| above (as a double-double),
// c = k / 4 for k = rint(4 a) and t = (a - c) / (1 + a c), |t| <= 1/8: atan a = atan c + atan t
// (the addition theorem), atan c from a table of double-doubles, t to double-double accuracy
// (its quotient's residual by fused multiply-adds), atan t = t + t^3 (-1/3 + t^2/5 - ...) cut
// after t^17/17 (first dropped term below 4e-18 of t); the sum with its rounding error carried
// (Dekker's fast two-sum); for |x| > 1, pi/2 - atan a with pi/2 as a double-double; the sign
// of x last.  atan(+-inf) = +-pi/2, atan(+-0) = +-0, NaN in, NaN out.
GPHS_MATH_FN double atan_radians(double x) {
  const double ax = __builtin_fabs(x);
  const bool swap = ax > 1.0;
  const double a = swap ? 1.0 / ax : ax;
  // 1 / ax = a + a_lo (the residual of the division; 0 where 1 / ax is no longer normal)
  const double a_lo = swap && ax < 0x1p1000 ? __builtin_fma(-a, ax, 1.0) / ax : 0.0;
  const double k = __builtin_rint(4.0 * a);
  const double c = 0.25 * k;
  const double num = a - c;                                  // exact (Sterbenz)
  // den = 1 + (a + a_lo) c = den_hi + den_lo; a c <= 1, so the fast two-sum holds
  const double pc = a * c;
  const double pce = __builtin_fma(a, c, -pc) + a_lo * c;
  const double den_hi = 1.0 + pc;
  const double den_lo = ((1.0 - den_hi) + pc) + pce;
  const double t = num / den_hi;
  const double t_lo = (__builtin_fma(-t, den_hi, num) + (a_lo - t * den_lo)) / den_hi;
  const double z = t * t;
  double p = GPHS_LIN(z, 1.0 / 17.0, -1.0 / 15.0);
  p = GPHS_FMA(p, z, 1.0 / 13.0);
  p = GPHS_FMA(p, z, -1.0 / 11.0);
  p = GPHS_FMA(p, z, 1.0 / 9.0);
  p = GPHS_FMA(p, z, -1.0 / 7.0);
  p = GPHS_FMA(p, z, 1.0 / 5.0);
  p = GPHS_FMA(p, z, -1.0 / 3.0);
  double th = 0.0, tl = 0.0;   // atan(k / 4)
  if (k >= 4.0) { th = GPHS_PIN(kAtanR4, k); tl = GPHS_PIN(kAtanR4Lo, k); }
  else if (k >= 3.0) { th = GPHS_PIN(kAtanR3, k); tl = GPHS_PIN(kAtanR3Lo, k); }
  else if (k >= 2.0) { th = GPHS_PIN(kAtanR2, k); tl = GPHS_PIN(kAtanR2Lo, k); }
  else if (k >= 1.0) { th = GPHS_PIN(kAtanR1, k); tl = GPHS_PIN(kAtanR1Lo, k); }
  // atan a = th + t + (tl + t_lo + t z p); |t| <= 1/8 < th unless th = 0
  const double s = th + t;
  const double lo = (((th - s) + t) + tl) + __builtin_fma(t * z, p, t_lo);
  const double rh = s + lo;
  double r = rh;
  if (swap) {
    const double rl = (s - rh) + lo;
    const double ph = GPHS_PIN(kPio2Hi, rh);
    const double d = ph - rh;                               // pi/2 > rh: the fast two-sum holds
    r = d + (((ph - d) - rh) + (GPHS_PIN(kPio2Lo, rh) - rl));
  }
  return __builtin_copysign(r, x);
}

// atan2(y, x) / (2 pi), plus 1 where it is negative: the azimuth of (x, y) in turns, in [0, 1]
// (the inverse maps' x_phi = phi / 2 pi with phi = atan2(y, x) + 2 pi for phi < 0).  With
// a = min(|x|, |y|) / max(|x|, |y|) in [0, 1], c = k / 4 for k = rint(4 a) and
// t = (a - c) / (1 + a c), |t| <= 1/8, atan a = atan c + atan t (the addition theorem); atan c is
// a table of five turns, atan t the Taylor series t - t^3/3 + ... cut after t^17/17 (first
// dropped term below 1e-17 of t).  The octant, half-plane and sign then place the angle.
// atan2(+-0, +-0) gives 0 (x = +0) and 1/2 (x = -0), as atan2 does; NaN in, NaN out.
GPHS_MATH_FN double atan2_turns(double y, double x) {
  const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
  const bool swap = ay > ax;
  const double num = swap ? ax : ay, den = swap ? ay : ax;
  const double a = den > 0 ? quot(num, den) : (den == den ? 0.0 : den);
  const double k = __builtin_rint(4.0 * a);
  const double c = 0.25 * k;
  const double t = quot(a - c, __builtin_fma(a, c, 1.0));   // a - c is exact (Sterbenz)
  const double z = t * t;
  double p = GPHS_LIN(z, 1.0 / 17.0, -1.0 / 15.0);
  p = GPHS_FMA(p, z, 1.0 / 13.0);
  p = GPHS_FMA(p, z, -1.0 / 11.0);
  p = GPHS_FMA(p, z, 1.0 / 9.0);
  p = GPHS_FMA(p, z, -1.0 / 7.0);
  p = GPHS_FMA(p, z, 1.0 / 5.0);
  p = GPHS_FMA(p, z, -1.0 / 3.0);
  const double tt = t * GPHS_PIN(kInv2Pi, t);            // atan t / 2 pi = tt (1 + z p)
  // atan(k / 4) / 2 pi for k = 0 .. 4
  const double tk = k < 2.0 ? (k < 1.0 ? 0.0 : GPHS_PIN(kAtanQ1, k)) :
                    (k < 3.0 ? GPHS_PIN(kAtanQ2, k) : (k < 4.0 ? GPHS_PIN(kAtanQ3, k) : 0.125));
  double r = tk + __builtin_fma(tt * z, p, tt);          // in [0, 1/8]
  r = swap ? 0.25 - r : r;
  r = __builtin_signbit(x) ? 0.5 - r : r;
  return y < 0 ? 1.0 - r : r;
}

}  // namespace math
}  // namespace gphs
