// jsm.h -- the JS-compatible FP64 math of the dye schedules on the device: log, exp, sin and cos as fdlibm 5.3
// (Sun Microsystems, 1993; the published algorithms and coefficients) computes them in the form V8 runs for
// Math.log/exp/sin/cos, which the reference compiles as jsm:: (core/src/jsmath_fdlibm.cpp:137-1564). The results
// are the reference's bits: every step is an IEEE FP64 operation in fdlibm's order, the file is compiled without
// FP contraction, and division and square root are the correctly rounded ones.
//
// Coverage: sin/cos restate the argument reductions for |x| <= 2^19 pi/2 (823,549.6); larger finite arguments
// return NaN (the Payne-Hanek branch is not restated; the schedules only take cos(2 pi u), u < 1).
//
// The algorithms and coefficients restated here are fdlibm's, which carries this notice:
// ====================================================
// Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
// Developed at SunSoft, a Sun Microsystems, Inc. business.
// Permission to use, copy, modify, and distribute this
// software is freely granted, provided that this notice
// is preserved.
// ====================================================
// V8's form of them (src/base/ieee754.cc): Copyright 2016 the V8 project authors. All rights reserved.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace iscg {
namespace jsm {

constexpr double kPi = 3.141592653589793;  // jsm::PI

__device__ __forceinline__ int32_t hi_of(double x) { return __double2hiint(x); }
__device__ __forceinline__ uint32_t lo_of(double x) { return static_cast<uint32_t>(__double2loint(x)); }
__device__ __forceinline__ double of_words(uint32_t hi, uint32_t lo)
{
   return __hiloint2double(static_cast<int>(hi), static_cast<int>(lo));
}

// log(x) = k ln2 + log(1 + f): k and 1 + f in [sqrt(2)/2, sqrt(2)) from the exponent and mantissa, then
// log(1 + f) = f - (f^2/2 - s (f^2/2 + R)) with s = f / (2 + f) and R(s^2) the Remez polynomial Lg1..Lg7 (split into
// even and odd halves); |f| < 2^-20 takes the short series. log(+-0) = -inf, log(x < 0) = NaN.
__device__ inline double log(double x)
{
   const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
   const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01, Lg3 = 2.857142874366239149e-01,
                Lg4 = 2.222219843214978396e-01, Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                Lg7 = 1.479819860511658591e-01;
   int32_t hx = hi_of(x);
   int32_t k = 0;
   if (hx < 0x00100000)
   {
      if (((hx & 0x7fffffff) | static_cast<int32_t>(lo_of(x))) == 0)
         return -__builtin_inf();
      if (hx < 0)
         return __builtin_nan("");
      k -= 54;
      x *= 1.80143985094819840000e+16;  // 2^54: subnormals scaled up
      hx = hi_of(x);
   }
   if (hx >= 0x7ff00000)
      return x + x;
   k += (hx >> 20) - 1023;
   hx &= 0x000fffff;
   const int32_t norm = (hx + 0x95f64) & 0x100000;  // mantissa above sqrt(2): halve x, raise k
   x = of_words(static_cast<uint32_t>(hx | (norm ^ 0x3ff00000)), lo_of(x));
   k += norm >> 20;
   const double f = x - 1.0;
   const double dk = static_cast<double>(k);
   if ((0x000fffff & (2 + hx)) < 3)
   {
      if (f == 0.0)
         return k == 0 ? 0.0 : dk * ln2_hi + dk * ln2_lo;
      const double R = f * f * (0.5 - 0.33333333333333333 * f);
      return k == 0 ? f - R : dk * ln2_hi - ((R - dk * ln2_lo) - f);
   }
   const double s = f / (2.0 + f);
   const double z = s * s;
   const double w = z * z;
   const int32_t mid = (hx - 0x6147a) | (0x6b851 - hx);  // > 0: 1 + f away from 1, the f^2/2 split pays
   const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
   const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
   const double R = t2 + t1;
   if (mid > 0)
   {
      const double hfsq = 0.5 * f * f;
      if (k == 0)
         return f - (hfsq - s * (hfsq + R));
      return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
   }
   if (k == 0)
      return f - s * (f - R);
   return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

// exp(x): x = k ln2 + r with |r| <= ln2/2 (ln2 as hi + lo, k by rounding x/ln2), exp(r) = 1 + r + r c / (2 - c) with
// c = r - r^2 P(r^2) (P1..P5), then the 2^k scale as a multiply by a power of two built from exponent bits (split for
// k = 1024 and k < -1021). exp(1) is the double nearest e.
__device__ inline double exp(double x)
{
   const double ln2HI = 6.93147180369123816490e-01, ln2LO = 1.90821492927058770002e-10;
   const double invln2 = 1.44269504088896338700e+00;
   const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05,
                P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
   const double twom1000 = 9.33263618503218878990e-302, two1023 = 8.988465674311579539e307;
   double hi = 0.0, lo = 0.0;
   int32_t k = 0;
   uint32_t hx = static_cast<uint32_t>(hi_of(x));
   const bool neg = (hx >> 31) != 0;
   hx &= 0x7fffffff;
   if (hx >= 0x40862e42)
   {
      if (hx >= 0x7ff00000)
      {
         if (((hx & 0xfffff) | lo_of(x)) != 0)
            return x + x;
         return neg ? 0.0 : x;
      }
      if (x > 7.09782712893383973096e+02)
         return __builtin_inf();
      if (x < -7.45133219101941108420e+02)
         return 0.0;
   }
   if (hx > 0x3fd62e42)
   {
      if (hx < 0x3ff0a2b2)
      {
         if (x == 1.0)
            return 2.718281828459045;
         hi = neg ? x + ln2HI : x - ln2HI;
         lo = neg ? -ln2LO : ln2LO;
         k = neg ? -1 : 1;
      }
      else
      {
         k = static_cast<int32_t>(invln2 * x + (neg ? -0.5 : 0.5));
         const double t = k;
         hi = x - t * ln2HI;
         lo = t * ln2LO;
      }
      x = hi - lo;
   }
   else if (hx < 0x3e300000)
   {
      return 1.0 + x;  // |x| < 2^-28
   }
   const double t = x * x;
   const double c = x - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
   if (k == 0)
      return 1.0 - ((x * c) / (c - 2.0) - x);
   const double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
   if (k >= -1021)
   {
      if (k == 1024)
         return y * 2.0 * two1023;
      return y * of_words(0x3ff00000u + (static_cast<uint32_t>(k) << 20), 0);
   }
   return y * of_words(0x3ff00000u + (static_cast<uint32_t>(k + 1000) << 20), 0) * twom1000;
}

// x = n pi/2 + (y0 + y1) for |x| <= 2^19 pi/2: pi/2 in 33-bit pieces (pio2_1, pio2_2, pio2_3 with their tails),
// one, two or three applied as the cancellation in r = |x| - n pio2_1 demands. Larger |x|: NaN.
__device__ inline int32_t rem_pio2(double x, double& y0, double& y1)
{
   const double invpio2 = 6.36619772367581382433e-01;
   const double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11;
   const double pio2_2 = 6.07710050630396597660e-11, pio2_2t = 2.02226624879595063154e-21;
   const double pio2_3 = 2.02226624871116645580e-21, pio2_3t = 8.47842766036889956997e-32;
   const int32_t hx = hi_of(x);
   const int32_t ix = hx & 0x7fffffff;
   if (ix <= 0x3fe921fb)
   {
      y0 = x;
      y1 = 0.0;
      return 0;
   }
   if (ix < 0x4002d97c)  // |x| < 3 pi/4: n = +-1
   {
      const double s = hx > 0 ? 1.0 : -1.0;
      double z = x - s * pio2_1;
      const double tail = ix != 0x3ff921fb ? pio2_1t : pio2_2t;
      if (ix == 0x3ff921fb)
         z -= s * pio2_2;
      y0 = z - s * tail;
      y1 = (z - y0) - s * tail;
      return hx > 0 ? 1 : -1;
   }
   if (ix > 0x413921fb)
   {
      y0 = y1 = __builtin_nan("");
      return 0;
   }
   const double ax = __builtin_fabs(x);
   const int32_t n = static_cast<int32_t>(ax * invpio2 + 0.5);
   const double fn = static_cast<double>(n);
   double r = ax - fn * pio2_1;
   double w = fn * pio2_1t;
   y0 = r - w;
   // High words of n pi/2 for n = 1..32: such an x cancels, so it always takes the exponent check.
   const bool nearMultiple =
      n >= 32 || ix == hi_of(fn * 1.57079632679489655800e+00);
   if (nearMultiple)
   {
      const int32_t j = ix >> 20;
      int32_t i = j - ((hi_of(y0) >> 20) & 0x7ff);
      if (i > 16)
      {
         double t = r;
         w = fn * pio2_2;
         r = t - w;
         w = fn * pio2_2t - ((t - r) - w);
         y0 = r - w;
         i = j - ((hi_of(y0) >> 20) & 0x7ff);
         if (i > 49)
         {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y0 = r - w;
         }
      }
   }
   y1 = (r - y0) - w;
   if (hx < 0)
   {
      y0 = -y0;
      y1 = -y1;
      return -n;
   }
   return n;
}

// cos(x + y) on [-pi/4, pi/4]: 1 - x^2/2 + x^4 C(x^2) (C1..C6); for |x| >= 0.3 the x^2/2 term is taken against qx
// (x/4 from the high word, or 0.28125 above 0.78125) so 1 - qx is exact.
__device__ inline double kcos(double x, double y)
{
   const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03, C3 = 2.48015872894767294178e-05,
                C4 = -2.75573143513906633035e-07, C5 = 2.08757232129817482790e-09,
                C6 = -1.13596475577881948265e-11;
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix < 0x3e400000 && static_cast<int>(x) == 0)
      return 1.0;
   const double z = x * x;
   const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
   if (ix < 0x3fd33333)
      return 1.0 - (0.5 * z - (z * r - x * y));
   const double qx = ix > 0x3fe90000 ? 0.28125 : of_words(static_cast<uint32_t>(ix - 0x00200000), 0);
   const double hz = 0.5 * z - qx;
   const double a = 1.0 - qx;
   return a - (hz - (z * r - x * y));
}

// sin(x + y) on [-pi/4, pi/4]: x + x^3 (S1 + x^2 S(x^2)) (S1..S6); tail = false when y is 0.
__device__ inline double ksin(double x, double y, bool tail)
{
   const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03, S3 = -1.98412698298579493134e-04,
                S4 = 2.75573137070700676789e-06, S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix < 0x3e400000 && static_cast<int>(x) == 0)
      return x;
   const double z = x * x;
   const double v = z * x;
   const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
   if (!tail)
      return x + v * (S1 + z * r);
   return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

__device__ inline double cos(double x)
{
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix <= 0x3fe921fb)
      return kcos(x, 0.0);
   if (ix >= 0x7ff00000)
      return x - x;
   double y0, y1;
   const int32_t n = rem_pio2(x, y0, y1);
   switch (n & 3)
   {
   case 0: return kcos(y0, y1);
   case 1: return -ksin(y0, y1, true);
   case 2: return -kcos(y0, y1);
   default: return ksin(y0, y1, true);
   }
}

__device__ inline double sin(double x)
{
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix <= 0x3fe921fb)
      return ksin(x, 0.0, false);
   if (ix >= 0x7ff00000)
      return x - x;
   double y0, y1;
   const int32_t n = rem_pio2(x, y0, y1);
   switch (n & 3)
   {
   case 0: return ksin(y0, y1, true);
   case 1: return kcos(y0, y1);
   case 2: return -ksin(y0, y1, true);
   default: return -kcos(y0, y1);
   }
}

// cbrt(x) as fdlibm's successor in FreeBSD msun computes it (the form V8 runs for Math.cbrt): a 5-bit estimate t
// from the exponent (hx / 3 + B1, or via 2^54 for subnormals), a polynomial P0..P4 in r = t^3 / x to 23 bits, t
// rounded away from zero to 23 bits, then one Newton step t + t (x / t^2 - t) / (2 t + x / t^2).
__device__ inline double cbrt(double x)
{
   constexpr uint32_t B1 = 715094163, B2 = 696219795;
   constexpr double P0 = 1.87595182427177009643, P1 = -1.88497979543377169875, P2 = 1.621429720105354466140,
                    P3 = -0.758397934778766047437, P4 = 0.145996192886612446982;
   int32_t hx = hi_of(x);
   const uint32_t low = lo_of(x);
   const uint32_t sign = static_cast<uint32_t>(hx) & 0x80000000u;
   hx = static_cast<int32_t>(static_cast<uint32_t>(hx) ^ sign);
   if (hx >= 0x7ff00000)
      return x + x;
   double t;
   if (hx < 0x00100000)
   {
      if ((static_cast<uint32_t>(hx) | low) == 0)
         return x;
      t = of_words(0x43500000u, 0u) * x;
      const uint32_t high = static_cast<uint32_t>(hi_of(t));
      t = of_words(sign | ((high & 0x7fffffffu) / 3 + B2), 0u);
   }
   else
   {
      t = of_words(sign | (static_cast<uint32_t>(hx) / 3 + B1), 0u);
   }
   double r = (t * t) * (t / x);
   t = t * ((P0 + r * (P1 + r * P2)) + ((r * r) * r) * (P3 + r * P4));
   const unsigned long long bits =
      (static_cast<unsigned long long>(__double_as_longlong(t)) + 0x80000000ull) & 0xffffffffc0000000ull;
   t = __longlong_as_double(static_cast<long long>(bits));
   const double s = t * t;
   r = x / s;
   const double w = t + t;
   r = (r - t) / (w + r);
   return t + t * r;
}

// V8's two-argument Math.hypot: both normalised by the larger, sqrt of the sum of squares, times it.
__device__ inline double hypot(double x, double y)
{
   const double a = fabs(x), b = fabs(y);
   if (isinf(a) || isinf(b))
      return __builtin_inf();
   if (isnan(a) || isnan(b))
      return __builtin_nan("");
   const double mx = a > b ? a : b;
   if (mx == 0)
      return 0;
   const double an = a / mx, bn = b / mx;
   return sqrt(an * an + bn * bn) * mx;
}

// V8's generic Math.hypot of three: normalised by the largest, a Kahan-compensated sum of squares in argument order.
__device__ inline double hypot(double x, double y, double z)
{
   const double v[3] = {fabs(x), fabs(y), fabs(z)};
   bool nan = false;
   double mx = 0;
   for (int i = 0; i < 3; ++i)
   {
      if (isnan(v[i]))
         nan = true;
      else if (v[i] > mx)
         mx = v[i];
   }
   if (isinf(mx))
      return __builtin_inf();
   if (nan)
      return __builtin_nan("");
   if (mx == 0)
      return 0;
   double sum = 0, comp = 0;
   for (int i = 0; i < 3; ++i)
   {
      const double n = v[i] / mx;
      const double summand = n * n - comp;
      const double pre = sum + summand;
      comp = (pre - sum) - summand;
      sum = pre;
   }
   return sqrt(sum) * mx;
}

}  // namespace jsm
}  // namespace iscg
