// gemhip_math.h -- constants and scalar arithmetic shared by every transport kernel.
//
// Everything here must round exactly like the reference build (gfortran for the
// x86-64 baseline: no FMA, IEEE divide and sqrt, glibc pow).  The library is
// compiled with -ffp-contract=off, so a*b+c stays two rounded operations; divides
// stay divides (never a reciprocal multiply).  The header also compiles as plain
// C++ on the host (GEMHIP_HD expands to nothing) so the CPU tests can check the
// scalar routines against libm.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define GEMHIP_HD __host__ __device__ __forceinline__
#else
#define GEMHIP_HD inline
#endif

namespace gemhip {

constexpr int LSP = 7;  // phys_consts.in.f90:41

// phys_consts.in.f90:24-26, 82-84
constexpr double kB = 1.38064852e-23;
constexpr double elchrg = 1.60217662e-19;
constexpr double amu = 1.660539040e-27;
constexpr double mindensdiv = 1.0e-5;
constexpr double mindens = 1.0e-100, mindensnull = 1.0e-20;

// ms = [16,30,28,32,14,1,5.485799090e-4]*amu (phys_consts.in.f90:43); the product is
// a single IEEE multiply, identical whether folded at compile time or not.
GEMHIP_HD double ms(int isp) {
  return (isp == 1 ? 16.0 : isp == 2 ? 30.0 : isp == 3 ? 28.0 : isp == 4 ? 32.0
          : isp == 5 ? 14.0 : isp == 6 ? 1.0 : 5.485799090e-4) * amu;
}
// qs = [1,1,1,1,1,1,-1]*elchrg (phys_consts.in.f90:45)
GEMHIP_HD double qs(int isp) { return (isp < LSP ? 1.0 : -1.0) * elchrg; }
// gammas (phys_consts.in.f90:47-53)
GEMHIP_HD double gammas(int isp) {
  return (isp == 2 || isp == 3 || isp == 4) ? 7.0 / 5.0 : 5.0 / 3.0;
}
// collisions.f90:51-53: default-real literals converted to real64, so each value is
// the single-precision number nearest the decimal, widened.
GEMHIP_HD double thermal_coeff(int isp) {
  return isp == 1 ? (double)0.1019e-12f : isp == 2 ? (double)0.0747e-12f
       : isp == 3 ? (double)0.0754e-12f : isp == 4 ? (double)0.0701e-12f
       : isp == 5 ? (double)0.1068e-12f : (double)0.3986e-12f;
}

// Fortran MIN/MAX as gfortran evaluates them on ordered operands.
GEMHIP_HD double fmin_f(double a, double b) { return (b < a) ? b : a; }
GEMHIP_HD double fmax_f(double a, double b) { return (b > a) ? b : a; }

// x / b from y = RN(1/b), correctly rounded: q0 = RN(x y) is within a few ulps of x/b,
// the FMA residual x - b q0 is exact, and one correction step q0 + r y, rounded once,
// gives RN(x/b) (Markstein's theorem for a correctly rounded reciprocal).  Where the
// residual is zero (exact quotients, zero numerators: q0 has the sign IEEE division gives)
// or not a number (infinite numerators) q0 is the quotient.
__device__ __forceinline__ double div_rcp(double x, double b, double y) {
  const double q0 = x * y;
  const double r = __builtin_fma(-b, q0, x);
  const double q1 = __builtin_fma(r, y, q0);
  return __builtin_amdgcn_class(r, 0x63) ? q0 : q1;    // sNaN qNaN -0 +0
}

// advec.f90:438-447
GEMHIP_HD double minmod(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  return (fabs(a) < fabs(b)) ? a : b;
}

GEMHIP_HD double fma_exact(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return fma(a, b, c);
#endif
}

// x**2.5 rounded to nearest, which is what glibc's pow returns for these arguments
// (its error bound is below 0.52 ulp and exact-or-correct for x^n.5 in practice; the
// CPU test checks agreement on millions of samples).  x^2.5 = x^2 * sqrt(x) is formed
// in double-double: x^2 exactly by an FMA product split, sqrt(x) to ~2^-104 by one
// Newton correction, then one final rounding.  Arguments outside the normal range the
// transport path sees fall back to the library pow.
GEMHIP_HD double pow25(double x) {
  if (!(x > 0x1p-400 && x < 0x1p+400)) return pow(x, 2.5);
  const double s = sqrt(x);                 // correctly rounded
  const double e = fma_exact(-s, s, x);     // x - s*s exactly
  const double sl = e / (2.0 * s);          // sqrt(x) ~= s + sl
  const double qh = x * x;
  const double ql = fma_exact(x, x, -qh);   // x*x = qh + ql exactly
  const double ph = qh * s;
  const double pl = fma_exact(qh, s, -ph) + (qh * sl + ql * s);
  return ph + pl;
}

// x**1.5 rounded to nearest in the same way: x * sqrt(x) in double-double, one rounding.
GEMHIP_HD double pow15(double x) {
  if (!(x > 0x1p-400 && x < 0x1p+400)) return pow(x, 1.5);
  const double s = sqrt(x);
  const double e = fma_exact(-s, s, x);
  const double sl = e / (2.0 * s);
  const double ph = x * s;
  const double pl = fma_exact(x, s, -ph) + x * sl;
  return ph + pl;
}

}  // namespace gemhip
