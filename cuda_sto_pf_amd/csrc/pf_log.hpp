// Natural logarithm of a double, short enough for per-voxel FP64 work.
//
// The chemical potentials take two logarithms per voxel and their finite
// difference four; the library routine spends about 100 vector instructions
// on each (double-double arithmetic).  This one spends about 40 and stays
// within one ulp, which the FP64 path's tolerance (1e-10 against the CPU
// restatement) leaves ample room for.
//
// Method (the classical atanh reduction, as in Sun's fdlibm e_log.c and
// Muller, "Elementary Functions", ch. 11), written out from the identities:
//   x = m 2^e with m in [1/sqrt2, sqrt2)       (frexp, then one doubling)
//   f = m - 1                                  (exact: Sterbenz)
//   s = f / (2 + f),  log(m) = 2 atanh(s) = 2s + s R(s^2),
//   R(z) = sum_{k>=1} 2 z^k / (2k + 1)         (the Taylor series of atanh)
//   2s = f - s f  =>  log(m) = f - (f^2/2 - s (f^2/2 + R)),
// so the exact f carries the leading bits and the rounding of s only reaches
// a term below f^3/4.  |s| <= 3 - 2 sqrt2 < 0.1716, z = s^2 < 0.02944; the
// coefficients 2/(2k+1) through k = 10 leave a truncation error below 1e-18
// relative.  e ln2 + f is carried as an exact pair of doubles and the small
// terms are added to it last, so the result is rounded about once.  Arguments
// outside [DBL_MIN, DBL_MAX] (zero, negative, subnormal, infinite, NaN) go to
// the library log.
//
// Every multiply-add that matters is an explicit fma, so the host and device
// compilers produce the same bits whatever their contraction settings; the
// host build of this header is what the accuracy test measures.
#pragma once
#include <math.h>

#ifndef PF_LOG_HD
#if defined(__HIPCC__)
#define PF_LOG_HD __host__ __device__ __forceinline__
#else
#define PF_LOG_HD inline
#endif
#endif

namespace pf {

PF_LOG_HD double pf_log(double x) {
  if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) return log(x);
  int e;
  double m = frexp(x, &e);   // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double r = 2.0 / 21.0;
  r = fma(r, z, 2.0 / 19.0);
  r = fma(r, z, 2.0 / 17.0);
  r = fma(r, z, 2.0 / 15.0);
  r = fma(r, z, 2.0 / 13.0);
  r = fma(r, z, 2.0 / 11.0);
  r = fma(r, z, 2.0 / 9.0);
  r = fma(r, z, 2.0 / 7.0);
  r = fma(r, z, 2.0 / 5.0);
  r = fma(r, z, 2.0 / 3.0);   // R(z) = z r
  const double hfsq = 0.5 * (f * f);
  const double t = fma(-s, fma(r, z, hfsq), hfsq);   // log(m) = f - t
  // e ln2 = e LN2_HI + e LN2_LO; LN2_HI has 42 significant bits, so e LN2_HI
  // is exact for every exponent of a double
  const double LN2_HI = 0x1.62e42fefa3800p-1;
  const double LN2_LO = 0x1.ef35793c76730p-45;
  const double de = double(e);
  const double hi = de * LN2_HI;
  // hi + f as an unevaluated pair (s1, err): Fast2Sum, exact because
  // |hi| >= ln2 > |f| whenever hi != 0
  const double s1 = hi + f;
  const double err = (hi - s1) + f;
  return s1 + (err - fma(-de, LN2_LO, t));
}

}  // namespace pf
