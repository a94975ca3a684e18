// Natural logarithm of a float Jaccard, in double, for the core/accessory
// regression (the fit of log J against k, src/dist/linear_regression.cpp:
// 16-60).  Host and device: the distance epilogue calls it on the GPU, and
// a host build of this same header checks its accuracy (exhaustively over
// the floats of [2^-12, 2], against a long-double log).
//
// x = m 2^e with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with
// s = (m - 1) / (m + 1), |s| <= 0.1716, summed through s^17 (the next term is
// below 3e-16), plus e ln 2 with ln 2 split into a short head and a tail
// (fdlibm's split) so that e * head is exact: about 1e-15 absolute, with a
// third of the instructions of the library log.  Zero, subnormal and
// non-finite values take the library log (ln 0 = -inf, as in the
// reference's fit).
#pragma once

#include <cmath>

__host__ __device__ inline double log_jaccard(float xf) {
  if (!(xf >= 1.17549435e-38f && xf <= 3.40282347e+38f)) return log((double)xf);
  int e = 0;
  double m = frexp((double)xf, &e);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double sv = (m - 1.0) / (m + 1.0);
  const double s2 = sv * sv;
  double poly = 2.0 / 17.0;
  poly = fma(poly, s2, 2.0 / 15.0);
  poly = fma(poly, s2, 2.0 / 13.0);
  poly = fma(poly, s2, 2.0 / 11.0);
  poly = fma(poly, s2, 2.0 / 9.0);
  poly = fma(poly, s2, 2.0 / 7.0);
  poly = fma(poly, s2, 2.0 / 5.0);
  poly = fma(poly, s2, 2.0 / 3.0);
  poly = fma(poly, s2, 2.0);
  const double lnm = sv * poly;
  const double ed = (double)e;
  return fma(ed, 6.93147180369123816490e-01, fma(ed, 1.90821492927058770002e-10, lnm));
}
