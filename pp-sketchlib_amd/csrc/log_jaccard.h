// Natural logarithm of a float Jaccard, in double, for the core/accessory
// regression (the fit of log J against k, src/dist/linear_regression.cpp:
// 16-60).  Host and device: the distance epilogue calls it on the GPU, and
// a host build of this same header checks its accuracy (exhaustively over
// the floats of [2^-12, 2], against a long-double log).
//
// x = m 2^e with m in [1, 2).  The top 8 fraction bits of m pick one of 256
// buckets [1 + i/256, 1 + (i+1)/256); the table holds c_i, the double nearest
// 1 / (1 + (i + 1/2) / 256), and t_i, the double nearest -ln c_i (bucket 0:
// c = 1, t = 0, so that ln 1 = 0 exactly).  Then ln m = t_i + ln(1 + u) with
// u = m c_i - 1 (one fma; |u| <= 2^-9, 2^-8 in bucket 0), whose series is
// summed through u^6 (the next term is below 2^-56 / 7), and
// ln x = e ln 2 + t_i + ln(1 + u) with ln 2 split into a short head and a tail
// (fdlibm's split) so that e * head is exact: about 1e-15 absolute over the
// Jaccard range, with no division and no branch.  Zero gives -inf, negative
// values and NaN give NaN, +inf gives +inf, as the library log does.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

// Entries of the table: c_i at [2 i], t_i at [2 i + 1].
constexpr int LOG_TAB_BUCKETS = 256;
constexpr int LOG_TAB_DOUBLES = 2 * LOG_TAB_BUCKETS;

// Host: fill the table (the device gets a copy of these same doubles).
inline void log_table_fill(double* tab) {
  for (int i = 0; i < LOG_TAB_BUCKETS; ++i) {
    const double c = i == 0 ? 1.0 : 1.0 / (1.0 + (i + 0.5) / LOG_TAB_BUCKETS);
    tab[2 * i] = c;
    tab[2 * i + 1] = i == 0 ? 0.0 : (double)(-std::log((long double)c));
  }
}

__host__ __device__ inline double log_jaccard(float xf, const double* tab) {
  const double x = (double)xf;
  int e = 0;
  const double m = 2.0 * frexp(x, &e);  // [1, 2) for finite nonzero x
  uint64_t bits;
  std::memcpy(&bits, &m, sizeof(bits));
  const int i = (int)((bits >> 44) & (LOG_TAB_BUCKETS - 1));  // in range for any x
  const double c = tab[2 * i];
  const double t = tab[2 * i + 1];
  const double u = fma(m, c, -1.0);
  double p = fma(-1.0 / 6.0, u, 0.2);
  p = fma(p, u, -0.25);
  p = fma(p, u, 1.0 / 3.0);
  p = fma(p, u, -0.5);
  p = fma(p, u, 1.0);
  const double lm = fma(u, p, t);  // ln m
  const double ed = (double)(e - 1);
  double y = fma(ed, 6.93147180369123816490e-01, fma(ed, 1.90821492927058770002e-10, lm));
  if (!(xf > 0.0f)) y = xf == 0.0f ? -INFINITY : NAN;
  if (xf == INFINITY) y = INFINITY;
  return y;
}
