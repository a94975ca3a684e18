// Bin fraction of a match count, c / bins with bins = 64 s64 (the observed
// Jaccard of src/reference.cpp:86-96 before the random-match correction),
// without a division.  Host and device: the distance epilogue calls it on the
// GPU, and a host build of this same header checks it against the IEEE
// division for every sketch size the library accepts (s64 in [1, 1023]) and
// every count 0 .. bins, 33.5M cases, all equal bit for bit.
//
// rbins = 1 / bins (one division per tile).  q = c rbins is within a few
// ulps of c / bins; the remainder c - q bins is exact in one fma, and one
// more fma adds its quotient: 3 instructions instead of the division's 10.
#pragma once

#include <cmath>
#include <cstdint>

__host__ __device__ inline double bin_fraction(uint32_t c, double bins, double rbins) {
  const double cd = (double)c;
  const double q = cd * rbins;
  return fma(fma(-q, bins, cd), rbins, q);
}
