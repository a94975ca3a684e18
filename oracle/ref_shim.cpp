/*
 * C entry point around the reference's calc_intersize (dist/dist.cpp:50-78),
 * which takes std::vector pointers.  Test infrastructure: lets the oracle's
 * restatement be cross-checked against the reference's own compiled loop.
 */
#include <cstddef>
#include <cstdint>
#include <vector>

#include "dist/dist.hpp"

extern "C" size_t ref_calc_intersize(const uint64_t *s1, const uint64_t *s2,
                                     size_t sketchsize64, size_t bbits) {
  std::vector<uint64_t> a(s1, s1 + sketchsize64 * bbits);
  std::vector<uint64_t> b(s2, s2 + sketchsize64 * bbits);
  return calc_intersize(&a, &b, sketchsize64, bbits);
}
