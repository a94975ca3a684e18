/*
 * C entry points around the reference's random number generators, built
 * from its sources where they lie (random/rng.cpp: xoshiro256** with its
 * seeding and jump; random/kmeans/mersenneTwister2002.c: the MT19937 behind
 * its k-means seeding).  Test infrastructure: pins the generators of the
 * Monte Carlo random-match construction.
 */
#include <cstdint>

#include "random/rng.hpp"
#include "random/kmeans/mersenneTwister2002.c"

// n outputs of Xoshiro(seed) after `jumps` calls of jump().
extern "C" void ref_xoshiro(uint64_t seed, int jumps, uint64_t n, uint64_t *out) {
  Xoshiro g(seed);
  for (int j = 0; j < jumps; ++j) g.jump();
  for (uint64_t i = 0; i < n; ++i) out[i] = g();
}

// n values of genrand_double() after init_genrand(seed).
extern "C" void ref_mt_double(uint32_t seed, uint64_t n, double *out) {
  init_genrand(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = genrand_double();
}
