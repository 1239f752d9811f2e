// Seeded random propagator fields: a splitmix64 finaliser per element, keyed
// by the global site, so the field does not depend on the process grid
// (src/provider.cpp:10-57).
#pragma once

#include "lattice.hpp"

namespace hlbl {

// splitmix64 output function: add the golden-ratio increment, then two
// xor-shift-multiply rounds and a final xor-shift.
HLBL_HD uint64_t hash64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// top 53 bits of the hash as a double in [-1, 1)
HLBL_HD double hash_to_unit(uint64_t x) {
  return (double)(hash64(x) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// the key of element (f, a, c) at the global site with linear key g
HLBL_HD uint64_t element_key(uint64_t g, int f, int a, int c, uint64_t seed) {
  return ((g * 2 + f) * 12 + a) * 24 + c + seed * 0x2545F4914F6CDD1DULL;
}

// linear key of a global site, base 1000003 (provider.cpp:38-43)
HLBL_HD uint64_t site_key(int t, int x, int y, int z) {
  return (((uint64_t)t * 1000003u + x) * 1000003u + y) * 1000003u + z;
}

}  // namespace hlbl
