// Device helpers shared by several kernel files (HIP sources only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qmbed {

// Entry i of the deterministic start vector (QMBED src/solve/mod.rs:729-742):
// the (i+1)-th state of the LCG x <- a x + c from x0 = seed | 1, mapped to
// [-0.5, 0.5), reached by jump-ahead (square-and-multiply of the affine map).
__device__ __forceinline__ double deterministic_entry(uint64_t i, uint64_t seed) {
  const uint64_t A = 6364136223846793005ULL, C = 1442695040888963407ULL;
  uint64_t steps = i + 1;
  uint64_t ra = 1, rc = 0;  // accumulated affine map
  uint64_t pa = A, pc = C;  // current power
  while (steps) {
    if (steps & 1) {  // r = p ∘ r
      rc = pa * rc + pc;
      ra = pa * ra;
    }
    pc = pa * pc + pc;
    pa = pa * pa;
    steps >>= 1;
  }
  const uint64_t x = ra * (seed | 1ULL) + rc;
  return double(x >> 11) / 9007199254740992.0 - 0.5;
}

}  // namespace qmbed
