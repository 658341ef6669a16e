// Philox4x32-10 counter-based random number generator.
//
// Written from the published description in J. K. Salmon, M. A. Moraes,
// R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11): a 4x32-bit counter block is mixed by ten rounds of two 32x32->64
// multiplies with a Weyl-sequence key schedule.  The function is pure integer
// arithmetic, so the host generator and any device generator built from this
// header produce bit-identical streams for the same (key, counter).
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define SGC_HD __host__ __device__ __forceinline__
#else
#define SGC_HD static inline
#endif

namespace sgc {

struct Philox4x32 {
  uint32_t v[4];
};

SGC_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

// Ten rounds; multipliers and key increments are the constants published with
// the Philox4x32 family.
SGC_HD Philox4x32 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c0, &hi0, &lo0);
    philox_mulhilo(M1, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  Philox4x32 out;
  out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
  return out;
}

// R-MAT quadrant thresholds in 2^-32 units: a draw r (uniform uint32) lands in
// quadrant (0,0) if r < ta, (0,1) if r < tab, (1,0) if r < tabc, else (1,1).
struct RmatThresholds {
  uint32_t ta, tab, tabc;
};

SGC_HD uint32_t rmat_prob_to_u32(double p) {
  if (p <= 0.0) return 0u;
  if (p >= 1.0) return 0xFFFFFFFFu;
  return (uint32_t)(p * 4294967296.0);
}

// Edge draw `i` of an R-MAT graph with 2^scale vertices.  Level L (0 = most
// significant bit) uses 32-bit word L%4 of Philox block
// (counter = {i_lo, i_hi, L/4, 0xC0FFEE00}, key = {seed_lo, seed_hi}).
SGC_HD void rmat_draw(uint64_t i, int scale, uint64_t seed, RmatThresholds t,
                      uint32_t* src, uint32_t* dst) {
  uint32_t u = 0, v = 0;
  Philox4x32 blk;
  blk.v[0] = blk.v[1] = blk.v[2] = blk.v[3] = 0;
  for (int L = 0; L < scale; ++L) {
    if ((L & 3) == 0)
      blk = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), (uint32_t)(L >> 2), 0xC0FFEE00u,
                          (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t r = blk.v[L & 3];
    uint32_t bit = 1u << (scale - 1 - L);
    if (r < t.ta) {
    } else if (r < t.tab) {
      v |= bit;
    } else if (r < t.tabc) {
      u |= bit;
    } else {
      u |= bit;
      v |= bit;
    }
  }
  *src = u;
  *dst = v;
}

}  // namespace sgc
