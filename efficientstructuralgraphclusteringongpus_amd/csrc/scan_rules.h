// Per-entry flag values and the degree-only pre-decision, shared by the
// kernels and by host code (the out-of-core chip planner must mark exactly
// the edges the prep kernel leaves undecided).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SGC_RULE __host__ __device__ __forceinline__
#else
#define SGC_RULE static inline
#endif

namespace sgc {

// Per-edge flag byte stored for every directed CSR entry.  Undecided edges are
// marked in both rows: kFlagUDown in the higher (degree, id) endpoint's row,
// which owns the computation, kFlagUUp in the other row, which receives the
// result.  After the similarity phase every flag is kFlagN or kFlagY.
// kFlagUSkip: an owned undecided edge that an out-of-core chip leaves to the
// chip holding both of its endpoints' blocks.
// kFlagDeferUp / kFlagDeferDown: an undecided edge the pruned core check
// never computed because both endpoints were decided (the rows as for
// kFlagUUp / kFlagUDown); decided later only if a later pass needs it.
enum : uint8_t {
  kFlagN = 0,
  kFlagY = 1,
  kFlagUUp = 2,
  kFlagUDown = 3,
  kFlagUSkip = 4,
  kFlagDeferUp = 5,
  kFlagDeferDown = 6
};

// Degree-only pre-decision of the directed entry (u, v) with degrees du, dv
// (GPUSCAN++/GPUScan.cu:143-169; orientation GPUSCAN++/Graph.cpp:75-83): the
// edge's lower (degree, id) endpoint a and the other endpoint b give
// A = d(a) + 1, B = d(b) + 1; N if A < B*eps*eps, Y if eps*eps*A*B <= 4,
// else undecided, owned by b.
SGC_RULE uint8_t prep_flag(uint32_t u, uint32_t v, uint32_t du, uint32_t dv, double eps) {
  const bool up = dv > du || (dv == du && v > u);  // u is the lower endpoint
  const uint32_t da = up ? du : dv, db = up ? dv : du;
  const double A = (double)(da + 1u);
  const double B = (double)(db + 1u);
  if (A < B * eps * eps) return kFlagN;
  if (eps * eps * A * B <= 4.0) return kFlagY;
  return up ? kFlagUUp : kFlagUDown;
}

}  // namespace sgc
