// Host half of the out-of-core tier.  The reference (GPUSCAN++O/Graph.cpp:
// 414-576) cuts edge ranges into chips and copies both endpoint rows of every
// edge into each chip; here rows are cut into blocks and a chip holds two
// blocks, copied in parallel into pinned staging.
#include "ooc_host.h"

#include <string.h>

#include <algorithm>

#include "scan_rules.h"

namespace sgc {

void ooc_degrees(const sgc_host_graph* g, uint32_t* deg) {
  const uint64_t* rp = g->row_ptr;
#pragma omp parallel for schedule(static)
  for (uint64_t x = 0; x < g->n; ++x) deg[x] = (uint32_t)(rp[x + 1] - rp[x]);
}

int ooc_plan_rows(const sgc_host_graph* g, uint64_t cap, uint64_t row_cap, std::vector<uint64_t>* bounds,
                  uint64_t* max_entries, uint64_t* max_rows) {
  const uint64_t* rp = g->row_ptr;
  if (row_cap < 1) row_cap = 1;
  bounds->assign(1, 0);
  uint64_t start = 0, worst = 0, most = 0;
  for (uint64_t x = 0; x < g->n; ++x) {
    const uint64_t d = rp[x + 1] - rp[x];
    if (d > cap)
      return fail(SGC_ENOMEM, "row %llu has %llu entries, more than a block of the device budget holds (%llu)",
                  (unsigned long long)x, (unsigned long long)d, (unsigned long long)cap);
    if (rp[x + 1] - rp[start] > cap || x + 1 - start > row_cap) {
      bounds->push_back(x);
      worst = std::max(worst, rp[x] - rp[start]);
      most = std::max(most, x - start);
      start = x;
    }
  }
  bounds->push_back(g->n);
  worst = std::max(worst, rp[g->n] - rp[start]);
  most = std::max(most, g->n - start);
  if (max_entries) *max_entries = worst;
  if (max_rows) *max_rows = most;
  return SGC_OK;
}

void ooc_copy_rows(const sgc_host_graph* g, uint64_t r0, uint64_t r1, uint32_t base, int32_t* dst, uint32_t* off) {
  const uint64_t* rp = g->row_ptr;
  const uint64_t b = rp[r0], m = rp[r1] - b;
  const uint64_t nr = r1 - r0;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i <= nr; ++i) off[i] = base + (uint32_t)(rp[r0 + i] - b);
  if (!dst) return;
  const uint64_t piece = 1 << 20;
#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t p = 0; p < m; p += piece)
    memcpy(dst + p, g->adj + b + p, (size_t)std::min(piece, m - p) * sizeof(int32_t));
}

void ooc_merge_rows(const sgc_host_graph* g, uint64_t r0, uint64_t r1, const uint8_t* chip_flag, uint8_t* flag) {
  const uint64_t b = g->row_ptr[r0], m = g->row_ptr[r1] - b;
  uint8_t* out = flag + b;
  // a blend rather than a conditional store, so that the loop vectorises
#pragma omp parallel for schedule(static)
  for (uint64_t k = 0; k < m; ++k) {
    const uint8_t c = chip_flag[k], o = out[k];
    out[k] = c <= kFlagY ? c : o;
  }
}

}  // namespace sgc
