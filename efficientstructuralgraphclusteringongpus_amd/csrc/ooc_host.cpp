// Host half of the out-of-core tier (GPUSCAN++O/Graph.cpp:414-576 plans
// chips by edge ranges with a per-edge byte cost and copies each chip's
// endpoint rows; here a chip is a range of owner rows plus the short rows
// their undecided edges read, assembled in parallel into pinned buffers).
#include "ooc_host.h"

#include <omp.h>
#include <string.h>

#include <algorithm>

#include "scan_rules.h"

namespace sgc {

void ooc_degrees(const sgc_host_graph* g, uint32_t* deg) {
  const uint64_t* rp = g->row_ptr;
#pragma omp parallel for schedule(static)
  for (uint64_t x = 0; x < g->n; ++x) deg[x] = (uint32_t)(rp[x + 1] - rp[x]);
}

// Entries a chip holding owner row b needs: the row itself plus the short
// row of every undecided edge it owns.
static uint64_t owner_cost(const sgc_host_graph* g, const uint32_t* deg, double eps, uint64_t b) {
  uint64_t c = deg[b];
  for (uint64_t k = g->row_ptr[b]; k < g->row_ptr[b + 1]; ++k) {
    const uint32_t a = (uint32_t)g->adj[k];
    if (prep_flag((uint32_t)b, a, deg[b], deg[a], eps) == kFlagUDown) c += deg[a];
  }
  return c;
}

static int greedy_ranges(const std::vector<uint64_t>& cost, uint64_t cap, std::vector<uint64_t>* bounds,
                         uint64_t* max_cost) {
  bounds->clear();
  bounds->push_back(0);
  uint64_t acc = 0, worst = 0;
  for (uint64_t x = 0; x < cost.size(); ++x) {
    if (cost[x] > cap)
      return fail(SGC_ENOMEM, "row %llu needs %llu entries in one chip, more than the device budget allows (%llu)",
                  (unsigned long long)x, (unsigned long long)cost[x], (unsigned long long)cap);
    if (acc + cost[x] > cap) {
      bounds->push_back(x);
      worst = std::max(worst, acc);
      acc = 0;
    }
    acc += cost[x];
  }
  bounds->push_back(cost.size());
  worst = std::max(worst, acc);
  if (max_cost) *max_cost = worst;
  return SGC_OK;
}

// Entries owner row b adds to the open chip [r0, b): its row plus the short
// rows of its undecided edges not stamped with the chip's tag and not owner
// rows already in the chip.  With take, stamps them.
static uint64_t owner_increment(const sgc_host_graph* g, const uint32_t* deg, double eps, uint64_t r0, uint64_t b,
                                int32_t* stamp, int32_t tag, bool take) {
  uint64_t c = deg[b];
  for (uint64_t k = g->row_ptr[b]; k < g->row_ptr[b + 1]; ++k) {
    const uint32_t a = (uint32_t)g->adj[k];
    if (a >= r0 && a < b) continue;
    if (prep_flag((uint32_t)b, a, deg[b], deg[a], eps) != kFlagUDown) continue;
    if (__atomic_load_n(&stamp[a], __ATOMIC_RELAXED) == tag) continue;
    c += deg[a];
    if (take) __atomic_store_n(&stamp[a], tag, __ATOMIC_RELAXED);
  }
  return c;
}

int ooc_plan_similarity(const sgc_host_graph* g, const uint32_t* deg, double eps, uint64_t cap,
                        std::vector<uint64_t>* bounds, uint64_t* max_cost) {
  // Rows are cut into one block per thread (by entries); each thread plans its
  // block greedily, counting a short row once per chip.  A stamp overwritten
  // by another thread only makes a row count twice, so every chip's count
  // stays an upper bound of what ooc_build_chip stores.
  const uint64_t n = g->n;
  if (g->m <= cap) {  // every row is an owner row: one chip, no short rows
    bounds->assign({0, n});
    if (max_cost) *max_cost = g->m;
    return SGC_OK;
  }
  const int nt = std::max(1, std::min<int>(omp_get_max_threads(), (int)std::max<uint64_t>(n / 4096, 1)));
  std::vector<uint64_t> cut(nt + 1);
  for (int t = 0; t <= nt; ++t) {
    const uint64_t target = g->row_ptr[n] / nt * t;
    cut[t] = t == nt ? n : (uint64_t)(std::lower_bound(g->row_ptr, g->row_ptr + n, target) - g->row_ptr);
  }
  std::vector<int32_t> stamp(n, -1);
  std::vector<std::vector<uint64_t>> part(nt);
  std::vector<uint64_t> worst(nt, 0);
  std::vector<int> err(nt, 0);
  std::vector<uint64_t> bad_row(nt, 0), bad_cost(nt, 0);
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num();
    int32_t tag = t << 20;  // 2^20 chips per thread before tags repeat
    uint64_t r0 = cut[t], acc = 0;
    for (uint64_t b = cut[t]; b < cut[t + 1] && !err[t]; ++b) {
      // count and stamp in one go (a stamp may change under another thread);
      // if the row does not fit, close the chip and count it again afresh
      uint64_t c = owner_increment(g, deg, eps, r0, b, stamp.data(), tag, true);
      if (acc + c > cap && b > r0) {
        part[t].push_back(b);
        worst[t] = std::max(worst[t], acc);
        r0 = b;
        acc = 0;
        tag = (tag & ~0xFFFFF) | ((tag + 1) & 0xFFFFF);
        c = owner_increment(g, deg, eps, r0, b, stamp.data(), tag, true);
      }
      if (c > cap) {
        err[t] = 1;
        bad_row[t] = b;
        bad_cost[t] = c;
      }
      acc += c;
    }
    worst[t] = std::max(worst[t], acc);
  }
  for (int t = 0; t < nt; ++t)
    if (err[t])
      return fail(SGC_ENOMEM, "row %llu needs %llu entries in one chip, more than the device budget allows (%llu)",
                  (unsigned long long)bad_row[t], (unsigned long long)bad_cost[t], (unsigned long long)cap);
  bounds->assign(1, 0);
  uint64_t w = 0;
  for (int t = 0; t < nt; ++t) {
    for (uint64_t b : part[t]) bounds->push_back(b);
    if (t + 1 < nt && cut[t + 1] > cut[t]) bounds->push_back(cut[t + 1]);
    w = std::max(w, worst[t]);
  }
  if (bounds->back() != n) bounds->push_back(n);
  // drop empty ranges (blocks without rows)
  bounds->erase(std::unique(bounds->begin(), bounds->end()), bounds->end());
  if (bounds->size() == 1) bounds->push_back(n);
  if (max_cost) *max_cost = w;
  return SGC_OK;
}

uint64_t ooc_max_row_cost(const sgc_host_graph* g, const uint32_t* deg, double eps) {
  uint64_t worst = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : worst)
  for (uint64_t b = 0; b < g->n; ++b) worst = std::max(worst, owner_cost(g, deg, eps, b));
  return worst;
}

int ooc_plan_rows(const sgc_host_graph* g, uint64_t cap, std::vector<uint64_t>* bounds, uint64_t* max_entries) {
  std::vector<uint64_t> cost(g->n);
#pragma omp parallel for schedule(static)
  for (uint64_t x = 0; x < g->n; ++x) cost[x] = g->row_ptr[x + 1] - g->row_ptr[x];
  return greedy_ranges(cost, cap, bounds, max_entries);
}

void ooc_build_chip(const sgc_host_graph* g, const uint32_t* deg, double eps, uint64_t r0, uint64_t r1,
                    int32_t* stamp, int32_t tag, int32_t* adj, uint32_t* off, OocChip* chip) {
  const uint64_t* rp = g->row_ptr;
  const uint64_t base = rp[r0];
  chip->r0 = r0;
  chip->r1 = r1;
  chip->m_own = rp[r1] - base;
  const uint64_t nr = r1 - r0;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i <= nr; ++i) off[i] = (uint32_t)(rp[r0 + i] - base);
  // short rows, each claimed once through its stamp
  const int nt = omp_get_max_threads();
  std::vector<std::vector<uint32_t>> found(nt);
#pragma omp parallel
  {
    std::vector<uint32_t>& mine = found[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
    for (uint64_t b = r0; b < r1; ++b) {
      for (uint64_t k = rp[b]; k < rp[b + 1]; ++k) {
        const uint32_t a = (uint32_t)g->adj[k];
        if (a >= r0 && a < r1) continue;
        if (prep_flag((uint32_t)b, a, deg[b], deg[a], eps) != kFlagUDown) continue;
        int32_t old = __atomic_load_n(&stamp[a], __ATOMIC_RELAXED);
        while (old != tag) {
          if (__atomic_compare_exchange_n(&stamp[a], &old, tag, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
            mine.push_back(a);
            break;
          }
        }
      }
    }
  }
  chip->ids.clear();
  for (auto& v : found) chip->ids.insert(chip->ids.end(), v.begin(), v.end());
  std::sort(chip->ids.begin(), chip->ids.end());  // deterministic layout
  chip->starts.resize(chip->ids.size());
  uint64_t pos = chip->m_own;
  for (size_t i = 0; i < chip->ids.size(); ++i) {
    chip->starts[i] = (uint32_t)pos;
    pos += deg[chip->ids[i]];
  }
  chip->m_chip = pos;
  memcpy(adj, g->adj + base, chip->m_own * sizeof(int32_t));
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t i = 0; i < chip->ids.size(); ++i) {
    const uint32_t a = chip->ids[i];
    memcpy(adj + chip->starts[i], g->adj + rp[a], (size_t)deg[a] * sizeof(int32_t));
  }
}

void ooc_merge_flags(const sgc_host_graph* g, const uint32_t* deg, const OocChip& chip, const uint8_t* chip_flag,
                     uint8_t* flag) {
  const uint64_t base = g->row_ptr[chip.r0];
#pragma omp parallel for schedule(static)
  for (uint64_t k = 0; k < chip.m_own; ++k)
    if (chip_flag[k] <= kFlagY) flag[base + k] = chip_flag[k];
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t i = 0; i < chip.ids.size(); ++i) {
    const uint32_t a = chip.ids[i];
    const uint8_t* cf = chip_flag + chip.starts[i];
    uint8_t* hf = flag + g->row_ptr[a];
    for (uint32_t j = 0; j < deg[a]; ++j)
      if (cf[j] <= kFlagY) hf[j] = cf[j];
  }
}

}  // namespace sgc
