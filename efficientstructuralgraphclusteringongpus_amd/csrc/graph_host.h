// Host CSR container shared by the loader, generators, the device upload and
// the out-of-core tier.
#pragma once

#include <stdint.h>

#include "common.h"

struct sgc_host_graph {
  uint64_t n = 0;
  uint64_t m = 0;
  uint64_t* row_ptr = nullptr;  // n + 1 offsets
  int32_t* adj = nullptr;       // m neighbour ids, rows strictly increasing
  bool pinned = false;          // adj allocated with hipHostMalloc
};

namespace sgc {

sgc_host_graph* new_host_graph(uint64_t n, uint64_t m, bool pinned);
// Full contract check: ids < n, rows strictly increasing, no self loops and,
// if requested, every (u,v) has its (v,u).
int validate_graph(const sgc_host_graph* g, bool check_symmetry);
// Builds a simple undirected graph from `count` directed draws: self loops
// dropped, both directions added, duplicates removed, rows sorted.  With
// `release` (the malloc'd src and dst), both are freed, and release[0..1]
// set to null, as soon as they are no longer read.
int build_from_pairs(uint64_t n, uint64_t count, const uint32_t* src, const uint32_t* dst,
                     sgc_host_graph** out, uint32_t** release = nullptr);
// g's row offsets narrowed to uint32 into out[0..n] (the caller guarantees
// m < 2^32) and its largest degree, over OpenMP threads.
void narrow_row_ptr(const sgc_host_graph* g, uint32_t* out, uint64_t* max_degree);
// memcpy over OpenMP threads, for large copies into page-locked staging.
void parallel_copy(void* dst, const void* src, size_t bytes);

}  // namespace sgc
