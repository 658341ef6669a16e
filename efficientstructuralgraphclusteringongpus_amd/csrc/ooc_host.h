// Host half of the out-of-core tier: row-block planning, block copies into
// pinned staging and the merge of chip flags into the whole-graph flag array
// (g++/OpenMP; the device half is in ooc.hip).
#pragma once

#include <stdint.h>

#include <vector>

#include "graph_host.h"

namespace sgc {

// Degrees as uint32 (every degree is < 2^31 for a graph the OOC tier takes).
void ooc_degrees(const sgc_host_graph* g, uint32_t* deg);

// Consecutive rows cut into blocks bounds[i] .. bounds[i+1] of at most cap
// entries and at most row_cap rows each.  *max_entries and *max_rows receive
// the largest block's entries and rows (either may be null).  Fails with
// SGC_ENOMEM if one row alone exceeds cap.
int ooc_plan_rows(const sgc_host_graph* g, uint64_t cap, uint64_t row_cap, std::vector<uint64_t>* bounds,
                  uint64_t* max_entries, uint64_t* max_rows);

// Copies the entries of rows [r0, r1) to dst (parallel; skipped for a null
// dst) and writes their offsets, shifted by `base`, to off[0 .. r1-r0].
void ooc_copy_rows(const sgc_host_graph* g, uint64_t r0, uint64_t r1, uint32_t base, int32_t* dst, uint32_t* off);

// Copies the final flags (N or Y) of rows [r0, r1), held in chip_flag, into
// the whole-graph flag array; entries the chip left undecided are skipped,
// so every entry is written by exactly the chip that decides it (entries
// decided by degree alone are written, identically, by every chip holding
// their row).
void ooc_merge_rows(const sgc_host_graph* g, uint64_t r0, uint64_t r1, const uint8_t* chip_flag, uint8_t* flag);

}  // namespace sgc
