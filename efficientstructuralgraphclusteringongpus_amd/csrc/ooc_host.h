// Host half of the out-of-core tier: chip planning, chip assembly and the
// merge of chip flags into the whole-graph flag array (g++/OpenMP; the device
// half is in ooc.hip).
#pragma once

#include <stdint.h>

#include <vector>

#include "graph_host.h"

namespace sgc {

// Degrees as uint32 (every degree is < 2^31 for a graph the OOC tier takes).
void ooc_degrees(const sgc_host_graph* g, uint32_t* deg);

// Similarity chips: consecutive owner-row ranges, bounds[i] .. bounds[i+1],
// each of at most `cap` entries counted as d(b) plus d(a) for every
// undecided edge (a, b) the range owns (an upper bound of the chip: short
// rows shared by several owners are stored once).  *max_cost receives the
// largest chip's bound.  Fails with SGC_ENOMEM if one row alone exceeds cap.
int ooc_plan_similarity(const sgc_host_graph* g, const uint32_t* deg, double eps, uint64_t cap,
                        std::vector<uint64_t>* bounds, uint64_t* max_cost);

// Largest similarity-chip bound of a single owner row (see above).
uint64_t ooc_max_row_cost(const sgc_host_graph* g, const uint32_t* deg, double eps);

// Row chips for the passes after similarity: consecutive rows whose entries
// sum to at most cap.  *max_entries receives the largest chip.
int ooc_plan_rows(const sgc_host_graph* g, uint64_t cap, std::vector<uint64_t>* bounds, uint64_t* max_entries);

// One similarity chip, assembled into caller-provided (pinned) buffers:
//   adj[0, m_own)         the owner rows r0 .. r1-1, as in the graph;
//   off[0, r1-r0]         their chip offsets (off[0] = 0, off[r1-r0] = m_own);
//   adj[m_own, m_chip)    each short row an undecided owned edge reads, once,
//                         at starts[i] for row ids[i] (rows in [r0, r1) are
//                         already present and are not repeated).
// `stamp` (n int32, owned by the caller, initially all -1) marks rows taken;
// `tag` must differ between chips.
struct OocChip {
  uint64_t r0 = 0, r1 = 0;
  uint64_t m_own = 0, m_chip = 0;
  std::vector<uint32_t> ids, starts;
};
void ooc_build_chip(const sgc_host_graph* g, const uint32_t* deg, double eps, uint64_t r0, uint64_t r1,
                    int32_t* stamp, int32_t tag, int32_t* adj, uint32_t* off, OocChip* chip);

// Copies the final flags (N or Y) of a chip's entries into the whole-graph
// array: owner rows entry by entry, short rows where the chip decided the
// mirror entry.  Entries the chip left undecided or never wrote are skipped,
// so every entry is written by exactly the chip that decides it.
void ooc_merge_flags(const sgc_host_graph* g, const uint32_t* deg, const OocChip& chip, const uint8_t* chip_flag,
                     uint8_t* flag);

}  // namespace sgc
