// The SCAN phases after the similarity phase as wave- and thread-level device
// functions over one graph's arrays: the single-graph kernels (kernels.hip)
// run them over grid-stride loops, the batched query kernels (serve.hip) over
// the tiles of many queries in one launch.  Semantics:
// GPUSCAN++/GPUScan.cu:276-367 (core), :465-652 (union-find, minimum-id
// roots), :654-730 (borders, canonical minimum candidate), :732-798 (hubs and
// outliers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"
#include "device_util.h"

namespace sgc {

// Arrays of one query over one graph (entries [0, m), vertices [0, n)).
struct PhaseView {
  uint32_t n, m;
  const uint32_t* src;  // m: row of each entry
  const int32_t* adj;   // m
  const uint32_t* rb;   // row starts and ends (null: labelled-entry counts are skipped)
  const uint32_t* re;
  const uint8_t* flag;  // m
  uint32_t* sd;         // n: similar entries per row
  uint8_t* core;        // n
  int32_t* parent;      // n
  int32_t* label;       // n
  int32_t* lo;          // n
  int32_t* hi;          // n
};

// ---- core check: core(x) <=> #similar neighbours >= mu (mu <= 0: >= 1).
// Returns x's row length when x is a core and row offsets are known (the
// labelled entries classify chooses its side by), else 0; *marked = core.
__device__ __forceinline__ uint32_t core_mark_vertex(const PhaseView& v, uint32_t x, int32_t mu, uint32_t* marked) {
  const uint32_t thr = mu >= 1 ? (uint32_t)mu : 1u;
  const bool c = v.sd[x] >= thr;
  v.core[x] = c;
  v.parent[x] = (int32_t)x;
  v.label[x] = kNoLabel;
  v.lo[x] = kNoLabel;  // classify's neighbour-label bounds
  v.hi[x] = -1;
  *marked = c;
  return c && v.rb ? v.re[x] - v.rb[x] : 0u;
}

// ---- union-find, parents always smaller than children
__device__ __forceinline__ int32_t uf_find(int32_t* parent, int32_t x) {
  int32_t p = ld_coherent(parent + x);
  while (p != x) {
    const int32_t gp = ld_coherent(parent + p);
    if (gp != p) st_coherent(parent + x, gp);  // path splitting
    x = p;
    p = gp;
  }
  return x;
}

__device__ __forceinline__ void uf_union(int32_t* parent, int32_t a, int32_t b) {
  for (;;) {
    const int32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra == rb) return;
    const int32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
    if (atomicCAS(parent + hi, hi, lo) == hi) return;
  }
}

// Joins the similar core-core edges among entries k_j = k0 + j * step
// (j < kEdgeChunks), each visited from its larger endpoint.
__device__ __forceinline__ void union_entries(const PhaseView& v, uint64_t k0, uint64_t step) {
  constexpr uint32_t C = kEdgeChunks;
  uint32_t u[C];
  int32_t w[C];
  bool y[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const uint64_t k = k0 + j * step;
    y[j] = k < v.m && v.flag[k] == kFlagY;
    u[j] = y[j] ? v.src[k] : 0u;
    w[j] = y[j] ? v.adj[k] : 0;
    y[j] = y[j] && (uint32_t)w[j] < u[j];
  }
  bool cc[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) cc[j] = y[j] && v.core[u[j]] && v.core[w[j]];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j)
    if (cc[j]) uf_union(v.parent, (int32_t)u[j], w[j]);
}

// Roots go to label: the finds keep splitting paths in parent (stores of
// ancestors only), which must not land on a finished label.
__device__ __forceinline__ void flatten_vertex(const PhaseView& v, uint32_t x) {
  if (v.core[x]) v.label[x] = uf_find(v.parent, (int32_t)x);
}

// ---- borders: the minimum candidate cluster id over similar edges to cores,
// for the wave chunk of kEdgeChunks x 64 entries at `base`.  Only similar
// entries read rows and neighbours; the others get a marker row, so one
// row's lanes may form several runs, whose minimum may also take lanes of
// the same row from another run (that row's candidates too).  Returns the
// entries of the border rows this lane labelled first (rows known only).
__device__ __forceinline__ uint32_t border_chunk(const PhaseView& v, uint64_t base) {
  constexpr uint32_t C = kEdgeChunks;
  const uint32_t lane = lane_id();
  uint32_t labeled = 0;
  bool y[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const uint64_t k = base + j * kWave + lane;
    y[j] = k < v.m && v.flag[k] == kFlagY;
  }
  uint32_t row[C];
  int32_t c[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const uint64_t k = base + j * kWave + lane;
    row[j] = y[j] ? v.src[k] : 0xFFFFFFFFu;
    c[j] = y[j] ? v.adj[k] : 0;
  }
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) y[j] = y[j] && v.core[c[j]] && !v.core[row[j]];
  int32_t cand[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) cand[j] = y[j] ? v.label[c[j]] : kNoLabel;
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    if (!__ballot(y[j])) continue;
    const uint32_t r = y[j] ? row[j] : 0xFFFFFFFFu;
    const int32_t best = wave_seg_reduce(r, cand[j], OpMin());
    if (wave_seg_head(r) && best != kNoLabel && atomicMin(&v.label[r], best) == kNoLabel && v.rb)
      labeled += v.re[r] - v.rb[r];
  }
  return labeled;
}

// ---- hubs and outliers: every unlabelled vertex's min and max neighbour
// label, for the wave chunk at `base`.  from_labeled: walk the labelled rows
// and hand each label to the unlabelled neighbours (few labelled entries),
// else walk the unlabelled rows and gather.  Two different labels <=> the
// reference finds a first labelled neighbour and a later one with another.
__device__ __forceinline__ void classify_chunk(const PhaseView& v, uint64_t base, bool from_labeled) {
  constexpr uint32_t C = kEdgeChunks;
  const uint32_t lane = lane_id();
  uint32_t row[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const uint64_t k = base + j * kWave + lane;
    row[j] = k < v.m ? v.src[k] : 0xFFFFFFFFu;
  }
  int32_t lr[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) lr[j] = row[j] != 0xFFFFFFFFu ? v.label[row[j]] : kNoLabel;
  // neighbours only where this side needs them
  int32_t nb[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j)
    nb[j] = row[j] != 0xFFFFFFFFu && (from_labeled ? lr[j] != kNoLabel : lr[j] == kNoLabel)
                ? v.adj[base + j * kWave + lane]
                : 0;
  if (from_labeled) {
    bool to[C];
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) to[j] = lr[j] != kNoLabel && v.label[nb[j]] == kNoLabel;
#pragma unroll
    for (uint32_t j = 0; j < C; ++j)
      if (to[j]) {
        atomicMin(&v.lo[nb[j]], lr[j]);
        atomicMax(&v.hi[nb[j]], lr[j]);
      }
    return;
  }
  int32_t ly[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) ly[j] = row[j] != 0xFFFFFFFFu && lr[j] == kNoLabel ? v.label[nb[j]] : kNoLabel;
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const bool has = ly[j] != kNoLabel;
    if (!__ballot(has)) continue;
    const int32_t vmin = wave_seg_reduce(row[j], has ? ly[j] : kNoLabel, OpMin());
    const int32_t vmax = wave_seg_reduce(row[j], has ? ly[j] : -1, OpMax());
    if (wave_seg_head(row[j]) && vmax >= 0) {
      atomicMin(&v.lo[row[j]], vmin);
      atomicMax(&v.hi[row[j]], vmax);
    }
  }
}

// Final label of vertex x and its role, added to the counts: cores, the
// clusters (roots), borders, hubs, outliers, and x's similar entries.
struct RoleCounts {
  uint32_t cores = 0, clusters = 0, borders = 0, hubs = 0, outliers = 0;
  uint64_t similar = 0;
};

__device__ __forceinline__ int32_t finalize_vertex(const PhaseView& v, uint32_t x, int32_t mu, RoleCounts* rc) {
  rc->similar += v.sd[x];
  int32_t l = v.label[x];
  if (v.core[x]) {
    ++rc->cores;
    rc->clusters += l == (int32_t)x;
  } else if (l != kNoLabel) {
    ++rc->borders;
  } else {
    // mu <= 0: the reference classifies only role -2 vertices and there are
    // none, so every non-core keeps role -1 (GPUScan.cu:978-984).
    const bool hub = mu <= 0 || (v.lo[x] != kNoLabel && v.lo[x] != v.hi[x]);
    l = hub ? -1 : -2;
    v.label[x] = l;
    rc->hubs += hub;
    rc->outliers += !hub;
  }
  return l;
}

}  // namespace sgc
