// SCAN kernels for gfx950 (wave64) after the similarity phase, plus graph
// preparation and the one-wave-per-edge similarity kernel.  Every kernel is a
// grid-stride loop over vertices or directed CSR entries; per-row aggregation
// inside a wave uses a segmented shuffle reduction (entries of one row are
// consecutive lanes) so each row issues at most one atomic per wave.
#include <hip/hip_runtime.h>

#include <hipcub/device/device_radix_sort.hpp>

#include "device.h"
#include "device_util.h"
#include "scan_phases.h"

namespace sgc {

namespace {

// ---------------------------------------------------------------------------
// Graph preparation
// ---------------------------------------------------------------------------

// src[k] = row of entry k; one wave per vertex, lanes stride the row.
__global__ void __launch_bounds__(kBlock) k_fill_src(uint32_t n, const uint32_t* __restrict__ rp,
                                                     uint32_t* __restrict__ src) {
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t x = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; x < n; x += waves) {
    const uint32_t b = rp[x], e = rp[x + 1];
    for (uint32_t k = b + lane; k < e; k += kWave) src[k] = (uint32_t)x;
  }
}

// mirror[k] = position of the reverse entry (v, u) of entry k = (u, v).  Of
// the two entries of an edge only the one whose neighbour row is the shorter,
// (d(v), v) <= (d(u), u), searches: a binary search of u in the sorted row of
// v, which writes both mirror[k] and mirror[p].  A pair is found only from
// entry k, so each word has one writer.  Entries whose reverse entry is
// missing (asymmetric input) get 0xFFFFFFFF.  Four searches per thread run
// in lockstep so that their loads overlap; each halves its window only down
// to eight entries and reads those with eight independent loads, which ends
// the searches in short rows (most rows) after one round trip.
__global__ void __launch_bounds__(kBlock) k_mirror(uint32_t m, const uint32_t* __restrict__ rp,
                                                   const uint32_t* __restrict__ src, const int32_t* __restrict__ adj,
                                                   uint32_t* __restrict__ mirror) {
  constexpr uint32_t C = 4;  // searches per thread
  constexpr uint32_t W = 7;  // windows [lo, hi] of at most W + 1 entries end in one round of loads
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k0 < m; k0 += stride * C) {
    // lower bound of key in [lo, end): it lies in [lo, hi], and entries from
    // hi on are >= key
    uint32_t lo[C], hi[C], end[C];
    int32_t key[C];
    bool on[C];
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      const uint64_t k = k0 + j * stride;
      lo[j] = hi[j] = end[j] = 0;
      key[j] = 0;
      on[j] = false;
      if (k < m) {
        const uint32_t v = (uint32_t)adj[k], u = src[k];
        const uint32_t vb = rp[v], ve = rp[v + 1], du = rp[u + 1] - rp[u], dv = ve - vb;
        key[j] = (int32_t)u;
        on[j] = dv < du || (dv == du && v <= u);
        if (on[j]) {
          lo[j] = vb;
          hi[j] = end[j] = ve;
        }
      }
    }
    // halve the windows longer than W, all searches in lockstep
    for (;;) {
      bool any = false;
#pragma unroll
      for (uint32_t j = 0; j < C; ++j) any |= hi[j] - lo[j] > W;
      if (!any) break;
      int32_t val[C];
      uint32_t mid[C];
#pragma unroll
      for (uint32_t j = 0; j < C; ++j) {
        mid[j] = lo[j] + ((hi[j] - lo[j]) >> 1);
        val[j] = hi[j] - lo[j] > W ? adj[mid[j]] : 0;
      }
#pragma unroll
      for (uint32_t j = 0; j < C; ++j)
        if (hi[j] - lo[j] > W) {
          if (val[j] < key[j])
            lo[j] = mid[j] + 1;
          else
            hi[j] = mid[j];
        }
    }
    // the last window [lo, hi] (clipped to the row) in one round of loads:
    // the key's position is lo + (entries below it)
    int32_t val[C][W + 1];
#pragma unroll
    for (uint32_t j = 0; j < C; ++j)
#pragma unroll
      for (uint32_t t = 0; t <= W; ++t)
        val[j][t] = on[j] && lo[j] + t <= hi[j] && lo[j] + t < end[j] ? adj[lo[j] + t] : 0x7FFFFFFF;
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      if (!on[j]) continue;
      uint32_t below = 0;
      bool found = false;
#pragma unroll
      for (uint32_t t = 0; t <= W; ++t) {
        below += val[j][t] < key[j];
        found |= val[j][t] == key[j];
      }
      const uint32_t k = (uint32_t)(k0 + j * stride), p = lo[j] + below;
      mirror[k] = found ? p : 0xFFFFFFFFu;
      if (found) mirror[p] = k;
    }
  }
}

// The reverse-entry index by sorting.  In a symmetric CSR with sorted rows,
// the entries stably sorted by neighbour id come in the order of their
// reverse entries: row v lists its neighbours u ascending, and the entries
// (u, v) ordered by v, then by position (so by u), are ordered the same way.
// So the entry (u, v) at k that the sort puts at j has its reverse at j
// exactly when entry j is (v, u), and then entry j's reverse is k: the index
// is written at j, in order, not scattered.  Each pair is checked, and an
// input where some is not falls back to the search (k_mirror).  The sort's
// values carry the entry's row beside its position, so the check reads no
// other row.
__global__ void __launch_bounds__(kBlock) k_rows_and_positions(uint32_t m, const uint32_t* __restrict__ src,
                                                               uint64_t* __restrict__ vals) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += stride)
    vals[k] = ((uint64_t)src[k] << 32) | k;
}

__global__ void __launch_bounds__(kBlock) k_mirror_from_sorted(uint32_t m, const uint32_t* __restrict__ src,
                                                               const int32_t* __restrict__ adj,
                                                               const uint32_t* __restrict__ keys,
                                                               const uint64_t* __restrict__ vals,
                                                               uint32_t* __restrict__ mirror,
                                                               unsigned long long* __restrict__ bad) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  unsigned long long miss = 0;
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < m; j += stride) {
    const uint64_t val = vals[j];
    const uint32_t k = (uint32_t)val, u = (uint32_t)(val >> 32), v = keys[j];
    const bool ok = src[j] == v && (uint32_t)adj[j] == u;
    mirror[j] = ok ? k : 0xFFFFFFFFu;
    miss += ok ? 0u : 1u;
  }
  block_atomic_add(bad, miss);
}

// ndeg[k] = d(adj[k]), so the prep kernel streams the neighbour degrees
// instead of gathering two row offsets per entry on every pass.
__global__ void __launch_bounds__(kBlock) k_neighbor_degrees(uint32_t m, const uint32_t* __restrict__ rp,
                                                             const int32_t* __restrict__ adj,
                                                             uint32_t* __restrict__ ndeg) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += stride) {
    const uint32_t v = (uint32_t)adj[k];
    ndeg[k] = rp[v + 1] - rp[v];
  }
}

// ---------------------------------------------------------------------------
// Per-pass initialisation
// ---------------------------------------------------------------------------

// lo / hi hold the pruned core check's similar / dissimilar counts and core
// its vertex states until the core pass (k_core_mark) gives them their
// final meaning.
__global__ void __launch_bounds__(kBlock) k_init(uint32_t n, uint32_t* __restrict__ sd, int32_t* __restrict__ lo,
                                                 int32_t* __restrict__ hi, uint8_t* __restrict__ core,
                                                 uint32_t* __restrict__ cnt_u,
                                                 uint64_t* __restrict__ work_u, uint64_t* __restrict__ cnt) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += stride) {
    sd[x] = 0;
    lo[x] = 0;
    hi[x] = 0;
    core[x] = kVUndecided;
    cnt_u[x] = 0;
    work_u[x] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < kCntSlots) cnt[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Similarity, one wave per undecided edge (GPUSCAN++/GPUScan.cu:174-222)
// ---------------------------------------------------------------------------

// Dense list of the owned undecided entries (kFlagUDown), in any order, for
// the one-wave-per-edge kernel: 16 consecutive entries per thread, one
// append per block iteration.
__global__ void __launch_bounds__(kBlock) k_list_owned(uint32_t m, const uint8_t* __restrict__ flag,
                                                       uint32_t* __restrict__ list, uint64_t* __restrict__ cnt) {
  constexpr uint32_t E = 16;
  unsigned long long* const ctr[1] = {(unsigned long long*)&cnt[kCntDense]};
  const uint64_t tile = (uint64_t)kBlock * E;
  for (uint64_t base = (uint64_t)blockIdx.x * tile; base < m; base += (uint64_t)gridDim.x * tile) {  // block-uniform
    const uint64_t k0 = base + (uint64_t)threadIdx.x * E;
    uint32_t sel = 0;
#pragma unroll
    for (uint32_t j = 0; j < E; ++j)
      if (k0 + j < m && flag[k0 + j] == kFlagUDown) sel |= 1u << j;
    uint32_t pos[1] = {(uint32_t)__popc(sel)};
    block_reserve<1>(pos, ctr);
    for (uint32_t j = 0; j < E; ++j)
      if ((sel >> j) & 1u) list[pos[0]++] = (uint32_t)(k0 + j);
  }
}

// One wave per undecided edge (a, b), s = d(a) <= l = d(b): lanes take the
// short row 64 elements at a time and binary-search each in the long row,
// keeping a per-lane lower bound (each lane's keys increase).  The count is
// checked against the exact threshold after every 64 elements: reached ->
// similar, unreachable -> not similar.  Writes the flag of both directed
// entries.  Kept beside the tiered kernels as an independent second
// implementation (selectable per context).
__global__ void __launch_bounds__(kBlock) k_similarity_wave(const uint32_t* __restrict__ rb,
                                                            const uint32_t* __restrict__ re,
                                                            const int32_t* __restrict__ adj,
                                                            const uint32_t* __restrict__ src, double eps,
                                                            uint8_t* __restrict__ flag,
                                                            const uint32_t* __restrict__ list,
                                                            uint64_t* __restrict__ cnt) {
  const uint32_t lane = lane_id();
  const uint64_t nU = cnt[kCntDense];
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  uint64_t probes = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; i < nU; i += waves) {
    const uint32_t k = list[i];
    const uint32_t b = src[k];
    const uint32_t a = (uint32_t)adj[k];
    const uint32_t ab = rb[a], ae = re[a], bb = rb[b], be = re[b];
    const uint32_t s = ae - ab, l = be - bb;
    const double rhs = predicate_rhs(eps, s, l);
    const bool thresholded = s <= kMaxThresholdRow;
    const uint32_t cstar = thresholded ? min_common_threshold(rhs, s) : 0u;
    uint32_t cn = 0;
    uint32_t lo = bb;
    uint32_t off = 0;
    bool decided = false, sim = false;
    if (thresholded && cstar == 0) {
      decided = true;
      sim = true;
    } else if (thresholded && cstar > s) {
      decided = true;
      sim = false;
    }
    while (!decided && off < s) {
      bool hit = false;
      if (off + lane < s) {
        const int32_t x = adj[ab + off + lane];
        uint32_t h = be;
        while (lo < h) {
          const uint32_t mid = (lo + h) >> 1;
          if (adj[mid] < x)
            lo = mid + 1;
          else
            h = mid;
        }
        hit = lo < be && adj[lo] == x;
      }
      cn += (uint32_t)__popcll(__ballot(hit));
      off += kWave;
      if (thresholded) {
        const uint32_t rest = off < s ? s - off : 0u;
        if (cn >= cstar) {
          decided = true;
          sim = true;
        } else if (cn + rest < cstar) {
          decided = true;
          sim = false;
        }
      }
    }
    probes += off < s ? off : s;
    if (!decided) sim = thresholded ? cn >= cstar : predicate_exact(cn, rhs);
    const uint32_t mpos = wave_lower_bound(adj, ab, ae, (int32_t)b);
    if (lane == 0) {
      const uint8_t f = sim ? kFlagY : kFlagN;
      flag[k] = f;
      if (mpos < ae && adj[mpos] == (int32_t)b) flag[mpos] = f;
    }
  }
  if (lane == 0 && probes) atomicAdd((unsigned long long*)&cnt[kCntProbes], (unsigned long long)probes);
}

// ---------------------------------------------------------------------------
// Core check (GPUSCAN++/GPUScan.cu:115-131, :276-367)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_core_count(uint32_t m, const uint32_t* __restrict__ src,
                                                       const uint8_t* __restrict__ flag, uint32_t* __restrict__ sd) {
  constexpr uint32_t C = kEdgeChunks;
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t base = ((uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * (kWave * C); base < m;
       base += waves * kWave * C) {
    uint32_t row[C], y[C];
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      const uint64_t k = base + j * kWave + lane;
      y[j] = k < m && flag[k] == kFlagY ? 1u : 0u;
    }
    // rows of similar entries only; the others get a marker row, so one
    // row's similar entries may form several runs of lanes: each run's first
    // lane adds the run's count (ballot over [first lane, next run))
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) row[j] = y[j] ? src[base + j * kWave + lane] : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      const uint64_t ym = __ballot(y[j]);
      if (!ym) continue;
      const bool head = wave_seg_head(row[j]);
      const uint64_t after = __ballot(head) & ~lanemask_le(lane);
      const uint64_t run = (after ? lanemask_lt((uint32_t)(__ffsll((long long)after) - 1)) : ~0ull) & ~lanemask_lt(lane);
      if (head && y[j]) atomicAdd(&sd[row[j]], (uint32_t)__popcll(ym & run));
    }
  }
}

// core(x) <=> #similar neighbours >= mu (mu <= 0: >= 1).  The reference's
// sd/ed state machine reaches exactly this fixed point: sd only counts
// similar edges, ed only discounts dissimilar ones, and every edge of a
// still-undecided vertex is evaluated.
// mu comes from mu_dev when given (parameters of a captured launch), else mu.
// Also adds the entries of the core rows to cnt[kCntLabeled] when the row
// offsets are given (rb null: views without them count nothing here).
__global__ void __launch_bounds__(kBlock) k_core_mark(uint32_t n, const uint32_t* __restrict__ sd, int32_t mu,
                                                      const int32_t* __restrict__ mu_dev,
                                                      const uint32_t* __restrict__ rb,
                                                      const uint32_t* __restrict__ re, uint8_t* __restrict__ core,
                                                      int32_t* __restrict__ parent, int32_t* __restrict__ label,
                                                      int32_t* __restrict__ lo, int32_t* __restrict__ hi,
                                                      uint64_t* __restrict__ cnt) {
  if (mu_dev) mu = *mu_dev;
  PhaseView v{};
  v.n = n;
  v.sd = const_cast<uint32_t*>(sd);
  v.rb = rb, v.re = re, v.core = core, v.parent = parent, v.label = label, v.lo = lo, v.hi = hi;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t labeled = 0, marked = 0;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += stride) {
    uint32_t mk;
    labeled += core_mark_vertex(v, (uint32_t)x, mu, &mk);
    marked += mk;
  }
  block_atomic_add((unsigned long long*)&cnt[kCntLabeled], labeled);
  block_atomic_add((unsigned long long*)&cnt[kCntMarked], marked);
}

// ---------------------------------------------------------------------------
// Clusters: lock-free union-find, parents always smaller than children
// (GPUSCAN++/GPUScan.cu:465-528, :632-652; scan_phases.h)
// ---------------------------------------------------------------------------

// gate >= 0: runs only when cnt[gate] is not 0 (the union-find pass repeated
// after the second similarity round of a pruned pass).
__global__ void __launch_bounds__(kBlock) k_union(uint32_t m, const uint32_t* __restrict__ src,
                                                  const int32_t* __restrict__ adj, const uint8_t* __restrict__ flag,
                                                  const uint8_t* __restrict__ core, int32_t* parent,
                                                  const uint64_t* __restrict__ cnt, int gate) {
  if (cnt[kCntMarked] == 0) return;  // no core: no core-core edge to join
  if (gate >= 0 && cnt[gate] == 0) return;
  PhaseView v{};
  v.m = m, v.src = src, v.adj = adj, v.flag = flag, v.core = const_cast<uint8_t*>(core), v.parent = parent;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k0 < m; k0 += stride * kEdgeChunks)
    union_entries(v, k0, stride);
}

__global__ void __launch_bounds__(kBlock) k_flatten(uint32_t n, const uint8_t* __restrict__ core, int32_t* parent,
                                                    int32_t* __restrict__ label) {
  PhaseView v{};
  v.n = n, v.core = const_cast<uint8_t*>(core), v.parent = parent, v.label = label;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += stride) flatten_vertex(v, (uint32_t)x);
}

// Deferred edges of a pruned pass that the later passes need, after the
// union-find pass over the known similar core-core edges: a core-core edge
// whose endpoints are still in different trees (GPUSCAN++/GPUScan.cu:
// 557-585 computes these on demand) and a core-non-core edge (a border
// candidate, :692-719).  They are marked undecided again (owner side
// kFlagUDown, the other kFlagUUp) for a second similarity round and counted
// in cnt[kCntRedo].  A non-core-non-core edge is never needed: classify
// reads labels only.  Edges deferred with both endpoints in one tree cannot
// join anything.
__global__ void __launch_bounds__(kBlock) k_redo_mark(uint32_t m, const uint32_t* __restrict__ src,
                                                      const int32_t* __restrict__ adj,
                                                      const uint32_t* __restrict__ mirror, uint8_t* flag,
                                                      const uint8_t* __restrict__ core, int32_t* parent,
                                                      uint64_t* __restrict__ cnt) {
  if (cnt[kCntMarked] == 0 || cnt[kCntDeferred] == 0) return;
  constexpr uint32_t C = kEdgeChunks;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint32_t redo = 0;
  for (uint64_t k0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k0 < m; k0 += stride * C) {
    bool d[C];
    uint32_t b[C], a[C];
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      const uint64_t k = k0 + j * stride;
      d[j] = k < m && flag[k] == kFlagDeferDown;
      b[j] = d[j] ? src[k] : 0u;
      a[j] = d[j] ? (uint32_t)adj[k] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      if (!d[j]) continue;
      const bool ca = core[a[j]], cb = core[b[j]];
      const bool need =
          ca && cb ? uf_find(parent, (int32_t)a[j]) != uf_find(parent, (int32_t)b[j]) : ca != cb;
      if (!need) continue;
      const uint64_t k = k0 + j * stride;
      flag[k] = kFlagUDown;
      const uint32_t mk = mirror[k];  // 0xFFFFFFFF: no reverse entry (asymmetric input)
      if (mk < m) flag[mk] = kFlagUUp;
      ++redo;
    }
  }
  block_atomic_add((unsigned long long*)&cnt[kCntRedo], redo);
}

// ---------------------------------------------------------------------------
// Borders (GPUSCAN++/GPUScan.cu:654-730): minimum candidate cluster id
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_border(uint32_t m, const uint32_t* __restrict__ src,
                                                   const int32_t* __restrict__ adj, const uint8_t* __restrict__ flag,
                                                   const uint8_t* __restrict__ core, const uint32_t* __restrict__ rb,
                                                   const uint32_t* __restrict__ re, int32_t* label,
                                                   uint64_t* __restrict__ cnt) {
  if (cnt[kCntMarked] == 0) return;  // no core: no border
  PhaseView v{};
  v.m = m, v.src = src, v.adj = adj, v.flag = flag, v.core = const_cast<uint8_t*>(core), v.rb = rb, v.re = re;
  v.label = label;
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  // entries of the border rows this lane labels first; the core rows'
  // entries were counted by k_core_mark
  uint32_t labeled = 0;
  for (uint64_t base = ((uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * (kWave * kEdgeChunks);
       base < m; base += waves * kWave * kEdgeChunks)
    labeled += border_chunk(v, base);
  block_atomic_add((unsigned long long*)&cnt[kCntLabeled], labeled);
}

// ---------------------------------------------------------------------------
// Hubs and outliers (GPUSCAN++/GPUScan.cu:732-798)
// ---------------------------------------------------------------------------

// Every launch over the graph's entries (m_all of them, maybe in several row
// chips) takes the same side: from the labelled rows when they hold fewer
// than a third of the entries.
__global__ void __launch_bounds__(kBlock) k_classify(uint32_t m, uint64_t m_all, const uint32_t* __restrict__ src,
                                                     const int32_t* __restrict__ adj,
                                                     const int32_t* __restrict__ label, int32_t* __restrict__ lo,
                                                     int32_t* __restrict__ hi, const uint64_t* __restrict__ cnt) {
  // no core means no label anywhere: lo and hi keep their initial values
  if (cnt[kCntMarked] == 0) return;
  PhaseView v{};
  v.m = m, v.src = src, v.adj = adj, v.label = const_cast<int32_t*>(label), v.lo = lo, v.hi = hi;
  const bool from_labeled = 3 * cnt[kCntLabeled] < m_all;
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t base = ((uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * (kWave * kEdgeChunks);
       base < m; base += waves * kWave * kEdgeChunks)
    classify_chunk(v, base, from_labeled);
}

// Also counts the similar entries: sd[x] holds those of row x.
__global__ void __launch_bounds__(kBlock) k_finalize(uint32_t n, int32_t mu, const int32_t* __restrict__ mu_dev,
                                                     const uint32_t* __restrict__ sd, const int32_t* __restrict__ lo,
                                                     const int32_t* __restrict__ hi, const uint8_t* __restrict__ core,
                                                     int32_t* __restrict__ label, uint64_t* __restrict__ cnt) {
  if (mu_dev) mu = *mu_dev;
  PhaseView v{};
  v.n = n, v.sd = const_cast<uint32_t*>(sd), v.lo = const_cast<int32_t*>(lo), v.hi = const_cast<int32_t*>(hi);
  v.core = const_cast<uint8_t*>(core), v.label = label;
  RoleCounts rc;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += stride)
    finalize_vertex(v, (uint32_t)x, mu, &rc);
  block_atomic_add((unsigned long long*)&cnt[kCntCores], rc.cores);
  block_atomic_add((unsigned long long*)&cnt[kCntClusters], rc.clusters);
  block_atomic_add((unsigned long long*)&cnt[kCntBorders], rc.borders);
  block_atomic_add((unsigned long long*)&cnt[kCntHubs], rc.hubs);
  block_atomic_add((unsigned long long*)&cnt[kCntOutliers], rc.outliers);
  block_atomic_add((unsigned long long*)&cnt[kCntSimilar], rc.similar);
}

// Plain streaming copy (16 bytes per lane per step), the reference point for
// the HBM byte counters.
__global__ void __launch_bounds__(kBlock) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

void launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s) {
  const uint64_t n = bytes / sizeof(uint4);
  if (n) k_copy<<<grid_for(n, kBlock), kBlock, 0, s>>>((const uint4*)src, (uint4*)dst, n);
}

void launch_fill_src(const sgc_dgraph* g, hipStream_t s) {
  if (g->n == 0 || g->m == 0) return;
  k_fill_src<<<grid_for(g->n, kBlock / kWave), kBlock, 0, s>>>(g->n, g->row_ptr, g->src);
}

void launch_neighbor_degrees(const sgc_dgraph* g, hipStream_t s) {
  if (g->m && g->ndeg)
    k_neighbor_degrees<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->row_ptr, g->adj, g->ndeg);
}

void launch_mirror(const sgc_dgraph* g, hipStream_t s) {
  if (!g->m || !g->mirror) return;
  // words of entries without a reverse entry are never written by k_mirror
  (void)hipMemsetAsync(g->mirror, 0xFF, (size_t)g->m * 4, s);
  k_mirror<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->row_ptr, g->src, g->adj, g->mirror);
}

hipError_t build_mirror(const sgc_dgraph* g, hipStream_t s) {
  if (!g->m || !g->mirror) return hipSuccess;
  const uint32_t m = g->m;
  int bits = 1;
  while (bits < 32 && (1ull << bits) < (uint64_t)g->n) ++bits;
  size_t tmp_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint64_t*)nullptr, (uint64_t*)nullptr, m, 0, bits, s);
  // keys out, values in and out, the sort's scratch and the mismatch count
  char* mem = nullptr;
  const size_t o_keys = 0, o_vin = ((size_t)m * 4 + 255) / 256 * 256, o_vout = o_vin + (size_t)m * 8,
               o_tmp = o_vout + (size_t)m * 8, o_bad = o_tmp + (tmp_bytes + 255) / 256 * 256;
  if (hipMalloc(&mem, o_bad + 256) != hipSuccess) {
    (void)hipGetLastError();
    launch_mirror(g, s);  // not enough room to sort: search each reverse entry
    return hipGetLastError();
  }
  uint32_t* keys = (uint32_t*)(mem + o_keys);
  uint64_t* vin = (uint64_t*)(mem + o_vin);
  uint64_t* vout = (uint64_t*)(mem + o_vout);
  unsigned long long* bad = (unsigned long long*)(mem + o_bad);
  k_rows_and_positions<<<grid_for(m, kBlock), kBlock, 0, s>>>(m, g->src, vin);
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(mem + o_tmp, tmp_bytes, (const uint32_t*)g->adj, keys, vin, vout,
                                                     m, 0, bits, s);
  if (e == hipSuccess) e = hipMemsetAsync(bad, 0, 8, s);
  if (e == hipSuccess) {
    k_mirror_from_sorted<<<grid_for(m, kBlock), kBlock, 0, s>>>(m, g->src, g->adj, keys, vout, g->mirror, bad);
    e = hipGetLastError();
  }
  unsigned long long h_bad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h_bad, bad, 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(mem);
  if (e == hipSuccess && h_bad) {  // some entry lacks its reverse: the search finds those that exist
    launch_mirror(g, s);
    e = hipGetLastError();
  }
  return e;
}

void launch_init(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s) {
  k_init<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(g->n, b.sd, b.lo, b.hi, b.core, b.cnt_u, b.work_u, b.cnt);
}

void launch_similarity_wave(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, hipStream_t s) {
  if (g->m == 0) return;
  const uint32_t blocks = (uint32_t)(cu_count > 0 ? cu_count : 256) * 8;
  // the list takes the tiered kernels' uedge array, which this kernel replaces
  uint32_t* list = reinterpret_cast<uint32_t*>(b.uedge);
  k_list_owned<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, b.flag, list, b.cnt);
  k_similarity_wave<<<blocks, kBlock, 0, s>>>(g->row_begin, g->row_end, g->adj, g->src, p.eps, b.flag, list,
                                              b.cnt);
}

// Edge passes take the entries [0, g->m) of g (a whole graph or one chip of
// rows); vertex passes cover all g->n vertices.
void launch_core_count(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s) {
  if (g->m) k_core_count<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->src, b.flag, b.sd);
}

void launch_core_mark(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s,
                      const int32_t* mu_dev) {
  k_core_mark<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(g->n, b.sd, p.mu, mu_dev, g->row_begin, g->row_end, b.core,
                                                       b.parent, b.label, b.lo, b.hi, b.cnt);
}

void launch_union(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s, int gate) {
  if (g->m)
    k_union<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->src, g->adj, b.flag, b.core, b.parent, b.cnt, gate);
}

void launch_redo_mark(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s) {
  if (g->m && g->mirror)
    k_redo_mark<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->src, g->adj, g->mirror, b.flag, b.core, b.parent,
                                                         b.cnt);
}

void launch_flatten(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s) {
  k_flatten<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(g->n, b.core, b.parent, b.label);
}

void launch_border(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s) {
  if (g->m) k_border<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->src, g->adj, b.flag, b.core, g->row_begin,
                                                     g->row_end, b.label, b.cnt);
}

void launch_classify_edges(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s, uint64_t m_all) {
  if (g->m)
    k_classify<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, m_all ? m_all : g->m, g->src, g->adj, b.label, b.lo,
                                                         b.hi, b.cnt);
}

void launch_finalize(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s,
                     const int32_t* mu_dev) {
  k_finalize<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(g->n, p.mu, mu_dev, b.sd, b.lo, b.hi, b.core, b.label, b.cnt);
}


void launch_core(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s) {
  launch_core_count(g, b, s);
  launch_core_mark(g, b, p, s);
}

void launch_cluster(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s) {
  launch_union(g, b, s);
  launch_flatten(g, b, s);
}

void launch_classify(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s) {
  launch_classify_edges(g, b, s);
  launch_finalize(g, b, p, s);
}

}  // namespace sgc
