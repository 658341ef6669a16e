// Similarity phase (GPUSCAN++/GPUScan.cu:133-222), tiered for gfx950.
//
// 1. k_prep: per directed entry, the reference's degree-only test; an
//    undecided edge (a, b), a the lower (degree, id) endpoint, is owned by b:
//    its entry in b's row is marked kFlagUDown, the one in a's row kFlagUUp.
//    Per owner b: number of owned edges and the sum of their short-row
//    lengths d(a) (the streaming work); the owned edges themselves go to
//    uedge[rb[b] .. rb[b] + count), inside b's own row range, so no scan
//    over the owners is needed.
// 2. k_plan: each owner's edges are cut into units of bounded work.
//    Owners with d(b) <= the tuned cut-off (at most kT1MaxDeg) go to the
//    small tier, the rest to the workgroup tier.
// 3. Counting (k_sim_units): the owner's row N(b) is put into LDS and the
//    short rows N(a) of the unit's edges are streamed through it in claims
//    that every wave of the workgroup takes dynamically (up to 16 coalesced
//    loads per lane in flight in the large tier).  Per edge the running
//    (hits, consumed) pair decides early: reached -> similar, unreachable ->
//    not similar (the reference's predicate in its exact threshold form).
//    The result is written to both directed entries.
//    Small tier: owners of degree <= 1024, 256-thread workgroups, the row as
//    a two-table cuckoo hash (linear probing if an insertion chain fails).
//    Large tier: 1024-thread workgroups, the row as a cuckoo hash (whole, or
//    in runs of 2^20-id windows) or one window at a time as an exact 128 KB
//    bitmap.
// 4. k_sim_tiny: owners whose edges' whole merge work is small, one edge per
//    lane, merging the two rows with the same threshold and early exit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <hipcub/device/device_scan.hpp>

#include "device.h"
#include "device_util.h"

namespace sgc {

namespace {

constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;

constexpr uint32_t kT1MaxDeg = 1024;  // small-tier tables are sized for this
// Edges per unit, bounded by LDS at 40 B per edge: six 256-thread small-tier
// workgroups of 25.6 KB share a CU; a workgroup-tier unit's 768 edges sit
// beside its 128 KB table (162 KB in all).
constexpr uint32_t kT1MaxEdges = 256;
constexpr uint32_t kT2MaxEdges = 768;

__device__ __forceinline__ uint32_t hash_slot(uint32_t key, uint32_t shift) {
  return (key * 0x9E3779B1u) >> shift;
}

// Linear-probing insert; keys of one row are distinct and the table is at
// most half full, so the probe always reaches an empty slot.
__device__ __forceinline__ void table_insert(uint32_t* tab, uint32_t mask, uint32_t shift, uint32_t key) {
  uint32_t s = hash_slot(key, shift);
  while (atomicCAS(&tab[s], kEmptySlot, key) != kEmptySlot) s = (s + 1) & mask;
}

__device__ __forceinline__ bool table_contains(const uint32_t* tab, uint32_t mask, uint32_t shift, uint32_t key) {
  uint32_t s = hash_slot(key, shift);
  for (;;) {
    const uint32_t v = tab[s];
    if (v == key) return true;
    if (v == kEmptySlot) return false;
    s = (s + 1) & mask;
  }
}

__device__ __forceinline__ uint32_t table_size_for(uint32_t count) {
  uint32_t t = 64;
  while (t < 2 * count) t <<= 1;
  return t;
}

__device__ __forceinline__ uint32_t log2_pow2(uint32_t t) { return 31 - __clz(t); }

// ---------------------------------------------------------------------------
// 1. Degree-only pre-decision (GPUSCAN++/GPUScan.cu:133-172)
// ---------------------------------------------------------------------------

// ny/nn (pruned passes): each row's degree-decided similar / dissimilar
// entries are added to its vertex's counts.  redo (the second round of a
// pruned pass): the degree-only test is not rerun; the entries the round
// marked kFlagUDown are planned as owned, their count goes to
// cnt[kCntRedoOwned], and the kernel returns at once when cnt[kCntRedo] is 0.
// Where k_prep puts each owner's edge records:
//   kRecRow     at the owner's own row range, uedge[rb[b] + slot];
//   kRecCount   nowhere: flags and per-owner counts only (first of two passes);
//   kRecCompact packed by owner: owner b's records start at rec_start[b],
//               its exclusive prefix of the counts (second pass: the flags of
//               the first are read back, owned = kFlagUDown).  One round
//               writes the owners whose rec_start lies in [r0, r0 + w), at
//               rec_cur[b] - r0 + slot (rec_cur: b's cursor, starting at
//               rec_start[b]); the round's records then occupy
//               [0, w + the largest owner's count).
enum RecMode { kRecRow = 0, kRecCount = 1, kRecCompact = 2 };

template <bool COUNTS, bool REDO, int REC = kRecRow>
__global__ void __launch_bounds__(kBlock) k_prep(uint32_t m, const uint32_t* __restrict__ rb,
                                                 const uint32_t* __restrict__ re, const int32_t* __restrict__ adj,
                                                 const uint32_t* __restrict__ src, const uint32_t* __restrict__ ndeg,
                                                 const uint32_t* __restrict__ mirror,
                                                 double eps, bool count_all, PairFilter pf,
                                                 uint8_t* __restrict__ flag,
                                                 uint32_t* __restrict__ cnt_u, uint64_t* __restrict__ work_u,
                                                 uint4* __restrict__ uedge, uint64_t* __restrict__ cnt,
                                                 uint32_t* __restrict__ ny, uint32_t* __restrict__ nn,
                                                 uint32_t* __restrict__ rec_cur = nullptr,
                                                 const uint32_t* __restrict__ rec_start = nullptr, uint32_t r0 = 0,
                                                 uint32_t w = 0) {
  constexpr bool redo = REDO;
  constexpr bool from_flags = REDO || REC == kRecCompact;  // owned entries from the flags already written
  if (redo && cnt[kCntRedo] == 0) return;
  if (REC == kRecCompact && r0 >= cnt[kCntUEdges]) return;  // a round past the chip's records
  constexpr uint32_t C = kEdgeChunks;
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  uint32_t owned_total = 0;
  for (uint64_t base = ((uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * (kWave * C); base < m;
       base += waves * kWave * C) {
    // all loads of the C chunks first, then the gathers they feed; with the
    // neighbour-degree index the neighbour's degree is streamed, not gathered
    uint32_t u[C], v[C], ub[C], du[C], dv[C];
    uint8_t f0[C];  // passes over written flags: the entry's flag
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      const uint64_t k = base + j * kWave + lane;
      u[j] = k < m ? src[k] : 0xFFFFFFFFu;  // rows must not decrease across lanes
      v[j] = k < m ? (uint32_t)adj[k] : 0u;
      dv[j] = k < m && ndeg ? ndeg[k] : 0u;
      f0[j] = from_flags && k < m ? flag[k] : kFlagUSkip;
    }
    // Only entries that can be owned here need the degree gathers: in a pass
    // over written flags the owned ones; in an out-of-core chip's first pass
    // the entries with one endpoint in each of the chip's two blocks (every
    // other entry is decided in the chip of its own two blocks and is only
    // marked skipped here).
    bool need[C];
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      need[j] = u[j] != 0xFFFFFFFFu;
      if (from_flags) {
        need[j] = need[j] && f0[j] == kFlagUDown;
      } else if (pf.on) {
        const bool u0 = u[j] - pf.lo0 < pf.hi0 - pf.lo0, u1 = u[j] - pf.lo1 < pf.hi1 - pf.lo1;
        const bool v0 = v[j] - pf.lo0 < pf.hi0 - pf.lo0, v1 = v[j] - pf.lo1 < pf.hi1 - pf.lo1;
        need[j] = need[j] && ((u0 && v1) || (u1 && v0));
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      ub[j] = need[j] ? rb[u[j]] : 0u;
      du[j] = need[j] ? re[u[j]] : 0u;
      if (!ndeg) dv[j] = need[j] ? re[v[j]] - rb[v[j]] : 0u;
    }
    uint32_t sb[C], rank[C];
    uint64_t om[C];
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      const uint64_t k = base + j * kWave + lane;
      const uint32_t row = u[j];
      bool owned = false;
      uint32_t yn = 0;  // this entry's degree decision: 1 << 16 similar, 1 dissimilar
      if (k < m) {
        const uint32_t x = v[j];
        du[j] -= ub[j];
        if (from_flags) {
          owned = need[j];
          if (REC == kRecCompact && owned) owned = rec_start[row] - r0 < w;  // this round's owners only
        } else {
          uint8_t f = kFlagUSkip;  // outside the chip's pair of blocks
          if (need[j]) {
            f = prep_flag(row, x, du[j], dv[j], eps);
            if (count_all && f <= kFlagY)  // every edge to its owner, for exact counts
              f = (dv[j] > du[j] || (dv[j] == du[j] && x > row)) ? kFlagUUp : kFlagUDown;
          }
          if (need[j] || !(pf.keep0 && row - pf.lo0 < pf.hi0 - pf.lo0)) flag[k] = f;
          owned = f == kFlagUDown;
          yn = f == kFlagY ? 0x10000u : (f == kFlagN ? 1u : 0u);
        }
      }
      if (COUNTS && __ballot(yn != 0)) {
        // the row's degree-decided entries, one pair of atomics per row segment
        const uint32_t sum = wave_seg_reduce(row, yn, OpAdd());
        if (wave_seg_head(row) && sum) {
          if (sum >> 16) atomicAdd(&ny[row], sum >> 16);
          if (sum & 0xFFFFu) atomicAdd(&nn[row], sum & 0xFFFFu);
        }
      }
      om[j] = __ballot(owned);
      sb[j] = 0;
      rank[j] = 0;
      if (om[j]) {
        if (lane == 0) owned_total += (uint32_t)__popcll(om[j]);
        // this lane's row segment [h, next head): owned count at the head,
        // this lane's rank among the owned entries below it
        const uint64_t ws = wave_seg_reduce(row, owned ? (uint64_t)dv[j] : 0ull, OpAdd());
        const bool head = wave_seg_head(row);
        const uint64_t hm = __ballot(head);
        const uint32_t h = 63 - __clzll((long long)(hm & lanemask_le(lane)));
        const uint64_t after = hm & ~lanemask_le(lane);
        const uint64_t seg =
            (after ? lanemask_lt((uint32_t)(__ffsll((long long)after) - 1)) : ~0ull) & ~lanemask_lt(h);
        const uint32_t c = (uint32_t)__popcll(om[j] & seg);
        rank[j] = (uint32_t)__popcll(om[j] & seg & lanemask_lt(lane)) | (h << 24);
        if (head && c) {
          if (REC == kRecCompact) {
            sb[j] = atomicAdd(&rec_cur[row], c);  // counts and work came with the first pass
          } else {
            sb[j] = atomicAdd(&cnt_u[row], c);
            atomicAdd((unsigned long long*)&work_u[row], (unsigned long long)ws);
          }
        }
      }
    }
    // owner b's undecided edges go to uedge[rb[b] + slot], slot < cnt_u[b],
    // as (entry, its reverse entry or kNone without the index, row start of
    // a, d(a))
#pragma unroll
    for (uint32_t j = 0; j < C; ++j) {
      if (REC == kRecCount || !om[j]) continue;
      const uint32_t slot = __shfl(sb[j], rank[j] >> 24) + (rank[j] & 0xFFFFFFu);
      if ((om[j] >> lane) & 1ull)
        uedge[(REC == kRecCompact ? 0u - r0 : ub[j]) + slot] =
            make_uint4((uint32_t)(base + j * kWave + lane), mirror ? mirror[base + j * kWave + lane] : kNone,
                       rb[v[j]], dv[j]);
    }
  }
  if (REC != kRecCompact)
    block_atomic_add((unsigned long long*)&cnt[redo ? kCntRedoOwned : kCntUEdges], owned_total);
}

// ---------------------------------------------------------------------------
// 2. Work units
// ---------------------------------------------------------------------------

// Tiny owners: the merge work of all their edges, sum of d(a) + d(b), is at
// most tiny_work.  They get no work unit; the planner lists them and
// k_sim_tiny merges their edges, one edge per lane, since a unit's fixed
// set-up would cost more than their whole work.
__device__ __forceinline__ bool tiny_owner(uint32_t owned, uint64_t work, uint32_t l, uint32_t tiny_work) {
  return work + (uint64_t)owned * l <= tiny_work;
}

// Edges per unit so that a unit streams about `target` short-row elements.
__device__ __forceinline__ uint32_t edges_per_unit(uint32_t cnt, uint64_t work, double target, uint32_t cap) {
  double k = work ? target * (double)cnt / (double)work : (double)cnt;
  uint32_t e = k >= (double)cnt ? cnt : (k < 1.0 ? 1u : (uint32_t)k);
  if (e > cap) e = cap;
  return e ? e : 1u;
}

// vstate (pruned passes, first round): every vertex's state from its
// degree-decided counts: core once ny >= mu, not core once d - nn < mu
// (GPUSCAN++/GPUScan.cu:140-169).  redo: the second round of a pruned pass
// (owned entries counted in cnt[kCntRedoOwned]; nothing to do when
// cnt[kCntRedo] is 0).
__global__ void __launch_bounds__(kBlock) k_plan(uint32_t x0, uint32_t x1, const uint32_t* __restrict__ rb,
                                                 const uint32_t* __restrict__ re, SimTuning tune,
                                                 const uint32_t* __restrict__ cnt_u,
                                                 const uint64_t* __restrict__ work_u,
                                                 uint4* __restrict__ units, uint32_t ucap,
                                                 uint32_t* __restrict__ tiny, uint64_t* __restrict__ cnt,
                                                 uint8_t* __restrict__ vstate, const uint32_t* __restrict__ ny,
                                                 const uint32_t* __restrict__ nn, int32_t mu, bool redo,
                                                 const uint32_t* __restrict__ rec_start = nullptr, uint32_t r0 = 0,
                                                 uint32_t w = 0) {
  constexpr uint32_t V = 4;  // vertices per thread per block iteration
  if (redo && cnt[kCntRedo] == 0) return;
  // Every unit holds at least one owned entry, so ucap units always suffice
  // when each undecided edge is owned once.  An input without a reverse
  // entry for every entry can own more: then no unit is written (the
  // regions from both ends would meet), and the host reports the count.
  // A round of packed records (rec_start) plans the owners whose records
  // start in [r0, r0 + w), which hold fewer than ucap records.
  const uint64_t owned_all = cnt[redo ? kCntRedoOwned : kCntUEdges];
  if (rec_start && r0 >= owned_all) return;
  if (!rec_start && owned_all > ucap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[kCntOverflow] = owned_all;
    return;
  }
  if (vstate) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t x = x0 + (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < x1; x += stride) {
      const uint32_t d = re[x] - rb[x], y = ny[x], no = nn[x];
      vstate[x] = y >= (uint32_t)mu ? kVCore : (d - no < (uint32_t)mu ? kVNonCore : kVUndecided);
    }
  }
  unsigned long long* const ctr[3] = {(unsigned long long*)&cnt[kCntUnits1], (unsigned long long*)&cnt[kCntUnits2],
                                      (unsigned long long*)&cnt[kCntTiny]};
  const uint64_t tile = (uint64_t)kBlock * V;
  for (uint64_t base = x0 + (uint64_t)blockIdx.x * tile; base < x1; base += (uint64_t)gridDim.x * tile) {  // block-uniform
    // records of owner x: [rs, rs + owned) in uedge, its row [start, stop) in adj
    uint32_t owned[V], c1[V], c2[V], c3[V], per[V], start[V], stop[V], rs[V];
#pragma unroll
    for (uint32_t j = 0; j < V; ++j) {
      const uint64_t x = base + j * kBlock + threadIdx.x;
      owned[j] = x < x1 ? cnt_u[x] : 0u;
      if (rec_start && owned[j] && rec_start[x] - r0 >= w) owned[j] = 0;  // another round's owner
    }
    uint32_t need[3] = {0, 0, 0};
#pragma unroll
    for (uint32_t j = 0; j < V; ++j) {
      const uint64_t x = base + j * kBlock + threadIdx.x;
      c1[j] = c2[j] = c3[j] = 0;
      per[j] = 1;
      start[j] = stop[j] = rs[j] = 0;
      if (owned[j]) {
        start[j] = rb[x];
        stop[j] = re[x];
        // the owner's edge records: at its own row range, or packed
        rs[j] = rec_start ? rec_start[x] - r0 : start[j];
        const uint32_t d = stop[j] - start[j];
        if (tiny_owner(owned[j], work_u[x], d, tune.tiny_work)) {
          c3[j] = 1;  // decided by k_sim_tiny, from the tiny-owner list
        } else if (d <= tune.small_max_degree) {
          per[j] = edges_per_unit(owned[j], work_u[x], tune.small_unit_work, kT1MaxEdges);
          c1[j] = (owned[j] + per[j] - 1) / per[j];
        } else {
          per[j] = edges_per_unit(owned[j], work_u[x], tune.large_unit_work, kT2MaxEdges);
          c2[j] = (owned[j] + per[j] - 1) / per[j];
        }
      }
      need[0] += c1[j];
      need[1] += c2[j];
      need[2] += c3[j];
    }
    block_reserve<3>(need, ctr);
    // small-tier units from the front of `units`, workgroup-tier units from
    // the back: (begin, end) in uedge, the owner's row [begin, end) in adj
#pragma unroll
    for (uint32_t j = 0; j < V; ++j) {
      const uint32_t cu = c1[j] + c2[j];
      const uint64_t end = (uint64_t)rs[j] + owned[j];
      for (uint32_t t = 0; t < cu; ++t) {
        const uint64_t j0 = (uint64_t)rs[j] + (uint64_t)t * per[j];
        const uint64_t j1 = j0 + per[j] < end ? j0 + per[j] : end;
        const uint32_t at = c1[j] ? need[0] + t : ucap - 1 - (need[1] + t);
        units[at] = make_uint4((uint32_t)j0, (uint32_t)j1, start[j], stop[j]);
      }
      if (c3[j]) tiny[need[2]] = (uint32_t)(base + j * kBlock + threadIdx.x);
      need[0] += c1[j];
      need[1] += c2[j];
      need[2] += c3[j];
    }
  }
}

// One undecided edge of a tiny owner x (row [bb, be)): merges the short row
// and the owner row with the same exact threshold and early exit as the
// tiered kernels, and writes both entries.
__device__ __forceinline__ void tiny_edge(uint32_t m, uint32_t x, uint32_t bb, uint32_t be, uint32_t rs,
                                          uint32_t i, const int32_t* __restrict__ adj,
                                          const uint32_t* __restrict__ mirror, double eps,
                                          const uint4* __restrict__ uedge, uint8_t* __restrict__ flag,
                                          uint32_t* __restrict__ cn_out) {
  const uint32_t l = be - bb;
  const uint4 ed = uedge[rs + i];  // (entry, reverse entry, row start of a, d(a)); rs: x's first record
  const uint32_t k = ed.x, ab = ed.z, s = ed.w, ae = ab + s;
  if (k >= m) return;
  const double rhs = predicate_rhs(eps, s, l);
  const uint32_t cstar = s <= kMaxThresholdRow && !cn_out ? min_common_threshold(rhs, s) : kNone;
  // merge in blocks of 4 entries per row: count the matches between the
  // two current blocks, then advance the block with the smaller last
  // value (both if equal).  A matching pair is current together exactly
  // once, and cn + min(entries left in either row) bounds the final
  // count, so the early exit is the tiered kernels'.
  uint32_t ia = ab, ib = bb, cn = 0, mpos = kNone;
  while (ia < ae && ib < be) {
    if (cstar != kNone) {
      const uint32_t rest = ae - ia < be - ib ? ae - ia : be - ib;
      if (cn >= cstar || cn + rest < cstar) break;
    }
    const uint32_t na = ae - ia < 4 ? ae - ia : 4, nb = be - ib < 4 ? be - ib : 4;
    int32_t va[4], vb[4];
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      va[q] = q < na ? adj[ia + q] : 0;
      vb[q] = q < nb ? adj[ib + q] : 0;
    }
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      if (q < na && va[q] == (int32_t)x) mpos = ia + q;
#pragma unroll
      for (uint32_t r = 0; r < 4; ++r) cn += (q < na && r < nb && va[q] == vb[r]) ? 1u : 0u;
    }
    const int32_t amax = va[na - 1], bmax = vb[nb - 1];
    if (amax <= bmax) ia += na;
    if (bmax <= amax) ib += nb;
  }
  const bool sim = cstar != kNone ? cn >= cstar : predicate_exact(cn, rhs);
  if (mirror) {
    mpos = ed.y;
  } else if (mpos == kNone) {
    // the merge stopped before reaching x in a's row
    uint32_t lo = ia, hi = ae;
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if (adj[mid] < (int32_t)x)
        lo = mid + 1;
      else
        hi = mid;
    }
    mpos = lo < ae && adj[lo] == (int32_t)x ? lo : kNone;
  }
  if (cn_out) {
    cn_out[k] = cn;
    if (mpos != kNone) cn_out[mpos] = cn;
  } else {
    const uint8_t f = sim ? kFlagY : kFlagN;
    flag[k] = f;
    if (mpos != kNone) flag[mpos] = f;
  }
}

// Tiny owners (tiny_owner), listed by the planner.  Each wave takes 64
// listed owners and spreads their edges over its lanes, one edge per lane
// per round (a lane finds its owner by binary search over the owners' edge
// counts, scanned into LDS), so a wave's time follows its longest edge, not
// its busiest owner.  The list keeps the lanes on owners with work, where a
// grid over all vertices left about one lane in twelve busy.
__global__ void __launch_bounds__(kBlock) k_sim_tiny(uint32_t m, const uint32_t* __restrict__ rb,
                                                     const uint32_t* __restrict__ re, const int32_t* __restrict__ adj,
                                                     const uint32_t* __restrict__ mirror, double eps,
                                                     const uint32_t* __restrict__ cnt_u,
                                                     const uint32_t* __restrict__ tiny,
                                                     const uint4* __restrict__ uedge, uint8_t* __restrict__ flag,
                                                     uint32_t* __restrict__ cn_out, const uint64_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ rec_start, uint32_t r0) {
  __shared__ uint32_t s_incl[kBlock], s_x[kBlock], s_bb[kBlock], s_be[kBlock], s_rs[kBlock];
  const uint32_t lane = lane_id(), w0 = (threadIdx.x / kWave) * kWave;
  const uint64_t count = cnt[kCntTiny];
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t g = ((uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * kWave; g < count;
       g += waves * kWave) {
    uint32_t x = 0, owned = 0, bb = 0, be = 0, rs = 0;
    if (g + lane < count) {
      x = tiny[g + lane];
      owned = cnt_u[x];
      bb = rb[x];
      be = re[x];
      rs = rec_start ? rec_start[x] - r0 : bb;  // the owner's first edge record
    }
    const uint32_t incl = wave_incl_scan(owned);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
    // wave-private LDS rows: this wave's own lanes write and read them, in order
    s_incl[w0 + lane] = incl;
    s_x[w0 + lane] = x;
    s_bb[w0 + lane] = bb;
    s_be[w0 + lane] = be;
    s_rs[w0 + lane] = rs;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
      const uint32_t idx = e0 + lane;
      if (idx >= total) continue;
      // the owner: the first lane whose inclusive edge count exceeds idx
      uint32_t lo = 0, hi = kWave - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_incl[w0 + mid] > idx)
          hi = mid;
        else
          lo = mid + 1;
      }
      const uint32_t first = lo ? s_incl[w0 + lo - 1] : 0u;
      tiny_edge(m, s_x[w0 + lo], s_bb[w0 + lo], s_be[w0 + lo], s_rs[w0 + lo], idx - first, adj, mirror, eps, uedge,
                flag, cn_out);
    }
    __builtin_amdgcn_wave_barrier();  // every lane is done with the rows before the next group overwrites them
  }
}

// ---------------------------------------------------------------------------
// 3. Counting
// ---------------------------------------------------------------------------

// Ids per bitmap window of the large tier: 2^20 bits = 128 KB of LDS, plus
// one word that stays zero and absorbs every probe outside the window.
constexpr uint32_t kWindowBits = 20;
constexpr uint32_t kWindowWords = (1u << kWindowBits) / 32;

// Tier shapes: threads per workgroup, LDS words for the owner-row table,
// whether the table is a window bitmap (else a hash table of the whole row,
// at most a quarter full), owned edges per unit, elements per claim at most
// (CHUNK / 64 loads in flight per lane).
template <int THREADS, uint32_t TABLE_WORDS, bool BITMAP, uint32_t MAX_EDGES, uint32_t CHUNK>
struct Tier {
  static constexpr int kThreads = THREADS;
  static constexpr uint32_t kTableWords = TABLE_WORDS;
  static constexpr bool kBitmap = BITMAP;
  static constexpr uint32_t kMaxEdges = MAX_EDGES;
  static constexpr uint32_t kChunk = CHUNK;
  static constexpr uint32_t kLoads = CHUNK / kWave;
  // waves per SIMD the register allocation must allow: six small-tier
  // workgroups per CU fit its LDS
  static constexpr int kMinWaves = BITMAP ? 1 : 6;
};
using SmallTier = Tier<256, 4 * kT1MaxDeg, false, kT1MaxEdges, 256>;
using LargeTier = Tier<1024, kWindowWords + 4, true, kT2MaxEdges, 1024>;

// Elements to claim next for an edge in state (hits, consumed): twice the
// fewest after which a decision can be possible, min(c* - hits, hits + rest -
// c* + 1), but at most 256 more than that, rounded up to whole wave loads and
// capped at one chunk.  A decision at the earliest possible point needs every
// claimed element to hit (or to miss), which is rare, so the extra elements
// are seldom wasted and keep more loads in flight per claim.  Rows counted
// fully (c* = kNone) claim whole chunks.
__device__ __forceinline__ uint32_t claim_size(unsigned long long st, uint32_t cstar, uint32_t s, uint32_t chunk) {
  if (cstar == kNone) return chunk;
  const uint32_t hits = (uint32_t)(st >> 32), used = (uint32_t)st;
  const uint32_t rest = s - used;
  const uint32_t to_reach = cstar - hits;              // hits < c* while undecided
  const uint32_t to_fail = hits + rest - cstar + 1;    // hits + rest >= c* while undecided
  uint32_t want = to_reach < to_fail ? to_reach : to_fail;
  want = want < 256 ? 2 * want : want + 256;
  want = (want + kWave - 1) / kWave * kWave;
  return want < chunk ? want : chunk;
}

// (hits << 32 | consumed) of an edge with threshold c* (kNone: count fully)
// and short-row length s: decided once hits >= c* or hits + (s - consumed)
// < c*.  hits only grows and consumed only grows, so any snapshot is safe.
__device__ __forceinline__ bool edge_decided(unsigned long long st, uint32_t cstar, uint32_t s) {
  if (cstar == kNone) return false;
  const uint32_t hits = (uint32_t)(st >> 32), used = (uint32_t)st;
  return hits >= cstar || hits + (s - used) < cstar;
}

// Hash probes of the wave's N loads x[j] (lanes inside a claim of len elements), all
// linear probes advancing together; each round issues its LDS reads before
// looking at any, so a round costs one LDS latency however many probes are
// still running.  Returns the wave's number of keys found (uniform).
template <uint32_t N>
__device__ __forceinline__ uint32_t probe_hash(const uint32_t* tab, uint32_t mask, uint32_t shift, const int32_t* x,
                                               uint32_t len) {
  uint32_t slot[N];
  uint32_t pend = 0;  // bit j: probe j still running
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) {
    slot[j] = hash_slot((uint32_t)x[j], shift);
    if (lane_id() + j * kWave < len) pend |= 1u << j;
  }
  uint32_t h = 0;
  while (__ballot(pend != 0)) {
    uint32_t v[N];
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) v[j] = tab[slot[j]];
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) {
      const bool on = (pend >> j) & 1u;
      const bool found = on && v[j] == (uint32_t)x[j];
      const bool stop = on && (found || v[j] == kEmptySlot);
      h += (uint32_t)__popcll(__ballot(found));
      if (stop) pend &= ~(1u << j);
      slot[j] = (slot[j] + 1) & mask;
    }
  }
  return h;
}

// Position of key b among the wave's N loads x[j] of the claim [s0, s0 + len),
// if there, into *mir (b's entry in the short row, the edge's mirror).
template <uint32_t N>
__device__ __forceinline__ void record_mirror(const int32_t* x, uint32_t len, uint32_t b, uint32_t s0, uint32_t* mir) {
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) {
    const uint64_t mm = __ballot(x[j] == (int32_t)b) & load_mask(j, len);
    if (mm && lane_id() == 0) *mir = s0 + j * kWave + (uint32_t)(__ffsll((long long)mm) - 1);
  }
}

// What a claim needs to know about the unit's current table.
struct ClaimTable {
  const int32_t* adj;
  const uint32_t* tab;
  uint32_t wlo, whi;    // large tier: the window [wlo, whi)
  uint32_t mask, shift;  // hash table shape (cuckoo: second table's offset, hash shift)
  uint32_t b;            // the owner
  bool last;             // no later window
  bool mir_here;         // the table's id range holds b
};

// Two-table cuckoo hashes in LDS: tables of 2^bits slots at offsets 0 and
// 2^bits, hashed by the top `bits` bits of two multiplicative hashes
// (shift = 32 - bits); every lookup is exactly two LDS reads.
//   Workgroup tier: 2 x 16384 slots.  Owners of degree <= kCuckooMaxKeys
//   hold their whole row (at most half full), longer rows hold runs of
//   consecutive windows with at most that many keys, so a claim is cut at
//   the run's end instead of at every window's.
//   Small tier: 2 x 2^bits slots with 2^bits >= 2 d(b) (at most a quarter
//   full), in place of linear probing, whose lock-step rounds last as long
//   as the wave's longest probe.
// Insertion evicts along a chain of at most kCuckooKicks; a failed chain
// (never observed at these loads) sends the unit to bitmap windows (large
// tier) or linear probing (small tier).
constexpr uint32_t kCuckooSlots = 16384;
constexpr uint32_t kCuckooShift = 18;  // 32 - log2(kCuckooSlots)
constexpr uint32_t kCuckooMaxKeys = 8192;
constexpr int kCuckooKicks = 128;

__device__ __forceinline__ uint32_t cuckoo_h0(uint32_t x, uint32_t shift) { return (x * 0x9E3779B1u) >> shift; }
__device__ __forceinline__ uint32_t cuckoo_h1(uint32_t x, uint32_t shift) {
  return ((x ^ (x >> 16)) * 0x85EBCA77u) >> shift;
}

__device__ __forceinline__ bool cuckoo_insert(uint32_t* tab, uint32_t key, uint32_t slots, uint32_t shift,
                                              uint32_t side = 0) {
  uint32_t k = key;
  for (int i = 0; i < kCuckooKicks; ++i) {
    const uint32_t slot = side ? slots + cuckoo_h1(k, shift) : cuckoo_h0(k, shift);
    const uint32_t old = atomicExch(&tab[slot], k);
    if (old == kEmptySlot) return true;
    k = old;  // the evicted key moves to its slot in the other table
    side ^= 1u;
  }
  return false;
}

// Inserts this thread's N keys (kEmptySlot: none): every first placement
// into table 0 is issued before any result is looked at, so a thread's keys
// cost one LDS round trip when nothing is evicted (the common case at these
// loads); an evicted key continues its chain from table 1.
template <uint32_t N>
__device__ __forceinline__ bool cuckoo_insert_all(uint32_t* tab, const uint32_t (&keys)[N], uint32_t slots,
                                                  uint32_t shift) {
  uint32_t ev[N];
#pragma unroll
  for (uint32_t j = 0; j < N; ++j)
    ev[j] = keys[j] != kEmptySlot ? atomicExch(&tab[cuckoo_h0(keys[j], shift)], keys[j]) : kEmptySlot;
  bool ok = true;
#pragma unroll
  for (uint32_t j = 0; j < N; ++j)
    if (ev[j] != kEmptySlot) ok = cuckoo_insert(tab, ev[j], slots, shift, 1u) && ok;
  return ok;
}

// One claim of NL = ceil(len / 64) wave loads, straight-line so that all its
// global loads, then all its LDS reads, are in flight together.  Buffer loads
// past the claim's end read 0 without touching memory; those lanes are kept
// out of every count.
// Returns the hits (uniform); *used = elements consumed (large tier: up to the
// window's end, which is also lowered into *ehi when the claim crosses it).
template <class T, uint32_t NL, bool CUCKOO>
__device__ __forceinline__ uint32_t count_claim(const ClaimTable& c, uint32_t s0, uint32_t len, uint32_t* mir,
                                                uint32_t* ehi, uint32_t* used_out) {
  const uint32_t lane = lane_id();
  const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(c.adj + s0, len * 4u);
  int32_t x[NL];
#pragma unroll
  for (uint32_t j = 0; j < NL; ++j) x[j] = __builtin_amdgcn_raw_buffer_load_b32(rs, (lane + j * kWave) * 4u, 0, 0);
  uint32_t h = 0, used = len;
  if constexpr (CUCKOO) {
    if constexpr (T::kBitmap) {
      // a run of windows without exact slices: the claim's elements from
      // whi on belong to later passes (none of them is a key of this run's
      // table, so only the consumed count stops there)
      if (!c.last) {
        const int32_t whs = (int32_t)c.whi;  // 0 < whi < 2^31 whenever a later window exists
#pragma unroll
        for (uint32_t j = 0; j < NL; ++j) {
          const uint64_t over = __ballot(x[j] >= whs) & load_mask(j, len);
          if (over && used == len) used = j * kWave + (uint32_t)(__ffsll((long long)over) - 1);
        }
        if (used < len && lane == 0) atomicMin(ehi, s0 + used);
      }
    }
    if (c.mir_here) record_mirror<NL>(x, len, c.b, s0, mir);
    uint32_t v0[NL], v1[NL];
#pragma unroll
    for (uint32_t j = 0; j < NL; ++j) {
      // workgroup tier: the fixed table shape as constants
      const uint32_t shift = T::kBitmap ? kCuckooShift : c.shift, second = T::kBitmap ? kCuckooSlots : c.mask;
      v0[j] = c.tab[cuckoo_h0((uint32_t)x[j], shift)];
      v1[j] = c.tab[second + cuckoo_h1((uint32_t)x[j], shift)];
    }
#pragma unroll
    for (uint32_t j = 0; j < NL; ++j) {
      uint64_t hm = __ballot(v0[j] == (uint32_t)x[j] || v1[j] == (uint32_t)x[j]);
      if (j + 1 == NL) hm &= load_mask(j, len);  // lanes past the claim read id 0
      h += (uint32_t)__popcll(hm);
    }
  } else if constexpr (T::kBitmap) {
    // first position whose value reaches whi (sorted row): from there on the
    // elements belong to later windows and probe the zero word
    if (!c.last) {
      const int32_t whs = (int32_t)c.whi;  // 0 < whi < 2^31 whenever a later window exists
#pragma unroll
      for (uint32_t j = 0; j < NL; ++j) {
        const uint64_t over = __ballot(x[j] >= whs);  // padding (0) never
        if (over && used == len) used = j * kWave + (uint32_t)(__ffsll((long long)over) - 1);
      }
      if (used < len && lane == 0) atomicMin(ehi, s0 + used);
    }
    if (c.mir_here) record_mirror<NL>(x, len, c.b, s0, mir);
    uint32_t v[NL];
#pragma unroll
    for (uint32_t j = 0; j < NL; ++j) {
      x[j] = (int32_t)((uint32_t)x[j] - c.wlo);  // offset: >= 2^20 outside the window
      const uint32_t wi = (uint32_t)x[j] >> 5;
      v[j] = c.tab[wi < kWindowWords ? wi : kWindowWords];
    }
#pragma unroll
    for (uint32_t j = 0; j < NL; ++j) h += (uint32_t)__popcll(__ballot((v[j] >> ((uint32_t)x[j] & 31)) & 1u));
    // lanes past the claim probed id 0: a hit only in window 0
    const uint32_t pad = NL * kWave - len;
    if (pad && c.wlo == 0 && (c.tab[0] & 1u)) h -= pad;
  } else {
    if (c.mir_here) record_mirror<NL>(x, len, c.b, s0, mir);
    h = probe_hash<NL>(c.tab, c.mask, c.shift, x, len);
  }
  *used_out = used;
  return h;
}

// count_claim<T, nl> for a runtime nl in [LO, HI], by binary dispatch.
template <class T, bool CUCKOO, uint32_t LO = 1, uint32_t HI = T::kLoads>
__device__ __forceinline__ uint32_t count_dispatch(uint32_t nl, const ClaimTable& c, uint32_t s0, uint32_t len,
                                                   uint32_t* mir, uint32_t* ehi, uint32_t* used) {
  if constexpr (LO == HI) {
    return count_claim<T, LO, CUCKOO>(c, s0, len, mir, ehi, used);
  } else {
    constexpr uint32_t MID = (LO + HI) / 2;
    if (nl <= MID) return count_dispatch<T, CUCKOO, LO, MID>(nl, c, s0, len, mir, ehi, used);
    return count_dispatch<T, CUCKOO, MID + 1, HI>(nl, c, s0, len, mir, ehi, used);
  }
}

// One unit = one owner b and up to MAX_EDGES of its undecided edges (a, b).
// The owner row N(b) goes into LDS (small tier: hash table; large tier: the
// whole row or runs of 2^20-id windows as a cuckoo hash, else one window at a
// time as an exact bitmap).  The short rows N(a) are
// streamed in claims that every wave of the workgroup takes dynamically: a
// wave walks one edge in order, claiming each time just enough elements for a
// decision to become possible, takes the next unstarted edge when its edge
// closes, and helps unfinished edges once all are started.  Per edge,
// (hits, consumed) live in one 64-bit LDS word updated by a single atomic
// add, so every read is a consistent snapshot for the early exit.  In the
// large tier a short row's slice for the current window ends where its
// values reach the window's end; the claim that crosses it records the
// boundary, so no search is needed, and claims are capped near the slice
// end predicted from the owner row's share of values below it, so little is
// loaded past the boundary.  Load, probe and count instructions run only for
// the claim's filled 64-element loads.  Rows too long for the threshold form
// (s > kMaxThresholdRow) are counted fully.
template <class T>
__global__ void __launch_bounds__(T::kThreads, T::kMinWaves)
    k_sim_units(uint32_t n, uint32_t m, const uint32_t* __restrict__ src, const int32_t* __restrict__ adj,
                const uint32_t* __restrict__ mirror, double eps, uint8_t* __restrict__ flag,
                uint32_t* __restrict__ cn_out, const uint32_t* __restrict__ woff, uint32_t wstride,
                const uint4* __restrict__ uedge, const uint4* __restrict__ units, uint32_t ucap,
                uint64_t* __restrict__ cnt, int units_slot, int queue_slot, bool hash_tables,
                uint8_t* __restrict__ vstate, uint32_t* __restrict__ ny, uint32_t* __restrict__ nn, int32_t mu) {
  __shared__ __align__(16) uint32_t tab[T::kTableWords];
  __shared__ unsigned long long e_state[T::kMaxEdges];  // hits << 32 | consumed
  // (c*, d(a), claim cap, slice end) of each edge: what a wave needs to size
  // a claim, in one 16-byte LDS read
  __shared__ __align__(16) uint4 e_q[T::kMaxEdges];
  // e_q.w also carries an edge's position between passes: where the next
  // pass's slice may start (the row start before the first pass)
  __shared__ uint32_t e_ab[T::kMaxEdges], e_pos[T::kMaxEdges], e_mir[T::kMaxEdges], e_a[T::kMaxEdges];
  __shared__ uint32_t s_unit, s_cursor, s_kend, s_fail;
  // pruned core check: the owner's similar / dissimilar edges decided in the
  // unit, and the unit's deferred edges
  __shared__ uint32_t s_by, s_bn, s_def;
  uint64_t deferred = 0;  // thread 0: deferred edges of this workgroup
  __shared__ uint4 s_next;  // the next unit's descriptor
  // the owner row's window boundaries, from the window-offset index
  // (n <= 2^28 for the boundaries in LDS; wstride = windows + 1)
  constexpr uint32_t kMaxOwnerWindows = 256;
  __shared__ uint32_t s_bw[T::kBitmap ? kMaxOwnerWindows + 1 : 1];
  const bool bw_fits = T::kBitmap && wstride > 2 && wstride <= kMaxOwnerWindows + 1;
  const uint32_t lane = lane_id();
  const uint64_t nunits = cnt[units_slot];
  uint64_t streamed = 0, loaded = 0, claims = 0;
#ifdef SGC_PHASE_TIMERS
  // debug builds: wall-clock ticks per wave in unit set-up, window set-up,
  // streaming, the end-of-window barrier and the final writes
  uint64_t tm[5] = {0, 0, 0, 0, 0}, t_ = wall_clock64();
#define SGC_TICK(i)                   \
  do {                                \
    const uint64_t now_ = wall_clock64(); \
    tm[i] += now_ - t_;               \
    t_ = now_;                        \
  } while (0)
#else
#define SGC_TICK(i) \
  do {              \
  } while (0)
#endif
  // Units come from a shared queue (their work varies widely).  Thread 0
  // takes the next unit while the current one runs and puts its descriptor
  // into s_next once its own streaming is done, so neither round trip sits
  // at the top of a unit.
  uint32_t nxt = 0;
  auto fetch_next = [&]() {
    if (nxt < nunits) s_next = units[T::kBitmap ? ucap - 1 - nxt : nxt];
  };
  if (threadIdx.x == 0) {
    nxt = (uint32_t)atomicAdd((unsigned long long*)&cnt[queue_slot], 1ull);
    fetch_next();
  }
  for (;;) {
    if (threadIdx.x == 0) {
      s_unit = nxt;
      if (nxt < nunits) nxt = (uint32_t)atomicAdd((unsigned long long*)&cnt[queue_slot], 1ull);
    }
    __syncthreads();
    const uint32_t u = s_unit;
    if (u >= nunits) break;
    const uint4 un = s_next;  // (begin, end) in uedge, the owner's row [bb, be) in adj
    const uint32_t ne = un.y - un.x, bb = un.z, be = un.w, l = be - bb;
    if (ne > T::kMaxEdges || ne == 0 || l == 0) {
      __syncthreads();  // never for consistent plans; every thread has read s_unit and s_next
      if (threadIdx.x == 0) fetch_next();
      continue;
    }
    const uint32_t b = src[bb];
    // small tier: the owner row in registers, loaded beside the edge records
    constexpr uint32_t KPT = T::kBitmap ? 1 : (kT1MaxDeg + T::kThreads - 1) / T::kThreads;
    uint32_t key[KPT];
    if constexpr (!T::kBitmap) {
#pragma unroll
      for (uint32_t j = 0; j < KPT; ++j) {
        const uint32_t t = threadIdx.x + j * T::kThreads;
        key[j] = t < l ? (uint32_t)adj[bb + t] : 0u;
      }
    }
    uint32_t my_k = kNone;  // this thread's edge entry (threads < ne), for the final writes
    // pruned core check (GPUSCAN++/GPUScan.cu:303): an edge whose endpoints
    // are both decided when the unit starts is deferred, not computed
    uint32_t my_a = 0;
    uint8_t my_sa = kVUndecided, sb = kVUndecided;
    bool my_defer = false;
    if (vstate && threadIdx.x == 0) s_by = s_bn = s_def = 0;
    if (threadIdx.x < ne) {
      const uint32_t t = threadIdx.x;
      const uint4 ed = uedge[un.x + t];  // (entry, reverse entry, row start of a, d(a))
      const uint32_t k = ed.x;
      uint32_t ab = 0, s = 0, cstar = 0, a = 0;
      if (k < m) {
        if (T::kBitmap && (woff || vstate)) a = (uint32_t)adj[k];  // window-offset index, vertex state
        ab = ed.z;
        s = ed.w;
        if (vstate) {
          my_a = a;
          my_sa = vstate[a];
          sb = vstate[b];
          my_defer = my_sa != kVUndecided && sb != kVUndecided;
        }
        const double rhs = predicate_rhs(eps, s, l);
        cstar = s <= kMaxThresholdRow && !cn_out ? min_common_threshold(rhs, s) : kNone;
        if (my_defer) s = cstar = 0;  // nothing to stream: decided at once, not written as a result
      }
      my_k = k;
      e_a[t] = a;
      e_ab[t] = ab;
      e_q[t].y = s;  // 0 for a slot without an edge
      e_q[t].w = ab;
      e_q[t].x = cstar;
      e_mir[t] = k < m ? ed.y : kNone;
      e_state[t] = 0ull;
    }
    // The owner row's window boundaries in LDS: from the window-offset index,
    // or, without it, by one binary search per boundary in the owner row (rows
    // longer than a whole-row hash only: the others are one pass)
    const bool whole = T::kBitmap && l <= kCuckooMaxKeys && n > (1u << kWindowBits);
    const bool bw = bw_fits && (woff || !whole);
    if (bw && threadIdx.x < wstride) {
      const uint32_t w = threadIdx.x;
      if (woff) {
        s_bw[w] = woff[(size_t)b * wstride + w];
      } else {
        uint32_t lo = bb, hi = be;
        if (w + 1 == wstride) {
          lo = be;
        } else if (w > 0) {
          const int32_t key = (int32_t)(w << kWindowBits);
          while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (adj[mid] < key)
              lo = mid + 1;
            else
              hi = mid;
          }
        }
        s_bw[w] = lo;
      }
    }
    if constexpr (T::kBitmap) __syncthreads();  // every thread picks the passes from s_bw
    SGC_TICK(0);
    // Large tier, per pass over the owner row: the whole row as a cuckoo
    // hash when it has at most kCuckooMaxKeys keys and the ids span more than
    // one window; else a cuckoo hash of the longest run of two or more
    // consecutive windows holding at most kCuckooMaxKeys keys; else the exact
    // bitmap of one window.  A failed cuckoo build sends the rest of the unit
    // to bitmap windows.
    bool hash_ok = hash_tables;
    const uint32_t nwin = bw ? wstride - 1 : 0;
    uint32_t kcur = bb;
    while (kcur < be) {
      // ---- LDS table for owner keys adj[kcur, kend) of the ids [wlo, whi)
      // (windows w0 .. w0 + wspan - 1).  Small tier: hash table of the whole row.
      uint32_t kend = be, wlo = 0, whi = 0xFFFFFFFFu, mask = 0, shift = 0, wspan = 1;
      uint32_t pre_lo = 0, pre_hi = 0;  // this thread's edge slice (exact passes)
      bool hashed = false;
      if constexpr (T::kBitmap) {
        uint32_t w0 = 0, w1 = 0;
        if (hash_ok && whole) {
          hashed = true;
        } else {
          if (bw) {
            // the window of the first key left, from the owner's window
            // boundaries in LDS: the last w with s_bw[w] <= kcur (s_bw[0] is
            // the row start, so one lane always qualifies; the boundaries do
            // not decrease, so the qualifying windows are a prefix)
            for (uint32_t base = 0; base < nwin; base += kWave) {
              const uint64_t le = __ballot(base + lane < nwin && s_bw[base + lane] <= kcur);
              if (!le) break;
              w0 = base + 63u - (uint32_t)__clzll((long long)le);
            }
          } else {
            w0 = (uint32_t)adj[kcur] >> kWindowBits;
          }
          w1 = w0 + 1;
          if (hash_ok && nwin > 1)
            while (w1 < nwin && s_bw[w1 + 1] - kcur <= kCuckooMaxKeys) ++w1;
          hashed = w1 >= w0 + 2;
          wlo = w0 << kWindowBits;
          whi = w1 < nwin || !bw ? wlo + ((w1 - w0) << kWindowBits) : 0xFFFFFFFFu;
          wspan = w1 - w0;
        }
        // the edges' slices of this pass from the window-offset index, loaded
        // before the table is built so that their latency overlaps it
        if (woff && !(hashed && whole) && threadIdx.x < ne && e_q[threadIdx.x].y != 0) {
          const uint32_t* wo = woff + (size_t)e_a[threadIdx.x] * wstride + (wlo >> kWindowBits);
          pre_lo = wo[0];
          pre_hi = wo[wspan];
        }
        if (hashed) {
          if (!whole) kend = s_bw[w1];
          // at most kCuckooMaxKeys keys: all of a thread's keys are loaded
          // before the table is cleared, then inserted together
          constexpr uint32_t KB = (kCuckooMaxKeys + T::kThreads - 1) / T::kThreads;
          uint32_t kk[KB];
#pragma unroll
          for (uint32_t j = 0; j < KB; ++j) {
            const uint32_t t = kcur + threadIdx.x + j * T::kThreads;
            kk[j] = t < kend ? (uint32_t)adj[t] : kEmptySlot;
          }
          uint4* tab4 = reinterpret_cast<uint4*>(tab);
          for (uint32_t t = threadIdx.x; t < 2 * kCuckooSlots / 4; t += T::kThreads)
            tab4[t] = make_uint4(kEmptySlot, kEmptySlot, kEmptySlot, kEmptySlot);
          if (threadIdx.x == 0) s_fail = 0;
          __syncthreads();
          mask = kCuckooSlots;
          shift = kCuckooShift;
          const bool ok = cuckoo_insert_all<KB>(tab, kk, kCuckooSlots, kCuckooShift);
          if (!ok) s_fail = 1;
          __syncthreads();
          if (s_fail) {  // redo this pass with bitmap windows
            hash_ok = false;
            continue;
          }
        } else {
          if (threadIdx.x < kWave && !bw) {
            const uint32_t e = whi > 0x7FFFFFFFu ? be : wave_lower_bound(adj, kcur, be, (int32_t)whi);
            if (lane == 0) s_kend = e;
          }
          uint4* tab4 = reinterpret_cast<uint4*>(tab);
          for (uint32_t t = threadIdx.x; t < T::kTableWords / 4; t += T::kThreads) tab4[t] = make_uint4(0, 0, 0, 0);
          __syncthreads();
          kend = bw ? s_bw[w0 + 1] : s_kend;
          for (uint32_t t = kcur + threadIdx.x; t < kend; t += T::kThreads) {
            const uint32_t x = (uint32_t)adj[t] - wlo;
            atomicOr(&tab[x >> 5], 1u << (x & 31));
          }
        }
      } else if (hash_ok) {
        // two tables of 2^bits >= 2l slots each
        const uint32_t slots = table_size_for(l);
        mask = slots;
        shift = 32 - log2_pow2(slots);
        uint4* tab4 = reinterpret_cast<uint4*>(tab);  // 2 * slots is a multiple of 128
        for (uint32_t t = threadIdx.x; t < 2 * slots / 4; t += T::kThreads)
          tab4[t] = make_uint4(kEmptySlot, kEmptySlot, kEmptySlot, kEmptySlot);
        if (threadIdx.x == 0) s_fail = 0;
        __syncthreads();
        bool ok = true;  // key by key: batched placements cost this tier register spills
#pragma unroll
        for (uint32_t j = 0; j < KPT; ++j)
          if (threadIdx.x + j * T::kThreads < l) ok = cuckoo_insert(tab, key[j], slots, shift) && ok;
        if (!ok) s_fail = 1;
        __syncthreads();
        if (s_fail) {  // redo with linear probing
          hash_ok = false;
          continue;
        }
        hashed = true;
      } else {
        const uint32_t TS = table_size_for(2 * l);
        mask = TS - 1;
        shift = 32 - log2_pow2(TS);
        uint4* tab4 = reinterpret_cast<uint4*>(tab);  // TS is a multiple of 64
        for (uint32_t t = threadIdx.x; t < TS / 4; t += T::kThreads)
          tab4[t] = make_uint4(kEmptySlot, kEmptySlot, kEmptySlot, kEmptySlot);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < KPT; ++j)
          if (threadIdx.x + j * T::kThreads < l) table_insert(tab, mask, shift, key[j]);
      }
      const bool last = kend >= be;
      // slices known from the window-offset index (every pass but a whole-row hash)
      const bool exact = T::kBitmap && woff && !(hashed && whole);
      // the table's id range holds b (its position in the short rows is
      // recorded while streaming when there is no mirror index)
      const bool mir_here = !mirror && ((hashed && whole) || b - wlo < whi - wlo);
      const ClaimTable cc{adj, tab, wlo, whi, mask, shift, b, last || exact, mir_here};
      if (threadIdx.x < ne) {
        const uint32_t t = threadIdx.x;
        const uint32_t cur = e_q[t].w, ab = e_ab[t], s = e_q[t].y;
        e_pos[t] = cur;
        e_q[t].w = ab + s;
        uint32_t cap = T::kChunk;
        if (exact) {
          // the edge's slice of this window; the elements skipped since the
          // last window lie in windows without owner keys: consumed, no hits
          uint32_t lo = cur, hi = cur;
          if (s != 0) {
            lo = pre_lo;
            hi = pre_hi;
            if (lo > cur) e_state[t] += (unsigned long long)(lo - cur);
          }
          e_pos[t] = lo;
          e_q[t].w = hi;
        } else if (!last) {
          // the short row's slice is taken to end where the owner row's
          // share of values below whi puts it
          const uint32_t pe = ab + (uint32_t)((uint64_t)s * (kend - bb) / l);
          const uint32_t pred = pe > cur ? pe - cur : 0u;
          const uint32_t c = (pred + pred / 4 + 2 * kWave - 1) / kWave * kWave;
          cap = c < T::kChunk ? c : T::kChunk;
        }
        e_q[t].z = cap;
      }
      if (threadIdx.x == 0) s_cursor = 0;
      __syncthreads();  // table and per-edge cursors ready
      SGC_TICK(1);
      // ---- stream the short rows
      uint32_t cur_e = kNone;
      for (;;) {
        uint32_t e = kNone, start = 0, end = 0;
        if (cur_e != kNone) {
          // one round of LDS reads, then the cursor atomic; a stale (higher) slice
          // end only lets a claim start past a bitmap window's end, where
          // it consumes nothing and finds no keys
          const unsigned long long st = e_state[cur_e];
          const uint4 q = e_q[cur_e];
          const uint32_t cstar = q.x, s_e = q.y, cap = q.z, ehi = q.w;
          uint32_t got = 0xFFFFFFFFu;
          if (!edge_decided(st, cstar, s_e)) {
            const uint32_t want = claim_size(st, cstar, s_e, cap);
            if (lane == 0) got = atomicAdd(&e_pos[cur_e], want);
            got = __builtin_amdgcn_readfirstlane(got);  // lane 0's, without an LDS permute
            if (got < ehi) {
              e = cur_e;
              start = got;
              end = got + want < ehi ? got + want : ehi;
            }
          }
          if (e == kNone) cur_e = kNone;
        }
        if (e == kNone) {
          uint32_t t = 0;
          if (lane == 0) t = atomicAdd(&s_cursor, 1u);
          t = __builtin_amdgcn_readfirstlane(t);
          if (t < ne) {
            cur_e = t;
            continue;
          }
          // every edge started: help an unfinished one
          uint32_t found = kNone;
          for (uint32_t base = 0; base < ne && found == kNone; base += kWave) {
            const uint32_t q = base + lane;
            const bool open = q < ne && e_pos[q] < e_q[q].w && !edge_decided(e_state[q], e_q[q].x, e_q[q].y);
            const uint64_t bm = __ballot(open);
            if (bm) found = base + (uint32_t)(__ffsll((long long)bm) - 1);
          }
          if (found == kNone) break;
          cur_e = found;
          continue;
        }
        // the claim [s0, s0 + len) as wave-uniform values; its loads are
        // issued inside count_claim, sized to the claim
        const uint32_t s0 = __builtin_amdgcn_readfirstlane(start);
        const uint32_t len = __builtin_amdgcn_readfirstlane(end - start);
        uint32_t used;
        const uint32_t nl = (len + kWave - 1) / kWave;
        uint32_t h;
        if (hashed)
          h = count_dispatch<T, true>(nl, cc, s0, len, &e_mir[e], &e_q[e].w, &used);
        else
          h = count_dispatch<T, false>(nl, cc, s0, len, &e_mir[e], &e_q[e].w, &used);
        if (lane == 0) atomicAdd(&e_state[e], ((unsigned long long)h << 32) | (unsigned long long)used);
        streamed += used;
        loaded += len;
        ++claims;
      }
      // the next unit's descriptor, while the other waves finish streaming
      // (every thread read s_next before the table barriers)
      if (threadIdx.x == 0 && kend >= be) fetch_next();
      SGC_TICK(2);
      __syncthreads();  // all streaming done: the table may be rebuilt, e_q.w holds each slice's end
      SGC_TICK(3);
      kcur = kend;
    }
    // ---- decide and write both directed entries of every edge
    if (threadIdx.x < ne) {
      const uint32_t t = threadIdx.x;
      const uint32_t k = my_k;
      if (k < m) {
        const uint32_t hits = (uint32_t)(e_state[t] >> 32);
        const uint32_t cstar = e_q[t].x;
        const uint32_t ab = e_ab[t], s = e_q[t].y;
        const bool sim = cstar != kNone ? hits >= cstar : predicate_exact(hits, predicate_rhs(eps, s, l));
        uint32_t mpos = e_mir[t];
        if (mpos == kNone && !mirror) {
          uint32_t lo = ab, hi = ab + s;
          while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (adj[mid] < (int32_t)b)
              lo = mid + 1;
            else
              hi = mid;
          }
          mpos = lo < ab + s && adj[lo] == (int32_t)b ? lo : kNone;
        }
        if (cn_out) {  // exact counts instead of flags
          cn_out[k] = hits;
          if (mpos != kNone) cn_out[mpos] = hits;
        } else if (my_defer) {
          flag[k] = kFlagDeferDown;
          if (mpos != kNone) flag[mpos] = kFlagDeferUp;
          atomicAdd(&s_def, 1u);
        } else {
          const uint8_t f = sim ? kFlagY : kFlagN;
          flag[k] = f;
          if (mpos != kNone) flag[mpos] = f;
          if (vstate) {
            // a's counts now; b's once per unit, below.  A vertex is core
            // once it has mu similar edges and not core once fewer than mu
            // edges are left that are not dissimilar (d(a) = s).
            if (my_sa == kVUndecided) {
              if (sim) {
                if (atomicAdd(&ny[my_a], 1u) + 1u == (uint32_t)mu) vstate[my_a] = kVCore;
              } else {
                if (s - atomicAdd(&nn[my_a], 1u) == (uint32_t)mu) vstate[my_a] = kVNonCore;
              }
            }
            atomicAdd(sim ? &s_by : &s_bn, 1u);
          }
        }
      }
    }
    __syncthreads();
    if (vstate && threadIdx.x == 0) deferred += s_def;
    if (vstate && threadIdx.x == 0 && ne && sb == kVUndecided) {
      const uint32_t y = s_by, no = s_bn, um = (uint32_t)mu;
      if (y) {
        const uint32_t old = atomicAdd(&ny[b], y);
        if (old < um && old + y >= um) vstate[b] = kVCore;
      }
      if (no) {
        const uint32_t old = atomicAdd(&nn[b], no);
        if (l - old >= um && l - old - no < um) vstate[b] = kVNonCore;
      }
    }
    SGC_TICK(4);
  }
#ifdef SGC_PHASE_TIMERS
  if (lane == 0)
    for (int i = 0; i < 5; ++i)
      atomicAdd((unsigned long long*)&cnt[kCntTimer0 + (T::kBitmap ? 0 : 5) + i], (unsigned long long)tm[i]);
#endif
#undef SGC_TICK
  if (threadIdx.x == 0 && deferred) atomicAdd((unsigned long long*)&cnt[kCntDeferred], (unsigned long long)deferred);
  if (lane == 0 && claims) {
    atomicAdd((unsigned long long*)&cnt[kCntProbes], (unsigned long long)streamed);
    atomicAdd((unsigned long long*)&cnt[kCntLoaded], (unsigned long long)loaded);
    atomicAdd((unsigned long long*)&cnt[kCntClaims], (unsigned long long)claims);
  }
}

// Window-offset index: woff[x * (W + 1) + w] = first position of row x whose
// value is >= w * 2^20 (w = 0: the row start, w = W: its end).  One thread
// per (row, window) boundary.
__global__ void __launch_bounds__(kBlock) k_window_offsets(uint32_t n, const uint32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ adj, uint32_t W,
                                                           uint32_t* __restrict__ woff) {
  const uint64_t total = (uint64_t)n * (W + 1);
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += stride) {
    const uint32_t x = (uint32_t)(i / (W + 1)), w = (uint32_t)(i % (W + 1));
    uint32_t lo = rp[x], hi = rp[x + 1];
    if (w == 0) {
      hi = lo;
    } else if (w < W) {
      const int32_t key = (int32_t)(w << kWindowBits);
      while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (adj[mid] < key)
          lo = mid + 1;
        else
          hi = mid;
      }
    } else {
      lo = hi;
    }
    woff[i] = lo;
  }
}

}  // namespace

uint32_t window_count(uint32_t n) { return (uint32_t)(((uint64_t)n + (1u << kWindowBits) - 1) >> kWindowBits); }

uint64_t window_index_bytes(uint64_t n, uint64_t m) {
  const uint32_t W = window_count((uint32_t)n);
  const uint64_t bytes = n * (W + 1) * 4;
  return W > 1 && bytes <= std::max<uint64_t>(16 * std::max<uint64_t>(m, 1), 1ull << 30) ? bytes : 0;
}

void launch_window_offsets(const sgc_dgraph* g, hipStream_t s) {
  if (!g->woff || g->n == 0) return;
  const uint64_t total = (uint64_t)g->n * (g->wcount + 1);
  k_window_offsets<<<grid_for(total, kBlock), kBlock, 0, s>>>(g->n, g->row_ptr, g->adj, g->wcount, g->woff);
}

size_t scan_scratch_bytes(uint32_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, n);
  return bytes;
}

void launch_prep(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s, bool count_all) {
  if (g->m == 0) return;
  const uint32_t grid = grid_for(g->m, kBlock);
  if (p.prune && !count_all)
    k_prep<true, false><<<grid, kBlock, 0, s>>>(g->m, g->row_begin, g->row_end, g->adj, g->src, g->ndeg, g->mirror,
                                                p.eps, count_all, g->pair, b.flag, b.cnt_u, b.work_u, b.uedge, b.cnt,
                                                b.ny, b.nn);
  else
    k_prep<false, false><<<grid, kBlock, 0, s>>>(g->m, g->row_begin, g->row_end, g->adj, g->src, g->ndeg, g->mirror,
                                                 p.eps, count_all, g->pair, b.flag, b.cnt_u, b.work_u, b.uedge, b.cnt,
                                                 nullptr, nullptr);
}

void launch_plan(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s) {
  if (g->m == 0) return;
  SimTuning t = g->tuning;
  if (t.small_max_degree > kT1MaxDeg) t.small_max_degree = kT1MaxDeg;
  k_plan<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(0, g->n, g->row_begin, g->row_end, t, b.cnt_u, b.work_u, b.units,
                                                   b.ucap, b.tiny, b.cnt, p.prune ? b.core : nullptr, b.ny, b.nn,
                                                   p.mu, false);
}

void launch_similarity(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, int impl, hipEvent_t mid,
                       hipStream_t s, uint32_t* cn_out) {
  if (g->m == 0) return;
  if (impl == kSimWavePerEdge && !cn_out) {
    launch_similarity_wave(g, b, p, cu_count, s);
    if (mid) (void)hipEventRecord(mid, s);
    return;
  }
  const uint32_t cus = (uint32_t)(cu_count > 0 ? cu_count : 256);
  // the core-check pruning runs in the workgroup tier, which launches first
  // and holds the long rows; the later tiers decide every edge they own
  const bool prune = p.prune && !cn_out && g->mirror;
  k_sim_units<LargeTier><<<cus, LargeTier::kThreads, 0, s>>>(g->n, g->m, g->src, g->adj, g->mirror, p.eps, b.flag,
                                                             cn_out, g->woff, window_count(g->n) + 1, b.uedge, b.units,
                                                             b.ucap, b.cnt, kCntUnits2, kCntQueue2,
                                                             g->tuning.hash_tables != 0, prune ? b.core : nullptr,
                                                             b.ny, b.nn, p.mu);
  if (mid) (void)hipEventRecord(mid, s);
  k_sim_units<SmallTier><<<cus * 8, SmallTier::kThreads, 0, s>>>(g->n, g->m, g->src, g->adj, g->mirror, p.eps,
                                                                 b.flag, cn_out, nullptr, 0, b.uedge, b.units, b.ucap,
                                                                 b.cnt, kCntUnits1, kCntQueue1, g->tuning.hash_tables != 0,
                                                                 nullptr, nullptr, nullptr, 0);
  if (g->tuning.tiny_work) {
    // the list's length is on the device: enough threads for every owner of
    // the chip's vertices, grid-striding beyond two per CU slot
    const uint64_t want = ((uint64_t)g->n + kBlock - 1) / kBlock, cap = (uint64_t)cus * 32;
    k_sim_tiny<<<(uint32_t)(want < cap ? want : cap), kBlock, 0, s>>>(g->m, g->row_begin, g->row_end, g->adj,
                                                                     g->mirror, p.eps, b.cnt_u, b.tiny, b.uedge,
                                                                     b.flag, cn_out, b.cnt, nullptr, 0u);
  }
}

namespace {
// Unit, queue and tiny-owner counters before a round of packed records (so a
// round with no owners runs no unit of the previous one).
__global__ void k_round_reset(uint64_t* __restrict__ cnt) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt[kCntUnits1] = cnt[kCntUnits2] = 0;
    cnt[kCntQueue1] = cnt[kCntQueue2] = 0;
    cnt[kCntTiny] = 0;
  }
}
// The rows of a chip (its pair filter's one or two row ranges), or all rows
// of a whole graph: the only rows whose planner counts a pass touches.
struct RowRanges {
  uint32_t lo[2], hi[2];
  int count;
};
RowRanges chip_rows(const sgc_dgraph* g) {
  if (!g->pair.on) return RowRanges{{0, 0}, {g->n, 0}, 1};
  if (g->pair.lo1 == g->pair.lo0 && g->pair.hi1 == g->pair.hi0)
    return RowRanges{{g->pair.lo0, 0}, {g->pair.hi0, 0}, 1};
  return RowRanges{{g->pair.lo0, g->pair.lo1}, {g->pair.hi0, g->pair.hi1}, 2};
}

// A second scanned range continues the first: out[i] += the first range's
// last start + its last count.
__global__ void __launch_bounds__(kBlock) k_scan_carry(uint32_t* __restrict__ out, uint32_t count,
                                                       const uint32_t* __restrict__ last_start,
                                                       const uint32_t* __restrict__ last_count) {
  const uint32_t add = *last_start + *last_count;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) out[i] += add;
}
}  // namespace

void launch_prep_packed(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s) {
  if (g->m == 0) return;
  k_prep<false, false, kRecCount><<<grid_for(g->m, kBlock), kBlock, 0, s>>>(
      g->m, g->row_begin, g->row_end, g->adj, g->src, g->ndeg, g->mirror, p.eps, false, g->pair, b.flag, b.cnt_u,
      b.work_u, b.uedge, b.cnt, nullptr, nullptr);
  // each owner's first record: an exclusive scan of the owned counts over
  // the chip's rows, the second range continuing where the first ends
  const RowRanges rr = chip_rows(g);
  for (int q = 0; q < rr.count; ++q) {
    const uint32_t x0 = rr.lo[q], count = rr.hi[q] - rr.lo[q];
    if (!count) continue;
    size_t tmp_bytes = b.scan_tmp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, tmp_bytes, b.cnt_u + x0, b.rec_start + x0, count, s);
    if (q == 1 && rr.hi[0] > rr.lo[0])
      k_scan_carry<<<grid_for(count, kBlock), kBlock, 0, s>>>(b.rec_start + x0, count, b.rec_start + rr.hi[0] - 1,
                                                              b.cnt_u + rr.hi[0] - 1);
    (void)hipMemcpyAsync(b.rec_cur + x0, b.rec_start + x0, (size_t)count * 4, hipMemcpyDeviceToDevice, s);
  }
}

void launch_similarity_packed(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, int impl,
                              hipEvent_t mid, hipStream_t s, uint32_t w, uint32_t rounds) {
  if (g->m == 0) return;
  if (impl == kSimWavePerEdge) {  // lists owned entries from the flags; no records
    launch_similarity_wave(g, b, p, cu_count, s);
    if (mid) (void)hipEventRecord(mid, s);
    return;
  }
  const uint32_t cus = (uint32_t)(cu_count > 0 ? cu_count : 256);
  SimTuning t = g->tuning;
  if (t.small_max_degree > kT1MaxDeg) t.small_max_degree = kT1MaxDeg;
  const RowRanges rr = chip_rows(g);
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t r0 = r * w;
    k_round_reset<<<1, kWave, 0, s>>>(b.cnt);
    k_prep<false, false, kRecCompact><<<grid_for(g->m, kBlock), kBlock, 0, s>>>(
        g->m, g->row_begin, g->row_end, g->adj, g->src, g->ndeg, g->mirror, p.eps, false, g->pair, b.flag, b.cnt_u,
        b.work_u, b.uedge, b.cnt, nullptr, nullptr, b.rec_cur, b.rec_start, r0, w);
    for (int q = 0; q < rr.count; ++q)
      if (rr.hi[q] > rr.lo[q])
        k_plan<<<grid_for(rr.hi[q] - rr.lo[q], kBlock), kBlock, 0, s>>>(
            rr.lo[q], rr.hi[q], g->row_begin, g->row_end, t, b.cnt_u, b.work_u, b.units, b.ucap, b.tiny, b.cnt, nullptr,
            nullptr, nullptr, p.mu, false, b.rec_start, r0, w);
    k_sim_units<LargeTier><<<cus, LargeTier::kThreads, 0, s>>>(g->n, g->m, g->src, g->adj, g->mirror, p.eps, b.flag,
                                                               nullptr, g->woff, window_count(g->n) + 1, b.uedge, b.units,
                                                               b.ucap, b.cnt, kCntUnits2, kCntQueue2,
                                                               g->tuning.hash_tables != 0, nullptr, nullptr, nullptr,
                                                               0);
    if (r == 0 && mid) (void)hipEventRecord(mid, s);
    k_sim_units<SmallTier><<<cus * 8, SmallTier::kThreads, 0, s>>>(g->n, g->m, g->src, g->adj, g->mirror, p.eps,
                                                                   b.flag, nullptr, nullptr, 0, b.uedge, b.units,
                                                                   b.ucap, b.cnt, kCntUnits1, kCntQueue1,
                                                                   g->tuning.hash_tables != 0, nullptr, nullptr,
                                                                   nullptr, 0);
    if (t.tiny_work) {
      const uint64_t want = ((uint64_t)g->n + kBlock - 1) / kBlock, cap = (uint64_t)cus * 32;
      k_sim_tiny<<<(uint32_t)(want < cap ? want : cap), kBlock, 0, s>>>(g->m, g->row_begin, g->row_end, g->adj,
                                                                       g->mirror, p.eps, b.cnt_u, b.tiny, b.uedge,
                                                                       b.flag, nullptr, b.cnt, b.rec_start, r0);
    }
  }
}

namespace {
// Planner state for the second similarity round of a pruned pass: unit and
// queue counters always (so that the round's kernels find no unit when it
// has nothing to do), per-owner counts only when it has edges.
__global__ void __launch_bounds__(kBlock) k_redo_init(uint32_t n, uint32_t* __restrict__ cnt_u,
                                                      uint64_t* __restrict__ work_u, uint64_t* __restrict__ cnt) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt[kCntUnits1] = cnt[kCntUnits2] = 0;
    cnt[kCntQueue1] = cnt[kCntQueue2] = 0;
    cnt[kCntTiny] = 0;
    cnt[kCntRedoOwned] = 0;
  }
  if (cnt[kCntRedo] == 0) return;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += stride) {
    cnt_u[x] = 0;
    work_u[x] = 0;
  }
}
}  // namespace

void launch_redo(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, hipStream_t s) {
  if (g->m == 0 || !g->mirror) return;
  launch_redo_mark(g, b, s);
  k_redo_init<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(g->n, b.cnt_u, b.work_u, b.cnt);
  k_prep<false, true><<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->row_begin, g->row_end, g->adj, g->src,
                                                                g->ndeg, g->mirror, p.eps, false, g->pair, b.flag,
                                                                b.cnt_u, b.work_u, b.uedge, b.cnt, nullptr, nullptr);
  // no tiny-owner list this round: it would live in the label array, which
  // the core pass has initialised by now
  SimTuning t = g->tuning;
  if (t.small_max_degree > kT1MaxDeg) t.small_max_degree = kT1MaxDeg;
  t.tiny_work = 0;
  k_plan<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(0, g->n, g->row_begin, g->row_end, t, b.cnt_u, b.work_u, b.units,
                                                   b.ucap, b.tiny, b.cnt, nullptr, nullptr, nullptr, p.mu, true);
  const uint32_t cus = (uint32_t)(cu_count > 0 ? cu_count : 256);
  k_sim_units<LargeTier><<<cus, LargeTier::kThreads, 0, s>>>(g->n, g->m, g->src, g->adj, g->mirror, p.eps, b.flag,
                                                             nullptr, g->woff, window_count(g->n) + 1, b.uedge, b.units,
                                                             b.ucap, b.cnt, kCntUnits2, kCntQueue2,
                                                             g->tuning.hash_tables != 0, nullptr, nullptr, nullptr, 0);
  k_sim_units<SmallTier><<<cus * 8, SmallTier::kThreads, 0, s>>>(g->n, g->m, g->src, g->adj, g->mirror, p.eps,
                                                                 b.flag, nullptr, nullptr, 0, b.uedge, b.units,
                                                                 b.ucap, b.cnt, kCntUnits1, kCntQueue1,
                                                                 g->tuning.hash_tables != 0, nullptr, nullptr, nullptr,
                                                                 0);
  launch_union(g, b, s, kCntRedo);
}

}  // namespace sgc
