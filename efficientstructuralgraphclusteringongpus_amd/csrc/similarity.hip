// Similarity phase (GPUSCAN++/GPUScan.cu:133-222), tiered for gfx950.
//
// 1. k_prep: per directed entry, the reference's degree-only test; an
//    undecided edge (a, b), a the lower (degree, id) endpoint, is owned by b:
//    its entry in b's row is marked kFlagUDown, the one in a's row kFlagUUp.
//    Per owner b: number of owned edges and the sum of their short-row
//    lengths d(a) (the streaming work).
// 2. exclusive scan of the owned-edge counts, scatter of the owned entries
//    into `ulist` grouped by owner.
// 3. k_plan: each owner's edges are cut into units of bounded work.
//    Owners with d(b) <= kT1MaxDeg go to the wave tier, the rest to the
//    workgroup tier.
// 4. Counting: the owner's row N(b) is put into an LDS hash table and each
//    owned edge streams the short row N(a) (4 coalesced loads per lane per
//    step) probing the table.  Every 256 elements the running count is
//    compared with the exact threshold of the reference's predicate:
//    reached -> similar, unreachable -> not similar.  The result is written
//    to both directed entries (the position of b in N(a) comes from the
//    stream itself or from a 64-way wave search).
//    Wave tier: each wave owns a private table (<= 2048 slots) and works one
//    unit at a time.  Workgroup tier: 1024 threads share a 32768-slot table;
//    rows longer than 16384 are split into value segments, each edge's short
//    row is streamed segment by segment with its count kept in LDS.
#include <hip/hip_runtime.h>

#include "device.h"
#include "device_util.h"

namespace sgc {

namespace {

constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;

constexpr uint32_t kT1MaxDeg = 1024;
constexpr uint32_t kT1Slots = 2048;
constexpr int kT1Waves = 4;
constexpr double kT1UnitWork = 65536.0;

constexpr uint32_t kT2SegCap = 16384;
constexpr uint32_t kT2Slots = 32768;
constexpr int kT2Threads = 1024;
constexpr uint32_t kT2MaxEdges = 512;
constexpr double kT2UnitWork = 1048576.0;

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

__global__ void __launch_bounds__(kBlock) k_prep(uint32_t m, const uint32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ adj, const uint32_t* __restrict__ src,
                                                 double eps, uint8_t* __restrict__ flag,
                                                 uint32_t* __restrict__ cnt_u, uint64_t* __restrict__ work_u,
                                                 uint64_t* __restrict__ cnt) {
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint32_t owned_total = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < m; base += stride) {
    const uint64_t k = base + lane;
    bool owned = false;
    uint32_t row = 0xFFFFFFFFu;
    uint64_t w = 0;
    if (k < m) {
      const uint32_t u = src[k];
      const uint32_t v = (uint32_t)adj[k];
      const uint32_t du = deg_of(rp, u), dv = deg_of(rp, v);
      // orientation of GPUSCAN++/Graph.cpp:75-83: a = lower (degree, id)
      const bool up = dv > du || (dv == du && v > u);
      const uint32_t da = up ? du : dv, db = up ? dv : du;
      const double A = (double)(da + 1u);
      const double B = (double)(db + 1u);
      uint8_t f;
      if (A < B * eps * eps)
        f = kFlagN;
      else if (eps * eps * A * B <= 4.0)
        f = kFlagY;
      else
        f = up ? kFlagUUp : kFlagUDown;
      flag[k] = f;
      if (f == kFlagUDown) {
        owned = true;
        row = u;
        w = dv;
      }
    }
    const uint64_t om = __ballot(owned);
    if (om) {
      owned_total += (uint32_t)__popcll(om);
      const uint32_t c = wave_seg_reduce(row, owned ? 1u : 0u, OpAdd());
      const uint64_t ws = wave_seg_reduce(row, w, OpAdd());
      const bool head = wave_seg_head(row);  // all lanes: the shuffle reads the previous lane
      if (owned && head) {
        atomicAdd(&cnt_u[row], c);
        atomicAdd((unsigned long long*)&work_u[row], (unsigned long long)ws);
      }
    }
  }
  if (lane == 0 && owned_total) atomicAdd((unsigned long long*)&cnt[kCntUEdges], (unsigned long long)owned_total);
}

// ---------------------------------------------------------------------------
// 2. Exclusive scan (three kernels) and scatter into ulist
// ---------------------------------------------------------------------------

constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) sh[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < kBlock / kWave; ++i) {
      const uint32_t t = sh[i];
      sh[i] = acc;
      acc += t;
    }
    sh[kBlock / kWave] = acc;
  }
  __syncthreads();
  const uint32_t r = sh[w] + incl - v;
  *total = sh[kBlock / kWave];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(uint32_t nitems, const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ bsum) {
  __shared__ uint32_t sh[kBlock / kWave + 1];
  const uint64_t t0 = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t s = 0;
  for (uint32_t i = 0; i < kScanItems; ++i)
    if (t0 + i < nitems) s += in[t0 + i];
  uint32_t total;
  block_excl_scan(s, sh, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Single block: exclusive scan of nb block sums in place; out[nitems] gets the
// grand total.
__global__ void __launch_bounds__(kBlock) k_scan_blocks(uint32_t nb, uint32_t* __restrict__ bsum, uint32_t nitems,
                                                        uint32_t* __restrict__ out) {
  __shared__ uint32_t sh[kBlock / kWave + 1];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0u;
    uint32_t total;
    const uint32_t e = block_excl_scan(v, sh, &total);
    if (i < nb) bsum[i] = carry + e;
    carry += total;
  }
  if (threadIdx.x == 0) out[nitems] = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(uint32_t nitems, const uint32_t* __restrict__ in,
                                                       const uint32_t* __restrict__ bsum, uint32_t* __restrict__ out) {
  __shared__ uint32_t sh[kBlock / kWave + 1];
  const uint64_t t0 = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t s = 0;
  for (uint32_t i = 0; i < kScanItems; ++i) {
    vals[i] = t0 + i < nitems ? in[t0 + i] : 0u;
    s += vals[i];
  }
  uint32_t total;
  uint32_t run = bsum[blockIdx.x] + block_excl_scan(s, sh, &total);
  for (uint32_t i = 0; i < kScanItems; ++i) {
    if (t0 + i < nitems) out[t0 + i] = run;
    run += vals[i];
  }
}

// Owned entries to ulist[ustart[b] + cursor]: one atomic per row segment of
// each wave; order inside a row is whatever the atomics give (each entry is
// counted exactly once, which is all the planner needs).
__global__ void __launch_bounds__(kBlock) k_scatter(uint32_t m, const uint32_t* __restrict__ src,
                                                    const uint8_t* __restrict__ flag,
                                                    const uint32_t* __restrict__ cnt_u,
                                                    const uint32_t* __restrict__ ustart, uint32_t* __restrict__ fill,
                                                    uint32_t* __restrict__ ulist) {
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < m; base += stride) {
    const uint64_t k = base + lane;
    const bool sel = k < m && flag[k] == kFlagUDown;
    const uint64_t sm = __ballot(sel);
    if (!sm) continue;
    const uint32_t row = sel ? src[k] : 0xFFFFFFFFu;
    const bool head = wave_seg_head(row);
    const uint64_t hm = __ballot(head);
    // my segment's head lane and the next head after it
    const uint32_t h = 63 - __clzll((long long)(hm & lanemask_le(lane)));
    const uint64_t after = hm & ~lanemask_le(h);
    const uint32_t next = after ? (uint32_t)(__ffsll((long long)after) - 1) : (uint32_t)kWave;
    uint32_t segbase = 0;
    if (sel && head) segbase = atomicAdd(&fill[row], next - h);
    segbase = __shfl(segbase, h);
    if (sel) {
      const uint32_t off = segbase + (lane - h);
      if (off < cnt_u[row]) ulist[ustart[row] + off] = (uint32_t)k;
    }
  }
}

// ---------------------------------------------------------------------------
// 3. Work units
// ---------------------------------------------------------------------------

// Edges per unit so that a unit streams about `target` short-row elements.
__device__ __forceinline__ uint32_t edges_per_unit(uint32_t cnt, uint64_t work, double target, uint32_t cap) {
  double k = work ? target * (double)cnt / (double)work : (double)cnt;
  uint32_t e = k >= (double)cnt ? cnt : (k < 1.0 ? 1u : (uint32_t)k);
  if (e > cap) e = cap;
  return e ? e : 1u;
}

__global__ void __launch_bounds__(kBlock) k_plan(uint32_t n, const uint32_t* __restrict__ rp,
                                                 const uint32_t* __restrict__ cnt_u,
                                                 const uint64_t* __restrict__ work_u,
                                                 const uint32_t* __restrict__ ustart, uint2* __restrict__ units1,
                                                 uint2* __restrict__ units2, uint64_t* __restrict__ cnt) {
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < n; base += stride) {
    const uint64_t x = base + lane;
    uint32_t c1 = 0, c2 = 0, per = 1, owned = 0, start = 0;
    if (x < n) {
      owned = cnt_u[x];
      if (owned) {
        start = ustart[x];
        const uint32_t d = deg_of(rp, (uint32_t)x);
        if (d <= kT1MaxDeg) {
          per = edges_per_unit(owned, work_u[x], kT1UnitWork, owned);
          c1 = (owned + per - 1) / per;
        } else {
          per = edges_per_unit(owned, work_u[x], kT2UnitWork, kT2MaxEdges);
          c2 = (owned + per - 1) / per;
        }
      }
    }
    const uint32_t i1 = wave_incl_scan(c1), i2 = wave_incl_scan(c2);
    const uint32_t t1 = __shfl(i1, kWave - 1), t2 = __shfl(i2, kWave - 1);
    uint32_t b1 = 0, b2 = 0;
    if (lane == 0) {
      if (t1) b1 = (uint32_t)atomicAdd((unsigned long long*)&cnt[kCntUnits1], (unsigned long long)t1);
      if (t2) b2 = (uint32_t)atomicAdd((unsigned long long*)&cnt[kCntUnits2], (unsigned long long)t2);
    }
    b1 = __shfl(b1, 0);
    b2 = __shfl(b2, 0);
    uint2* dst = c1 ? units1 + b1 + (i1 - c1) : units2 + b2 + (i2 - c2);
    const uint32_t cu = c1 ? c1 : c2;
    const uint64_t end = (uint64_t)start + owned;
    for (uint32_t t = 0; t < cu; ++t) {
      const uint64_t j0 = (uint64_t)start + (uint64_t)t * per;
      const uint64_t j1 = j0 + per < end ? j0 + per : end;
      dst[t] = make_uint2((uint32_t)j0, (uint32_t)j1);
    }
  }
}

// ---------------------------------------------------------------------------
// 4. Counting
// ---------------------------------------------------------------------------

// Result of streaming part of one short row against the table.
struct StreamState {
  uint32_t cn;       // common neighbours found so far
  uint32_t pos;      // next unstreamed index of the short row
  uint32_t mirror;   // index of b in the short row, kNone if not seen yet
  bool decided;
  bool sim;
};

// Streams adj[pos, hi) of the short row (which ends at ae) against the table,
// 256 elements per step, checking the exact threshold after each step.
// Wave-uniform arguments; returns the updated state.
__device__ __forceinline__ StreamState stream_probe(const int32_t* __restrict__ adj, const uint32_t* tab,
                                                    uint32_t mask, uint32_t shift, uint32_t hi, uint32_t ae,
                                                    int32_t b, bool thresholded, uint32_t cstar, StreamState st) {
  const uint32_t lane = lane_id();
  uint32_t my_mirror = kNone;
  while (!st.decided && st.pos < hi) {
    const uint32_t i0 = st.pos + lane;
    const uint32_t i1 = i0 + 64, i2 = i0 + 128, i3 = i0 + 192;
    const int32_t x0 = i0 < hi ? adj[i0] : -1;
    const int32_t x1 = i1 < hi ? adj[i1] : -1;
    const int32_t x2 = i2 < hi ? adj[i2] : -1;
    const int32_t x3 = i3 < hi ? adj[i3] : -1;
    const bool h0 = x0 >= 0 && table_contains(tab, mask, shift, (uint32_t)x0);
    const bool h1 = x1 >= 0 && table_contains(tab, mask, shift, (uint32_t)x1);
    const bool h2 = x2 >= 0 && table_contains(tab, mask, shift, (uint32_t)x2);
    const bool h3 = x3 >= 0 && table_contains(tab, mask, shift, (uint32_t)x3);
    if (x0 == b) my_mirror = i0;
    if (x1 == b) my_mirror = i1;
    if (x2 == b) my_mirror = i2;
    if (x3 == b) my_mirror = i3;
    st.cn += (uint32_t)(__popcll(__ballot(h0)) + __popcll(__ballot(h1)) + __popcll(__ballot(h2)) +
                        __popcll(__ballot(h3)));
    st.pos = st.pos + 256 < hi ? st.pos + 256 : hi;
    if (thresholded) {
      const uint32_t rest = ae - st.pos;
      if (st.cn >= cstar) {
        st.decided = true;
        st.sim = true;
      } else if (st.cn + rest < cstar) {
        st.decided = true;
        st.sim = false;
      }
    }
  }
  const uint64_t mm = __ballot(my_mirror != kNone);
  if (mm) st.mirror = __shfl(my_mirror, __ffsll((long long)mm) - 1);
  return st;
}

// Writes the decision to the owned entry k and to the mirror entry in the
// short row [ab, ae) (searching the part not streamed when it was not seen).
__device__ __forceinline__ void write_result(const int32_t* __restrict__ adj, uint8_t* __restrict__ flag, uint32_t k,
                                             int32_t b, uint32_t ab, uint32_t ae, const StreamState& st) {
  uint32_t mpos = st.mirror;
  if (mpos == kNone) {
    const uint32_t p = wave_lower_bound(adj, ab, ae, b);
    mpos = p < ae && adj[p] == b ? p : kNone;
  }
  if (lane_id() == 0) {
    const uint8_t f = st.sim ? kFlagY : kFlagN;
    flag[k] = f;
    if (mpos != kNone) flag[mpos] = f;
  }
}

__device__ __forceinline__ StreamState initial_state(bool thresholded, uint32_t cstar, uint32_t s, uint32_t ab) {
  StreamState st;
  st.cn = 0;
  st.pos = ab;
  st.mirror = kNone;
  st.decided = false;
  st.sim = false;
  if (thresholded && cstar == 0) {
    st.decided = true;
    st.sim = true;
  } else if (thresholded && cstar > s) {
    st.decided = true;
  }
  return st;
}

// Wave tier: one wave per unit, private table.
__global__ void __launch_bounds__(kT1Waves * kWave) k_sim_wave_tier(uint32_t m, const uint32_t* __restrict__ rp,
                                                                    const int32_t* __restrict__ adj,
                                                                    const uint32_t* __restrict__ src, double eps,
                                                                    uint8_t* __restrict__ flag,
                                                                    const uint32_t* __restrict__ ulist,
                                                                    const uint2* __restrict__ units,
                                                                    uint64_t* __restrict__ cnt) {
  __shared__ uint32_t tabs[kT1Waves][kT1Slots];
  const uint32_t lane = lane_id();
  uint32_t* tab = tabs[threadIdx.x / kWave];
  const uint64_t nunits = cnt[kCntUnits1];
  uint64_t streamed = 0;
  for (;;) {
    uint64_t u = 0;
    if (lane == 0) u = atomicAdd((unsigned long long*)&cnt[kCntQueue1], 1ull);
    u = __shfl(u, 0);
    if (u >= nunits) break;
    const uint2 un = units[u];
    const uint32_t k0 = ulist[un.x];
    if (k0 >= m) continue;  // never for consistent counts; keeps every index in range
    const uint32_t b = src[k0];
    const uint32_t bb = rp[b], be = rp[b + 1], l = be - bb;
    if (l > kT1MaxDeg) continue;  // the planner sends only short owner rows here
    const uint32_t T = table_size_for(l);
    const uint32_t mask = T - 1, shift = 32 - log2_pow2(T);
    for (uint32_t t = lane; t < T; t += kWave) tab[t] = kEmptySlot;
    wave_lds_sync();
    for (uint32_t t = bb + lane; t < be; t += kWave) table_insert(tab, mask, shift, (uint32_t)adj[t]);
    wave_lds_sync();
    for (uint32_t j0 = un.x; j0 < un.y; j0 += kWave) {
      // lane t fetches edge j0 + t
      uint32_t my_k = kNone, my_ab = 0, my_ae = 0;
      if (j0 + lane < un.y) {
        my_k = ulist[j0 + lane];
        if (my_k < m) {
          const uint32_t a = (uint32_t)adj[my_k];
          my_ab = rp[a];
          my_ae = rp[a + 1];
        }
      }
      const uint32_t cnt_here = un.y - j0 < (uint32_t)kWave ? un.y - j0 : (uint32_t)kWave;
      for (uint32_t e = 0; e < cnt_here; ++e) {
        const uint32_t k = __shfl(my_k, e), ab = __shfl(my_ab, e), ae = __shfl(my_ae, e);
        if (k >= m) continue;
        const uint32_t s = ae - ab;
        const double rhs = predicate_rhs(eps, s, l);
        const bool thresholded = s <= kMaxThresholdRow;
        const uint32_t cstar = thresholded ? min_common_threshold(rhs, s) : 0u;
        StreamState st = initial_state(thresholded, cstar, s, ab);
        st = stream_probe(adj, tab, mask, shift, ae, ae, (int32_t)b, thresholded, cstar, st);
        if (!st.decided) st.sim = thresholded ? st.cn >= cstar : predicate_exact(st.cn, rhs);
        streamed += st.pos - ab;
        write_result(adj, flag, k, (int32_t)b, ab, ae, st);
      }
    }
    wave_lds_sync();
  }
  if (lane == 0 && streamed) atomicAdd((unsigned long long*)&cnt[kCntProbes], (unsigned long long)streamed);
}

// Workgroup tier: 1024 threads per unit, shared table, value segments of the
// owner's row for rows longer than kT2SegCap.
__global__ void __launch_bounds__(kT2Threads) k_sim_group_tier(uint32_t m, const uint32_t* __restrict__ rp,
                                                               const int32_t* __restrict__ adj,
                                                               const uint32_t* __restrict__ src, double eps,
                                                               uint8_t* __restrict__ flag,
                                                               const uint32_t* __restrict__ ulist,
                                                               const uint2* __restrict__ units,
                                                               uint64_t* __restrict__ cnt) {
  __shared__ uint32_t tab[kT2Slots];
  __shared__ uint32_t e_k[kT2MaxEdges], e_cur[kT2MaxEdges], e_cn[kT2MaxEdges], e_mir[kT2MaxEdges];
  __shared__ uint32_t s_unit, s_next;
  const uint32_t lane = lane_id();
  const uint64_t nunits = cnt[kCntUnits2];
  uint64_t streamed = 0;
  for (;;) {
    if (threadIdx.x == 0) s_unit = (uint32_t)atomicAdd((unsigned long long*)&cnt[kCntQueue2], 1ull);
    __syncthreads();
    const uint64_t u = s_unit;
    if (u >= nunits) break;
    const uint2 un = units[u];
    const uint32_t ne = un.y - un.x;
    const uint32_t k0 = ulist[un.x];
    if (k0 >= m || ne > kT2MaxEdges) {  // never for consistent counts
      __syncthreads();  // every thread has read s_unit before it is rewritten
      continue;
    }
    const uint32_t b = src[k0];
    const uint32_t bb = rp[b], be = rp[b + 1], l = be - bb;
    for (uint32_t t = threadIdx.x; t < ne; t += kT2Threads) {
      const uint32_t k = ulist[un.x + t];
      const bool ok = k < m;
      e_k[t] = k;
      e_cur[t] = ok ? rp[(uint32_t)adj[k]] : 0u;
      e_cn[t] = ok ? 0u : kNone;
      e_mir[t] = kNone;
    }
    const uint32_t nseg = (l + kT2SegCap - 1) / kT2SegCap;
    for (uint32_t seg = 0; seg < nseg; ++seg) {
      const uint32_t sb = bb + seg * kT2SegCap;
      const uint32_t se = sb + kT2SegCap < be ? sb + kT2SegCap : be;
      const bool last = seg + 1 == nseg;
      const int32_t vhi = last ? 0x7FFFFFFF : adj[se];  // this segment owns values < vhi
      const uint32_t T = table_size_for(se - sb);
      const uint32_t mask = T - 1, shift = 32 - log2_pow2(T);
      for (uint32_t t = threadIdx.x; t < T; t += kT2Threads) tab[t] = kEmptySlot;
      __syncthreads();
      for (uint32_t t = sb + threadIdx.x; t < se; t += kT2Threads) table_insert(tab, mask, shift, (uint32_t)adj[t]);
      if (threadIdx.x == 0) s_next = 0;
      __syncthreads();
      for (;;) {
        uint32_t e = 0;
        if (lane == 0) e = atomicAdd(&s_next, 1u);
        e = __shfl(e, 0);
        if (e >= ne) break;
        const uint32_t cn0 = e_cn[e];
        if (cn0 == kNone) continue;  // decided in an earlier segment
        const uint32_t k = e_k[e];
        const uint32_t a = (uint32_t)adj[k];
        const uint32_t ab = rp[a], ae = rp[a + 1], s = ae - ab;
        const double rhs = predicate_rhs(eps, s, l);
        const bool thresholded = s <= kMaxThresholdRow;
        const uint32_t cstar = thresholded ? min_common_threshold(rhs, s) : 0u;
        StreamState st = initial_state(thresholded, cstar, s, ab);
        if (!st.decided) {
          st.cn = cn0;
          st.pos = e_cur[e];
          st.mirror = e_mir[e];
          const uint32_t hi = last ? ae : wave_lower_bound(adj, st.pos, ae, vhi);
          const uint32_t p0 = st.pos;
          st = stream_probe(adj, tab, mask, shift, hi, ae, (int32_t)b, thresholded, cstar, st);
          streamed += st.pos - p0;
          if (!st.decided && last) {
            st.decided = true;
            st.sim = thresholded ? st.cn >= cstar : predicate_exact(st.cn, rhs);
          }
        }
        if (st.decided) {
          write_result(adj, flag, k, (int32_t)b, ab, ae, st);
          if (lane == 0) e_cn[e] = kNone;
        } else if (lane == 0) {
          e_cn[e] = st.cn;
          e_cur[e] = st.pos;
          e_mir[e] = st.mirror;
        }
      }
      __syncthreads();
    }
  }
  if (lane == 0 && streamed) atomicAdd((unsigned long long*)&cnt[kCntProbes], (unsigned long long)streamed);
}

}  // namespace

void launch_prep(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s) {
  if (g->m == 0) return;
  k_prep<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->row_ptr, g->adj, g->src, p.eps, b.flag, b.cnt_u,
                                                   b.work_u, b.cnt);
}

static void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t nitems, uint32_t* tmp, hipStream_t s) {
  const uint32_t nb = (nitems + kScanTile - 1) / kScanTile;
  const uint32_t grid = nb ? nb : 1;
  k_scan_reduce<<<grid, kBlock, 0, s>>>(nitems, in, tmp);
  k_scan_blocks<<<1, kBlock, 0, s>>>(grid, tmp, nitems, out);
  k_scan_apply<<<grid, kBlock, 0, s>>>(nitems, in, tmp, out);
}

size_t scan_tmp_words(uint64_t nitems) { return (nitems + kScanTile - 1) / kScanTile + 1; }

void launch_plan(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s) {
  if (g->m == 0) return;
  exclusive_scan(b.cnt_u, b.ustart, g->n, b.scan_tmp, s);
  k_scatter<<<grid_for(g->m, kBlock), kBlock, 0, s>>>(g->m, g->src, b.flag, b.cnt_u, b.ustart, b.fill, b.ulist);
  k_plan<<<grid_for(g->n, kBlock), kBlock, 0, s>>>(g->n, g->row_ptr, b.cnt_u, b.work_u, b.ustart, b.units1,
                                                   b.units2, b.cnt);
}

void launch_similarity(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, int impl,
                       hipStream_t s) {
  if (g->m == 0) return;
  if (impl == kSimWavePerEdge) {
    launch_similarity_wave(g, b, p, cu_count, s);
    return;
  }
  const uint32_t cus = (uint32_t)(cu_count > 0 ? cu_count : 256);
  k_sim_group_tier<<<cus, kT2Threads, 0, s>>>(g->m, g->row_ptr, g->adj, g->src, p.eps, b.flag, b.ulist, b.units2,
                                              b.cnt);
  k_sim_wave_tier<<<cus * 5, kT1Waves * kWave, 0, s>>>(g->m, g->row_ptr, g->adj, g->src, p.eps, b.flag, b.ulist,
                                                       b.units1, b.cnt);
}

}  // namespace sgc
