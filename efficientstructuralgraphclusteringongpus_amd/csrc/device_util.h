// Wave64 device helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"

namespace sgc {

constexpr int kBlock = 256;
// Edge passes: 64-entry chunks a wave takes per iteration, their loads all
// issued before the gathers they feed (memory-level parallelism per wave).
constexpr uint32_t kEdgeChunks = 4;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ uint32_t deg_of(const uint32_t* rp, uint32_t x) { return rp[x + 1] - rp[x]; }

__device__ __forceinline__ uint64_t lanemask_lt(uint32_t lane) { return (1ull << lane) - 1ull; }
__device__ __forceinline__ uint64_t lanemask_le(uint32_t lane) { return lane == 63 ? ~0ull : (2ull << lane) - 1ull; }

// Raw buffer resource over `bytes` bytes at p (gfx9 descriptor word 3): a
// buffer load at or past the end returns 0 instead of touching memory, so a
// partial wave load needs neither an exec mask nor a clamped address.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// Lanes of load j (elements j*64 .. j*64+63) that lie inside a claim of len
// elements, as a ballot mask.
__device__ __forceinline__ uint64_t load_mask(uint32_t j, uint32_t len) {
  if (j * kWave >= len) return 0ull;
  const uint32_t r = len - j * kWave;
  return r >= (uint32_t)kWave ? ~0ull : (1ull << r) - 1ull;
}

// Loads and stores that bypass the non-coherent per-CU cache: union-find reads
// parents that other CUs are changing.
__device__ __forceinline__ int32_t ld_coherent(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Orders one wave's LDS accesses across lanes (a wave's DS operations execute
// in order; this keeps the compiler from moving them across the point).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Segmented suffix reduction over a wave: `row` must be non-decreasing across
// lanes (equal rows then form contiguous segments; a lane compares only with
// the lane `off` away, so an interleaved pattern u, v, u would merge); on
// return the first lane of each segment holds the reduction of its segment.
// All 64 lanes must call it.
template <typename T, typename Op>
__device__ __forceinline__ T wave_seg_reduce(uint32_t row, T val, Op op) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T ov = __shfl_down(val, off);
    uint32_t orow = __shfl_down(row, off);
    if (lane + off < kWave && orow == row) val = op(val, ov);
  }
  return val;
}

__device__ __forceinline__ bool wave_seg_head(uint32_t row) {
  uint32_t prev = __shfl_up(row, 1);
  return lane_id() == 0 || prev != row;
}

// Inclusive prefix sum over the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t o = __shfl_up(v, off);
    if (lane >= (uint32_t)off) v += o;
  }
  return v;
}

// Block-wide slot reservation: thread t asks for v[i] slots of counter
// ctr[i] (i < N); on return v[i] is its first slot.  One atomic per counter
// per call; every thread of the kBlock-thread block must call it.
template <int N>
__device__ __forceinline__ void block_reserve(uint32_t (&v)[N], unsigned long long* const (&ctr)[N]) {
  __shared__ uint32_t s_part[N][kBlock / kWave];
  __shared__ uint32_t s_base[N];
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  uint32_t incl[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    incl[i] = wave_incl_scan(v[i]);
    if (lane == kWave - 1) s_part[i][w] = incl[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < N; ++i) {
      uint32_t t = 0;
      for (int q = 0; q < kBlock / kWave; ++q) {
        const uint32_t c = s_part[i][q];
        s_part[i][q] = t;
        t += c;
      }
      s_base[i] = t ? (uint32_t)atomicAdd(ctr[i], (unsigned long long)t) : 0u;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = s_base[i] + s_part[i][w] + incl[i] - v[i];
  __syncthreads();
}

struct OpAdd {
  template <typename T>
  __device__ T operator()(T a, T b) const { return a + b; }
};
struct OpMin {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return a < b ? a : b; }
};
struct OpMax {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
};

// First index in the sorted range adj[lo, hi) whose value is >= key (hi if
// none), by the whole wave: 64 evenly spaced probes narrow the range 64-fold
// per round.  lo, hi and key must be wave-uniform.
__device__ __forceinline__ uint32_t wave_lower_bound(const int32_t* __restrict__ adj, uint32_t lo, uint32_t hi,
                                                     int32_t key) {
  const uint32_t lane = lane_id();
  while (hi - lo > (uint32_t)kWave) {
    const uint32_t step = (hi - lo + kWave - 1) / kWave;
    const uint32_t p = lo + lane * step;
    const int32_t v = p < hi ? adj[p] : 0x7FFFFFFF;
    const uint64_t below = __ballot(v < key);
    if (below == 0) return lo;
    const uint32_t last = 63 - __clzll((long long)below);
    const uint32_t nlo = lo + last * step + 1;
    const uint32_t nhi = lo + (last + 1) * step;
    lo = nlo;
    hi = nhi < hi ? nhi : hi;
  }
  const uint32_t p = lo + lane;
  const int32_t v = p < hi ? adj[p] : 0x7FFFFFFF;
  return lo + (uint32_t)__popcll(__ballot(v < key));
}

// Grid for a grid-stride loop: enough blocks to fill every CU several times
// over, few enough that per-block atomics on shared counters stay cheap.
inline uint32_t grid_for(uint64_t items, uint32_t per_block) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (uint32_t)g;
}

// Adds v over the whole block to *dst with one atomic (all threads call it).
__device__ __forceinline__ void block_atomic_add(unsigned long long* dst, unsigned long long v) {
  __shared__ unsigned long long part[1024 / kWave];
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off);
  if (lane == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t i = 0; i < (blockDim.x + kWave - 1) / kWave; ++i) t += part[i];
    if (t) atomicAdd(dst, t);
  }
  __syncthreads();
}

// Smallest c in [0, s] with (double)((c+2)^2) >= rhs, or s + 1 if none; the
// reference's predicate `(cn+2)*(cn+2) < rhs` (int product, double compare,
// GPUSCAN++/GPUScan.cu:220-221) is then `cn < c` while (s+2)^2 < 2^31.
__device__ __forceinline__ uint32_t min_common_threshold(double rhs, uint32_t s) {
  if (!(rhs > 4.0)) return 0;  // also NaN: (double)lhs < NaN is false -> similar
  double r = sqrt(rhs);
  double c0 = ceil(r) - 2.0;
  uint32_t c = c0 <= 0.0 ? 0u : (c0 >= (double)(s + 1) ? s + 1 : (uint32_t)c0);
  while (c > 0 && (double)((uint64_t)(c + 1) * (uint64_t)(c + 1)) >= rhs) --c;
  while (c <= s && (double)((uint64_t)(c + 2) * (uint64_t)(c + 2)) < rhs) ++c;
  return c;
}

// The reference's predicate with its int32 wrap, for rows too long for the
// exact threshold.
__device__ __forceinline__ bool predicate_exact(uint32_t cn, double rhs) {
  const uint32_t t = cn + 2u;
  const int32_t lhs = (int32_t)(t * t);
  return !((double)lhs < rhs);
}

// Rows up to this length use the threshold form: (s + 2)^2 < 2^31.
constexpr uint32_t kMaxThresholdRow = 46338;

// rhs of the predicate for first-list degree s and second-list degree l, in
// the reference's operand order.
__device__ __forceinline__ double predicate_rhs(double eps, uint32_t s, uint32_t l) {
  return eps * eps * (double)(int32_t)(s + 1u) * (double)(int32_t)(l + 1u);
}

}  // namespace sgc
