// Device-side structures shared by the C-ABI implementation and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "scan_rules.h"

#define SGC_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return ::sgc::fail(SGC_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                               \
  } while (0)

namespace sgc {

constexpr int kWave = 64;
constexpr int32_t kNoLabel = 0x7FFFFFFF;


// Counters written by the kernels (one 64-bit slot each).
enum Counter {
  kCntUEdges = 0,
  kCntSimilar,
  kCntCores,
  kCntClusters,
  kCntBorders,
  kCntHubs,
  kCntOutliers,
  kCntProbes,
  kCntUnits1,   // small-tier work units
  kCntUnits2,   // workgroup-tier work units
  kCntQueue1,   // next small-tier unit to claim
  kCntQueue2,   // next workgroup-tier unit to claim
  kCntLoaded,   // short-row elements loaded by the claims (consumed or not)
  kCntClaims,   // claims taken
  kCntDense,    // owned undecided entries listed for the one-wave-per-edge kernel
  kCntLabeled,  // entries of labelled rows (cores, borders): picks the classify pass's side
  kCntTiny,     // tiny owners listed by the planner for the tiny-owner kernel
  kCntOverflow, // owned undecided entries, when more than the unit array holds (asymmetric input)
  kCntMarked,   // cores marked by the core pass: none lets the later edge passes return at once
  kCntRedo,     // deferred edges the union-find and border passes need: decided after all
  kCntRedoOwned,  // owned undecided entries of that second similarity round
  kCntDeferred,   // edges deferred by the core-check pruning and never needed
  kCntTimer0 = 32,  // phase timers of debug builds (SGC_PHASE_TIMERS): 5 per tier
  kCntSlots = 48
};

// Similarity kernels selectable per context.
enum SimImpl { kSimTiered = 0, kSimWavePerEdge = 1 };

// Grow-only device allocation owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace sgc

struct sgc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  char name[256] = {0};
  int cu_count = 0;
  sgc::DevBuf work;      // per-pass arrays
  sgc::DevBuf counters;  // kCntSlots x uint64
  uint64_t* host_counters = nullptr;  // pinned mirror
  hipEvent_t ev[8] = {};
  int sim_impl = sgc::kSimTiered;
  void* scratch = nullptr;  // this GPU's static scratch (sgc_scan_device_noalloc)
  // page-locked staging for graph uploads from pageable memory (two
  // buffers, allocated at the first such upload and kept until destroy)
  void* bounce[2] = {nullptr, nullptr};
  hipEvent_t bounce_done[2] = {nullptr, nullptr};
  bool prune = false;       // core-check pruning in passes that return no similarity flags (sgc_ctx_set_pruning)
};

namespace sgc {
// Planner knobs of the tiered similarity kernels (see sgc_tuning).
struct SimTuning {
  uint32_t small_max_degree = 1024;  // owners up to this degree take the small tier (<= 1024)
  double small_unit_work = 262144.0;  // short-row elements per small-tier unit
  double large_unit_work = 8388608.0;  // per workgroup-tier unit
  uint32_t hash_tables = 1;  // 1: cuckoo hashes where they fit; 0: linear probing and bitmap windows only
  uint32_t tiny_work = 2048;  // owners with sum over their edges of d(a) + d(b) <= this: merged, one edge per lane
};
}  // namespace sgc

namespace sgc {
// Out-of-core chip of two row blocks [lo0, hi0) and [lo1, hi1) (equal for a
// chip of one block): the prep kernel keeps only owned edges whose endpoints
// lie one in each block.
// keep0: entries of rows in [lo0, hi0) outside the pair keep their flag (an
// earlier chip of the same group of blocks decided them); every other entry
// outside the pair is marked kFlagUSkip.
struct PairFilter {
  uint32_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
  bool on = false;
  bool keep0 = false;
};
}  // namespace sgc

struct sgc_dgraph {
  sgc_ctx* ctx = nullptr;
  uint32_t n = 0;
  uint32_t m = 0;
  uint32_t max_degree = 0;
  uint32_t* row_ptr = nullptr;  // n + 1
  int32_t* adj = nullptr;       // m
  uint32_t* src = nullptr;      // m: row of each entry
  // m: position of the reverse entry (v, u) of each entry (u, v); resident
  // graphs only (null: the similarity kernels find it themselves)
  uint32_t* mirror = nullptr;
  // m: degree of each entry's neighbour, d(adj[k]); resident graphs only
  // (null: the prep kernel gathers it from the row offsets)
  uint32_t* ndeg = nullptr;
  // Row x is adj[row_begin[x], row_end[x]) for every row the similarity
  // kernels read; row_end[x] - row_begin[x] is x's degree for every x.  A
  // resident graph uses row_ptr and row_ptr + 1; an out-of-core chip holds
  // only some rows (absent rows: a range of the right length that is never
  // read).
  const uint32_t* row_begin = nullptr;
  const uint32_t* row_end = nullptr;
  sgc::SimTuning tuning;
  sgc::PairFilter pair;
  // Window-offset index of a resident graph with n > 2^20 (null otherwise):
  // (wcount + 1) positions per row, where the row crosses each 2^20-id
  // window boundary; the workgroup tier then streams exact window slices.
  uint32_t* woff = nullptr;
  uint32_t wcount = 0;
};

namespace sgc {

// Arrays one SCAN pass needs on the device.
struct PassBuffers {
  uint8_t* flag;     // m
  uint32_t* sd;      // n: similar-degree counts
  int32_t* parent;   // n: union-find forest over cores (parent < child)
  int32_t* label;    // n: cluster id (roots written by the flatten kernel)
  int32_t* lo;       // n: classify min neighbour label
  int32_t* hi;       // n: classify max neighbour label
  uint8_t* core;     // n
  uint32_t* cnt_u;   // n: undecided edges owned by each vertex (kFlagUDown entries)
  uint64_t* work_u;  // n: sum of the short-row lengths of those edges
  // m: owner b's undecided edges at [rb[b], rb[b] + cnt_u[b]) as (entry,
  // its reverse entry, row start of a, d(a))
  uint4* uedge;
  // ucap work units (begin, end in uedge, the owner's row begin and end):
  // small tier from the front, workgroup tier from the back
  uint4* units;
  // n: ids of the tiny owners (cnt[kCntTiny] of them).  Shares label's
  // memory, which is free until the core pass writes it.
  uint32_t* tiny;
  // Pruned core check, until the core pass: similar (ny) and dissimilar (nn)
  // edges seen per vertex, in lo's and hi's memory, and each vertex's
  // state (kVUndecided / kVCore / kVNonCore) in core's.
  uint32_t* ny;
  uint32_t* nn;
  // Packed edge records of out-of-core chips (launch_prep_packed): each
  // owner's first record (n, the exclusive prefix of its owned count) and
  // its cursor (n), and the scan's scratch.
  uint32_t* rec_start = nullptr;
  uint32_t* rec_cur = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  uint32_t ucap;
  uint64_t* cnt;     // kCntSlots
};

struct ScanParams {
  double eps;
  int32_t mu;
  // core-check pruning (GPUSCAN++/GPUScan.cu:303): the workgroup tier defers
  // an edge whose endpoints are both decided (core or not) when it starts
  bool prune = false;
};

// Per-vertex state of the pruned core check, kept in PassBuffers::core until
// the core pass writes the final flags there.
enum : uint8_t { kVUndecided = 0, kVCore = 1, kVNonCore = 2 };

// Carves the arrays of one pass over g from the context's work buffer.
int carve_pass_buffers(sgc_ctx* ctx, const sgc_dgraph* g, PassBuffers* pb);

// Units of both tiers together are bounded by the number of owned undecided
// edges: every edge is owned once and has two entries, so at most m / 2.
inline size_t unit_capacity(uint64_t m) { return (size_t)(m / 2) + 2; }

// Status for a pass whose planner found more owned undecided entries than
// unit_capacity (counters as read back from the device): an input in which
// some entry has no reverse entry.  SGC_OK otherwise.
int check_unit_overflow(const uint64_t* cnt_host, uint64_t m);

// Kernel launchers (kernels.hip, similarity.hip).  All run on `s` and never
// synchronise.
void launch_fill_src(const sgc_dgraph* g, hipStream_t s);
// fills g->mirror (after launch_fill_src) by a search per entry; entries
// without a reverse entry get 0xFFFFFFFF
void launch_mirror(const sgc_dgraph* g, hipStream_t s);
// fills g->mirror (after launch_fill_src) by a stable sort of the entries by
// neighbour, or by launch_mirror where the input is not symmetric or the
// sort's memory is not there; synchronises the stream
hipError_t build_mirror(const sgc_dgraph* g, hipStream_t s);
// fills g->ndeg
void launch_neighbor_degrees(const sgc_dgraph* g, hipStream_t s);
// 2^20-id windows of the large-tier bitmap covering n ids, and the kernel
// filling g->woff (see sgc_dgraph).
uint32_t window_count(uint32_t n);
// Bytes of the window-offset index a resident graph of n vertices and m
// entries keeps: n * (W + 1) words when there is more than one window and
// that is at most max(16 m, 1 GiB) bytes (four times the adjacency); else 0.
uint64_t window_index_bytes(uint64_t n, uint64_t m);
void launch_window_offsets(const sgc_dgraph* g, hipStream_t s);
void launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s);
void launch_init(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s);
// count_all: every edge is left to its owner (for exact common-neighbour
// counts) instead of the degree-only decision.
// With p.prune, also counts each row's degree-decided similar and dissimilar
// entries (ny, nn).
void launch_prep(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s, bool count_all = false);
// Scratch bytes of the exclusive scan over n per-owner counts (packed records).
size_t scan_scratch_bytes(uint32_t n);
// Packed edge records, for chips whose record array holds fewer records than
// the chip has entries: the prep kernel writes flags and per-owner counts,
// the counts are scanned into b.rec_start (and copied to b.rec_cur); then
// launch_similarity_packed decides the owned edges in `rounds` rounds, round
// r taking the owners whose records start in [r w, (r + 1) w), packed into
// b.uedge (which must hold w plus the largest owned count); rounds past the
// chip's records return at once.
void launch_prep_packed(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s);
void launch_similarity_packed(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, int impl,
                              hipEvent_t mid, hipStream_t s, uint32_t w, uint32_t rounds);
// With p.prune, also sets every vertex's state from those counts.
void launch_plan(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s);
// Second similarity round of a pruned pass (after the union-find pass): the
// deferred edges the cluster and border passes need (core-core edges whose
// roots differ, core-non-core edges) are marked undecided again, planned and
// decided, and the union-find pass repeats.  Every kernel returns at once
// when no deferred edge is needed.
void launch_redo(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, hipStream_t s);
// `mid` (optional) is recorded between the workgroup-tier and wave-tier
// kernels.  With cn_out the tiered kernels count every owned edge fully and
// write |N(a) & N(b)| to both of its entries instead of flags.
void launch_similarity(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, int impl, hipEvent_t mid,
                       hipStream_t s, uint32_t* cn_out = nullptr);
void launch_similarity_wave(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, int cu_count, hipStream_t s);
void launch_core(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s);
void launch_cluster(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s);
void launch_border(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s);
void launch_classify(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s);
// The single kernels the four launchers above are made of (edge passes over
// entries [0, g->m), vertex passes over all g->n vertices); the out-of-core
// path runs the edge passes once per chip.
void launch_core_count(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s);
// mu_dev (optional): read mu from device memory instead of p (captured launches)
void launch_core_mark(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s,
                      const int32_t* mu_dev = nullptr);
// gate >= 0: the kernel runs only when cnt[gate] is not 0
void launch_union(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s, int gate = -1);
// marks the deferred edges a pruned pass needs after its union-find pass
// (count in cnt[kCntRedo]); needs g->mirror
void launch_redo_mark(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s);
void launch_flatten(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s);
// m_all: entries of the whole graph when g is one of several row chips
void launch_classify_edges(const sgc_dgraph* g, const PassBuffers& b, hipStream_t s, uint64_t m_all = 0);
void launch_finalize(const sgc_dgraph* g, const PassBuffers& b, ScanParams p, hipStream_t s,
                     const int32_t* mu_dev = nullptr);

}  // namespace sgc
