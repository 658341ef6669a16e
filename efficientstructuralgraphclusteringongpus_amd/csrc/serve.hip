// Serving engine: many SCAN queries (graph, eps, mu) against resident graphs.
// No counterpart in the reference, which runs one query per process
// (GPUSCAN++/main.cpp:20-111); the per-query semantics are the reference's
// (GPUSCAN++/GPUScan.cu:133-798, canonical borders).
//
// A graph is admitted if it fits the engine's HBM budget: its CSR, the row of
// each entry, cn[] and a fixed number of query slots.  At admission every
// edge's common-neighbour count |N(u) & N(v)| is computed once by the tiered
// similarity kernels in their counting mode.  A query is then a threshold
// pass over cn[] (the degree-only rule, then the reference's predicate in
// its int32/double form) followed by the core, union-find, border and
// classify kernels, all O(m) streaming passes on the slot's stream.
// Continuous batching: queries take free slots of their graph as they arrive
// and are retired as their streams finish; queued queries start in the freed
// slots, so several queries of different graphs overlap on the device.
// Each slot's work is captured once into a hipGraph whose kernels read eps
// and mu from the slot's device parameter block: a query is one 16-byte
// parameter copy and one graph launch (the graph holds no memset nodes; the
// per-query clearing is a kernel).  Eager launches remain selectable.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <unordered_map>
#include <vector>

#include "device.h"
#include "device_util.h"
#include "graph_host.h"
#include "scan_phases.h"

namespace sgc {
namespace {

constexpr int kStreams = 8;

// Query parameters in device memory (captured launches read them).
struct QueryParams {
  double eps;
  int32_t mu;
  int32_t pad;
};

// flag[k] for one query from cn[]: the degree-only rule first
// (GPUSCAN++/GPUScan.cu:143-169), then for undecided edges the predicate
// `(cn+2)*(cn+2) < eps*eps*(s+1)*(l+1)` -> not similar (GPUScan.cu:220-221).
// The same pass counts each row's similar entries into sd (the core count).
// The neighbour degree comes from ndeg when the graph has it.  One wave,
// kEdgeChunks x 64 entries at `base`.
__device__ __forceinline__ void threshold_chunk(uint32_t m, uint64_t base, const uint32_t* __restrict__ rp,
                                                const uint32_t* __restrict__ src, const int32_t* __restrict__ adj,
                                                const uint32_t* __restrict__ ndeg, const uint32_t* __restrict__ cn,
                                                double eps, uint8_t* __restrict__ flag, uint32_t* __restrict__ sd) {
  constexpr uint32_t C = kEdgeChunks;
  const uint32_t lane = lane_id();
  uint32_t u[C], v[C], du[C], dv[C];
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const uint64_t k = base + j * kWave + lane;
    u[j] = k < m ? src[k] : 0xFFFFFFFFu;  // rows must not decrease across lanes
    v[j] = k < m ? (uint32_t)adj[k] : 0u;
    dv[j] = k < m && ndeg ? ndeg[k] : 0u;
  }
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const bool in = u[j] != 0xFFFFFFFFu;
    du[j] = in ? rp[u[j] + 1] - rp[u[j]] : 0u;
    if (!ndeg) dv[j] = in ? rp[v[j] + 1] - rp[v[j]] : 0u;
  }
#pragma unroll
  for (uint32_t j = 0; j < C; ++j) {
    const uint64_t k = base + j * kWave + lane;
    uint32_t y = 0;
    if (k < m) {
      uint8_t f = prep_flag(u[j], v[j], du[j], dv[j], eps);
      if (f >= kFlagUUp) {
        const uint32_t s = f == kFlagUUp ? du[j] : dv[j], l = f == kFlagUUp ? dv[j] : du[j];
        f = predicate_exact(cn[k], predicate_rhs(eps, s, l)) ? kFlagY : kFlagN;
      }
      flag[k] = f;
      y = f == kFlagY;
    }
    if (!__ballot(y)) continue;
    const uint32_t sum = wave_seg_reduce(u[j], y, OpAdd());
    if (wave_seg_head(u[j]) && sum) atomicAdd(&sd[u[j]], sum);
  }
}

// eps comes from prm when given (captured launches), else eps.
__global__ void __launch_bounds__(kBlock) k_threshold(uint32_t m, const uint32_t* __restrict__ rp,
                                                      const uint32_t* __restrict__ src,
                                                      const int32_t* __restrict__ adj,
                                                      const uint32_t* __restrict__ ndeg,
                                                      const uint32_t* __restrict__ cn, double eps,
                                                      const QueryParams* __restrict__ prm,
                                                      uint8_t* __restrict__ flag, uint32_t* __restrict__ sd) {
  if (prm) eps = prm->eps;
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t base = ((uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * (kWave * kEdgeChunks);
       base < m; base += waves * kWave * kEdgeChunks)
    threshold_chunk(m, base, rp, src, adj, ndeg, cn, eps, flag, sd);
}

__global__ void __launch_bounds__(kBlock) k_query_init(uint32_t n, uint32_t* __restrict__ sd, int32_t* __restrict__ lo,
                                                       int32_t* __restrict__ hi, uint64_t* __restrict__ cnt) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += stride) {
    sd[x] = 0;
    lo[x] = kNoLabel;
    hi[x] = -1;
  }
  if (blockIdx.x == 0 && threadIdx.x < kCntSlots) cnt[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Batched queries: one launch of each phase walks a table of queries
// ---------------------------------------------------------------------------

// Queries per batch, and the tiles the batched kernels cut each query into:
// a block takes 1024 entries (four waves of kEdgeChunks x 64) or 256
// vertices of one query, so block-wide reductions never mix two queries.
constexpr uint32_t kBatchMax = 64;
constexpr uint32_t kEdgeTile = (kBlock / kWave) * kWave * kEdgeChunks;
constexpr uint32_t kVertexTile = kBlock;
// Fixed grid of every batched kernel (grid-stride over the batch's tiles).
constexpr uint32_t kBatchGrid = 1024;

// One query of a batch: its graph, its slot's buffers, its parameters, the
// host-visible outputs, and its first tile in the batch's edge and vertex
// tile spaces.
struct BatchItem {
  const uint32_t* rp;
  const uint32_t* src;
  const int32_t* adj;
  const uint32_t* ndeg;
  const uint32_t* cn;
  uint8_t* flag;
  uint32_t* sd;
  int32_t* parent;
  int32_t* label;
  int32_t* lo;
  int32_t* hi;
  uint8_t* core;
  uint64_t* cnt;
  int32_t* out_labels;  // the slot's page-locked label buffer, written by the device
  uint64_t* out_cnt;    // the slot's page-locked counters
  double eps;
  int32_t mu;
  uint32_t n, m;
  uint32_t e_tile0, v_tile0;
};

struct BatchTable {
  uint32_t count;    // queries in the batch
  uint32_t e_tiles;  // edge tiles of all queries
  uint32_t v_tiles;  // vertex tiles of all queries
  uint32_t pad;
  BatchItem items[kBatchMax];
};

// The query that tile t belongs to: the last one whose first tile is <= t
// (first tiles increase with the query index), by one wave ballot.
__device__ __forceinline__ uint32_t batch_item(const BatchTable* bt, uint32_t t, bool edge) {
  const uint32_t lane = lane_id();
  const bool in = lane < bt->count && (edge ? bt->items[lane].e_tile0 : bt->items[lane].v_tile0) <= t;
  return (uint32_t)__popcll(__ballot(in)) - 1u;
}

__device__ __forceinline__ PhaseView item_view(const BatchItem& q) {
  PhaseView v;
  v.n = q.n, v.m = q.m, v.src = q.src, v.adj = q.adj, v.rb = q.rp, v.re = q.rp + 1, v.flag = q.flag;
  v.sd = q.sd, v.core = q.core, v.parent = q.parent, v.label = q.label, v.lo = q.lo, v.hi = q.hi;
  return v;
}

// Loops of the batched kernels: tile t of the batch, its query q and the
// tile's first entry / vertex within q (block-uniform).
#define SGC_BATCH_TILES(bt, tiles_field, edge, TILE)                                 \
  for (uint32_t t = blockIdx.x; t < (bt)->tiles_field; t += gridDim.x)              \
    if (const uint32_t q_ = batch_item((bt), t, (edge)); true)                      \
      if (const BatchItem& Q = (bt)->items[q_]; true)                               \
        if (const uint64_t T0 = (uint64_t)(t - ((edge) ? Q.e_tile0 : Q.v_tile0)) * (TILE); true)

__global__ void __launch_bounds__(kBlock) k_batch_init(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, v_tiles, false, kVertexTile) {
    const uint64_t x = T0 + threadIdx.x;
    if (x < Q.n) {
      Q.sd[x] = 0;
      Q.lo[x] = kNoLabel;
      Q.hi[x] = -1;
    }
    if (T0 == 0 && threadIdx.x < kCntSlots) Q.cnt[threadIdx.x] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) k_batch_threshold(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, e_tiles, true, kEdgeTile) {
    const uint64_t base = T0 + (threadIdx.x / kWave) * (kWave * kEdgeChunks);
    threshold_chunk(Q.m, base, Q.rp, Q.src, Q.adj, Q.ndeg, Q.cn, Q.eps, Q.flag, Q.sd);
  }
}

__global__ void __launch_bounds__(kBlock) k_batch_core_mark(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, v_tiles, false, kVertexTile) {
    const PhaseView v = item_view(Q);
    const uint64_t x = T0 + threadIdx.x;
    uint32_t marked = 0, labeled = 0;
    if (x < Q.n) labeled = core_mark_vertex(v, (uint32_t)x, Q.mu, &marked);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntLabeled], labeled);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntMarked], marked);
  }
}

__global__ void __launch_bounds__(kBlock) k_batch_union(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, e_tiles, true, kEdgeTile) {
    if (Q.cnt[kCntMarked] == 0) continue;  // no core: no core-core edge to join
    union_entries(item_view(Q), T0 + threadIdx.x, kBlock);
  }
}

__global__ void __launch_bounds__(kBlock) k_batch_flatten(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, v_tiles, false, kVertexTile) {
    const uint64_t x = T0 + threadIdx.x;
    if (x < Q.n) flatten_vertex(item_view(Q), (uint32_t)x);
  }
}

__global__ void __launch_bounds__(kBlock) k_batch_border(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, e_tiles, true, kEdgeTile) {
    if (Q.cnt[kCntMarked] == 0) continue;  // no core: no border
    const uint32_t labeled = border_chunk(item_view(Q), T0 + (threadIdx.x / kWave) * (kWave * kEdgeChunks));
    block_atomic_add((unsigned long long*)&Q.cnt[kCntLabeled], labeled);
  }
}

__global__ void __launch_bounds__(kBlock) k_batch_classify(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, e_tiles, true, kEdgeTile) {
    if (Q.cnt[kCntMarked] == 0) continue;  // no label anywhere
    classify_chunk(item_view(Q), T0 + (threadIdx.x / kWave) * (kWave * kEdgeChunks), 3 * Q.cnt[kCntLabeled] < Q.m);
  }
}

// Final labels straight into the slot's page-locked buffer, and the counts.
__global__ void __launch_bounds__(kBlock) k_batch_finalize(const BatchTable* __restrict__ bt) {
  SGC_BATCH_TILES(bt, v_tiles, false, kVertexTile) {
    const PhaseView v = item_view(Q);
    const uint64_t x = T0 + threadIdx.x;
    RoleCounts rc;
    if (x < Q.n) Q.out_labels[x] = finalize_vertex(v, (uint32_t)x, Q.mu, &rc);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntCores], rc.cores);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntClusters], rc.clusters);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntBorders], rc.borders);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntHubs], rc.hubs);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntOutliers], rc.outliers);
    block_atomic_add((unsigned long long*)&Q.cnt[kCntSimilar], rc.similar);
  }
}

// Each query's counters to its slot's page-locked copy (block q: query q).
__global__ void __launch_bounds__(kBlock) k_batch_publish(const BatchTable* __restrict__ bt) {
  for (uint32_t q = blockIdx.x; q < bt->count; q += gridDim.x)
    if (threadIdx.x < kCntSlots) bt->items[q].out_cnt[threadIdx.x] = bt->items[q].cnt[threadIdx.x];
}

#undef SGC_BATCH_TILES

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device bytes of one query slot for a graph of n vertices, m entries.
size_t slot_bytes(uint64_t n, uint64_t m) {
  return align256(std::max<uint64_t>(m, 1)) + 5 * align256(n * 4 + 4) + align256(n + 1) + align256(kCntSlots * 8) +
         align256(sizeof(QueryParams));
}

struct Slot {
  bool busy = false;
  uint64_t ticket = 0;
  int stream = 0;
  PassBuffers pb{};
  int32_t* h_labels = nullptr;  // pinned
  uint64_t* h_cnt = nullptr;    // pinned
  hipEvent_t start = nullptr, done = nullptr;
  double t_submit = 0;
  QueryParams* d_params = nullptr;  // device
  QueryParams* h_params = nullptr;  // pinned
  hipGraphExec_t exec = nullptr;
  // batched: the slot's query runs in a batch lane (no per-slot events);
  // the device addresses of h_labels / h_cnt, which batched kernels write
  bool batched = false;
  int32_t* d_out_labels = nullptr;
  uint64_t* d_out_cnt = nullptr;
};

struct Graph {
  uint32_t id = 0;
  sgc_dgraph* dg = nullptr;
  uint32_t* cn = nullptr;
  void* slot_mem = nullptr;
  size_t bytes = 0;
  std::vector<Slot> slots;
  double precompute_ms = 0;
};

struct Query {
  uint64_t ticket;
  uint32_t graph;
  double eps;
  int32_t mu;
  double t_submit;
};

struct Result {
  std::vector<int32_t> labels;
  sgc_query_stats st;
};

// A lane runs one batch at a time: the table is written to page-locked
// memory, and the lane's hipGraph (captured once: table upload, then every
// batched kernel) is launched on the lane's stream.
struct BatchMember {
  Graph* G;
  Slot* sl;
  uint64_t ticket;
  double t_submit;
};
struct BatchLane {
  hipStream_t stream = nullptr;
  BatchTable* h_table = nullptr;  // page-locked
  BatchTable* d_table = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t start = nullptr, done = nullptr;
  bool busy = false;
  std::vector<BatchMember> members;
};
constexpr int kBatchLanes = 2;

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.f;
}

}  // namespace
}  // namespace sgc

struct sgc_server {
  sgc_ctx* ctx = nullptr;
  uint64_t budget = 0, used = 0;
  uint32_t slots_per_graph = 0;
  uint32_t next_graph = 1;
  uint64_t next_ticket = 1;
  hipStream_t streams[sgc::kStreams] = {};
  hipStream_t capture_stream = nullptr;  // query graphs are captured here, at admission
  int next_stream = 0;
  std::unordered_map<uint32_t, sgc::Graph*> graphs;
  std::deque<sgc::Query> pending;
  std::unordered_map<uint64_t, sgc::Result> done;
  std::deque<uint64_t> fresh;  // retired tickets not yet listed by sgc_server_completed
  std::unordered_map<uint64_t, uint32_t> inflight;  // ticket -> graph
  bool graph_launch = true;
  // batching: queries on graphs of at most batch_cutoff entries run in
  // batches of up to batch_max (0: off)
  uint64_t batch_cutoff = 0;
  uint32_t batch_max = sgc::kBatchMax;
  bool lanes_ready = false;
  sgc::BatchLane lanes[sgc::kBatchLanes];
};

namespace sgc {
namespace {

void free_graph(Graph* G) {
  for (Slot& s : G->slots) {
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.h_params) (void)hipHostFree(s.h_params);
    if (s.h_labels) (void)hipHostFree(s.h_labels);
    if (s.h_cnt) (void)hipHostFree(s.h_cnt);
    if (s.start) (void)hipEventDestroy(s.start);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  if (G->slot_mem) (void)hipFree(G->slot_mem);
  if (G->cn) (void)hipFree(G->cn);
  if (G->dg) sgc_dgraph_free(G->dg);
  delete G;
}

// The work of one query on stream st: eager (p given) or reading eps and mu
// from the slot's parameter block (p null, for capture).
void enqueue_query(const Graph* G, const Slot& sl, const ScanParams* p, hipStream_t st) {
  const sgc_dgraph* g = G->dg;
  const ScanParams pv = p ? *p : ScanParams{0.0, 0};
  const QueryParams* prm = p ? nullptr : sl.d_params;
  const int32_t* mu_dev = p ? nullptr : &sl.d_params->mu;
  k_query_init<<<grid_for(g->n, kBlock), kBlock, 0, st>>>(g->n, sl.pb.sd, sl.pb.lo, sl.pb.hi, sl.pb.cnt);
  if (g->m)
    k_threshold<<<grid_for(g->m, kBlock), kBlock, 0, st>>>(g->m, g->row_ptr, g->src, g->adj, g->ndeg, G->cn, pv.eps,
                                                            prm, sl.pb.flag, sl.pb.sd);
  launch_core_mark(g, sl.pb, pv, st, mu_dev);
  launch_union(g, sl.pb, st);
  launch_flatten(g, sl.pb, st);
  launch_border(g, sl.pb, st);
  launch_classify_edges(g, sl.pb, st);
  launch_finalize(g, sl.pb, pv, st, mu_dev);
  (void)hipMemcpyAsync(sl.h_labels, sl.pb.label, (size_t)g->n * 4, hipMemcpyDeviceToHost, st);
  (void)hipMemcpyAsync(sl.h_cnt, sl.pb.cnt, kCntSlots * 8, hipMemcpyDeviceToHost, st);
}

// Captures the slot's query work into an executable graph.  Capture runs at
// admission on the server's own capture stream, which carries no other work:
// never on a serving stream, where other graphs' queries may be running.
int capture_slot(sgc_server* S, const Graph* G, Slot& sl) {
  hipStream_t st = S->capture_stream;
  SGC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  enqueue_query(G, sl, nullptr, st);
  const hipError_t le = hipGetLastError();
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(st, &graph);
  if (le != hipSuccess || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(SGC_EDEVICE, "capture of the query graph failed: %s",
                hipGetErrorString(le != hipSuccess ? le : ce));
  }
  const hipError_t ie = hipGraphInstantiate(&sl.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(SGC_EDEVICE, "query graph instantiation failed: %s", hipGetErrorString(ie));
  return SGC_OK;
}

// Creates the batch lanes on first use: stream, table buffers, events, and
// the lane's graph (table upload, then each batched kernel over the table's
// tiles with a fixed grid), captured on the idle capture stream.
int ensure_lanes(sgc_server* S) {
  if (S->lanes_ready) return SGC_OK;
  for (BatchLane& L : S->lanes) {
    if (!L.stream) SGC_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    if (!L.h_table) SGC_HIP(hipHostMalloc((void**)&L.h_table, sizeof(BatchTable), hipHostMallocDefault));
    if (!L.d_table) {
      if (hipMalloc((void**)&L.d_table, sizeof(BatchTable)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SGC_ENOMEM, "batch table allocation failed");
      }
    }
    if (!L.start) SGC_HIP(hipEventCreate(&L.start));
    if (!L.done) SGC_HIP(hipEventCreate(&L.done));
    memset(L.h_table, 0, sizeof(BatchTable));
    if (!L.exec) {
      hipStream_t st = S->capture_stream;
      SGC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      (void)hipMemcpyAsync(L.d_table, L.h_table, sizeof(BatchTable), hipMemcpyHostToDevice, st);
      k_batch_init<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_threshold<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_core_mark<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_union<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_flatten<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_border<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_classify<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_finalize<<<kBatchGrid, kBlock, 0, st>>>(L.d_table);
      k_batch_publish<<<kBatchMax, kBlock, 0, st>>>(L.d_table);
      const hipError_t le = hipGetLastError();
      hipGraph_t graph = nullptr;
      const hipError_t ce = hipStreamEndCapture(st, &graph);
      if (le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(SGC_EDEVICE, "capture of the batch graph failed: %s",
                    hipGetErrorString(le != hipSuccess ? le : ce));
      }
      const hipError_t ie = hipGraphInstantiate(&L.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) return fail(SGC_EDEVICE, "batch graph instantiation failed: %s", hipGetErrorString(ie));
    }
  }
  S->lanes_ready = true;
  return SGC_OK;
}

void free_lanes(sgc_server* S) {
  for (BatchLane& L : S->lanes) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (L.exec) (void)hipGraphExecDestroy(L.exec);
    if (L.h_table) (void)hipHostFree(L.h_table);
    if (L.d_table) (void)hipFree(L.d_table);
    if (L.start) (void)hipEventDestroy(L.start);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    L = BatchLane();
  }
  S->lanes_ready = false;
}

// Launches query q into free slot s of graph G.
int launch_query(sgc_server* S, Graph* G, Slot& sl, const Query& q) {
  hipStream_t st = S->streams[sl.stream];
  SGC_HIP(hipEventRecord(sl.start, st));
  if (S->graph_launch) {
    sl.h_params->eps = q.eps;
    sl.h_params->mu = q.mu;
    SGC_HIP(hipMemcpyAsync(sl.d_params, sl.h_params, sizeof(QueryParams), hipMemcpyHostToDevice, st));
    SGC_HIP(hipGraphLaunch(sl.exec, st));
  } else {
    const ScanParams p{q.eps, q.mu};
    enqueue_query(G, sl, &p, st);
    SGC_HIP(hipGetLastError());
  }
  SGC_HIP(hipEventRecord(sl.done, st));
  sl.busy = true;
  sl.ticket = q.ticket;
  sl.t_submit = q.t_submit;
  S->inflight[q.ticket] = G->id;
  return SGC_OK;
}

bool batchable(const sgc_server* S, const Graph* G) { return S->batch_cutoff && G->dg->m <= S->batch_cutoff; }

// Fills lane's table with its members' queries and launches the lane.
int launch_batch(sgc_server* S, BatchLane& L, const std::vector<double>& eps, const std::vector<int32_t>& mu) {
  BatchTable* T = L.h_table;
  uint32_t et = 0, vt = 0;
  T->count = (uint32_t)L.members.size();
  for (uint32_t q = 0; q < T->count; ++q) {
    const BatchMember& b = L.members[q];
    const sgc_dgraph* g = b.G->dg;
    const PassBuffers& pb = b.sl->pb;
    BatchItem& it = T->items[q];
    it.rp = g->row_ptr, it.src = g->src, it.adj = g->adj, it.ndeg = g->ndeg, it.cn = b.G->cn;
    it.flag = pb.flag, it.sd = pb.sd, it.parent = pb.parent, it.label = pb.label, it.lo = pb.lo, it.hi = pb.hi;
    it.core = pb.core, it.cnt = pb.cnt;
    it.out_labels = b.sl->d_out_labels, it.out_cnt = b.sl->d_out_cnt;
    it.eps = eps[q], it.mu = mu[q], it.n = g->n, it.m = g->m;
    it.e_tile0 = et, it.v_tile0 = vt;
    et += (uint32_t)(((uint64_t)g->m + kEdgeTile - 1) / kEdgeTile);
    vt += (uint32_t)std::max<uint64_t>(1, ((uint64_t)g->n + kVertexTile - 1) / kVertexTile);  // its counters are cleared there
  }
  T->e_tiles = et;
  T->v_tiles = vt;
  SGC_HIP(hipEventRecord(L.start, L.stream));
  SGC_HIP(hipGraphLaunch(L.exec, L.stream));
  SGC_HIP(hipEventRecord(L.done, L.stream));
  L.busy = true;
  for (const BatchMember& b : L.members) S->inflight[b.ticket] = b.G->id;
  return SGC_OK;
}

// Starts queued queries: queries on batchable graphs go to free batch lanes
// (each batch takes the oldest queued ones whose graph has a free slot, up to
// batch_max), the others to free slots of their graph (FIFO per graph,
// skipping queries whose graph is full so other graphs keep the device busy).
int dispatch(sgc_server* S) {
  for (BatchLane& L : S->lanes) {
    if (!S->batch_cutoff || !S->lanes_ready || L.busy) continue;
    L.members.clear();
    std::vector<double> eps;
    std::vector<int32_t> mu;
    for (auto it = S->pending.begin(); it != S->pending.end() && L.members.size() < S->batch_max;) {
      auto gi = S->graphs.find(it->graph);
      Graph* G = gi == S->graphs.end() ? nullptr : gi->second;
      Slot* free_slot = nullptr;
      if (G && batchable(S, G))
        for (Slot& sl : G->slots)
          if (!sl.busy) {
            free_slot = &sl;
            break;
          }
      if (!free_slot) {
        ++it;
        continue;
      }
      free_slot->busy = true;
      free_slot->batched = true;
      free_slot->ticket = it->ticket;
      free_slot->t_submit = it->t_submit;
      L.members.push_back(BatchMember{G, free_slot, it->ticket, it->t_submit});
      eps.push_back(it->eps);
      mu.push_back(it->mu);
      it = S->pending.erase(it);
    }
    if (L.members.empty()) break;
    if (int rc = launch_batch(S, L, eps, mu)) return rc;
  }
  for (auto it = S->pending.begin(); it != S->pending.end();) {
    auto gi = S->graphs.find(it->graph);
    if (gi == S->graphs.end()) {  // graph removed: drop with an error result
      Result r;
      memset(&r.st, 0, sizeof(r.st));
      r.st.status = SGC_EINVAL;
      S->done.emplace(it->ticket, std::move(r));
      S->fresh.push_back(it->ticket);  // listed by sgc_server_completed like any retired query
      it = S->pending.erase(it);
      continue;
    }
    Graph* G = gi->second;
    if (batchable(S, G) && S->lanes_ready) {  // waits for a batch lane
      ++it;
      continue;
    }
    Slot* free_slot = nullptr;
    for (Slot& s : G->slots)
      if (!s.busy) {
        free_slot = &s;
        break;
      }
    if (!free_slot) {
      ++it;
      continue;
    }
    int rc = launch_query(S, G, *free_slot, *it);
    if (rc) return rc;
    it = S->pending.erase(it);
  }
  return SGC_OK;
}

// Retires finished slots; with wait, blocks for the oldest in-flight query
// first.  Returns the number retired.
Result slot_result(const Graph* G, const Slot& s, double ms_device) {
  Result r;
  r.labels.assign(s.h_labels, s.h_labels + G->dg->n);
  memset(&r.st, 0, sizeof(r.st));
  r.st.status = SGC_OK;
  r.st.graph = G->id;
  r.st.n = G->dg->n;
  r.st.ms_device = ms_device;
  r.st.ms_latency = now_ms() - s.t_submit;
  const uint64_t* c = s.h_cnt;
  r.st.similar_edges = c[kCntSimilar] / 2;
  r.st.cores = c[kCntCores];
  r.st.clusters = c[kCntClusters];
  r.st.borders = c[kCntBorders];
  r.st.hubs = c[kCntHubs];
  r.st.outliers = c[kCntOutliers];
  return r;
}

int retire(sgc_server* S, bool wait, int* retired) {
  *retired = 0;
  if (wait) {
    // the in-flight query submitted first (its slot's, or its batch's, event)
    hipEvent_t oldest = nullptr;
    uint64_t t_old = ~0ull;
    for (auto& kv : S->graphs)
      for (Slot& s : kv.second->slots)
        if (s.busy && !s.batched && s.ticket < t_old) t_old = s.ticket, oldest = s.done;
    for (BatchLane& L : S->lanes)
      if (L.busy)
        for (const BatchMember& b : L.members)
          if (b.ticket < t_old) t_old = b.ticket, oldest = L.done;
    if (oldest) SGC_HIP(hipEventSynchronize(oldest));
  }
  for (BatchLane& L : S->lanes) {
    if (!L.busy) continue;
    const hipError_t e = hipEventQuery(L.done);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) return fail(SGC_EDEVICE, "batch failed: %s", hipGetErrorString(e));
    const double ms = elapsed(L.start, L.done);
    for (const BatchMember& b : L.members) {
      S->done.emplace(b.ticket, slot_result(b.G, *b.sl, ms));
      S->fresh.push_back(b.ticket);
      S->inflight.erase(b.ticket);
      b.sl->busy = false;
      b.sl->batched = false;
      ++*retired;
    }
    L.members.clear();
    L.busy = false;
  }
  for (auto& kv : S->graphs) {
    Graph* G = kv.second;
    for (Slot& s : G->slots) {
      if (!s.busy || s.batched) continue;
      const hipError_t e = hipEventQuery(s.done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return fail(SGC_EDEVICE, "query failed: %s", hipGetErrorString(e));
      S->done.emplace(s.ticket, slot_result(G, s, elapsed(s.start, s.done)));
      S->fresh.push_back(s.ticket);
      S->inflight.erase(s.ticket);
      s.busy = false;
      ++*retired;
    }
  }
  return SGC_OK;
}

}  // namespace
}  // namespace sgc

using namespace sgc;

extern "C" int sgc_server_create(sgc_ctx* ctx, uint64_t hbm_budget, uint32_t slots_per_graph, sgc_server** out) {
  clear_error();
  if (!ctx || !out || slots_per_graph == 0 || slots_per_graph > 64) return fail(SGC_EINVAL, "bad argument");
  SGC_HIP(hipSetDevice(ctx->device));
  sgc_server* S = new sgc_server();
  S->ctx = ctx;
  S->budget = hbm_budget;
  S->slots_per_graph = slots_per_graph;
  for (int i = 0; i <= kStreams; ++i) {
    hipStream_t* st = i < kStreams ? &S->streams[i] : &S->capture_stream;
    if (hipStreamCreateWithFlags(st, hipStreamNonBlocking) != hipSuccess) {
      for (int j = 0; j < i; ++j) (void)hipStreamDestroy(S->streams[j]);
      delete S;
      return fail(SGC_EDEVICE, "stream creation failed");
    }
  }
  *out = S;
  return SGC_OK;
}

extern "C" void sgc_server_destroy(sgc_server* S) {
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  for (int i = 0; i < kStreams; ++i) (void)hipStreamSynchronize(S->streams[i]);
  free_lanes(S);
  for (auto& kv : S->graphs) free_graph(kv.second);
  for (int i = 0; i < kStreams; ++i) (void)hipStreamDestroy(S->streams[i]);
  if (S->capture_stream) (void)hipStreamDestroy(S->capture_stream);
  delete S;
}

extern "C" uint64_t sgc_server_graph_bytes(const sgc_host_graph* g, uint32_t slots_per_graph) {
  if (!g) return 0;
  const uint64_t n = g->n, m = g->m;
  // row_ptr; adj, src, the reverse-entry and neighbour-degree indexes and
  // cn; the window-offset index (as the upload sizes it); the query slots.
  // The admission pass's similarity scratch is the context's and is not
  // held per graph.
  const uint64_t m_bytes = align256(std::max<uint64_t>(m, 1) * 4);
  const uint64_t woff = window_index_bytes(n, m);
  const uint64_t graph = align256((n + 1) * 4) + 5 * m_bytes + (woff ? align256(woff) : 0);
  return graph + slots_per_graph * slot_bytes(n, m);
}

extern "C" int sgc_server_add_graph(sgc_server* S, const sgc_host_graph* hg, uint32_t* graph_id,
                                    double* precompute_ms) {
  clear_error();
  if (!S || !hg || !graph_id) return fail(SGC_EINVAL, "null argument");
  const uint64_t need = sgc_server_graph_bytes(hg, S->slots_per_graph);
  if (S->used + need > S->budget)
    return fail(SGC_ENOMEM, "admission refused: graph needs %llu bytes, %llu of %llu in use",
                (unsigned long long)need, (unsigned long long)S->used, (unsigned long long)S->budget);
  sgc_ctx* ctx = S->ctx;
  SGC_HIP(hipSetDevice(ctx->device));
  Graph* G = new Graph();
  int rc = sgc_dgraph_upload(ctx, hg, &G->dg);
  if (rc) {
    delete G;
    return rc;
  }
  const uint32_t n = G->dg->n, m = G->dg->m;
  if (hipMalloc(&G->cn, std::max<size_t>(m, 1) * 4) != hipSuccess ||
      hipMalloc(&G->slot_mem, S->slots_per_graph * slot_bytes(n, m)) != hipSuccess) {
    (void)hipGetLastError();
    free_graph(G);
    return fail(SGC_ENOMEM, "device allocation for graph of %u vertices failed", n);
  }
  // exact common-neighbour counts of every edge, once
  const double t0 = now_ms();
  if (m) {
    PassBuffers pb;
    rc = carve_pass_buffers(ctx, G->dg, &pb);
    if (rc) {
      free_graph(G);
      return rc;
    }
    const ScanParams p{1.0, 2};
    hipStream_t s = ctx->stream;
    launch_init(G->dg, pb, s);
    launch_prep(G->dg, pb, p, s, true);
    launch_plan(G->dg, pb, p, s);
    launch_similarity(G->dg, pb, p, ctx->cu_count, kSimTiered, nullptr, s, G->cn);
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess || hipGetLastError() != hipSuccess) {
      free_graph(G);
      return fail(SGC_EDEVICE, "common-neighbour pass failed: %s", hipGetErrorString(e));
    }
    uint64_t c[kCntSlots];
    if (hipMemcpy(c, pb.cnt, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) {
      (void)hipGetLastError();
      free_graph(G);
      return fail(SGC_EDEVICE, "reading the common-neighbour pass counters failed");
    }
    if ((rc = check_unit_overflow(c, m))) {
      free_graph(G);
      return rc;
    }
  }
  G->precompute_ms = now_ms() - t0;
  // query slots
  G->slots.resize(S->slots_per_graph);
  char* base = (char*)G->slot_mem;
  for (uint32_t i = 0; i < S->slots_per_graph; ++i) {
    Slot& sl = G->slots[i];
    char* p = base + i * slot_bytes(n, m);
    sl.pb.flag = (uint8_t*)p;
    p += align256(std::max<uint32_t>(m, 1));
    sl.pb.sd = (uint32_t*)p;
    p += align256((size_t)n * 4 + 4);
    sl.pb.parent = (int32_t*)p;
    p += align256((size_t)n * 4 + 4);
    sl.pb.label = (int32_t*)p;
    p += align256((size_t)n * 4 + 4);
    sl.pb.lo = (int32_t*)p;
    p += align256((size_t)n * 4 + 4);
    sl.pb.hi = (int32_t*)p;
    p += align256((size_t)n * 4 + 4);
    sl.pb.core = (uint8_t*)p;
    p += align256((size_t)n + 1);
    sl.pb.cnt = (uint64_t*)p;
    p += align256(kCntSlots * 8);
    sl.d_params = (QueryParams*)p;
    sl.stream = S->next_stream;
    S->next_stream = (S->next_stream + 1) % kStreams;
    if (hipHostMalloc((void**)&sl.h_labels, std::max<size_t>(n, 1) * 4) != hipSuccess ||
        hipHostMalloc((void**)&sl.h_cnt, kCntSlots * 8) != hipSuccess ||
        hipHostMalloc((void**)&sl.h_params, sizeof(QueryParams)) != hipSuccess ||
        hipEventCreate(&sl.start) != hipSuccess || hipEventCreate(&sl.done) != hipSuccess) {
      (void)hipGetLastError();
      free_graph(G);
      return fail(SGC_ENOMEM, "query slot set-up failed");
    }
    // the device's view of the slot's page-locked outputs (batched kernels
    // write the labels and counters there directly)
    if (hipHostGetDevicePointer((void**)&sl.d_out_labels, sl.h_labels, 0) != hipSuccess ||
        hipHostGetDevicePointer((void**)&sl.d_out_cnt, sl.h_cnt, 0) != hipSuccess) {
      (void)hipGetLastError();
      free_graph(G);
      return fail(SGC_EDEVICE, "device address of a query slot's host buffers failed");
    }
    // the slot's query work as one replayable graph (used when graph launch
    // is on), captured now on the idle capture stream
    if ((rc = capture_slot(S, G, sl))) {
      free_graph(G);
      return rc;
    }
  }
  G->id = S->next_graph++;
  G->bytes = need;
  S->used += need;
  S->graphs[G->id] = G;
  *graph_id = G->id;
  if (precompute_ms) *precompute_ms = G->precompute_ms;
  return SGC_OK;
}

extern "C" int sgc_server_remove_graph(sgc_server* S, uint32_t graph_id) {
  clear_error();
  if (!S) return fail(SGC_EINVAL, "null server");
  auto it = S->graphs.find(graph_id);
  if (it == S->graphs.end()) return fail(SGC_EINVAL, "unknown graph %u", graph_id);
  Graph* G = it->second;
  for (Slot& s : G->slots)
    if (s.busy && !s.batched) SGC_HIP(hipEventSynchronize(s.done));
  for (BatchLane& L : S->lanes)
    if (L.busy) SGC_HIP(hipEventSynchronize(L.done));
  int retired = 0;
  int rc = retire(S, false, &retired);  // keep the finished results
  if (rc) return rc;
  S->used -= G->bytes;
  S->graphs.erase(it);
  free_graph(G);
  return SGC_OK;
}

extern "C" uint64_t sgc_server_bytes_in_use(const sgc_server* S) { return S ? S->used : 0; }

extern "C" int sgc_server_set_graph_launch(sgc_server* S, int enable) {
  clear_error();
  if (!S) return fail(SGC_EINVAL, "null server");
  S->graph_launch = enable != 0;
  return SGC_OK;
}

extern "C" int sgc_server_set_batching(sgc_server* S, uint64_t cutoff_entries, uint32_t max_queries) {
  clear_error();
  if (!S) return fail(SGC_EINVAL, "null server");
  if (max_queries < 1 || max_queries > kBatchMax)
    return fail(SGC_EINVAL, "max_queries %u outside [1, %u]", max_queries, kBatchMax);
  SGC_HIP(hipSetDevice(S->ctx->device));
  if (cutoff_entries) {
    if (int rc = ensure_lanes(S)) return rc;
  } else {
    // queries already in lanes finish there; queued ones go to slots
  }
  S->batch_cutoff = cutoff_entries;
  S->batch_max = max_queries;
  return SGC_OK;
}

extern "C" int sgc_server_get_batching(const sgc_server* S, uint64_t* cutoff_entries, uint32_t* max_queries) {
  clear_error();
  if (!S) return fail(SGC_EINVAL, "null server");
  if (cutoff_entries) *cutoff_entries = S->batch_cutoff;
  if (max_queries) *max_queries = S->batch_max;
  return SGC_OK;
}

extern "C" int sgc_server_autotune_batching(sgc_server* S, double eps, int32_t mu, uint64_t* cutoff_entries) {
  clear_error();
  if (!S) return fail(SGC_EINVAL, "null server");
  if (!S->pending.empty() || !S->inflight.empty()) return fail(SGC_EBUSY, "queries are still open");
  SGC_HIP(hipSetDevice(S->ctx->device));
  if (int rc = ensure_lanes(S)) return rc;
  const bool launch_mode = S->graph_launch;
  S->graph_launch = true;
  BatchLane& L = S->lanes[0];
  // per graph: its entries and how much less device time per query a batch
  // takes than a replay (negative: the batch is faster)
  std::vector<std::pair<uint64_t, double>> gain;
  int rc = SGC_OK;
  constexpr int kRuns = 3;  // the first of each warms up; the faster of the others counts
  for (auto& kv : S->graphs) {
    Graph* G = kv.second;
    const Query q{0, G->id, eps, mu, now_ms()};
    double t_replay = 1e30, t_batch = 1e30;
    // k queries on the graph's slots, per query, by wall time (host launch
    // work included): replayed at once on the slots' streams, which overlap
    // as they do under load, and as one batch
    const uint32_t k = std::min<uint32_t>((uint32_t)G->slots.size(), S->batch_max);
    for (int rep = 0; rep < kRuns && !rc; ++rep) {
      const double t0 = now_ms();
      for (uint32_t i = 0; i < k && !rc; ++i) rc = launch_query(S, G, G->slots[i], q);
      for (uint32_t i = 0; i < k; ++i) {
        Slot& sl = G->slots[i];
        if (sl.busy && hipEventSynchronize(sl.done) != hipSuccess && !rc)
          rc = fail(SGC_EDEVICE, "autotune replay failed");
        sl.busy = false;
      }
      S->inflight.erase(0);
      if (rep) t_replay = std::min<double>(t_replay, (now_ms() - t0) / k);
    }
    for (int rep = 0; rep < kRuns && !rc; ++rep) {
      const double t0 = now_ms();
      L.members.clear();
      for (uint32_t i = 0; i < k; ++i) L.members.push_back(BatchMember{G, &G->slots[i], 0, now_ms()});
      rc = launch_batch(S, L, std::vector<double>(k, eps), std::vector<int32_t>(k, mu));
      if (!rc && hipEventSynchronize(L.done) != hipSuccess) rc = fail(SGC_EDEVICE, "autotune batch failed");
      S->inflight.erase(0);
      L.busy = false;
      if (rep) t_batch = std::min<double>(t_batch, (now_ms() - t0) / k);
    }
    L.members.clear();
    if (rc) break;
    gain.emplace_back(G->dg->m, t_batch - t_replay);
  }
  S->graph_launch = launch_mode;
  if (rc) return rc;
  // the cut-off with the least device time summed over the graphs (one query
  // each): the graphs up to it batched, the rest replayed.  A single graph
  // whose batch wins by noise above a run of graphs where it loses does not
  // move the cut-off past them.
  std::sort(gain.begin(), gain.end());
  uint64_t best = 0;
  double sum = 0, best_sum = 0;
  for (size_t i = 0; i < gain.size(); ++i) {
    sum += gain[i].second;
    if ((i + 1 == gain.size() || gain[i + 1].first != gain[i].first) && sum < best_sum) {
      best_sum = sum;
      best = gain[i].first;
    }
  }
  S->batch_cutoff = best;
  if (cutoff_entries) *cutoff_entries = best;
  return SGC_OK;
}

extern "C" int sgc_server_submit(sgc_server* S, uint32_t graph_id, double eps, int32_t mu, uint64_t* ticket) {
  clear_error();
  if (!S || !ticket) return fail(SGC_EINVAL, "null argument");
  if (!S->graphs.count(graph_id)) return fail(SGC_EINVAL, "unknown graph %u", graph_id);
  SGC_HIP(hipSetDevice(S->ctx->device));
  Query q{S->next_ticket++, graph_id, eps, mu, now_ms()};
  S->pending.push_back(q);
  *ticket = q.ticket;
  return dispatch(S);
}

extern "C" int sgc_server_progress(sgc_server* S, int wait, uint32_t* completed) {
  clear_error();
  if (!S) return fail(SGC_EINVAL, "null server");
  SGC_HIP(hipSetDevice(S->ctx->device));
  int retired = 0;
  int rc = retire(S, wait && !S->inflight.empty(), &retired);
  if (rc) return rc;
  rc = dispatch(S);
  if (completed) *completed = (uint32_t)retired;
  return rc;
}

extern "C" int sgc_server_completed(sgc_server* S, uint64_t* tickets, uint32_t capacity, uint32_t* count) {
  clear_error();
  if (!S || !count || (capacity && !tickets)) return fail(SGC_EINVAL, "null argument");
  uint32_t c = 0;
  while (c < capacity && !S->fresh.empty()) {
    const uint64_t t = S->fresh.front();
    S->fresh.pop_front();
    if (S->done.count(t)) tickets[c++] = t;  // not collected already
  }
  *count = c;
  return SGC_OK;
}

extern "C" uint32_t sgc_server_pending(const sgc_server* S) {
  return S ? (uint32_t)(S->pending.size() + S->inflight.size()) : 0;
}

extern "C" int sgc_server_result(sgc_server* S, uint64_t ticket, int32_t* labels, uint64_t capacity,
                                 sgc_query_stats* st) {
  clear_error();
  if (!S) return fail(SGC_EINVAL, "null server");
  auto it = S->done.find(ticket);
  if (it == S->done.end()) {
    if (S->inflight.count(ticket)) return SGC_EBUSY;
    for (const Query& q : S->pending)
      if (q.ticket == ticket) return SGC_EBUSY;
    return fail(SGC_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
  }
  const Result& r = it->second;
  if (labels) {
    if (capacity < r.labels.size()) return fail(SGC_EINVAL, "labels buffer holds %llu < %zu", (unsigned long long)capacity, r.labels.size());
    memcpy(labels, r.labels.data(), r.labels.size() * 4);
  }
  if (st) *st = r.st;
  const int status = r.st.status;
  S->done.erase(it);
  return status == SGC_OK ? SGC_OK : fail(status, "query on a removed graph");
}
