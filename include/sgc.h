/*
 * sgc.h - C ABI of the MI355X SCAN structural graph clusterer (libsgc.so).
 *
 * Every entry point returns an int status (SGC_OK on success) and never calls
 * exit(); the message for the last failure on the calling thread is returned
 * by sgc_last_error().  Pointers are plain host or device pointers; nothing in
 * this header depends on C++ or on any framework type.
 *
 * Label encoding is the reference's (GPUSCAN++/GPUScan.cu:124-127, :649,
 * :780): label >= 0 is the cluster id (the minimum core id of the cluster),
 * -1 a hub, -2 an outlier.
 */
#ifndef SGC_H
#define SGC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGC_OK 0
#define SGC_EINVAL 1  /* bad argument */
#define SGC_EIO 2     /* file missing, unreadable or truncated */
#define SGC_EFORMAT 3 /* input violates the CSR contract */
#define SGC_ENOMEM 4  /* host/device allocation failed or budget too small */
#define SGC_EDEVICE 5 /* HIP runtime error */
#define SGC_ERANGE 6  /* size beyond what this entry point supports */
#define SGC_EBUSY 7   /* a served query has not finished yet */

#define SGC_HUB (-1)
#define SGC_OUTLIER (-2)

/* Similarity flag values written per directed CSR entry. */
#define SGC_SIM_NO 0
#define SGC_SIM_YES 1

const char* sgc_last_error(void);
const char* sgc_version(void);

/* ------------------------------------------------------------------------ */
/* Host CSR graphs                                                           */
/* ------------------------------------------------------------------------ */

typedef struct sgc_host_graph sgc_host_graph;

/* Replaces Graph::Graph(char* dir) (GPUSCAN++/Graph.cpp:147-153):
 * ReadDegree (:157-174), ReadAdjacencyList (:176-223), CheckInputGraph
 * (:270-288).  Reads <dir>/b_degree.bin = {int32 int_size, uint32 n,
 * uint32 m, int32 deg[n]} and <dir>/b_adj.bin = int32 adj[sum(deg)].
 * 64-bit extension: when sum(deg) >= 2^32 the header field holds
 * sum(deg) mod 2^32.  Self loops and rows that are not strictly increasing
 * fail with SGC_EFORMAT and the reference's message; ids >= n, header/size
 * mismatches and (check_symmetry != 0) missing reverse entries also fail. */
int sgc_graph_load(const char* dir, int check_symmetry, sgc_host_graph** out);

/* Writes the two files read by sgc_graph_load (the format written by
 * csr_convert/EdgeListToBinsConverter.cpp:85-121). */
int sgc_graph_save(const sgc_host_graph* g, const char* dir);

/* Copies a caller CSR (row_ptr[n+1], adj[m]) and validates it as
 * sgc_graph_load does. */
int sgc_graph_from_csr(uint64_t n, uint64_t m, const uint64_t* row_ptr, const int32_t* adj,
                       int check_symmetry, sgc_host_graph** out);

/* Builds a simple undirected graph from an edge list: drops self loops,
 * symmetrises, removes duplicates, sorts rows. */
int sgc_graph_from_edges(uint64_t n, uint64_t num_edges, const int32_t* src, const int32_t* dst,
                         sgc_host_graph** out);

/* Replaces csr_convert (EdgeListToBinsConverter.cpp:123-165): text edge
 * list ("u v" per line, '#'/'%' comments) -> <out_dir>/b_degree.bin and
 * b_adj.bin.  Unlike the reference converter it symmetrises and
 * de-duplicates.  n = 1 + max id. */
int sgc_graph_convert_edge_list(const char* txt_path, const char* out_dir);

/* Seeded R-MAT: round(edge_factor * 2^scale) draws with quadrant
 * probabilities a, b, c (d = 1-a-b-c) from the Philox4x32-10 stream, then
 * self loops dropped, symmetrised, de-duplicated, rows sorted.  n = 2^scale. */
int sgc_graph_rmat(int scale, double edge_factor, double a, double b, double c, uint64_t seed,
                   sgc_host_graph** out);

/* Seeded planted-partition graph with hubs.  Community sizes: i.i.d. Zipf
 * draws floor(U^(-1/(zipf_exponent-1))) capped at n/20, scaled to sum to n
 * (largest remainder, at least one vertex each); communities occupy
 * consecutive id ranges.  Intra-community edge probability
 * min(p_in, max_intra_degree/(size-1)); `inter_per_vertex` uniform draws per
 * vertex into other communities; `hubs` distinct hub vertices each joined to
 * every vertex with probability hub_fraction. */
int sgc_graph_planted(uint64_t n, uint32_t communities, double zipf_exponent, double p_in,
                      double max_intra_degree, uint32_t inter_per_vertex, uint32_t hubs,
                      double hub_fraction, uint64_t seed, sgc_host_graph** out);

/* The counter-based stream behind both generators: Philox4x32-10 of
 * counter ctr[4] under key[2] (Salmon et al., SC'11), written to out[4]. */
void sgc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

uint64_t sgc_graph_n(const sgc_host_graph* g);
uint64_t sgc_graph_m(const sgc_host_graph* g);
const uint64_t* sgc_graph_row_ptr(const sgc_host_graph* g);
const int32_t* sgc_graph_adj(const sgc_host_graph* g);
uint32_t sgc_graph_max_degree(const sgc_host_graph* g);
/* FNV-1a 64 over n, m, row_ptr and adj (little-endian words). */
uint64_t sgc_graph_digest(const sgc_host_graph* g);
void sgc_graph_free(sgc_host_graph* g);
/* Moves g's adjacency into page-locked host memory (hipHostMalloc), so that
 * sgc_scan_ooc sends row blocks to the device straight from the graph
 * instead of through a staging copy.  No counterpart in the reference, whose
 * out-of-core tier copies from pageable memory (GPUSCAN++O/SCANoffload.cu:
 * 922-1464).  SGC_EDEVICE without a device or when the allocation fails; g
 * is unchanged then.  Calling it on a pinned graph does nothing. */
int sgc_graph_pin(sgc_host_graph* g);
int sgc_graph_pinned(const sgc_host_graph* g);

/* Replaces getEidAndEdgeList (GPUSCAN++/Graph.cpp:52-141): orients every
 * undirected edge from its lower (degree, id) endpoint to the other one.
 * edge_u/edge_v[m/2] list the edges ordered by edge_u, then by edge_v;
 * eid[m] maps each directed CSR entry to its undirected edge id.  Any output
 * pointer may be NULL.  Requires m < 2^32. */
int sgc_orient(const sgc_host_graph* g, int32_t* edge_u, int32_t* edge_v, uint32_t* eid);

/* Work model of the similarity phase at `eps` (host, OpenMP): the number of
 * edges the degree-only test leaves undecided and their algorithmic bytes
 * sum 4 * (s + min(l, s * ceil(log2(l + 1)))) over those edges (s <= l the
 * endpoint degrees): the short row streamed once, the long row streamed
 * once or probed once per binary-search level, whichever is smaller.
 * short_elements = sum of s; alg_bytes_large = the part of alg_bytes of
 * edges whose higher endpoint has degree > 1024 (the workgroup tier). */
int sgc_graph_similarity_work(const sgc_host_graph* g, double eps, uint64_t* u_edges, double* alg_bytes,
                              double* short_elements, double* alg_bytes_large);

/* ------------------------------------------------------------------------ */
/* Device context and device-resident graphs                                 */
/* ------------------------------------------------------------------------ */

typedef struct sgc_ctx sgc_ctx;

/* One context per (device, stream).  The context owns a HIP stream created
 * at the default priority and a grow-only workspace. */
int sgc_ctx_create(int device, sgc_ctx** out);
void sgc_ctx_destroy(sgc_ctx* ctx);
int sgc_ctx_sync(sgc_ctx* ctx);
/* Name of the device the context runs on (e.g. "AMD Instinct MI355X"). */
const char* sgc_ctx_device_name(const sgc_ctx* ctx);
/* Free and total device memory in bytes (hipMemGetInfo on the context's
 * device); either pointer may be NULL.  The reference's out-of-core run
 * sizes its chips from the free memory in the same way
 * (GPUSCAN++O/SCANoffload.cu:975-978). */
int sgc_ctx_memory(sgc_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* Common-neighbour kernels of the similarity phase.  TIERED (default): owner
 * rows staged in LDS hash tables, short rows streamed against them.
 * WAVE_PER_EDGE: one wavefront per undecided edge, binary search of the short
 * row's elements in the long row (GPUSCAN++/GPUScan.cu:174-222 restated for
 * wave64); kept as an independent second implementation.  Results are
 * identical. */
#define SGC_SIM_KERNEL_TIERED 0
#define SGC_SIM_KERNEL_WAVE_PER_EDGE 1
int sgc_ctx_set_similarity_kernel(sgc_ctx* ctx, int kind);

/* Core-check pruning for passes that return no similarity flags (default
 * off): the reference's skip of edges whose endpoints are both decided
 * (GPUSCAN++/GPUScan.cu:303), with the edges the cluster and border passes
 * need decided on demand (:557-585, :692-719).  Passes that return flags,
 * passes with mu < 1 and the one-wave-per-edge kernel decide every edge.
 * Results do not depend on it.  Off by default because on R-MAT it defers
 * too few edges to pay for its bookkeeping (about 4 % slower at scale 24). */
int sgc_ctx_set_pruning(sgc_ctx* ctx, int enable);

typedef struct sgc_dgraph sgc_dgraph;

/* Copies a host CSR into HBM (uint32 offsets; requires m < 2^32, else
 * SGC_ERANGE - use sgc_scan_ooc for larger graphs) and builds the device
 * indexes (entry rows, reverse entries, neighbour degrees, window offsets).
 * A graph not in page-locked memory (sgc_graph_pin) is staged through two
 * 64 MB page-locked buffers that the context allocates at its first such
 * upload and keeps until sgc_ctx_destroy.  Building the reverse-entry index
 * sorts the entries in about 20 bytes of temporary device memory per entry;
 * without that room, or for an input whose entries lack their reverse, the
 * index is built by a search instead.  Synchronises the context stream. */
int sgc_dgraph_upload(sgc_ctx* ctx, const sgc_host_graph* g, sgc_dgraph** out);
void sgc_dgraph_free(sgc_dgraph* dg);

/* Planner knobs of the tiered similarity kernels for one resident graph.  No
 * counterpart in the reference, whose launch shape is fixed (one warp per
 * edge, TOTALTHDCOUNT threads, GPUSCAN++/GPUScan.cu:276-367).  Results do not
 * depend on them, only the time. */
typedef struct {
  uint32_t small_max_degree; /* owners of degree <= this take the small tier (1..1024; default 1024) */
  double small_unit_work;    /* short-row elements per small-tier work unit (default 262144) */
  double large_unit_work;    /* per workgroup-tier work unit (default 8388608) */
  uint32_t hash_tables;      /* 1 (default): owner rows as cuckoo hashes where they fit (small tier; whole
                                rows and runs of 2^20-id windows of <= 8192 keys in the workgroup tier);
                                0: linear probing (small tier) and one bitmap window at a time only */
  uint32_t tiny_work;        /* owners whose edges' merge work, sum of d(a) + d(b), is at most this are
                                decided by one thread each, without a work unit (0..65536; default 2048;
                                0 turns this off) */
  double ms_similarity;      /* sgc_dgraph_autotune: similarity time with these knobs */
  double ms_default;         /* sgc_dgraph_autotune: similarity time with the defaults */
  uint32_t trials;           /* sgc_dgraph_autotune: passes timed */
} sgc_tuning;
int sgc_dgraph_get_tuning(const sgc_dgraph* dg, sgc_tuning* t);
int sgc_dgraph_set_tuning(sgc_dgraph* dg, const sgc_tuning* t);
/* Times the similarity phase of dg at eps under candidate knobs (coordinate
 * descent over large_unit_work, small_max_degree, small_unit_work,
 * hash_tables, tiny_work; best of two passes each), keeps the fastest on dg
 * and reports it. */
int sgc_dgraph_autotune(sgc_ctx* ctx, sgc_dgraph* dg, double eps, sgc_tuning* best);
uint64_t sgc_dgraph_n(const sgc_dgraph* dg);
uint64_t sgc_dgraph_m(const sgc_dgraph* dg);

/* ------------------------------------------------------------------------ */
/* SCAN                                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  /* Device time from resident CSR to resident labels, hipEvent-timed on the
   * context stream, and its phases. */
  double ms_total;
  double ms_prep;       /* orientation + degree-only pre-decision */
  double ms_similarity; /* common-neighbour kernels */
  double ms_core;       /* core check */
  double ms_cluster;    /* union-find + flatten */
  double ms_border;     /* border assignment */
  double ms_classify;   /* hub / outlier */
  double ms_similarity_large; /* the part of ms_similarity spent on owners of degree > 1024 */
  uint64_t edges;       /* E = m/2 */
  uint64_t u_edges;     /* edges the degree-only test could not decide */
  uint64_t similar_edges;
  uint64_t cores;
  uint64_t clusters;
  uint64_t borders;
  uint64_t hubs;
  uint64_t outliers;
  double alg_bytes;      /* algorithmic bytes of the similarity kernels */
  double probe_elements; /* list elements consumed by them */
  double loaded_elements; /* list elements loaded by them (consumed or read past a window's end) */
  double claims;          /* wave claims of the tiered kernels */
  /* Core-check pruning (GPUSCAN++/GPUScan.cu:303): in a pruned pass the
   * workgroup-tier similarity kernel defers an edge whose endpoints are both
   * already decided (core, or not core); after the union-find pass the
   * deferred edges the clusters or borders need are decided after all
   * (redo_edges), the rest are never decided (deferred_edges) and are not
   * counted in similar_edges.  Labels and core flags are the same either
   * way. */
  uint64_t deferred_edges;
  uint64_t redo_edges;
  uint32_t pruned;        /* 1 if the pass ran pruned */
} sgc_stats;

/* Replaces GPUScan(graph_t*, int*, int*, double, int) (GPUSCAN++/GPUScan.h:7,
 * GPUScan.cu:800-1009) for a device-resident graph.  Outputs are host
 * arrays, each optional (NULL = not copied back): labels[n] (int32),
 * core[n] (1 = core), sim[m] (SGC_SIM_YES/NO per directed CSR entry).
 * Threshold arithmetic is the reference's own FP64 expression
 * (GPUScan.cu:143-156, :220-221).  Borders take the minimum candidate
 * cluster id. */
int sgc_scan(sgc_ctx* ctx, sgc_dgraph* dg, double eps, int32_t mu, int32_t* labels, uint8_t* core,
             uint8_t* sim, sgc_stats* stats);

/* Same pass with device output pointers (each optional); nothing is copied
 * to the host and the call returns once the pass completed on the stream. */
int sgc_scan_device(sgc_ctx* ctx, sgc_dgraph* dg, double eps, int32_t mu, int32_t* d_labels,
                    uint8_t* d_core, uint8_t* d_sim, sgc_stats* stats);

/* Same pass as sgc_scan_device, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference, whose GPUScan allocates and frees all its
 * device memory per call (GPUSCAN++/GPUScan.cu:800-1009).  The pass's work
 * arrays live in a static device array of sgc_noalloc_max_bytes() bytes
 * (64 MiB) that each GPU holds from the library's load until the process
 * ends; the call makes no hipMalloc, hipFree or allocating library call.
 * The caller passes the resident graph (uploaded beforehand) and the
 * optional device outputs, as for sgc_scan_device.  SGC_ERANGE when the
 * pass needs more than the scratch holds (sgc_noalloc_bytes(dg) bytes: about
 * 25 B per entry and 33 B per vertex).  Every call on one GPU shares that
 * GPU's scratch, whatever the context: calls that overlap on one GPU race;
 * each pass clears what it reads before writing. */
int sgc_scan_device_noalloc(sgc_ctx* ctx, sgc_dgraph* dg, double eps, int32_t mu, int32_t* d_labels,
                            uint8_t* d_core, uint8_t* d_sim, sgc_stats* stats);
/* Bytes of the static scratch behind sgc_scan_device_noalloc. */
uint64_t sgc_noalloc_max_bytes(void);
/* Work bytes one pass over dg needs (sgc_scan_device_noalloc accepts dg when
 * this is at most sgc_noalloc_max_bytes()). */
uint64_t sgc_noalloc_bytes(const sgc_dgraph* dg);

/* Calibration aid with no counterpart in the reference: copies `bytes`
 * between two fresh device buffers with a plain streaming kernel (k_copy),
 * twice, and reports the second copy's time.  Its HBM byte counters
 * (2 x bytes moved) are the yardstick for the counters of the SCAN kernels. */
int sgc_copy_bench(sgc_ctx* ctx, uint64_t bytes, double* ms);

/* Out-of-core run details. */
typedef struct {
  uint64_t device_bytes;     /* device memory the run allocated (<= the budget) */
  uint32_t similarity_chips; /* chips of the similarity phase (a group of row blocks, alone or with one later block) */
  uint32_t row_chips;        /* row chips of the core / cluster / border / classify passes */
  double bytes_h2d;          /* host-to-device bytes */
  double bytes_d2h;          /* device-to-host bytes */
  double ms_host;            /* host time planning, assembling chips and merging flags */
  double ms_phases;          /* wall time of the passes after similarity */
  uint32_t row_blocks;       /* row blocks of the similarity phase */
  uint32_t blocks_per_group; /* consecutive row blocks per group (4 where the budget allows, else 2 or 1) */
} sgc_ooc_stats;

/* Replaces the out-of-core GPUScan of GPUSCAN++O (SCANoffload.cu:922-1464,
 * chips GPUSCAN++O/Graph.cpp:414-576) for a host-resident graph of any size
 * (m may exceed 2^32; n < 2^31): at most `device_budget` bytes of device
 * memory are allocated.  Rows are cut into blocks and consecutive blocks into
 * groups; the similarity phase runs over each group alone and over each group
 * with every later block (each undecided edge in the chip holding both of its
 * rows), the later phases over chips of rows, with the per-vertex state
 * resident throughout.  Outputs as sgc_scan (host arrays,
 * each optional) and bit-identical to it.  stats->ms_similarity fields are
 * device time summed over chips; the per-phase fields are filled when all
 * rows fit one chip; ms_total is wall time.  SGC_ENOMEM if the budget is
 * below sgc_ooc_min_budget. */
int sgc_scan_ooc(sgc_ctx* ctx, const sgc_host_graph* g, double eps, int32_t mu, uint64_t device_budget,
                 int32_t* labels, uint8_t* core, uint8_t* sim, sgc_stats* stats, sgc_ooc_stats* ooc);

/* Smallest device budget sgc_scan_ooc accepts for g: the per-vertex state
 * plus a chip of two row blocks that each hold the longest row (eps is
 * accepted for future plans and does not change the result today).  No
 * counterpart in the reference, whose chip planner takes a fixed budget and
 * fails inside. */
int sgc_ooc_min_budget(const sgc_host_graph* g, double eps, uint64_t* bytes);

/* ------------------------------------------------------------------------ */
/* Serving engine (no counterpart in the reference, which answers one query  */
/* per process, GPUSCAN++/main.cpp:20-111)                                   */
/* ------------------------------------------------------------------------ */

typedef struct sgc_server sgc_server;

typedef struct {
  int32_t status;        /* SGC_OK, or the error of a query whose graph was removed */
  uint32_t graph;
  uint64_t n;
  double ms_device;      /* device time of the query on its stream */
  double ms_latency;     /* submit -> retired by sgc_server_progress */
  uint64_t similar_edges;
  uint64_t cores;
  uint64_t clusters;
  uint64_t borders;
  uint64_t hubs;
  uint64_t outliers;
} sgc_query_stats;

/* An engine holding resident graphs within `hbm_budget` bytes of device
 * memory, each with `slots_per_graph` concurrent query slots (1..64). */
int sgc_server_create(sgc_ctx* ctx, uint64_t hbm_budget, uint32_t slots_per_graph, sgc_server** out);
void sgc_server_destroy(sgc_server* srv);
/* Device bytes a graph takes in an engine (CSR, cn[], query slots). */
uint64_t sgc_server_graph_bytes(const sgc_host_graph* g, uint32_t slots_per_graph);
uint64_t sgc_server_bytes_in_use(const sgc_server* srv);
/* Launch mode of the query slots: 1 (default) = each slot's work is captured
 * once into a hipGraph and replayed per query with eps/mu read from device
 * memory; 0 = eager kernel launches. */
int sgc_server_set_graph_launch(sgc_server* srv, int enable);
/* Batching: queries on graphs of at most `cutoff_entries` directed entries
 * are not replayed one per slot but gathered, oldest first, into batches of
 * up to `max_queries` (1..64), each run by one launch of every phase kernel
 * that walks a table of (graph, eps, mu) queries (a hipGraph captured once
 * per batch lane; two lanes run at once).  Results are the same as a slot's.
 * cutoff 0 (the default) turns batching off.  No counterpart in the
 * reference. */
int sgc_server_set_batching(sgc_server* srv, uint64_t cutoff_entries, uint32_t max_queries);
int sgc_server_get_batching(const sgc_server* srv, uint64_t* cutoff_entries, uint32_t* max_queries);
/* Picks the batching cut-off from measurement: for every admitted graph, the
 * wall time per query of one (eps, mu) query on each of its slots (up to the
 * batch size), replayed at once on the slots' streams, against one batch of
 * the same queries (the faster of two runs after a warm-up each); the
 * cut-off becomes the graph size (in entries) that minimises the summed time
 * over all graphs when the graphs up to it are batched and the others
 * replayed, 0 if batching none is best.  Needs an idle server (SGC_EBUSY
 * otherwise). */
int sgc_server_autotune_batching(sgc_server* srv, double eps, int32_t mu, uint64_t* cutoff_entries);
/* Admission: uploads g and computes every edge's common-neighbour count once
 * (precompute_ms: that pass).  SGC_ENOMEM if the budget cannot hold it; the
 * caller may remove graphs and retry. */
int sgc_server_add_graph(sgc_server* srv, const sgc_host_graph* g, uint32_t* graph_id, double* precompute_ms);
/* Waits for the graph's running queries; queued ones end with SGC_EINVAL
 * (sgc_server_result returns it; sgc_server_completed lists them). */
int sgc_server_remove_graph(sgc_server* srv, uint32_t graph_id);
/* Queues a query (eps, mu as sgc_scan) and starts it at once if one of the
 * graph's slots is free. */
int sgc_server_submit(sgc_server* srv, uint32_t graph_id, double eps, int32_t mu, uint64_t* ticket);
/* Retires finished queries and starts queued ones in the freed slots; with
 * wait, first blocks until the oldest running query finishes. */
int sgc_server_progress(sgc_server* srv, int wait, uint32_t* completed);
/* Up to `capacity` tickets retired since the last call (each listed once,
 * unless its result was already taken), into tickets; *count of them.  Lets
 * a caller with many open queries collect results without polling each. */
int sgc_server_completed(sgc_server* srv, uint64_t* tickets, uint32_t capacity, uint32_t* count);
/* Queries submitted and not yet retired. */
uint32_t sgc_server_pending(const sgc_server* srv);
/* A retired query's labels (n int32, as sgc_scan) and statistics; the result
 * is then released.  SGC_EBUSY while it is queued or running. */
int sgc_server_result(sgc_server* srv, uint64_t ticket, int32_t* labels, uint64_t capacity, sgc_query_stats* stats);

/* ------------------------------------------------------------------------ */
/* Result files                                                              */
/* ------------------------------------------------------------------------ */

#define SGC_WRITE_TEXT 1

/* Writes a run's per-vertex result into directory `dir` (created if needed):
 * labels.bin int32[n] (cluster id = min core id, -1 hub, -2 outlier),
 * core.bin uint8[n], role.bin uint8[n] (0 core, 1 border, 2 hub, 3 outlier),
 * with SGC_WRITE_TEXT also labels.txt ("vertex label role" per line), and
 * summary.txt: the reference's "core number: %d" and "Cluster Number: %d"
 * lines (GPUSCAN++/GPUScan.cu:914, main.cpp:111) followed by eps, mu, role
 * counts, optional pass statistics (st may be NULL) and an FNV-1a digest of
 * labels.bin.  The reference writes no per-vertex output. */
int sgc_write_result(const char* dir, uint64_t n, const int32_t* labels, const uint8_t* core, double eps, int32_t mu,
                     const sgc_stats* st, int flags);

/* Reference-layout graph (GPUSCAN++/Graph.h:13-20). */
typedef struct {
  long n;
  long m;
  int32_t* adj;
  uint32_t* num_edges;
  uint32_t* eid;
} sgc_graph_t;

/* C-linkage twin of the drop-in int* GPUScan(graph_t* g, int* Edge_u,
 * int* Edge_v, double eps, int miu) (GPUSCAN++/GPUScan.h:7), which
 * libsgc.so also exports with the reference's C++ linkage (GPUScan.h in this
 * include folder).  Both keep the reference's contract and are the only entry
 * points that end the process: they return a malloc'd int32[n] the caller
 * frees, and on any error (null or out-of-range graph_t, null num_edges or
 * adj, a CSR violating the input contract, no device, allocation or HIP
 * failure) print "GPUassert: <message> <file> <line>" on stderr and exit
 * with the SGC_* status as the exit code (GPUSCAN++/GPUScan.cu:24-32).
 * Edge_u/Edge_v/eid are accepted for signature compatibility and are not
 * read: the device pass orients edges itself. */
int32_t* sgc_GPUScan(sgc_graph_t* g, int32_t* Edge_u, int32_t* Edge_v, double eps, int32_t miu);

#ifdef __cplusplus
}
#endif

#endif /* SGC_H */
