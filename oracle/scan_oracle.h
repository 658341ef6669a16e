/*
 * scan_oracle.h - CPU restatement of the GPUSCAN++ SCAN semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * it.  It is never linked into libsgc.so and never used as a fallback.
 */
#ifndef SCAN_ORACLE_H
#define SCAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_N = 0, ORC_Y = 1, ORC_U = 2 };

typedef struct {
  uint64_t edges;
  uint64_t prep_n, prep_y, prep_u;  /* degree-only pre-decision */
  uint64_t check_core_computed;     /* U edges computed by the core phase */
  uint64_t similar_edges;           /* after completing every U edge */
  uint64_t cores_after_prep;
  uint64_t cores;                   /* "core number" line */
  uint64_t clusters;                /* "Cluster Number" line */
  uint64_t borders, hubs, outliers;
} orc_stats;

/* Degree-only pre-decision of edge (a, b) with d(a) <= d(b):
 * GPUSCAN++/GPUScan.cu:143-169. */
int orc_prep(uint32_t deg_a, uint32_t deg_b, double eps);

/* Live similarity predicate of GPUSCAN++/GPUScan.cu:220-221 for cn common
 * neighbours, first-list degree s and second-list degree l: 1 = similar. */
int orc_predicate(uint32_t cn, uint32_t s, uint32_t l, double eps);

/* |N(u) ∩ N(v)| over sorted open rows. */
uint32_t orc_common(const uint64_t* row_ptr, const int32_t* adj, uint32_t u, uint32_t v);

/* getEidAndEdgeList restated (GPUSCAN++/Graph.cpp:52-141). */
void orc_orient(uint64_t n, const uint64_t* row_ptr, const int32_t* adj, int32_t* edge_u,
                int32_t* edge_v, uint32_t* eid);

/* Whole GPUScan() pass (GPUSCAN++/GPUScan.cu:800-1009) executed as one
 * sequential interleaving of its kernels, with the canonical border rule
 * (minimum candidate cluster id).  Outputs: labels[n]; core[n] (optional);
 * sim[m] (optional, every edge completed, 1 = similar, per directed CSR
 * entry).  Returns 0, or -1 if m >= 2^32. */
int orc_scan(uint64_t n, const uint64_t* row_ptr, const int32_t* adj, double eps, int32_t mu,
             int32_t* labels, uint8_t* core, uint8_t* sim, orc_stats* stats);

/* Similarity of `count` undirected edges given by one directed entry each
 * (entry[i] indexes adj), OpenMP-parallel; out[i] = 1 if similar.  This is
 * the CPU baseline's timed loop.  Returns the number of similar edges. */
uint64_t orc_similarity_entries(uint64_t n, const uint64_t* row_ptr, const int32_t* adj,
                                double eps, const uint64_t* entry, uint64_t count, uint8_t* out,
                                int threads);

/* The phases after the similarity phase (GPUSCAN++/GPUScan.cu:276-798)
 * given the final similarity of every directed entry (sim[m], 1 = similar,
 * symmetric): core flags (the fixed point of the sd/ed state machine,
 * :316-363), clusters by union-find over similar core-core edges labelled
 * by minimum core id (:465-652), borders by the canonical minimum candidate
 * (:654-730), hubs and outliers (:732-798).  Same outputs as orc_scan when
 * sim is orc_scan's own.  For checking a device pass at sizes where
 * orc_scan's sequential similarity is too slow (any m; 64-bit offsets).
 * stats: cores, clusters, borders, hubs, outliers, similar_edges.  Returns
 * 0, or -1 if n >= 2^31. */
int orc_labels_from_sim(uint64_t n, const uint64_t* row_ptr, const int32_t* adj, const uint8_t* sim,
                        int32_t mu, int32_t* labels, uint8_t* core, orc_stats* stats);

/* Degree-only pre-decision counts over every undirected edge
 * (GPUSCAN++/GPUScan.cu:133-172), OpenMP-parallel: edges decided N, Y and
 * left undecided (U). */
void orc_prep_counts(uint64_t n, const uint64_t* row_ptr, const int32_t* adj, double eps, uint64_t* prep_n,
                     uint64_t* prep_y, uint64_t* prep_u);

#ifdef __cplusplus
}
#endif

#endif
