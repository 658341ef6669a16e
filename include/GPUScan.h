/*
 * GPUScan.h - C++ drop-in for the reference's entry point.
 *
 * Declares the same function, with the same C++ linkage and the same
 * argument types, as GPUSCAN++/GPUScan.h:7:
 *
 *     int* GPUScan(graph_t* g, int* Edge_u, int* Edge_v, double eps, int miu);
 *
 * and the graph_t layout of GPUSCAN++/Graph.h:13-20 (a typedef'd unnamed
 * struct, so the mangled symbol is GPUScan(graph_t*, int*, int*, double, int)
 * exactly as the reference's callers bind it).  libsgc.so exports that
 * symbol: the reference's main.cpp and Graph.cpp link against libsgc.so
 * unchanged, in place of GPUScan.cu.
 *
 * Contract (the reference's, GPUSCAN++/GPUScan.cu:24-32, :800-1009):
 *  - g->n, g->m, g->adj[m] and g->num_edges[n+1] (uint32 row offsets) are
 *    read; Edge_u, Edge_v and g->eid are accepted and not read (the device
 *    pass orients every edge itself);
 *  - the result is a malloc'd int[n] the caller frees: cluster id (minimum
 *    core id of the cluster), -1 hub, -2 outlier;
 *  - device 0, one pass, blocking;
 *  - on any error (invalid graph_t, CSR that violates the input contract,
 *    no device, allocation or HIP failure) it prints
 *    "GPUassert: <message> <file> <line>" on stderr and exits the process
 *    with a non-zero status (the SGC_* code of sgc.h).
 *
 * A caller that already has the reference's Graph.h defines SGC_HAVE_GRAPH_T
 * before including this header (or uses the reference's own GPUScan.h).
 */
#ifndef SGC_GPUSCAN_H
#define SGC_GPUSCAN_H

#ifndef __cplusplus
#error "GPUScan.h is the C++ drop-in; C callers use sgc_GPUScan from sgc.h"
#endif

#ifndef SGC_HAVE_GRAPH_T
#define SGC_HAVE_GRAPH_T
typedef struct {
  long n;
  long m;
  int* adj;
  unsigned int* num_edges;
  unsigned int* eid;
} graph_t;
#endif

int* GPUScan(graph_t* g, int* Edge_u, int* Edge_v, double eps, int miu);

#endif /* SGC_GPUSCAN_H */
