// C entry points over the reference's own host code (GPUSCAN++/Graph.cpp),
// compiled from its sources by oracle/Makefile into oracle/_ref/.  TEST
// INFRASTRUCTURE ONLY: used by tests/ to pin the loader and the edge
// orientation of libsgc.so against the reference itself.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "Graph.h"

extern "C" {

// Graph::Graph(dir) (GPUSCAN++/Graph.cpp:147-153), then copies the CSR out.
// The reference exits the process on invalid input, so callers run this in a
// child process when the input may be invalid.
int ref_load(const char* dir, uint64_t* n, uint64_t* m, uint32_t** row_ptr, int32_t** adj) {
  Graph g(const_cast<char*>(dir));
  *n = g.nodemax;
  *m = g.edgemax;
  *row_ptr = (uint32_t*)malloc(sizeof(uint32_t) * (g.nodemax + 1));
  *adj = (int32_t*)malloc(sizeof(int32_t) * (g.edgemax ? g.edgemax : 1));
  memcpy(*row_ptr, g.node_off, sizeof(uint32_t) * (g.nodemax + 1));
  memcpy(*adj, g.edge_dst, sizeof(int32_t) * g.edgemax);
  free(g.node_off);
  free(g.edge_dst);
  return 0;
}

// getEidAndEdgeList (GPUSCAN++/Graph.cpp:52-141) on a caller CSR.
int ref_orient(long n, long m, int32_t* adj, uint32_t* num_edges, int32_t* edge_u, int32_t* edge_v,
               uint32_t* eid) {
  graph_t g;
  g.n = n;
  g.m = m;
  g.adj = adj;
  g.num_edges = num_edges;
  g.eid = nullptr;
  getEidAndEdgeList(&g, edge_u, edge_v);
  memcpy(eid, g.eid, sizeof(uint32_t) * m);
  free(g.eid);
  return 0;
}

void ref_free(void* p) { free(p); }
}
