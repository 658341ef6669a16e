// A caller of the C++ drop-in written the way GPUSCAN++/main.cpp:20-123 calls
// the reference: graph_t filled from a loaded CSR, Edge_u/Edge_v built by
// the orientation step, then int* GPUScan(graph_t*, int*, int*, double, int)
// through its C++ symbol, the "Cluster Number" count of main.cpp:83-111 and
// free() of the result.  Test program: tests/test_compat_entry.py runs it.
//
//   gpuscan_caller <folder> <eps> <miu> [--labels <file>] [--bad <kind>]
//
// --labels writes the int32[n] result; --bad corrupts graph_t before the
// call (kinds: n, null-offsets, offsets, self-loop) to check that the drop-in
// prints a GPUassert line and exits non-zero.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "GPUScan.h"
#include "sgc.h"

int main(int argc, char* argv[]) {
  if (argc < 4) {
    fprintf(stderr, "%s <Graph file> , parameter eps and miu\n", argv[0]);
    return 1;
  }
  const double eps = atof(argv[2]);
  const int miu = atoi(argv[3]);
  const char* labels_path = nullptr;
  const char* bad = nullptr;
  for (int i = 4; i + 1 < argc; i += 2) {
    if (!strcmp(argv[i], "--labels")) labels_path = argv[i + 1];
    if (!strcmp(argv[i], "--bad")) bad = argv[i + 1];
  }
  sgc_host_graph* hg = nullptr;
  if (sgc_graph_load(argv[1], 0, &hg) != SGC_OK) {
    fprintf(stderr, "load failed: %s\n", sgc_last_error());
    return 3;
  }
  const uint64_t n = sgc_graph_n(hg), m = sgc_graph_m(hg);
  graph_t g;
  g.n = (long)n;
  g.m = (long)m;
  g.adj = (int*)malloc(sizeof(int) * (m ? m : 1));
  g.num_edges = (unsigned*)malloc(sizeof(unsigned) * (n + 1));
  g.eid = (unsigned*)malloc(sizeof(unsigned) * (m ? m : 1));
  memcpy(g.adj, sgc_graph_adj(hg), sizeof(int) * m);
  for (uint64_t u = 0; u <= n; ++u) g.num_edges[u] = (unsigned)sgc_graph_row_ptr(hg)[u];
  int max_degree = 0;
  for (uint64_t i = 0; i < n; ++i)
    if ((int)(g.num_edges[i + 1] - g.num_edges[i]) > max_degree) max_degree = g.num_edges[i + 1] - g.num_edges[i];
  printf("max degree: %d\n", max_degree);
  int* Edge_u = (int*)malloc(sizeof(int) * (m / 2 + 1));
  int* Edge_v = (int*)malloc(sizeof(int) * (m / 2 + 1));
  if (sgc_orient(hg, Edge_u, Edge_v, g.eid) != SGC_OK) {
    fprintf(stderr, "orientation failed: %s\n", sgc_last_error());
    return 3;
  }
  sgc_graph_free(hg);
  fprintf(stderr, "node number and edge number: %ld, %ld\n", g.n, g.m);
  unsigned* saved_offsets = g.num_edges;
  if (bad) {
    if (!strcmp(bad, "n")) g.n = -1;
    if (!strcmp(bad, "null-offsets")) g.num_edges = nullptr;
    if (!strcmp(bad, "offsets")) g.num_edges[n] += 1;
    if (!strcmp(bad, "self-loop") && m) g.adj[0] = (int)0;  // row 0's first entry becomes 0 itself (or a duplicate)
  }
  fprintf(stderr, "Start GPUScan\n");
  int* cluster = GPUScan(&g, Edge_u, Edge_v, eps, miu);
  fprintf(stderr, "Finish GPUScan\n");
  std::vector<char> seen(n, 0);
  int sum = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (cluster[i] >= 0 && !seen[cluster[i]]) seen[cluster[i]] = 1, ++sum;
  fprintf(stderr, "Cluster Number: %d\n", sum);
  if (labels_path) {
    FILE* f = fopen(labels_path, "wb");
    if (!f || fwrite(cluster, sizeof(int), n, f) != n) return 4;
    fclose(f);
  }
  free(cluster);
  free(Edge_u);
  free(Edge_v);
  free(g.adj);
  free(saved_offsets);
  free(g.eid);
  fprintf(stderr, "Work Complete!\n");
  return 0;
}
