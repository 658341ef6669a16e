// Drives the host-side C++ of libsgc (graph_host.cpp, ooc_host.cpp,
// result_io.cpp) and the oracle (oracle/scan_oracle.c) under a CPU sanitizer
// build: ThreadSanitizer with LLVM's OpenMP runtime and its Archer tool
// (which tells the sanitizer about OpenMP's own synchronisation), or
// AddressSanitizer + UndefinedBehaviorSanitizer.  Built and run by
// tests/test_sanitizers.py; the GPU is never touched (the three HIP host
// calls graph_host.cpp can make are stubbed below and fail as on a machine
// without a device).
//
//   host_driver run <scratch dir>   every host path once; exit 0 when all
//                                   results agree
//   host_driver race                a loop with a planted data race, to show
//                                   that the thread sanitizer is live
//   host_driver overflow            a planted heap read out of bounds, for
//                                   the address sanitizer
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "graph_host.h"
#include "ooc_host.h"
#include "scan_oracle.h"
#include "sgc.h"

// ---- the HIP host calls of graph_host.cpp (sgc_graph_pin's allocation)
hipError_t hipHostMalloc(void** p, size_t, unsigned int) {
  *p = nullptr;
  return hipErrorNoDevice;
}
hipError_t hipHostFree(void*) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static void write_bytes(const std::string& path, const void* p, size_t n) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) {
    ++failures;
    return;
  }
  if (n) fwrite(p, 1, n, f);
  fclose(f);
}

// The two input files for a graph of row lengths deg[] and entries adj[],
// with the header's n and m as given (a malformed case may lie in them).
static void write_graph_files(const std::string& dir, uint32_t hn, uint32_t hm, const std::vector<int32_t>& deg,
                              const std::vector<int32_t>& adj) {
  std::vector<uint8_t> d(12 + 4 * deg.size());
  const int32_t sz = 4;
  memcpy(d.data(), &sz, 4);
  memcpy(d.data() + 4, &hn, 4);
  memcpy(d.data() + 8, &hm, 4);
  if (!deg.empty()) memcpy(d.data() + 12, deg.data(), 4 * deg.size());
  write_bytes(dir + "/b_degree.bin", d.data(), d.size());
  write_bytes(dir + "/b_adj.bin", adj.data(), 4 * adj.size());
}

static void check_graph(const sgc_host_graph* g, const std::string& dir) {
  const uint64_t n = sgc_graph_n(g), m = sgc_graph_m(g);
  const uint64_t* rp = sgc_graph_row_ptr(g);
  const int32_t* adj = sgc_graph_adj(g);

  // validation of a copy, with the symmetry check
  sgc_host_graph* c = nullptr;
  CHECK(sgc_graph_from_csr(n, m, rp, adj, 1, &c) == SGC_OK);
  if (c) CHECK(sgc_graph_digest(c) == sgc_graph_digest(g));
  sgc_graph_free(c);

  // orientation: library against the oracle's restatement
  std::vector<int32_t> eu(m / 2), ev(m / 2), ou(m / 2), ov(m / 2);
  std::vector<uint32_t> eid(m), oid(m);
  CHECK(sgc_orient(g, eu.data(), ev.data(), eid.data()) == SGC_OK);
  orc_orient(n, rp, adj, ou.data(), ov.data(), oid.data());
  CHECK(eu == ou && ev == ov && eid == oid);

  // the oracle's whole pass, its phases after similarity, prep counts and
  // the threaded per-edge similarity, all against each other
  for (double eps : {0.2, 0.5}) {
    for (int32_t mu : {2, 4}) {
      std::vector<int32_t> lab(n), lab2(n);
      std::vector<uint8_t> core(n), core2(n), sim(m);
      orc_stats st{}, st2{};
      CHECK(orc_scan(n, rp, adj, eps, mu, lab.data(), core.data(), sim.data(), &st) == 0);
      CHECK(orc_labels_from_sim(n, rp, adj, sim.data(), mu, lab2.data(), core2.data(), &st2) == 0);
      CHECK(lab == lab2 && core == core2 && st.clusters == st2.clusters && st.hubs == st2.hubs);
      uint64_t pn = 0, py = 0, pu = 0;
      orc_prep_counts(n, rp, adj, eps, &pn, &py, &pu);
      CHECK(pn == st.prep_n && py == st.prep_y && pu == st.prep_u);
      std::vector<uint64_t> entry;
      for (uint64_t x = 0; x < n; ++x)
        for (uint64_t k = rp[x]; k < rp[x + 1]; ++k)
          if ((uint64_t)adj[k] > x) entry.push_back(k);
      std::vector<uint8_t> out(entry.size());
      const uint64_t similar = orc_similarity_entries(n, rp, adj, eps, entry.data(), entry.size(), out.data(), 4);
      CHECK(similar == st.similar_edges);
      for (size_t i = 0; i < entry.size(); ++i) CHECK(out[i] == sim[entry[i]]);

      // the per-vertex result files
      CHECK(sgc_write_result((dir + "/result").c_str(), n, lab.data(), core.data(), eps, mu, nullptr,
                             SGC_WRITE_TEXT) == SGC_OK);
    }
  }

  // out-of-core host work: degrees, row blocks, block copies, flag merges
  std::vector<uint32_t> deg(n);
  sgc::ooc_degrees(g, deg.data());
  for (uint64_t x = 0; x < n; ++x) CHECK(deg[x] == rp[x + 1] - rp[x]);
  std::vector<uint64_t> bounds;
  uint64_t most = 0, rows = 0;
  const uint64_t row_cap = n / 7 + 1;
  CHECK(sgc::ooc_plan_rows(g, m / 5 + 1, row_cap, &bounds, &most, &rows) == SGC_OK);
  CHECK(!bounds.empty() && bounds.front() == 0 && bounds.back() == n);
  CHECK(rows <= row_cap);
  for (size_t b = 0; b + 1 < bounds.size(); ++b) CHECK(bounds[b + 1] - bounds[b] <= rows);
  std::vector<uint8_t> flag(m, 0xFF), chip;
  for (size_t b = 0; b + 1 < bounds.size(); ++b) {
    const uint64_t r0 = bounds[b], r1 = bounds[b + 1], e0 = rp[r0], me = rp[r1] - e0;
    CHECK(me <= most);
    std::vector<int32_t> rows(me);
    std::vector<uint32_t> off(r1 - r0 + 1);
    sgc::ooc_copy_rows(g, r0, r1, 7, rows.data(), off.data());
    CHECK(me == 0 || memcmp(rows.data(), adj + e0, 4 * me) == 0);
    CHECK(off[0] == 7 && off[r1 - r0] == 7 + me);
    chip.assign(me, 0);
    for (uint64_t k = 0; k < me; ++k) chip[k] = (uint8_t)(k % 4);  // N, Y and two undecided values
    sgc::ooc_merge_rows(g, r0, r1, chip.data(), flag.data());
    for (uint64_t k = 0; k < me; ++k) CHECK(flag[e0 + k] == (k % 4 <= 1 ? k % 4 : 0xFF));
  }

  // save and load back
  CHECK(sgc_graph_save(g, dir.c_str()) == SGC_OK);
  sgc_host_graph* back = nullptr;
  CHECK(sgc_graph_load(dir.c_str(), 1, &back) == SGC_OK);
  if (back) CHECK(sgc_graph_digest(back) == sgc_graph_digest(g));
  sgc_graph_free(back);
}

static void malformed_inputs(const std::string& dir) {
  // each must fail with a status, never crash or read out of bounds
  const std::vector<int32_t> deg = {2, 2, 2};
  const std::vector<int32_t> ok = {1, 2, 0, 2, 0, 1};
  sgc_host_graph* g = nullptr;
  write_graph_files(dir, 3, 6, deg, ok);
  CHECK(sgc_graph_load(dir.c_str(), 1, &g) == SGC_OK);
  sgc_graph_free(g);
  struct Case {
    uint32_t hn, hm;
    std::vector<int32_t> deg, adj;
  };
  const Case bad[] = {
      {3, 7, deg, ok},                                  // header m is not the degree sum
      {3, 6, deg, {1, 2, 0, 2, 0}},                     // adjacency file one entry short
      {3, 6, deg, {1, 2, 0, 2, 0, 1, 2}},               // one entry too many
      {3, 6, deg, {1, 3, 0, 2, 0, 1}},                  // id >= n
      {3, 6, deg, {0, 2, 0, 2, 0, 1}},                  // self loop
      {3, 6, deg, {2, 1, 0, 2, 0, 1}},                  // row not increasing
      {3, 6, {2, -1, 5}, ok},                           // negative degree
      {4, 6, deg, ok},                                  // header n beyond the degree list
      {3, 6, deg, {1, 2, 0, 2, 1, 0}},                  // row of vertex 2 decreasing
  };
  for (const Case& c : bad) {
    write_graph_files(dir, c.hn, c.hm, c.deg, c.adj);
    g = nullptr;
    const int rc = sgc_graph_load(dir.c_str(), 1, &g);
    if (rc == SGC_OK) fprintf(stderr, "malformed case %d accepted\n", (int)(&c - bad));
    CHECK(rc != SGC_OK);
    CHECK(g == nullptr);
    sgc_graph_free(g);
  }
  // entry (0, 1) without (1, 0): accepted without the symmetry check,
  // refused with it
  write_graph_files(dir, 3, 3, {2, 0, 1}, {1, 2, 0});
  g = nullptr;
  CHECK(sgc_graph_load(dir.c_str(), 0, &g) == SGC_OK);
  sgc_graph_free(g);
  g = nullptr;
  CHECK(sgc_graph_load(dir.c_str(), 1, &g) != SGC_OK);
  // a truncated degree file
  write_bytes(dir + "/b_degree.bin", "\x04\x00\x00\x00\x03", 5);
  CHECK(sgc_graph_load(dir.c_str(), 0, &g) != SGC_OK);
  // missing files
  CHECK(sgc_graph_load((dir + "/absent").c_str(), 0, &g) != SGC_OK);
  // an edge-list text file, with comments, duplicates and a self loop
  const char* txt = "# comment\n% comment\n0 1\n1 0\n1 2\n2 2\n3 1\n0 1\n";
  write_bytes(dir + "/edges.txt", txt, strlen(txt));
  CHECK(sgc_graph_convert_edge_list((dir + "/edges.txt").c_str(), dir.c_str()) == SGC_OK);
  CHECK(sgc_graph_load(dir.c_str(), 1, &g) == SGC_OK);
  if (g) CHECK(sgc_graph_n(g) == 4 && sgc_graph_m(g) == 6);
  sgc_graph_free(g);
}

static int run(const std::string& dir) {
  sgc_host_graph* g = nullptr;
  CHECK(sgc_graph_rmat(11, 8.0, 0.57, 0.19, 0.19, 3, &g) == SGC_OK);
  if (g) check_graph(g, dir);
  sgc_graph_free(g);
  g = nullptr;
  CHECK(sgc_graph_planted(3000, 40, 1.5, 0.05, 60.0, 4, 2, 0.05, 5, &g) == SGC_OK);
  if (g) check_graph(g, dir);
  sgc_graph_free(g);
  const int32_t es[] = {0, 1, 2, 2, 5}, ed[] = {1, 2, 0, 2, 4};
  g = nullptr;
  CHECK(sgc_graph_from_edges(6, 5, es, ed, &g) == SGC_OK);
  if (g) check_graph(g, dir);
  // without a device, page-locking fails and leaves the graph as it was
  if (g) CHECK(sgc_graph_pin(g) == SGC_EDEVICE && !sgc_graph_pinned(g) && sgc_graph_adj(g) != nullptr);
  sgc_graph_free(g);
  CHECK(sgc_graph_pin(nullptr) != SGC_OK);
  malformed_inputs(dir);
  printf("host driver: %d failed checks\n", failures);
  return failures ? 1 : 0;
}

// A shared counter bumped from every thread without synchronisation.
static int race() {
  long total = 0;
#pragma omp parallel for
  for (int i = 0; i < 100000; ++i) total += i;
  printf("racy total %ld\n", total);
  return 0;
}

// A read one element past a heap array.
static int overflow() {
  int32_t* a = new int32_t[4]();
  volatile int i = 4;
  const int v = a[i];
  delete[] a;
  printf("read %d\n", v);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && strcmp(argv[1], "race") == 0) return race();
  if (argc >= 2 && strcmp(argv[1], "overflow") == 0) return overflow();
  if (argc >= 3 && strcmp(argv[1], "run") == 0) return run(argv[2]);
  fprintf(stderr, "usage: %s run <dir> | race | overflow\n", argv[0]);
  return 2;
}
