// GPUScan <folder> <eps> <miu> [options]
//
// Drop-in for the reference command line (GPUSCAN++/main.cpp:20-123): same
// positional arguments (eps by atof, miu by atoi, :26-27), same usage error,
// same stdout/stderr lines for the parity-relevant outputs ("max degree",
// "node number and edge number", "core number", "Cluster Number").
// Options beyond the reference:
//   --out <file>      write labels as int32[n] followed by core flags uint8[n]
//   --out-dir <dir>   write labels.bin, core.bin, role.bin and summary.txt
//   --text            with --out-dir, also labels.txt (one vertex per line)
//   --device <id>     HIP device (default 0)
//   --check-symmetry  also verify that every (u,v) has (v,u)
//   --device-budget <GiB>  run out of core within that much device memory
// A graph that the device cannot hold (m >= 2^32 entries, or too little
// memory for the resident pass) runs out of core within the free device
// memory less a reserve, as the reference's out-of-core command line does
// with the same arguments (GPUSCAN++O/main.cpp, SCANoffload.cu:975-978).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "sgc.h"

using namespace std::chrono;

static double seconds_since(high_resolution_clock::time_point t0) {
  return duration_cast<milliseconds>(high_resolution_clock::now() - t0).count() / 1000.0;
}

int main(int argc, char* argv[]) {
  if (argc < 4) {
    fprintf(stderr, "%s <Graph file> , parameter eps and miu\n", argv[0]);
    return 1;
  }
  const double eps = atof(argv[2]);
  const int miu = atoi(argv[3]);
  const char* out_path = nullptr;
  const char* out_dir = nullptr;
  int text = 0;
  int device = 0;
  int check_symmetry = 0;
  double budget_gib = 0;
  for (int i = 4; i < argc; ++i) {
    if (!strcmp(argv[i], "--out") && i + 1 < argc)
      out_path = argv[++i];
    else if (!strcmp(argv[i], "--out-dir") && i + 1 < argc)
      out_dir = argv[++i];
    else if (!strcmp(argv[i], "--text"))
      text = 1;
    else if (!strcmp(argv[i], "--device") && i + 1 < argc)
      device = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--check-symmetry"))
      check_symmetry = 1;
    else if (!strcmp(argv[i], "--device-budget") && i + 1 < argc) {
      budget_gib = atof(argv[++i]);
      if (!(budget_gib > 0)) {
        fprintf(stderr, "--device-budget needs a positive size in GiB\n");
        return 1;
      }
    } else {
      fprintf(stderr, "unknown option %s\n", argv[i]);
      return 1;
    }
  }

  auto t_load = high_resolution_clock::now();
  sgc_host_graph* g = nullptr;
  if (sgc_graph_load(argv[1], check_symmetry, &g) != SGC_OK) {
    fprintf(stderr, "%s\n", sgc_last_error());
    return 1;
  }
  fprintf(stderr, "read graph time: %.3lf s\n", seconds_since(t_load));
  const uint64_t n = sgc_graph_n(g), m = sgc_graph_m(g);
  printf("max degree: %u\n", sgc_graph_max_degree(g));
  fflush(stdout);

  auto t_scan = high_resolution_clock::now();
  fprintf(stderr, "node number and edge number: %ld, %ld\n", (long)n, (long)m);
  fprintf(stderr, "Start GPUScan\n");
  sgc_ctx* ctx = nullptr;
  sgc_dgraph* dg = nullptr;
  if (sgc_ctx_create(device, &ctx) != SGC_OK) {
    fprintf(stderr, "GPUassert: %s\n", sgc_last_error());
    return 1;
  }
  std::vector<int32_t> labels(n ? n : 1);
  std::vector<uint8_t> core(n ? n : 1);
  sgc_stats st;
  uint64_t budget = (uint64_t)(budget_gib * (double)(1ull << 30));
  if (!budget) {
    auto t_mem = high_resolution_clock::now();
    const int rc = sgc_dgraph_upload(ctx, g, &dg);
    if (rc == SGC_ERANGE || rc == SGC_ENOMEM) {
      // too large for the resident pass: out of core within the free memory
      uint64_t free_bytes = 0;
      if (sgc_ctx_memory(ctx, &free_bytes, nullptr) != SGC_OK) {
        fprintf(stderr, "GPUassert: %s\n", sgc_last_error());
        return 1;
      }
      const uint64_t reserve = 256ull << 20;
      budget = free_bytes > 2 * reserve ? free_bytes - reserve : free_bytes / 2;
      fprintf(stderr, "graph does not fit the resident pass (%s); out of core\n", sgc_last_error());
    } else if (rc != SGC_OK) {
      fprintf(stderr, "GPUassert: %s\n", sgc_last_error());
      return 1;
    } else {
      fprintf(stderr, "Create GPU Memory time: %.3lf s\n", seconds_since(t_mem));
    }
  }
  if (budget) {
    sgc_ooc_stats oi;
    fprintf(stderr, "out of core within %.2f GiB of device memory\n", (double)budget / (double)(1ull << 30));
    if (sgc_scan_ooc(ctx, g, eps, miu, budget, labels.data(), core.data(), nullptr, &st, &oi) != SGC_OK) {
      fprintf(stderr, "GPUassert: %s\n", sgc_last_error());
      return 1;
    }
    fprintf(stderr,
            "similarity chips: %u (%u row blocks in groups of %u), row chips: %u, host to device %.3lf GB, back "
            "%.3lf GB\n",
            oi.similarity_chips, oi.row_blocks, oi.blocks_per_group, oi.row_chips, oi.bytes_h2d / 1e9,
            oi.bytes_d2h / 1e9);
  } else if (sgc_scan(ctx, dg, eps, miu, labels.data(), core.data(), nullptr, &st) != SGC_OK) {
    fprintf(stderr, "GPUassert: %s\n", sgc_last_error());
    return 1;
  }
  fprintf(stderr, "\nPrep time: %.3lf s\n", st.ms_prep / 1000.0);
  fprintf(stderr, "\nCheck Core time: %.3lf s\n", (st.ms_similarity + st.ms_core) / 1000.0);
  fprintf(stderr, "core number: %llu\n", (unsigned long long)st.cores);
  fprintf(stderr, "\nFirst Link Phase time: %.3lf s\n", st.ms_cluster / 1000.0);
  fprintf(stderr, "Cluster non-core time: %.3lf s\n", st.ms_border / 1000.0);
  fprintf(stderr, "\nClassify Others time: %.3lf s\n", st.ms_classify / 1000.0);
  fprintf(stderr, "device pass time: %.3lf ms (similarity %.3lf ms, %llu undecided edges)\n", st.ms_total,
          st.ms_similarity, (unsigned long long)st.u_edges);
  fprintf(stderr, "Finish GPUScan\n");
  fprintf(stderr, "gpuscan time: %.3lf s\n", seconds_since(t_scan));

  // Distinct labels >= 0 (GPUSCAN++/main.cpp:83-111).
  std::vector<char> seen(n ? n : 1, 0);
  long clusters = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (labels[i] >= 0 && !seen[labels[i]]) {
      seen[labels[i]] = 1;
      ++clusters;
    }
  fprintf(stderr, "Cluster Number: %ld\n", clusters);

  if (out_path) {
    FILE* f = fopen(out_path, "wb");
    if (!f || fwrite(labels.data(), 4, n, f) != n || fwrite(core.data(), 1, n, f) != n) {
      fprintf(stderr, "cannot write %s\n", out_path);
      return 1;
    }
    fclose(f);
  }
  if (out_dir && sgc_write_result(out_dir, n, labels.data(), core.data(), eps, miu, &st, text ? SGC_WRITE_TEXT : 0)) {
    fprintf(stderr, "%s\n", sgc_last_error());
    return 1;
  }
  sgc_dgraph_free(dg);
  sgc_ctx_destroy(ctx);
  sgc_graph_free(g);
  fprintf(stderr, "Work Complete!\n");
  return 0;
}
