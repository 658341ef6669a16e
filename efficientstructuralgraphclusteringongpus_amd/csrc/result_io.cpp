// Per-vertex result files and the run summary.  The reference prints only
// counts (GPUSCAN++/main.cpp:83-111, GPUScan.cu:914); these files make runs
// diffable vertex by vertex.
#include <errno.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <string>
#include <vector>

#include "common.h"
#include "sgc.h"

using namespace sgc;

namespace {

int write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return fail(SGC_EIO, "cannot create %s: %s", path.c_str(), strerror(errno));
  const size_t w = bytes ? fwrite(data, 1, bytes, f) : 0;
  const int c = fclose(f);
  if (w != bytes || c != 0) return fail(SGC_EIO, "cannot write %s", path.c_str());
  return SGC_OK;
}

uint64_t fnv1a(const void* p, size_t bytes) {
  const unsigned char* b = (const unsigned char*)p;
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

const char* kRoleName[] = {"core", "border", "hub", "outlier"};

}  // namespace

extern "C" int sgc_write_result(const char* dir, uint64_t n, const int32_t* labels, const uint8_t* core, double eps,
                                int32_t mu, const sgc_stats* st, int flags) {
  clear_error();
  if (!dir || (n && (!labels || !core))) return fail(SGC_EINVAL, "null argument");
  if (mkdir(dir, 0755) != 0 && errno != EEXIST) return fail(SGC_EIO, "cannot create %s: %s", dir, strerror(errno));
  const std::string d(dir);
  std::vector<uint8_t> role(n);
  uint64_t counts[4] = {0, 0, 0, 0};
  std::vector<char> seen(n, 0);
  uint64_t clusters = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const int32_t l = labels[i];
    uint8_t r;
    if (core[i])
      r = 0;
    else if (l >= 0)
      r = 1;
    else if (l == SGC_HUB)
      r = 2;
    else if (l == SGC_OUTLIER)
      r = 3;
    else
      return fail(SGC_EINVAL, "vertex %llu has label %d", (unsigned long long)i, l);
    if (l >= (int64_t)n) return fail(SGC_EINVAL, "vertex %llu has label %d >= n", (unsigned long long)i, l);
    if (l >= 0 && !seen[l]) {
      seen[l] = 1;
      ++clusters;
    }
    role[i] = r;
    ++counts[r];
  }
  int rc = write_file(d + "/labels.bin", labels, n * 4);
  if (!rc) rc = write_file(d + "/core.bin", core, n);
  if (!rc) rc = write_file(d + "/role.bin", role.data(), n);
  if (rc) return rc;
  if (flags & SGC_WRITE_TEXT) {
    const std::string p = d + "/labels.txt";
    FILE* f = fopen(p.c_str(), "w");
    if (!f) return fail(SGC_EIO, "cannot create %s: %s", p.c_str(), strerror(errno));
    bool ok = fprintf(f, "# vertex label role\n") > 0;
    for (uint64_t i = 0; i < n && ok; ++i)
      ok = fprintf(f, "%llu %d %s\n", (unsigned long long)i, labels[i], kRoleName[role[i]]) > 0;
    if (fclose(f) != 0 || !ok) return fail(SGC_EIO, "cannot write %s", p.c_str());
  }
  const std::string p = d + "/summary.txt";
  FILE* f = fopen(p.c_str(), "w");
  if (!f) return fail(SGC_EIO, "cannot create %s: %s", p.c_str(), strerror(errno));
  // the reference's count lines first, then the rest as "key: value"
  fprintf(f, "core number: %llu\n", (unsigned long long)counts[0]);
  fprintf(f, "Cluster Number: %llu\n", (unsigned long long)clusters);
  fprintf(f, "eps: %.17g\nmu: %d\nvertices: %llu\n", eps, mu, (unsigned long long)n);
  fprintf(f, "borders: %llu\nhubs: %llu\noutliers: %llu\n", (unsigned long long)counts[1],
          (unsigned long long)counts[2], (unsigned long long)counts[3]);
  if (st) {
    fprintf(f, "edges: %llu\nundecided edges: %llu\nsimilar edges: %llu\n", (unsigned long long)st->edges,
            (unsigned long long)st->u_edges, (unsigned long long)st->similar_edges);
    fprintf(f, "device pass ms: %.3f\nsimilarity ms: %.3f\n", st->ms_total, st->ms_similarity);
  }
  fprintf(f, "labels fnv1a: %016llx\n", (unsigned long long)fnv1a(labels, n * 4));
  if (fclose(f) != 0) return fail(SGC_EIO, "cannot write %s", p.c_str());
  return SGC_OK;
}
