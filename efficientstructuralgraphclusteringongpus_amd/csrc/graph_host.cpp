// Host graph tier: b_degree.bin/b_adj.bin reader and writer, validation,
// edge-list conversion, seeded generators, edge orientation.
#include "graph_host.h"

#include <errno.h>
#include <omp.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "philox.h"

namespace sgc {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

void progress(const char* fmt, ...) {
  static const bool on = getenv("SGC_PROGRESS") != nullptr;
  if (!on) return;
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "[sgc %.1f s] ", now_ms() / 1e3);
  vfprintf(stderr, fmt, ap);
  fputc('\n', stderr);
  fflush(stderr);
  va_end(ap);
}

void* host_alloc(size_t bytes, bool pinned) {
  if (bytes == 0) bytes = 8;
  if (pinned) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
  }
  return malloc(bytes);
}

void host_free(void* p, bool pinned) {
  if (!p) return;
  if (pinned)
    (void)hipHostFree(p);
  else
    free(p);
}

sgc_host_graph* new_host_graph(uint64_t n, uint64_t m, bool pinned) {
  sgc_host_graph* g = new (std::nothrow) sgc_host_graph();
  if (!g) return nullptr;
  g->n = n;
  g->m = m;
  g->pinned = pinned;
  g->row_ptr = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
  g->adj = (int32_t*)host_alloc(m * sizeof(int32_t), pinned);
  if (!g->row_ptr || !g->adj) {
    free(g->row_ptr);
    host_free(g->adj, pinned);
    delete g;
    return nullptr;
  }
  return g;
}

static void free_graph(sgc_host_graph* g) {
  if (!g) return;
  free(g->row_ptr);
  host_free(g->adj, g->pinned);
  delete g;
}

// Lower bound of x in the sorted range a[lo, hi).
static inline uint64_t lower_bound_u64(const int32_t* a, uint64_t lo, uint64_t hi, int32_t x) {
  while (lo < hi) {
    uint64_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

void narrow_row_ptr(const sgc_host_graph* g, uint32_t* out, uint64_t* max_degree) {
  const uint64_t* rp = g->row_ptr;
  const int64_t n = (int64_t)g->n;
  uint64_t maxd = 0;
#pragma omp parallel for schedule(static) reduction(max : maxd)
  for (int64_t u = 0; u <= n; ++u) {
    out[u] = (uint32_t)rp[u];
    if (u < n) maxd = std::max<uint64_t>(maxd, rp[u + 1] - rp[u]);
  }
  *max_degree = maxd;
}

void parallel_copy(void* dst, const void* src, size_t bytes) {
  const size_t piece = 1 << 20;
  const int64_t pieces = (int64_t)((bytes + piece - 1) / piece);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < pieces; ++p) {
    const size_t off = (size_t)p * piece;
    memcpy((char*)dst + off, (const char*)src + off, std::min(piece, bytes - off));
  }
}

int validate_graph(const sgc_host_graph* g, bool check_symmetry) {
  const uint64_t n = g->n;
  if (g->row_ptr[0] != 0 || g->row_ptr[n] != g->m)
    return fail(SGC_EFORMAT, "row offsets do not span the adjacency array (row_ptr[0]=%llu, row_ptr[n]=%llu, m=%llu)",
                (unsigned long long)g->row_ptr[0], (unsigned long long)g->row_ptr[n],
                (unsigned long long)g->m);
  // 0 ok, 1 self loop, 2 unsorted, 3 id out of range, 4 asymmetric, 5 bad offsets.
  // `err` is read by every thread to stop early and set once by the first
  // failing row: an atomic, so the early-exit reads race with nothing; the
  // (code, row) pair itself is stored under the critical section.
  std::atomic<int> err{0};
  int err_code = 0;
  uint64_t err_u = 0;
  auto report = [&](int code, uint64_t u) {
#pragma omp critical(sgc_validate)
    {
      if (!err_code) {
        err_code = code;
        err_u = u;
        err.store(code, std::memory_order_relaxed);
      }
    }
  };
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t u = 0; u < n; ++u) {
    if (err.load(std::memory_order_relaxed)) continue;
    uint64_t b = g->row_ptr[u], e = g->row_ptr[u + 1];
    if (e < b || e > g->m) {
      report(5, u);
      continue;
    }
    for (uint64_t k = b; k < e; ++k) {
      int32_t v = g->adj[k];
      int code = 0;
      if (v < 0 || (uint64_t)v >= n)
        code = 3;
      else if ((uint64_t)v == u)
        code = 1;
      else if (k > b && v <= g->adj[k - 1])
        code = 2;
      else if (check_symmetry) {
        uint64_t vb = g->row_ptr[v], ve = g->row_ptr[v + 1];
        uint64_t p = lower_bound_u64(g->adj, vb, ve, (int32_t)u);
        if (p == ve || g->adj[p] != (int32_t)u) code = 4;
      }
      if (code) {
        report(code, u);
        break;
      }
    }
  }
  switch (err_code) {
    case 1: return fail(SGC_EFORMAT, "Self loop (vertex %llu)", (unsigned long long)err_u);
    case 2: return fail(SGC_EFORMAT, "Edges not sorted in increasing id order! (vertex %llu)", (unsigned long long)err_u);
    case 3: return fail(SGC_EFORMAT, "neighbour id out of range [0, n) in row %llu", (unsigned long long)err_u);
    case 4: return fail(SGC_EFORMAT, "adjacency is not symmetric (row %llu)", (unsigned long long)err_u);
    case 5: return fail(SGC_EFORMAT, "row offsets decrease or exceed m at vertex %llu", (unsigned long long)err_u);
  }
  return SGC_OK;
}

int build_from_pairs(uint64_t n, uint64_t count, const uint32_t* src, const uint32_t* dst,
                     sgc_host_graph** out, uint32_t** release) {
  *out = nullptr;
  if (n == 0 || n > (1ull << 31)) return fail(SGC_EINVAL, "vertex count %llu outside [1, 2^31]", (unsigned long long)n);
  std::vector<uint64_t> cnt(n + 1, 0);
  std::atomic<int> bad{0};
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < count; ++i) {
    uint32_t u = src[i], v = dst[i];
    if (u >= n || v >= n) {
      bad.store(1, std::memory_order_relaxed);
      continue;
    }
    if (u == v) continue;
    __atomic_fetch_add(&cnt[u], 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&cnt[v], 1, __ATOMIC_RELAXED);
  }
  if (bad.load()) return fail(SGC_EINVAL, "edge endpoint >= n");
  progress("graph build: degrees counted (%llu pairs)", (unsigned long long)count);
  std::vector<uint64_t> off(n + 1);
  off[0] = 0;
  for (uint64_t u = 0; u < n; ++u) off[u + 1] = off[u] + cnt[u];
  const uint64_t total = off[n];
  uint32_t* buf = (uint32_t*)malloc(std::max<uint64_t>(total, 1) * sizeof(uint32_t));
  if (!buf) return fail(SGC_ENOMEM, "host allocation of %llu entries failed", (unsigned long long)total);
  std::vector<uint64_t>& fill = cnt;  // reuse as scatter cursors
#pragma omp parallel for schedule(static)
  for (uint64_t u = 0; u < n; ++u) fill[u] = off[u];
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < count; ++i) {
    uint32_t u = src[i], v = dst[i];
    if (u == v) continue;
    buf[__atomic_fetch_add(&fill[u], 1, __ATOMIC_RELAXED)] = v;
    buf[__atomic_fetch_add(&fill[v], 1, __ATOMIC_RELAXED)] = u;
  }
  progress("graph build: entries scattered (%llu)", (unsigned long long)total);
  // the pairs are no longer read: a caller that handed them over gets them
  // freed before the final adjacency is allocated (peak host memory: pairs +
  // scatter buffer, then scatter buffer + adjacency)
  if (release) {
    free(release[0]);
    free(release[1]);
    release[0] = release[1] = nullptr;
  }
  std::vector<uint64_t> deg(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t u = 0; u < n; ++u) {
    uint32_t* b = buf + off[u];
    uint32_t* e = buf + off[u + 1];
    std::sort(b, e);
    deg[u] = (uint64_t)(std::unique(b, e) - b);
  }
  uint64_t m = 0;
  for (uint64_t u = 0; u < n; ++u) m += deg[u];
  progress("graph build: rows sorted, m = %llu", (unsigned long long)m);
  sgc_host_graph* g = new_host_graph(n, m, false);
  if (!g) {
    free(buf);
    return fail(SGC_ENOMEM, "host allocation of graph with m=%llu failed", (unsigned long long)m);
  }
  g->row_ptr[0] = 0;
  for (uint64_t u = 0; u < n; ++u) g->row_ptr[u + 1] = g->row_ptr[u] + deg[u];
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t u = 0; u < n; ++u) {
    const uint32_t* b = buf + off[u];
    int32_t* o = g->adj + g->row_ptr[u];
    for (uint64_t k = 0; k < deg[u]; ++k) o[k] = (int32_t)b[k];
  }
  free(buf);
  *out = g;
  return SGC_OK;
}

}  // namespace sgc

using namespace sgc;

// ---------------------------------------------------------------------------
// File format
// ---------------------------------------------------------------------------

static int read_exact(FILE* f, void* dst, size_t bytes) {
  char* p = (char*)dst;
  while (bytes) {
    size_t chunk = bytes > (1u << 30) ? (1u << 30) : bytes;
    size_t got = fread(p, 1, chunk, f);
    if (got != chunk) return -1;
    p += got;
    bytes -= got;
  }
  return 0;
}

static int write_exact(FILE* f, const void* src, size_t bytes) {
  const char* p = (const char*)src;
  while (bytes) {
    size_t chunk = bytes > (1u << 30) ? (1u << 30) : bytes;
    if (fwrite(p, 1, chunk, f) != chunk) return -1;
    p += chunk;
    bytes -= chunk;
  }
  return 0;
}

extern "C" int sgc_graph_load(const char* dir, int check_symmetry, sgc_host_graph** out) {
  clear_error();
  if (!dir || !out) return fail(SGC_EINVAL, "null argument");
  *out = nullptr;
  std::string dpath = std::string(dir) + "/b_degree.bin";
  std::string apath = std::string(dir) + "/b_adj.bin";
  FILE* fd = fopen(dpath.c_str(), "rb");
  if (!fd) return fail(SGC_EIO, "cannot open %s: %s", dpath.c_str(), strerror(errno));
  int32_t int_size = 0;
  uint32_t n32 = 0, m32 = 0;
  if (read_exact(fd, &int_size, 4) || read_exact(fd, &n32, 4) || read_exact(fd, &m32, 4)) {
    fclose(fd);
    return fail(SGC_EIO, "%s: truncated header", dpath.c_str());
  }
  if (int_size != 4) {
    fclose(fd);
    return fail(SGC_EFORMAT, "%s: int_size field is %d, expected 4", dpath.c_str(), int_size);
  }
  const uint64_t n = n32;
  std::vector<int32_t> deg(n);
  if (n && read_exact(fd, deg.data(), n * 4)) {
    fclose(fd);
    return fail(SGC_EIO, "%s: truncated degree array (n=%llu)", dpath.c_str(), (unsigned long long)n);
  }
  fclose(fd);
  uint64_t m = 0;
  for (uint64_t u = 0; u < n; ++u) {
    if (deg[u] < 0) return fail(SGC_EFORMAT, "%s: negative degree at vertex %llu", dpath.c_str(), (unsigned long long)u);
    m += (uint64_t)deg[u];
  }
  if ((uint32_t)m != m32)
    return fail(SGC_EFORMAT, "%s: header m=%u does not match the degree sum %llu", dpath.c_str(), m32,
                (unsigned long long)m);
  struct stat st;
  if (stat(apath.c_str(), &st) != 0) return fail(SGC_EIO, "cannot open %s: %s", apath.c_str(), strerror(errno));
  // exactly the degree sum: shorter is a truncated file, longer a malformed one
  if ((uint64_t)st.st_size != m * 4)
    return fail((uint64_t)st.st_size < m * 4 ? SGC_EIO : SGC_EFORMAT,
                "%s: %llu bytes, expected %llu (4 per entry of the degree sum)", apath.c_str(),
                (unsigned long long)st.st_size, (unsigned long long)(m * 4));
  sgc_host_graph* g = new_host_graph(n, m, false);
  if (!g) return fail(SGC_ENOMEM, "host allocation failed (n=%llu, m=%llu)", (unsigned long long)n, (unsigned long long)m);
  g->row_ptr[0] = 0;
  for (uint64_t u = 0; u < n; ++u) g->row_ptr[u + 1] = g->row_ptr[u] + (uint64_t)deg[u];
  FILE* fa = fopen(apath.c_str(), "rb");
  if (!fa) {
    free_graph(g);
    return fail(SGC_EIO, "cannot open %s: %s", apath.c_str(), strerror(errno));
  }
  if (m && read_exact(fa, g->adj, m * 4)) {
    fclose(fa);
    free_graph(g);
    return fail(SGC_EIO, "%s: short read", apath.c_str());
  }
  fclose(fa);
  int rc = validate_graph(g, check_symmetry != 0);
  if (rc) {
    free_graph(g);
    return rc;
  }
  *out = g;
  return SGC_OK;
}

extern "C" int sgc_graph_save(const sgc_host_graph* g, const char* dir) {
  clear_error();
  if (!g || !dir) return fail(SGC_EINVAL, "null argument");
  if (g->n > 0xFFFFFFFFull) return fail(SGC_ERANGE, "n=%llu does not fit the uint32 header", (unsigned long long)g->n);
  std::string dpath = std::string(dir) + "/b_degree.bin";
  std::string apath = std::string(dir) + "/b_adj.bin";
  FILE* fd = fopen(dpath.c_str(), "wb");
  if (!fd) return fail(SGC_EIO, "cannot create %s: %s", dpath.c_str(), strerror(errno));
  int32_t int_size = 4;
  uint32_t n32 = (uint32_t)g->n, m32 = (uint32_t)g->m;  // m mod 2^32 (64-bit extension)
  std::vector<int32_t> deg(g->n);
  for (uint64_t u = 0; u < g->n; ++u) {
    uint64_t d = g->row_ptr[u + 1] - g->row_ptr[u];
    if (d > 0x7FFFFFFFull) {
      fclose(fd);
      return fail(SGC_ERANGE, "degree of vertex %llu exceeds int32", (unsigned long long)u);
    }
    deg[u] = (int32_t)d;
  }
  int bad = write_exact(fd, &int_size, 4) || write_exact(fd, &n32, 4) || write_exact(fd, &m32, 4) ||
            write_exact(fd, deg.data(), g->n * 4);
  bad |= fclose(fd) != 0;
  if (bad) return fail(SGC_EIO, "write to %s failed", dpath.c_str());
  FILE* fa = fopen(apath.c_str(), "wb");
  if (!fa) return fail(SGC_EIO, "cannot create %s: %s", apath.c_str(), strerror(errno));
  bad = write_exact(fa, g->adj, g->m * 4);
  bad |= fclose(fa) != 0;
  if (bad) return fail(SGC_EIO, "write to %s failed", apath.c_str());
  return SGC_OK;
}

extern "C" int sgc_graph_from_csr(uint64_t n, uint64_t m, const uint64_t* row_ptr, const int32_t* adj,
                                  int check_symmetry, sgc_host_graph** out) {
  clear_error();
  if (!out || !row_ptr || (m && !adj)) return fail(SGC_EINVAL, "null argument");
  *out = nullptr;
  if (n > 0xFFFFFFFFull) return fail(SGC_ERANGE, "n=%llu exceeds uint32", (unsigned long long)n);
  sgc_host_graph* g = new_host_graph(n, m, false);
  if (!g) return fail(SGC_ENOMEM, "host allocation failed");
  memcpy(g->row_ptr, row_ptr, (n + 1) * sizeof(uint64_t));
  if (m) memcpy(g->adj, adj, m * sizeof(int32_t));
  int rc = validate_graph(g, check_symmetry != 0);
  if (rc) {
    free_graph(g);
    return rc;
  }
  *out = g;
  return SGC_OK;
}

extern "C" int sgc_graph_from_edges(uint64_t n, uint64_t num_edges, const int32_t* src, const int32_t* dst,
                                    sgc_host_graph** out) {
  clear_error();
  if (!out || (num_edges && (!src || !dst))) return fail(SGC_EINVAL, "null argument");
  for (uint64_t i = 0; i < num_edges; ++i)
    if (src[i] < 0 || dst[i] < 0) return fail(SGC_EINVAL, "negative vertex id in edge %llu", (unsigned long long)i);
  return build_from_pairs(n, num_edges, (const uint32_t*)src, (const uint32_t*)dst, out);
}

extern "C" int sgc_graph_convert_edge_list(const char* txt_path, const char* out_dir) {
  clear_error();
  if (!txt_path || !out_dir) return fail(SGC_EINVAL, "null argument");
  FILE* f = fopen(txt_path, "r");
  if (!f) return fail(SGC_EIO, "cannot open %s: %s", txt_path, strerror(errno));
  std::vector<uint32_t> src, dst;
  char line[512];
  uint64_t lineno = 0;
  uint64_t maxid = 0;
  bool any = false;
  while (fgets(line, sizeof(line), f)) {
    ++lineno;
    char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '#' || *p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
    char* end = nullptr;
    errno = 0;
    unsigned long long u = strtoull(p, &end, 10);
    if (end == p || errno) {
      fclose(f);
      return fail(SGC_EFORMAT, "%s:%llu: cannot parse edge", txt_path, (unsigned long long)lineno);
    }
    p = end;
    unsigned long long v = strtoull(p, &end, 10);
    if (end == p || errno) {
      fclose(f);
      return fail(SGC_EFORMAT, "%s:%llu: cannot parse edge", txt_path, (unsigned long long)lineno);
    }
    if (u >= 0x7FFFFFFFull || v >= 0x7FFFFFFFull) {
      fclose(f);
      return fail(SGC_ERANGE, "%s:%llu: vertex id exceeds int32", txt_path, (unsigned long long)lineno);
    }
    src.push_back((uint32_t)u);
    dst.push_back((uint32_t)v);
    maxid = std::max<uint64_t>(maxid, std::max(u, v));
    any = true;
  }
  fclose(f);
  sgc_host_graph* g = nullptr;
  int rc = build_from_pairs(any ? maxid + 1 : 1, src.size(), src.data(), dst.data(), &g);
  if (rc) return rc;
  rc = sgc_graph_save(g, out_dir);
  free_graph(g);
  return rc;
}

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------

extern "C" int sgc_graph_rmat(int scale, double edge_factor, double a, double b, double c, uint64_t seed,
                              sgc_host_graph** out) {
  clear_error();
  if (!out) return fail(SGC_EINVAL, "null argument");
  *out = nullptr;
  if (scale < 1 || scale > 30) return fail(SGC_EINVAL, "scale %d outside [1, 30]", scale);
  if (!(edge_factor > 0) || !(a >= 0) || !(b >= 0) || !(c >= 0) || a + b + c > 1.0)
    return fail(SGC_EINVAL, "bad R-MAT parameters");
  const uint64_t n = 1ull << scale;
  const uint64_t draws = (uint64_t)std::llround(edge_factor * (double)n);
  RmatThresholds t;
  t.ta = rmat_prob_to_u32(a);
  t.tab = rmat_prob_to_u32(a + b);
  t.tabc = rmat_prob_to_u32(a + b + c);
  uint32_t* src = (uint32_t*)malloc(std::max<uint64_t>(draws, 1) * 4);
  uint32_t* dst = (uint32_t*)malloc(std::max<uint64_t>(draws, 1) * 4);
  if (!src || !dst) {
    free(src);
    free(dst);
    return fail(SGC_ENOMEM, "host allocation of %llu draws failed", (unsigned long long)draws);
  }
#pragma omp parallel for schedule(static, 65536)
  for (uint64_t i = 0; i < draws; ++i) rmat_draw(i, scale, seed, t, &src[i], &dst[i]);
  progress("rmat scale %d: %llu draws", scale, (unsigned long long)draws);
  uint32_t* pairs[2] = {src, dst};
  int rc = build_from_pairs(n, draws, src, dst, out, pairs);
  free(pairs[0]);  // null once build_from_pairs has released them
  free(pairs[1]);
  return rc;
}

// Uniform double in [0, 1) from the top 53 bits of two Philox words.
static inline double u01(const Philox4x32& r, int w) {
  uint64_t x = ((uint64_t)r.v[w] << 32) | r.v[(w + 1) & 3];
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

extern "C" int sgc_graph_planted(uint64_t n, uint32_t communities, double zipf_exponent, double p_in,
                                 double max_intra_degree, uint32_t inter_per_vertex, uint32_t hubs,
                                 double hub_fraction, uint64_t seed, sgc_host_graph** out) {
  clear_error();
  if (!out) return fail(SGC_EINVAL, "null argument");
  *out = nullptr;
  if (n < 2 || n > (1ull << 31) || communities == 0 || communities > n || !(zipf_exponent > 1.0) ||
      !(p_in >= 0 && p_in <= 1) || !(hub_fraction >= 0 && hub_fraction <= 1) || hubs > n)
    return fail(SGC_EINVAL, "bad planted-partition parameters");
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  // Community sizes: raw Zipf draws floor(U^(-1/(s-1))) capped at n/20, then
  // scaled to sum to n by largest remainder with a floor of one vertex.
  std::vector<double> raw(communities);
  const double cap = std::max(1.0, (double)n / 20.0);
  double rsum = 0;
  for (uint32_t i = 0; i < communities; ++i) {
    Philox4x32 r = philox4x32_10(i, 0, 0x5A1F0001u, 0, k0, k1);
    double u = 1.0 - u01(r, 0);  // (0, 1]
    double x = std::floor(std::pow(u, -1.0 / (zipf_exponent - 1.0)));
    raw[i] = std::min(std::max(x, 1.0), cap);
    rsum += raw[i];
  }
  std::vector<uint64_t> size(communities);
  std::vector<std::pair<double, uint32_t>> rem(communities);
  uint64_t assigned = 0;
  const uint64_t free_v = n - communities;  // one vertex reserved per community
  for (uint32_t i = 0; i < communities; ++i) {
    double share = raw[i] / rsum * (double)free_v;
    uint64_t fl = (uint64_t)std::floor(share);
    size[i] = 1 + fl;
    assigned += size[i];
    rem[i] = {share - (double)fl, i};
  }
  std::sort(rem.begin(), rem.end(), [](const std::pair<double, uint32_t>& x, const std::pair<double, uint32_t>& y) {
    return x.first > y.first || (x.first == y.first && x.second < y.second);
  });
  for (uint64_t j = 0; assigned < n; ++j, ++assigned) size[rem[j % communities].second] += 1;
  std::vector<uint64_t> cstart(communities + 1, 0);
  for (uint32_t i = 0; i < communities; ++i) cstart[i + 1] = cstart[i] + size[i];
  std::vector<uint32_t> comm_of(n);
  for (uint32_t i = 0; i < communities; ++i)
    for (uint64_t v = cstart[i]; v < cstart[i + 1]; ++v) comm_of[v] = i;

  // Per-community intra edges by geometric skipping over the upper-triangle
  // pair index (Batagelj & Brandes, 2005).
  std::vector<std::vector<uint32_t>> ts(omp_get_max_threads()), td(omp_get_max_threads());
#pragma omp parallel for schedule(dynamic, 16)
  for (uint32_t ci = 0; ci < communities; ++ci) {
    std::vector<uint32_t>& S = ts[omp_get_thread_num()];
    std::vector<uint32_t>& D = td[omp_get_thread_num()];
    const uint64_t k = size[ci];
    if (k < 2) continue;
    const double p = std::min(p_in, max_intra_degree / (double)(k - 1));
    if (!(p > 0)) continue;
    const uint64_t pairs = k * (k - 1) / 2;
    const double lq = std::log1p(-std::min(p, 1.0 - 1e-15));
    uint64_t idx = 0;
    uint64_t ctr = 0;
    for (;;) {
      Philox4x32 r = philox4x32_10((uint32_t)ctr, (uint32_t)(ctr >> 32), ci, 0x5A1F0002u, k0, k1);
      ++ctr;
      double u = 1.0 - u01(r, 0);
      double skip = p >= 1.0 ? 0.0 : std::floor(std::log(u) / lq);
      if (skip > (double)(pairs - idx)) break;
      idx += (uint64_t)skip;
      if (idx >= pairs) break;
      // Decode idx -> (i, j), i < j, row-major over the upper triangle.
      double kk = (double)k;
      uint64_t i = (uint64_t)std::floor((2 * kk - 1 - std::sqrt((2 * kk - 1) * (2 * kk - 1) - 8.0 * (double)idx)) / 2);
      auto row_start = [k](uint64_t row) { return row * (2 * k - row - 1) / 2; };
      while (i > 0 && row_start(i) > idx) --i;
      while (i + 1 < k && row_start(i + 1) <= idx) ++i;
      uint64_t j = i + 1 + (idx - row_start(i));
      S.push_back((uint32_t)(cstart[ci] + i));
      D.push_back((uint32_t)(cstart[ci] + j));
      ++idx;
    }
  }
  // Inter-community draws: uniform partner outside the vertex's community.
#pragma omp parallel for schedule(static, 4096)
  for (uint64_t v = 0; v < n; ++v) {
    std::vector<uint32_t>& S = ts[omp_get_thread_num()];
    std::vector<uint32_t>& D = td[omp_get_thread_num()];
    for (uint32_t j = 0; j < inter_per_vertex; ++j) {
      for (uint32_t attempt = 0; attempt < 64; ++attempt) {
        Philox4x32 r = philox4x32_10((uint32_t)v, (uint32_t)(v >> 32), j, 0x5A1F0003u + (attempt << 8), k0, k1);
        uint64_t w = (((uint64_t)r.v[0] << 32) | r.v[1]) % n;
        if (comm_of[w] != comm_of[v]) {
          S.push_back((uint32_t)v);
          D.push_back((uint32_t)w);
          break;
        }
      }
    }
  }
  // Hubs: distinct ids drawn by rejection, each joined to every other vertex
  // with probability hub_fraction (geometric skipping).
  std::vector<uint32_t> hub_ids;
  {
    std::vector<char> taken(n, 0);
    uint64_t ctr = 0;
    while (hub_ids.size() < hubs) {
      Philox4x32 r = philox4x32_10((uint32_t)ctr, (uint32_t)(ctr >> 32), 0, 0x5A1F0004u, k0, k1);
      ++ctr;
      uint64_t w = (((uint64_t)r.v[0] << 32) | r.v[1]) % n;
      if (!taken[w]) {
        taken[w] = 1;
        hub_ids.push_back((uint32_t)w);
      }
    }
  }
  if (hub_fraction > 0) {
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t h = 0; h < hubs; ++h) {
      std::vector<uint32_t>& S = ts[omp_get_thread_num()];
      std::vector<uint32_t>& D = td[omp_get_thread_num()];
      const double lq = std::log1p(-std::min(hub_fraction, 1.0 - 1e-15));
      uint64_t w = 0, ctr = 0;
      for (;;) {
        Philox4x32 r = philox4x32_10((uint32_t)ctr, (uint32_t)(ctr >> 32), h, 0x5A1F0005u, k0, k1);
        ++ctr;
        double u = 1.0 - u01(r, 0);
        double skip = hub_fraction >= 1.0 ? 0.0 : std::floor(std::log(u) / lq);
        if (skip >= (double)(n - w)) break;
        w += (uint64_t)skip;
        if (w >= n) break;
        S.push_back(hub_ids[h]);
        D.push_back((uint32_t)w);
        ++w;
      }
    }
  }
  // Thread buffers are concatenated in thread order; the CSR build sorts and
  // de-duplicates every row, so the result does not depend on scheduling.
  uint64_t total = 0;
  for (auto& s : ts) total += s.size();
  std::vector<uint32_t> src(total), dst(total);
  uint64_t pos = 0;
  for (size_t t = 0; t < ts.size(); ++t) {
    std::copy(ts[t].begin(), ts[t].end(), src.begin() + pos);
    std::copy(td[t].begin(), td[t].end(), dst.begin() + pos);
    pos += ts[t].size();
    std::vector<uint32_t>().swap(ts[t]);
    std::vector<uint32_t>().swap(td[t]);
  }
  return build_from_pairs(n, total, src.data(), dst.data(), out);
}

// ---------------------------------------------------------------------------
// Accessors
// ---------------------------------------------------------------------------

extern "C" uint64_t sgc_graph_n(const sgc_host_graph* g) { return g ? g->n : 0; }
extern "C" uint64_t sgc_graph_m(const sgc_host_graph* g) { return g ? g->m : 0; }
extern "C" const uint64_t* sgc_graph_row_ptr(const sgc_host_graph* g) { return g ? g->row_ptr : nullptr; }
extern "C" const int32_t* sgc_graph_adj(const sgc_host_graph* g) { return g ? g->adj : nullptr; }
extern "C" void sgc_graph_free(sgc_host_graph* g) { free_graph(g); }

extern "C" int sgc_graph_pin(sgc_host_graph* g) {
  clear_error();
  if (!g) return fail(SGC_EINVAL, "null graph");
  if (g->pinned) return SGC_OK;
  int32_t* p = (int32_t*)host_alloc(g->m * sizeof(int32_t), true);
  if (!p) {
    (void)hipGetLastError();
    return fail(SGC_EDEVICE, "page-locked host allocation of %llu bytes failed",
                (unsigned long long)(g->m * sizeof(int32_t)));
  }
  const uint64_t piece = 1 << 20;
#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t q = 0; q < g->m; q += piece)
    memcpy(p + q, g->adj + q, (size_t)std::min(piece, g->m - q) * sizeof(int32_t));
  host_free(g->adj, false);
  g->adj = p;
  g->pinned = true;
  return SGC_OK;
}

extern "C" int sgc_graph_pinned(const sgc_host_graph* g) { return g && g->pinned ? 1 : 0; }

extern "C" uint32_t sgc_graph_max_degree(const sgc_host_graph* g) {
  if (!g) return 0;
  uint64_t best = 0;
#pragma omp parallel for reduction(max : best)
  for (uint64_t u = 0; u < g->n; ++u) best = std::max<uint64_t>(best, g->row_ptr[u + 1] - g->row_ptr[u]);
  return (uint32_t)best;
}

extern "C" uint64_t sgc_graph_digest(const sgc_host_graph* g) {
  if (!g) return 0;
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](uint32_t w) {
    h ^= w;
    h *= 1099511628211ull;
  };
  mix((uint32_t)g->n);
  mix((uint32_t)(g->n >> 32));
  mix((uint32_t)g->m);
  mix((uint32_t)(g->m >> 32));
  for (uint64_t u = 0; u <= g->n; ++u) {
    mix((uint32_t)g->row_ptr[u]);
    mix((uint32_t)(g->row_ptr[u] >> 32));
  }
  for (uint64_t k = 0; k < g->m; ++k) mix((uint32_t)g->adj[k]);
  return h;
}

// ---------------------------------------------------------------------------
// Orientation (edge ids)
// ---------------------------------------------------------------------------

extern "C" int sgc_orient(const sgc_host_graph* g, int32_t* edge_u, int32_t* edge_v, uint32_t* eid) {
  clear_error();
  if (!g) return fail(SGC_EINVAL, "null graph");
  if (g->m >= (1ull << 32)) return fail(SGC_ERANGE, "m=%llu needs 64-bit edge ids", (unsigned long long)g->m);
  const uint64_t n = g->n;
  const uint64_t* rp = g->row_ptr;
  const int32_t* adj = g->adj;
  // rank[k] = 1-based rank of entry k among the "up" entries of its row (the
  // neighbour has larger (degree, id)), 0 for "down" entries.
  std::vector<uint32_t> rank(g->m);
  std::vector<uint64_t> base(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t u = 0; u < n; ++u) {
    const uint64_t du = rp[u + 1] - rp[u];
    uint32_t r = 0;
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      const uint64_t v = (uint64_t)adj[k];
      const uint64_t dv = rp[v + 1] - rp[v];
      const bool up = dv > du || (dv == du && v > u);
      rank[k] = up ? ++r : 0;
    }
    base[u + 1] = r;
  }
  for (uint64_t u = 0; u < n; ++u) base[u + 1] += base[u];
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t u = 0; u < n; ++u) {
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      const uint64_t v = (uint64_t)adj[k];
      uint64_t id;
      if (rank[k]) {
        id = base[u] + rank[k] - 1;
        if (edge_u) edge_u[id] = (int32_t)u;
        if (edge_v) edge_v[id] = (int32_t)v;
      } else {
        const uint64_t p = lower_bound_u64(adj, rp[v], rp[v + 1], (int32_t)u);
        id = base[v] + rank[p] - 1;
      }
      if (eid) eid[k] = (uint32_t)id;
    }
  }
  return SGC_OK;
}

extern "C" int sgc_graph_similarity_work(const sgc_host_graph* g, double eps, uint64_t* u_edges,
                                         double* alg_bytes, double* short_elements, double* alg_bytes_large) {
  clear_error();
  if (!g) return fail(SGC_EINVAL, "null graph");
  uint64_t cnt = 0;
  double bytes = 0, shorts = 0, large = 0;
  const uint64_t* rp = g->row_ptr;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : cnt, bytes, shorts, large)
  for (uint64_t u = 0; u < g->n; ++u) {
    const uint64_t du = rp[u + 1] - rp[u];
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      const uint64_t v = (uint64_t)g->adj[k];
      const uint64_t dv = rp[v + 1] - rp[v];
      if (!(dv > du || (dv == du && v > u))) continue;  // count each edge from its lower endpoint
      const double A = (double)(du + 1), B = (double)(dv + 1);
      if (A < B * eps * eps || eps * eps * A * B <= 4.0) continue;
      ++cnt;
      double levels = std::ceil(std::log2((double)dv + 1.0));
      const double eb = 4.0 * ((double)du + std::min((double)dv, (double)du * levels));
      bytes += eb;
      if (dv > 1024) large += eb;  // owner v in the workgroup tier
      shorts += (double)du;
    }
  }
  if (u_edges) *u_edges = cnt;
  if (alg_bytes) *alg_bytes = bytes;
  if (short_elements) *short_elements = shorts;
  if (alg_bytes_large) *alg_bytes_large = large;
  return SGC_OK;
}

extern "C" void sgc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  Philox4x32 r = philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

extern "C" const char* sgc_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* sgc_version(void) { return "sgc 0.1 (gfx950)"; }
