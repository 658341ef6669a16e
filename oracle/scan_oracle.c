/*
 * scan_oracle.c - CPU restatement of GPUSCAN++ (TEST INFRASTRUCTURE ONLY).
 *
 * The checker for libsgc.so: loaded only by tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg.  Plain C11 + OpenMP, built by
 * oracle/Makefile with -O2 -ffp-contract=off (no reassociation, no FMA), so
 * every double expression below is evaluated in IEEE binary64, left to right,
 * exactly as written in the reference.
 *
 * Each function cites the reference lines it restates.  How each part is
 * pinned (closed forms, brute force, the reference's own compiled host code)
 * is listed in tests/test_oracle.py and in the repository's design notes.
 */
#include "scan_oracle.h"

#include <limits.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t deg_of(const uint64_t* rp, uint64_t x) { return rp[x + 1] - rp[x]; }

/* GPUSCAN++/GPUScan.cu:143-169.  deg_u/deg_v there are the closed degrees of
 * Edge_u (lower (degree, id) endpoint) and Edge_v, converted to double. */
int orc_prep(uint32_t deg_a, uint32_t deg_b, double eps) {
  double A = (double)(deg_a + 1u);
  double B = (double)(deg_b + 1u);
  if (A < B * eps * eps) return ORC_N;        /* :145 evaluates (B*eps)*eps */
  if (eps * eps * A * B <= 4.0) return ORC_Y; /* :156 evaluates ((eps*eps)*A)*B */
  return ORC_U;
}

/* GPUSCAN++/GPUScan.cu:220-221: `(T + 2) * (T + 2)` is an int product (it
 * wraps modulo 2^32 once T + 2 > 46340) compared, after conversion to
 * double, against eps*eps*(s+1)*(l+1) evaluated left to right with int
 * operands converted to double. */
int orc_predicate(uint32_t cn, uint32_t s, uint32_t l, double eps) {
  uint32_t t = cn + 2u;
  int32_t lhs = (int32_t)(t * t);
  double rhs = eps * eps * (double)(int32_t)(s + 1u) * (double)(int32_t)(l + 1u);
  int unsimilar = (double)lhs < rhs;
  return !unsimilar;
}

/* cn = |N(u) ∩ N(v)| (GPUScan.cu:193-218 counts elements of the first row
 * found in the second); rows are sorted, so a merge gives the same count. */
uint32_t orc_common(const uint64_t* rp, const int32_t* adj, uint32_t u, uint32_t v) {
  uint64_t i = rp[u], ie = rp[u + 1], j = rp[v], je = rp[v + 1];
  uint32_t c = 0;
  while (i < ie && j < je) {
    int32_t x = adj[i], y = adj[j];
    if (x < y)
      ++i;
    else if (x > y)
      ++j;
    else {
      ++c;
      ++i;
      ++j;
    }
  }
  return c;
}

static uint64_t lower_bound_i32(const int32_t* a, uint64_t lo, uint64_t hi, int32_t x) {
  while (lo < hi) {
    uint64_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* GPUSCAN++/Graph.cpp:52-141.  Entry k of row u is "higher" when its
 * neighbour has a larger degree, or an equal degree and a larger id
 * (:75-83); those entries are numbered 1, 2, ... in row order (:76-81), the
 * per-row counts are prefix-summed (:87-92), and an edge's id is its lower
 * endpoint's base plus its rank (:109); lower entries find their id through
 * the position of u in the neighbour's row (:99-107). */
void orc_orient(uint64_t n, const uint64_t* rp, const int32_t* adj, int32_t* edge_u, int32_t* edge_v,
                uint32_t* eid) {
  const uint64_t m = rp[n];
  uint32_t* rank = (uint32_t*)malloc((m ? m : 1) * sizeof(uint32_t));
  uint64_t* base = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
  for (uint64_t u = 0; u < n; ++u) {
    uint64_t du = deg_of(rp, u);
    uint32_t r = 0;
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      uint64_t v = (uint64_t)adj[k];
      uint64_t dv = deg_of(rp, v);
      if (dv > du || (dv == du && v > u))
        rank[k] = ++r;
      else
        rank[k] = 0;
    }
    base[u + 1] = base[u] + r;
  }
  for (uint64_t u = 0; u < n; ++u) {
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      uint64_t v = (uint64_t)adj[k];
      uint64_t id;
      if (rank[k]) {
        id = base[u] + rank[k] - 1;
        if (edge_u) edge_u[id] = (int32_t)u;
        if (edge_v) edge_v[id] = (int32_t)v;
      } else {
        uint64_t p = lower_bound_i32(adj, rp[v], rp[v + 1], (int32_t)u);
        id = base[v] + rank[p] - 1;
      }
      if (eid) eid[k] = (uint32_t)id;
    }
  }
  free(rank);
  free(base);
}

/* findroot (GPUScan.cu:465-477): parents always point to smaller ids, so the
 * root of a tree is its minimum; compression re-points to the grandparent. */
static int32_t find_root(int32_t* role, int32_t x) {
  while (role[x] != x) {
    int32_t p = role[x];
    role[x] = role[p];
    x = p;
  }
  return x;
}

/* Union by hooking the larger root under the smaller (GPUScan.cu:503-521). */
static void unite(int32_t* role, int32_t a, int32_t b) {
  int32_t ra = find_root(role, a), rb = find_root(role, b);
  if (ra == rb) return;
  if (ra < rb)
    role[rb] = ra;
  else
    role[ra] = rb;
}

/* Updates of sd/ed/role after an edge is decided (prep: GPUScan.cu:146-166;
 * check_core: :316-363). */
static void note_edge(int sim, int32_t x, int32_t mu, int32_t* role, int32_t* sd, int32_t* ed) {
  if (role[x] != -1) return;
  if (sim) {
    sd[x] += 1;
    if (sd[x] >= mu) role[x] = x;
  } else {
    ed[x] -= 1;
    if (ed[x] < mu) role[x] = -2;
  }
}

int orc_scan(uint64_t n, const uint64_t* rp, const int32_t* adj, double eps, int32_t mu, int32_t* labels,
             uint8_t* core, uint8_t* sim, orc_stats* st) {
  const uint64_t m = rp[n];
  if (m >= (1ull << 32)) return -1;
  const uint64_t E = m / 2;
  orc_stats s;
  memset(&s, 0, sizeof(s));
  s.edges = E;

  int32_t* Eu = (int32_t*)malloc((E ? E : 1) * sizeof(int32_t));
  int32_t* Ev = (int32_t*)malloc((E ? E : 1) * sizeof(int32_t));
  uint32_t* eid = (uint32_t*)malloc((m ? m : 1) * sizeof(uint32_t));
  orc_orient(n, rp, adj, Eu, Ev, eid);

  int32_t* role = labels;
  int32_t* sd = (int32_t*)malloc((n ? n : 1) * sizeof(int32_t));
  int32_t* ed = (int32_t*)malloc((n ? n : 1) * sizeof(int32_t));
  char* flag = (char*)malloc(E ? E : 1);

  /* __init_input (GPUScan.cu:115-131) */
  for (uint64_t x = 0; x < n; ++x) {
    sd[x] = 0;
    ed[x] = (int32_t)deg_of(rp, x);
    role[x] = ed[x] < mu ? -2 : -1;
  }

  /* __prep_edge (GPUScan.cu:133-172) */
  for (uint64_t e = 0; e < E; ++e) {
    int32_t u = Eu[e], v = Ev[e];
    int p = orc_prep((uint32_t)deg_of(rp, u), (uint32_t)deg_of(rp, v), eps);
    if (p == ORC_N) {
      flag[e] = 'N';
      ++s.prep_n;
      note_edge(0, u, mu, role, sd, ed);
      note_edge(0, v, mu, role, sd, ed);
    } else if (p == ORC_Y) {
      flag[e] = 'Y';
      ++s.prep_y;
      note_edge(1, u, mu, role, sd, ed);
      note_edge(1, v, mu, role, sd, ed);
    } else {
      flag[e] = 'U';
      ++s.prep_u;
    }
  }
  for (uint64_t x = 0; x < n; ++x) s.cores_after_prep += role[x] >= 0;

  /* __check_core (GPUScan.cu:276-367): U edges with an undecided endpoint. */
  for (uint64_t e = 0; e < E; ++e) {
    int32_t u = Eu[e], v = Ev[e];
    if (flag[e] == 'U' && (role[u] == -1 || role[v] == -1)) {
      uint32_t cn = orc_common(rp, adj, (uint32_t)u, (uint32_t)v);
      int y = orc_predicate(cn, (uint32_t)deg_of(rp, u), (uint32_t)deg_of(rp, v), eps);
      ++s.check_core_computed;
      flag[e] = y ? 'Y' : 'N';
      note_edge(y, u, mu, role, sd, ed);
      note_edge(y, v, mu, role, sd, ed);
    }
  }
  for (uint64_t x = 0; x < n; ++x) s.cores += role[x] >= 0;
  if (core)
    for (uint64_t x = 0; x < n; ++x) core[x] = role[x] >= 0;

  /* Core list: thrust::copy_if of role values, which equal the ids here
   * (GPUScan.cu:919-920); ascending. */
  uint64_t ncore = 0;
  int32_t* cores = (int32_t*)malloc((s.cores ? s.cores : 1) * sizeof(int32_t));
  for (uint64_t x = 0; x < n; ++x)
    if (role[x] >= 0) cores[ncore++] = (int32_t)x;

  /* __cluster_core1 (GPUScan.cu:479-528): known-similar core-core edges,
   * visited from the larger endpoint. */
  for (uint64_t i = 0; i < ncore; ++i) {
    int32_t u = cores[i];
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      int32_t v = adj[k];
      if (u < v) break;
      if (role[v] >= 0 && flag[eid[k]] == 'Y') unite(role, u, v);
    }
  }
  /* __cluster_core2 (GPUScan.cu:530-630): undecided core-core edges whose
   * endpoints are still in different trees; the shorter row goes first
   * (:577-585). */
  for (uint64_t i = 0; i < ncore; ++i) {
    int32_t u = cores[i];
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      int32_t v = adj[k];
      if (v >= u) break;
      if (role[v] >= 0 && flag[eid[k]] == 'U' && find_root(role, u) != find_root(role, v)) {
        uint32_t du = (uint32_t)deg_of(rp, u), dv = (uint32_t)deg_of(rp, v);
        uint32_t cn = orc_common(rp, adj, (uint32_t)u, (uint32_t)v);
        int y = dv > du ? orc_predicate(cn, du, dv, eps) : orc_predicate(cn, dv, du, eps);
        flag[eid[k]] = y ? 'Y' : 'N';
        if (y) unite(role, u, v);
      }
    }
  }
  /* __cluster_core (GPUScan.cu:632-652): flatten to the root = min core id. */
  for (uint64_t i = 0; i < ncore; ++i) role[cores[i]] = find_root(role, cores[i]);
  for (uint64_t i = 0; i < ncore; ++i) s.clusters += role[cores[i]] == cores[i];

  /* __cluster_noncore (GPUScan.cu:654-730): a non-core neighbour v of a core
   * u joins u's cluster over a similar edge (U edges computed on demand,
   * :695-719).  Canonical rule: the minimum candidate cluster id, decided
   * after every core has been visited (the reference keeps whichever write
   * lands last, :716/:722). */
  int32_t* cand = (int32_t*)malloc((n ? n : 1) * sizeof(int32_t));
  for (uint64_t x = 0; x < n; ++x) cand[x] = INT32_MAX;
  for (uint64_t i = 0; i < ncore; ++i) {
    int32_t u = cores[i];
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      int32_t v = adj[k];
      if (role[v] >= 0) continue;  /* cores only: non-core roles are untouched here */
      if (flag[eid[k]] == 'U') {
        uint32_t du = (uint32_t)deg_of(rp, u), dv = (uint32_t)deg_of(rp, v);
        uint32_t cn = orc_common(rp, adj, (uint32_t)u, (uint32_t)v);
        int y = dv > du ? orc_predicate(cn, du, dv, eps) : orc_predicate(cn, dv, du, eps);
        flag[eid[k]] = y ? 'Y' : 'N';
      }
      if (flag[eid[k]] == 'Y' && role[u] < cand[v]) cand[v] = role[u];
    }
  }
  for (uint64_t x = 0; x < n; ++x)
    if (role[x] < 0 && cand[x] != INT32_MAX) {
      role[x] = cand[x];
      ++s.borders;
    }

  /* __classify_others (GPUScan.cu:732-798) over the first count(role == -2)
   * entries of the ascending list of role < 0 vertices (:978-984). */
  uint64_t others = 0;
  for (uint64_t x = 0; x < n; ++x) others += role[x] == -2;
  uint64_t taken = 0;
  for (uint64_t x = 0; x < n && taken < others; ++x) {
    if (role[x] >= 0) continue;
    ++taken;
    uint64_t b = rp[x], e = rp[x + 1];
    if (e - b <= 1) continue; /* :753 */
    uint64_t p = b;
    int32_t ne = -1;
    for (; p < e; ++p)
      if (role[adj[p]] >= 0) {
        ne = role[adj[p]];
        break;
      }
    if (ne >= 0 && p + 1 < e) { /* :770 */
      for (uint64_t w = p + 1; w < e; ++w) {
        int32_t r = role[adj[w]];
        if (r >= 0 && r != ne) {
          role[x] = -1;
          break;
        }
      }
    }
  }
  for (uint64_t x = 0; x < n; ++x) {
    s.hubs += role[x] == -1;
    s.outliers += role[x] == -2;
  }

  /* Complete every edge the pass left undecided, for the similarity output. */
  for (uint64_t e = 0; e < E; ++e) {
    if (flag[e] == 'U') {
      int32_t u = Eu[e], v = Ev[e];
      uint32_t cn = orc_common(rp, adj, (uint32_t)u, (uint32_t)v);
      flag[e] = orc_predicate(cn, (uint32_t)deg_of(rp, u), (uint32_t)deg_of(rp, v), eps) ? 'Y' : 'N';
    }
    s.similar_edges += flag[e] == 'Y';
  }
  if (sim)
    for (uint64_t k = 0; k < m; ++k) sim[k] = flag[eid[k]] == 'Y';

  if (st) *st = s;
  free(Eu);
  free(Ev);
  free(eid);
  free(sd);
  free(ed);
  free(flag);
  free(cores);
  free(cand);
  return 0;
}

uint64_t orc_similarity_entries(uint64_t n, const uint64_t* rp, const int32_t* adj, double eps,
                                const uint64_t* entry, uint64_t count, uint8_t* out, int threads) {
  uint64_t similar = 0;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : similar)
  for (uint64_t i = 0; i < count; ++i) {
    uint64_t k = entry[i];
    /* row of entry k: last u with rp[u] <= k */
    uint64_t lo = 0, hi = n;
    while (hi - lo > 1) {
      uint64_t mid = (lo + hi) >> 1;
      if (rp[mid] <= k)
        lo = mid;
      else
        hi = mid;
    }
    uint64_t u = lo, v = (uint64_t)adj[k];
    uint64_t du = deg_of(rp, u), dv = deg_of(rp, v);
    int up = dv > du || (dv == du && v > u);
    uint64_t a = up ? u : v, b = up ? v : u;
    uint32_t da = (uint32_t)deg_of(rp, a), db = (uint32_t)deg_of(rp, b);
    int p = orc_prep(da, db, eps);
    int y;
    if (p == ORC_U)
      y = orc_predicate(orc_common(rp, adj, (uint32_t)a, (uint32_t)b), da, db, eps);
    else
      y = p == ORC_Y;
    if (out) out[i] = (uint8_t)y;
    similar += (uint64_t)y;
  }
  return similar;
}

int orc_labels_from_sim(uint64_t n, const uint64_t* rp, const int32_t* adj, const uint8_t* sim, int32_t mu,
                        int32_t* labels, uint8_t* core, orc_stats* st) {
  const uint64_t m = rp[n];
  if (n >= (1ull << 31)) return -1;
  orc_stats s;
  memset(&s, 0, sizeof(s));
  s.edges = m / 2;
  int32_t* role = labels;
  uint64_t similar = 0;
  /* core check: the state machine's fixed point once every edge is known:
   * sd = #similar, ed = degree - #dissimilar = #similar, so core <=>
   * #similar >= mu (mu <= 0: no vertex starts at -2, one similar edge
   * makes a core, every other vertex stays -1). */
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : similar)
  for (uint64_t x = 0; x < n; ++x) {
    int32_t y = 0;
    for (uint64_t k = rp[x]; k < rp[x + 1]; ++k) y += sim[k] != 0;
    similar += (uint64_t)y;
    const int is_core = mu >= 1 ? y >= mu : y >= 1;
    role[x] = is_core ? (int32_t)x : (mu >= 1 ? -2 : -1);
  }
  s.similar_edges = similar / 2;
  uint64_t ncore = 0;
  for (uint64_t x = 0; x < n; ++x) ncore += role[x] >= 0;
  s.cores = ncore;
  if (core)
    for (uint64_t x = 0; x < n; ++x) core[x] = role[x] >= 0;
  int32_t* cores = (int32_t*)malloc((ncore ? ncore : 1) * sizeof(int32_t));
  ncore = 0;
  for (uint64_t x = 0; x < n; ++x)
    if (role[x] >= 0) cores[ncore++] = (int32_t)x;
  /* clusters: similar core-core edges (GPUScan.cu:479-630), then flatten
   * to the minimum core id (:632-652) */
  for (uint64_t i = 0; i < ncore; ++i) {
    int32_t u = cores[i];
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      int32_t v = adj[k];
      if (v >= u) break;
      if (role[v] >= 0 && sim[k]) unite(role, u, v);
    }
  }
  for (uint64_t i = 0; i < ncore; ++i) role[cores[i]] = find_root(role, cores[i]);
  for (uint64_t i = 0; i < ncore; ++i) s.clusters += role[cores[i]] == cores[i];
  /* borders: minimum candidate cluster over similar edges to cores */
  int32_t* cand = (int32_t*)malloc((n ? n : 1) * sizeof(int32_t));
  for (uint64_t x = 0; x < n; ++x) cand[x] = INT32_MAX;
  for (uint64_t i = 0; i < ncore; ++i) {
    int32_t u = cores[i];
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      int32_t v = adj[k];
      if (role[v] < 0 && sim[k] && role[u] < cand[v]) cand[v] = role[u];
    }
  }
  for (uint64_t x = 0; x < n; ++x)
    if (role[x] < 0 && cand[x] != INT32_MAX) {
      role[x] = cand[x];
      ++s.borders;
    }
  /* hubs and outliers (GPUScan.cu:732-798): a -2 vertex of degree >= 2 whose
   * neighbours carry two different labels >= 0 becomes -1 */
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t x = 0; x < n; ++x) {
    if (role[x] != -2) continue;
    uint64_t b = rp[x], e = rp[x + 1];
    if (e - b <= 1) continue;
    int32_t first = -1;
    for (uint64_t p = b; p < e; ++p) {
      /* a neighbour's final label: cores and borders are settled above, and
       * no vertex turned hub in this loop is read as a label (-1 < 0) */
      int32_t r = role[adj[p]];
      if (r < 0) continue;
      if (first < 0) {
        first = r;
      } else if (r != first) {
        cand[x] = -1; /* decided after the loop, so every read sees final labels */
        break;
      }
    }
  }
  for (uint64_t x = 0; x < n; ++x)
    if (role[x] == -2 && cand[x] == -1) role[x] = -1;
  for (uint64_t x = 0; x < n; ++x) {
    s.hubs += role[x] == -1;
    s.outliers += role[x] == -2;
  }
  if (st) *st = s;
  free(cores);
  free(cand);
  return 0;
}

void orc_prep_counts(uint64_t n, const uint64_t* rp, const int32_t* adj, double eps, uint64_t* pn, uint64_t* py,
                     uint64_t* pu) {
  uint64_t cn = 0, cy = 0, cu = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : cn, cy, cu)
  for (uint64_t x = 0; x < n; ++x) {
    const uint64_t dx = deg_of(rp, x);
    for (uint64_t k = rp[x]; k < rp[x + 1]; ++k) {
      const uint64_t v = (uint64_t)adj[k], dv = deg_of(rp, v);
      /* each undirected edge once, from its lower (degree, id) endpoint a
       * (GPUSCAN++/Graph.cpp:75-83) */
      if (!(dv > dx || (dv == dx && v > x))) continue;
      const int p = orc_prep((uint32_t)dx, (uint32_t)dv, eps);
      cn += p == ORC_N;
      cy += p == ORC_Y;
      cu += p == ORC_U;
    }
  }
  *pn = cn;
  *py = cy;
  *pu = cu;
}
