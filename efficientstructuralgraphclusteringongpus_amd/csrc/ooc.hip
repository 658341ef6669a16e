// Out-of-core SCAN for graphs larger than a device budget (GPUSCAN++O:
// chips GPUSCAN++O/Graph.h:70-85, Graph.cpp:414-576; phase-major
// orchestration SCANoffload.cu:922-1464).
//
// Per-vertex state stays resident (degrees, the chip row table, similar-degree
// counts, union-find forest, labels, classify bounds, planner arrays: about
// 57 bytes per vertex); the rest of the budget holds one chip at a time.
//  * Similarity: a chip is a range of owner rows plus every short row its
//    undecided edges read, stored once.  The chip's row table maps those
//    rows into the chip; absent rows keep only their degree (a range of the
//    right length that no kernel reads).  The in-memory prep, planner and
//    similarity kernels run on the chip unchanged; mirror entries land where
//    the short row sits in the chip.  Final flags go back into a pinned
//    whole-graph flag array.  The next chip is assembled on the host while
//    the device works on the current one, and the previous chip's flags are
//    merged meanwhile.
//  * Core check, union-find, border and classify: chips of whole rows with
//    their flags, one pass per phase (each phase needs the previous one's
//    global result); when all rows fit in one chip it is uploaded once.
// Labels, core flags and similarity flags are bit-identical to the resident
// path's.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "device.h"
#include "device_util.h"
#include "ooc_host.h"

namespace sgc {
namespace {

__global__ void __launch_bounds__(kBlock) k_rows_absent(uint32_t n, const uint32_t* __restrict__ deg,
                                                        uint32_t* __restrict__ rb, uint32_t* __restrict__ re) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += stride) {
    rb[x] = 0;
    re[x] = deg[x];
  }
}

__global__ void __launch_bounds__(kBlock) k_rows_range(uint32_t r0, uint32_t nr, const uint32_t* __restrict__ off,
                                                       uint32_t* __restrict__ rb, uint32_t* __restrict__ re) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < nr; i += stride) {
    rb[r0 + i] = off[i];
    re[r0 + i] = off[i + 1];
  }
}

__global__ void __launch_bounds__(kBlock) k_rows_list(uint32_t cnt, const uint32_t* __restrict__ ids,
                                                      const uint32_t* __restrict__ starts,
                                                      const uint32_t* __restrict__ deg, uint32_t* __restrict__ rb,
                                                      uint32_t* __restrict__ re) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += stride) {
    const uint32_t x = ids[i];
    rb[x] = starts[i];
    re[x] = starts[i] + deg[x];
  }
}

// src[k] = r0 + i for the entries [off[i], off[i+1]) of chip row i; one wave
// per row.
__global__ void __launch_bounds__(kBlock) k_fill_src_rows(uint32_t r0, uint32_t nr, const uint32_t* __restrict__ off,
                                                          uint32_t* __restrict__ src) {
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t i = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; i < nr; i += waves) {
    const uint32_t b = off[i], e = off[i + 1];
    for (uint32_t k = b + lane; k < e; k += kWave) src[k] = r0 + (uint32_t)i;
  }
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.f;
}

struct Layout {
  size_t bytes = 0;
  size_t take(size_t b) {
    const size_t o = bytes;
    bytes = (bytes + std::max<size_t>(b, 1) + 255) & ~(size_t)255;
    return o;
  }
};

// Bytes per entry of the two chip kinds (adj, src, flag, ulist, two unit
// arrays, short-row ids and starts / adj, src, flag).
constexpr uint64_t kSimChipBytesPerEntry = 4 + 4 + 1 + 4 + 8 + 8 + 4 + 4;
constexpr uint64_t kRowChipBytesPerEntry = 4 + 4 + 1;
constexpr uint64_t kMaxChipEntries = 0xFFFFFFF0ull;

struct OocRun {
  sgc_ctx* ctx = nullptr;
  const sgc_host_graph* g = nullptr;
  uint32_t n = 0;
  void* dev = nullptr;
  // resident
  uint32_t *deg = nullptr, *rb = nullptr, *re = nullptr, *off = nullptr;
  PassBuffers pb{};
  // chip region
  char* chip = nullptr;
  size_t chip_bytes = 0;
  // pinned host staging, two of each for overlap
  int32_t* h_adj[2] = {nullptr, nullptr};
  uint32_t* h_off[2] = {nullptr, nullptr};
  uint8_t* h_flag[2] = {nullptr, nullptr};
  uint64_t* h_cnt[2] = {nullptr, nullptr};
  uint8_t* flag = nullptr;  // whole-graph flags, pinned
  hipEvent_t done[2] = {nullptr, nullptr};
  double bytes_h2d = 0, bytes_d2h = 0, ms_host = 0;

  ~OocRun() {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 2; ++i) {
      if (h_adj[i]) (void)hipHostFree(h_adj[i]);
      if (h_off[i]) (void)hipHostFree(h_off[i]);
      if (h_flag[i]) (void)hipHostFree(h_flag[i]);
      if (h_cnt[i]) (void)hipHostFree(h_cnt[i]);
      if (done[i]) (void)hipEventDestroy(done[i]);
    }
    if (flag) (void)hipHostFree(flag);
    if (dev) (void)hipFree(dev);
  }
};

// Similarity-chip view: the owner entries [0, m_own) are the ones the
// kernels decide; short rows follow in the same arrays.
struct SimChipArrays {
  int32_t* adj;
  uint32_t* src;
  uint8_t* flag;
  uint32_t* ulist;
  uint2* units1;
  uint2* units2;
  uint32_t* ids;
  uint32_t* starts;
};

SimChipArrays sim_arrays(char* base, uint64_t cap) {
  Layout L;
  const size_t o_adj = L.take(cap * 4), o_src = L.take(cap * 4), o_flag = L.take(cap),
               o_ulist = L.take(cap * 4), o_u1 = L.take(unit_capacity(cap) * sizeof(uint2)),
               o_u2 = L.take(unit_capacity(cap) * sizeof(uint2)), o_ids = L.take(cap * 4), o_st = L.take(cap * 4);
  return SimChipArrays{(int32_t*)(base + o_adj),  (uint32_t*)(base + o_src), (uint8_t*)(base + o_flag),
                       (uint32_t*)(base + o_ulist), (uint2*)(base + o_u1),    (uint2*)(base + o_u2),
                       (uint32_t*)(base + o_ids), (uint32_t*)(base + o_st)};
}

size_t sim_chip_bytes(uint64_t cap) {
  Layout L;
  L.take(cap * 4), L.take(cap * 4), L.take(cap), L.take(cap * 4), L.take(unit_capacity(cap) * sizeof(uint2)),
      L.take(unit_capacity(cap) * sizeof(uint2)), L.take(cap * 4), L.take(cap * 4);
  return L.bytes;
}

size_t row_chip_bytes(uint64_t cap) {
  Layout L;
  L.take(cap * 4), L.take(cap * 4), L.take(cap);
  return L.bytes;
}

size_t resident_bytes(uint32_t n) {
  Layout L;
  for (int i = 0; i < 3; ++i) L.take((size_t)n * 4);  // deg, row table
  L.take(((size_t)n + 1) * 4);                        // owner offsets
  for (int i = 0; i < 5; ++i) L.take((size_t)n * 4);  // sd, parent, label, lo, hi
  L.take(n);                                          // core
  L.take((size_t)n * 4), L.take((size_t)n * 8), L.take((size_t)n * 4), L.take(((size_t)n + 1) * 4);
  L.take(scan_tmp_words(n) * 4), L.take(kCntSlots * 8);
  return L.bytes;
}

int run_ooc(OocRun& R, double eps, int32_t mu, uint64_t budget, int32_t* labels, uint8_t* core_out, uint8_t* sim_out,
            sgc_stats* st, sgc_ooc_stats* oi) {
  sgc_ctx* ctx = R.ctx;
  const sgc_host_graph* g = R.g;
  const uint32_t n = R.n;
  hipStream_t s = ctx->stream;
  const double t_start = now_ms();

  // ---- resident layout
  Layout L;
  const size_t o_deg = L.take((size_t)n * 4), o_rb = L.take((size_t)n * 4), o_re = L.take((size_t)n * 4),
               o_off = L.take(((size_t)n + 1) * 4), o_sd = L.take((size_t)n * 4), o_par = L.take((size_t)n * 4),
               o_lab = L.take((size_t)n * 4), o_lo = L.take((size_t)n * 4), o_hi = L.take((size_t)n * 4),
               o_core = L.take(n), o_cntu = L.take((size_t)n * 4), o_work = L.take((size_t)n * 8),
               o_fill = L.take((size_t)n * 4), o_ust = L.take(((size_t)n + 1) * 4),
               o_scan = L.take(scan_tmp_words(n) * 4), o_cnt = L.take(kCntSlots * 8);
  const size_t resident = L.bytes;  // == resident_bytes(n)
  if (budget <= resident)
    return fail(SGC_ENOMEM, "device budget %llu bytes is below the %llu resident bytes the graph's %u vertices need",
                (unsigned long long)budget, (unsigned long long)resident, n);
  const uint64_t room = budget - resident - 4096;
  const uint64_t cap_sim = std::min<uint64_t>(room / kSimChipBytesPerEntry, kMaxChipEntries);
  const uint64_t cap_row = std::min<uint64_t>(room / kRowChipBytesPerEntry, kMaxChipEntries);

  // ---- plans
  double t0 = now_ms();
  std::vector<uint32_t> deg(n);
  ooc_degrees(g, deg.data());
  std::vector<uint64_t> sim_bounds, row_bounds;
  uint64_t sim_max = 0, row_max = 0;
  int rc = ooc_plan_similarity(g, deg.data(), eps, cap_sim, &sim_bounds, &sim_max);
  if (rc) return rc;
  rc = ooc_plan_rows(g, cap_row, &row_bounds, &row_max);
  if (rc) return rc;
  const size_t chip_need = std::max(sim_chip_bytes(std::max<uint64_t>(sim_max, 1)),
                                    row_chip_bytes(std::max<uint64_t>(row_max, 1)));
  R.ms_host += now_ms() - t0;

  // ---- device and pinned host memory
  R.chip_bytes = chip_need;
  if (hipMalloc(&R.dev, resident + chip_need) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SGC_ENOMEM, "device allocation of %llu bytes failed", (unsigned long long)(resident + chip_need));
  }
  char* base = (char*)R.dev;
  R.deg = (uint32_t*)(base + o_deg);
  R.rb = (uint32_t*)(base + o_rb);
  R.re = (uint32_t*)(base + o_re);
  R.off = (uint32_t*)(base + o_off);
  PassBuffers& pb = R.pb;
  pb.sd = (uint32_t*)(base + o_sd);
  pb.parent = (int32_t*)(base + o_par);
  pb.label = (int32_t*)(base + o_lab);
  pb.lo = (int32_t*)(base + o_lo);
  pb.hi = (int32_t*)(base + o_hi);
  pb.core = (uint8_t*)(base + o_core);
  pb.cnt_u = (uint32_t*)(base + o_cntu);
  pb.work_u = (uint64_t*)(base + o_work);
  pb.fill = (uint32_t*)(base + o_fill);
  pb.ustart = (uint32_t*)(base + o_ust);
  pb.scan_tmp = (uint32_t*)(base + o_scan);
  pb.cnt = (uint64_t*)(base + o_cnt);
  R.chip = base + resident;

  const uint64_t sim_entries = std::max<uint64_t>(sim_max, 1);
  for (int i = 0; i < 2; ++i) {
    if (hipHostMalloc((void**)&R.h_adj[i], sim_entries * 4) != hipSuccess ||
        hipHostMalloc((void**)&R.h_off[i], ((size_t)n + 1) * 4) != hipSuccess ||
        hipHostMalloc((void**)&R.h_flag[i], sim_entries) != hipSuccess ||
        hipHostMalloc((void**)&R.h_cnt[i], kCntSlots * 8) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SGC_ENOMEM, "pinned host staging of %llu bytes failed", (unsigned long long)(sim_entries * 5));
    }
    SGC_HIP(hipEventCreateWithFlags(&R.done[i], hipEventDisableTiming));
  }
  if (hipHostMalloc((void**)&R.flag, std::max<uint64_t>(g->m, 1)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SGC_ENOMEM, "pinned host flag array of %llu bytes failed", (unsigned long long)g->m);
  }
  SGC_HIP(hipMemcpyAsync(R.deg, deg.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
  R.bytes_h2d += (double)n * 4;

  const ScanParams p{eps, mu};
  uint64_t totals[kCntSlots] = {0};
  double ms_sim = 0, ms_sim_large = 0;

  // ---- similarity chips
  const size_t nsim = sim_bounds.size() - 1;
  std::vector<int32_t> stamp(n, -1);
  OocChip chips[2];
  auto build = [&](size_t i) {
    const double tb = now_ms();
    ooc_build_chip(g, deg.data(), eps, sim_bounds[i], sim_bounds[i + 1], stamp.data(), (int32_t)i, R.h_adj[i & 1],
                   R.h_off[i & 1], &chips[i & 1]);
    R.ms_host += now_ms() - tb;
  };
  auto merge = [&](size_t i) -> int {
    SGC_HIP(hipEventSynchronize(R.done[i & 1]));
    const double tb = now_ms();
    ooc_merge_flags(g, deg.data(), chips[i & 1], R.h_flag[i & 1], R.flag);
    for (int c = 0; c < kCntSlots; ++c) totals[c] += R.h_cnt[i & 1][c];
    R.ms_host += now_ms() - tb;
    return SGC_OK;
  };
  const SimChipArrays A = sim_arrays(R.chip, sim_entries);
  if (g->m) build(0);
  for (size_t i = 0; g->m && i < nsim; ++i) {
    const OocChip& c = chips[i & 1];
    const uint32_t nr = (uint32_t)(c.r1 - c.r0), nshort = (uint32_t)c.ids.size();
    // the device may still read the previous chip: stream order serialises
    SGC_HIP(hipMemcpyAsync(A.adj, R.h_adj[i & 1], c.m_chip * 4, hipMemcpyHostToDevice, s));
    SGC_HIP(hipMemcpyAsync(R.off, R.h_off[i & 1], ((size_t)nr + 1) * 4, hipMemcpyHostToDevice, s));
    if (nshort) {
      SGC_HIP(hipMemcpyAsync(A.ids, c.ids.data(), (size_t)nshort * 4, hipMemcpyHostToDevice, s));
      SGC_HIP(hipMemcpyAsync(A.starts, c.starts.data(), (size_t)nshort * 4, hipMemcpyHostToDevice, s));
    }
    R.bytes_h2d += (double)c.m_chip * 4 + (double)(nr + 1) * 4 + (double)nshort * 8;
    k_rows_absent<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, R.deg, R.rb, R.re);
    if (nr) {
      k_rows_range<<<grid_for(nr, kBlock), kBlock, 0, s>>>((uint32_t)c.r0, nr, R.off, R.rb, R.re);
      k_fill_src_rows<<<grid_for(nr, kBlock / kWave), kBlock, 0, s>>>((uint32_t)c.r0, nr, R.off, A.src);
    }
    if (nshort) k_rows_list<<<grid_for(nshort, kBlock), kBlock, 0, s>>>(nshort, A.ids, A.starts, R.deg, R.rb, R.re);
    if (c.m_chip > c.m_own) SGC_HIP(hipMemsetAsync(A.flag + c.m_own, 0xFF, c.m_chip - c.m_own, s));
    sgc_dgraph view;
    view.ctx = ctx;
    view.n = n;
    view.m = (uint32_t)c.m_own;
    view.adj = A.adj;
    view.src = A.src;
    view.row_begin = R.rb;
    view.row_end = R.re;
    PassBuffers cb = pb;
    cb.flag = A.flag;
    cb.ulist = A.ulist;
    cb.units1 = A.units1;
    cb.units2 = A.units2;
    launch_init(&view, cb, s);
    launch_prep(&view, cb, p, s);
    launch_plan(&view, cb, s);
    SGC_HIP(hipEventRecord(ctx->ev[1], s));
    launch_similarity(&view, cb, p, ctx->cu_count, ctx->sim_impl, ctx->ev[7], s);
    SGC_HIP(hipEventRecord(ctx->ev[2], s));
    SGC_HIP(hipGetLastError());
    SGC_HIP(hipMemcpyAsync(R.h_flag[i & 1], A.flag, c.m_chip, hipMemcpyDeviceToHost, s));
    SGC_HIP(hipMemcpyAsync(R.h_cnt[i & 1], pb.cnt, kCntSlots * 8, hipMemcpyDeviceToHost, s));
    SGC_HIP(hipEventRecord(R.done[i & 1], s));
    R.bytes_d2h += (double)c.m_chip;
    // overlap: merge the previous chip and assemble the next one on the host
    if (i > 0) {
      rc = merge(i - 1);
      if (rc) return rc;
    }
    if (i + 1 < nsim) build(i + 1);
    SGC_HIP(hipEventSynchronize(ctx->ev[2]));
    ms_sim += elapsed(ctx->ev[1], ctx->ev[2]);
    ms_sim_large += c.m_own ? elapsed(ctx->ev[1], ctx->ev[7]) : 0.0;
  }
  if (g->m && nsim) {
    rc = merge(nsim - 1);
    if (rc) return rc;
  }

  // ---- core check, union-find, border, classify over row chips
  const double t_phase = now_ms();
  const size_t nrow = row_bounds.size() - 1;
  int32_t* c_adj = (int32_t*)R.chip;
  Layout RL;
  const size_t ro_adj = RL.take(std::max<uint64_t>(row_max, 1) * 4), ro_src = RL.take(std::max<uint64_t>(row_max, 1) * 4),
               ro_flag = RL.take(std::max<uint64_t>(row_max, 1));
  (void)ro_adj;
  uint32_t* c_src = (uint32_t*)(R.chip + ro_src);
  uint8_t* c_flag = (uint8_t*)(R.chip + ro_flag);
  sgc_dgraph whole;
  whole.ctx = ctx;
  whole.n = n;
  whole.m = 0;
  launch_init(&whole, pb, s);
  size_t loaded = (size_t)-1;
  auto load_rows = [&](size_t i, sgc_dgraph* v) -> int {
    const uint64_t r0 = row_bounds[i], r1 = row_bounds[i + 1];
    const uint64_t b0 = g->row_ptr[r0], mc = g->row_ptr[r1] - b0;
    v->ctx = ctx;
    v->n = n;
    v->m = (uint32_t)mc;
    v->adj = c_adj;
    v->src = c_src;
    if (loaded == i) return SGC_OK;
    const uint32_t nr = (uint32_t)(r1 - r0);
    uint32_t* ho = R.h_off[0];
    SGC_HIP(hipStreamSynchronize(s));  // staging and chip buffers are reused
    for (uint32_t k = 0; k <= nr; ++k) ho[k] = (uint32_t)(g->row_ptr[r0 + k] - b0);
    SGC_HIP(hipMemcpyAsync(R.off, ho, ((size_t)nr + 1) * 4, hipMemcpyHostToDevice, s));
    if (mc) {
      SGC_HIP(hipMemcpyAsync(c_adj, g->adj + b0, mc * 4, hipMemcpyHostToDevice, s));
      SGC_HIP(hipMemcpyAsync(c_flag, R.flag + b0, mc, hipMemcpyHostToDevice, s));
    }
    R.bytes_h2d += (double)mc * 5 + (double)(nr + 1) * 4;
    if (nr) k_fill_src_rows<<<grid_for(nr, kBlock / kWave), kBlock, 0, s>>>((uint32_t)r0, nr, R.off, c_src);
    loaded = i;
    return SGC_OK;
  };
  PassBuffers rb = pb;
  rb.flag = c_flag;
  sgc_dgraph v;
  double ms_core = 0, ms_cluster = 0, ms_border = 0, ms_classify = 0;
  SGC_HIP(hipEventRecord(ctx->ev[3], s));
  for (size_t i = 0; i < nrow; ++i) {
    if ((rc = load_rows(i, &v))) return rc;
    launch_core_count(&v, rb, s);
  }
  launch_core_mark(&whole, rb, p, s);
  SGC_HIP(hipEventRecord(ctx->ev[4], s));
  for (size_t i = 0; i < nrow; ++i) {
    if ((rc = load_rows(nrow - 1 - i, &v))) return rc;  // last chip first: it is still loaded
    launch_union(&v, rb, s);
  }
  launch_flatten(&whole, rb, s);
  SGC_HIP(hipEventRecord(ctx->ev[5], s));
  for (size_t i = 0; i < nrow; ++i) {
    if ((rc = load_rows(i, &v))) return rc;
    launch_border(&v, rb, s);
  }
  SGC_HIP(hipEventRecord(ctx->ev[6], s));
  for (size_t i = 0; i < nrow; ++i) {
    if ((rc = load_rows(nrow - 1 - i, &v))) return rc;
    launch_classify_edges(&v, rb, s);
    launch_count_similar(&v, rb, s);
  }
  launch_finalize(&whole, rb, p, s);
  SGC_HIP(hipEventRecord(ctx->ev[0], s));
  SGC_HIP(hipGetLastError());
  SGC_HIP(hipMemcpyAsync(R.h_cnt[0], pb.cnt, kCntSlots * 8, hipMemcpyDeviceToHost, s));
  if (labels) SGC_HIP(hipMemcpyAsync(labels, pb.label, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  if (core_out) SGC_HIP(hipMemcpyAsync(core_out, pb.core, n, hipMemcpyDeviceToHost, s));
  SGC_HIP(hipStreamSynchronize(s));
  R.bytes_d2h += (labels ? (double)n * 4 : 0) + (core_out ? (double)n : 0);
  if (nrow == 1) {
    ms_core = elapsed(ctx->ev[3], ctx->ev[4]);
    ms_cluster = elapsed(ctx->ev[4], ctx->ev[5]);
    ms_border = elapsed(ctx->ev[5], ctx->ev[6]);
    ms_classify = elapsed(ctx->ev[6], ctx->ev[0]);
  }
  const double ms_phases = now_ms() - t_phase;
  if (sim_out && g->m) memcpy(sim_out, R.flag, g->m);

  const uint64_t* c = R.h_cnt[0];
  if (st) {
    memset(st, 0, sizeof(*st));
    st->ms_total = now_ms() - t_start;
    st->ms_similarity = ms_sim;
    st->ms_similarity_large = ms_sim_large;
    st->ms_core = ms_core;
    st->ms_cluster = ms_cluster;
    st->ms_border = ms_border;
    st->ms_classify = ms_classify;
    st->edges = g->m / 2;
    st->u_edges = totals[kCntUEdges];
    st->similar_edges = c[kCntSimilar] / 2;
    st->cores = c[kCntCores];
    st->clusters = c[kCntClusters];
    st->borders = c[kCntBorders];
    st->hubs = c[kCntHubs];
    st->outliers = c[kCntOutliers];
    st->probe_elements = (double)totals[kCntProbes];
    st->loaded_elements = (double)totals[kCntLoaded];
    st->claims = (double)totals[kCntClaims];
  }
  if (oi) {
    memset(oi, 0, sizeof(*oi));
    oi->device_bytes = resident + chip_need;
    oi->similarity_chips = (uint32_t)nsim;
    oi->row_chips = (uint32_t)nrow;
    oi->bytes_h2d = R.bytes_h2d;
    oi->bytes_d2h = R.bytes_d2h;
    oi->ms_host = R.ms_host;
    oi->ms_phases = ms_phases;
  }
  return SGC_OK;
}

}  // namespace
}  // namespace sgc

extern "C" int sgc_ooc_min_budget(const sgc_host_graph* g, double eps, uint64_t* bytes) {
  using namespace sgc;
  clear_error();
  if (!g || !bytes) return fail(SGC_EINVAL, "null argument");
  if (g->n >= (1ull << 31)) return fail(SGC_ERANGE, "n=%llu: vertex ids must fit int32", (unsigned long long)g->n);
  std::vector<uint32_t> deg(g->n);
  ooc_degrees(g, deg.data());
  uint64_t maxd = 0;
  for (uint64_t x = 0; x < g->n; ++x) maxd = std::max<uint64_t>(maxd, deg[x]);
  const uint64_t cost = ooc_max_row_cost(g, deg.data(), eps);
  const uint64_t chip = std::max<uint64_t>(cost * kSimChipBytesPerEntry, maxd * kRowChipBytesPerEntry);
  *bytes = resident_bytes((uint32_t)g->n) + 4096 + chip + 4096;
  return SGC_OK;
}

extern "C" int sgc_scan_ooc(sgc_ctx* ctx, const sgc_host_graph* g, double eps, int32_t mu, uint64_t device_budget,
                            int32_t* labels, uint8_t* core, uint8_t* sim, sgc_stats* stats, sgc_ooc_stats* ooc) {
  using namespace sgc;
  clear_error();
  if (!ctx || !g) return fail(SGC_EINVAL, "null argument");
  if (g->n >= (1ull << 31)) return fail(SGC_ERANGE, "n=%llu: vertex ids must fit int32", (unsigned long long)g->n);
  SGC_HIP(hipSetDevice(ctx->device));
  OocRun R;
  R.ctx = ctx;
  R.g = g;
  R.n = (uint32_t)g->n;
  return run_ooc(R, eps, mu, device_budget, labels, core, sim, stats, ooc);
}
