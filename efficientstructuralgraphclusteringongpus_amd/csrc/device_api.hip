// C-ABI: device contexts, device-resident graphs and the plain SCAN pass.
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "GPUScan.h"
#include "device.h"
#include "graph_host.h"

using namespace sgc;

// Work memory of sgc_scan_device_noalloc: one copy per GPU, held from the
// library's load on that GPU until the process ends.  Host code never uses
// this name as an address: the device address of each GPU's copy is taken
// once per device by hipGetSymbolAddress (scratch_ptr_for) and passed on.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current; looked up
// once per device and kept.
int scratch_ptr_for(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return fail(SGC_EINVAL, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SGC_EDEVICE, "address of the static scratch on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return SGC_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int ensure(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return SGC_OK;
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
  hipError_t e = hipMalloc(&b.ptr, bytes);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    (void)hipGetLastError();
    return fail(SGC_ENOMEM, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  return SGC_OK;
}

// Byte offsets of one pass's arrays for graph g inside a work buffer; `bytes`
// is the buffer size they need.
struct PassLayout {
  size_t flag, sd, parent, label, lo, hi, core, cnt_u, work_u, uedge, units, bytes;
};

PassLayout pass_layout(const sgc_dgraph* g) {
  const size_t n = g->n, m = g->m;
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    size_t o = off;
    off = align_up(off + std::max<size_t>(bytes, 1), 256);
    return o;
  };
  PassLayout L;
  L.flag = take(m);
  L.sd = take(n * 4);
  L.parent = take(n * 4);
  L.label = take(n * 4);
  L.lo = take(n * 4);
  L.hi = take(n * 4);
  L.core = take(n);
  L.cnt_u = take(n * 4);
  L.work_u = take(n * 8);
  L.uedge = take(m * sizeof(uint4));
  L.units = take(unit_capacity(m) * sizeof(uint4));
  L.bytes = off;
  return L;
}

void bind_pass(char* base, const PassLayout& L, const sgc_dgraph* g, uint64_t* cnt, PassBuffers* pb) {
  pb->flag = (uint8_t*)(base + L.flag);
  pb->sd = (uint32_t*)(base + L.sd);
  pb->parent = (int32_t*)(base + L.parent);
  pb->label = (int32_t*)(base + L.label);
  pb->lo = (int32_t*)(base + L.lo);
  pb->hi = (int32_t*)(base + L.hi);
  pb->core = (uint8_t*)(base + L.core);
  pb->cnt_u = (uint32_t*)(base + L.cnt_u);
  pb->work_u = (uint64_t*)(base + L.work_u);
  pb->uedge = (uint4*)(base + L.uedge);
  pb->units = (uint4*)(base + L.units);
  pb->tiny = (uint32_t*)pb->label;
  pb->ny = (uint32_t*)pb->lo;
  pb->nn = (uint32_t*)pb->hi;
  pb->ucap = (uint32_t)unit_capacity(g->m);
  pb->cnt = cnt;
}

// Carves the per-pass arrays out of the context workspace (grown by
// hipMalloc when too small).
int carve(sgc_ctx* ctx, const sgc_dgraph* g, PassBuffers* pb) {
  const PassLayout L = pass_layout(g);
  int rc = ensure(ctx->work, L.bytes);
  if (rc) return rc;
  bind_pass((char*)ctx->work.ptr, L, g, (uint64_t*)ctx->counters.ptr, pb);
  return SGC_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

}  // namespace

int sgc::carve_pass_buffers(sgc_ctx* ctx, const sgc_dgraph* g, PassBuffers* pb) { return carve(ctx, g, pb); }

extern "C" int sgc_ctx_create(int device, sgc_ctx** out) {
  clear_error();
  if (!out) return fail(SGC_EINVAL, "null argument");
  *out = nullptr;
  int count = 0;
  SGC_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(SGC_EINVAL, "device %d not present (%d devices)", device, count);
  SGC_HIP(hipSetDevice(device));
  void* scratch_ptr = nullptr;
  if (int rc = scratch_ptr_for(device, &scratch_ptr)) return rc;
  sgc_ctx* ctx = new sgc_ctx();
  ctx->device = device;
  ctx->scratch = scratch_ptr;
  hipDeviceProp_t prop;
  SGC_HIP(hipGetDeviceProperties(&prop, device));
  strncpy(ctx->name, prop.name[0] ? prop.name : prop.gcnArchName, sizeof(ctx->name) - 1);
  ctx->cu_count = prop.multiProcessorCount;
  SGC_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  int rc = ensure(ctx->counters, kCntSlots * sizeof(uint64_t));
  if (rc) {
    delete ctx;
    return rc;
  }
  SGC_HIP(hipHostMalloc((void**)&ctx->host_counters, kCntSlots * sizeof(uint64_t), hipHostMallocDefault));
  for (auto& e : ctx->ev) SGC_HIP(hipEventCreate(&e));
  *out = ctx;
  return SGC_OK;
}

extern "C" void sgc_ctx_destroy(sgc_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->work.ptr) (void)hipFree(ctx->work.ptr);
  if (ctx->counters.ptr) (void)hipFree(ctx->counters.ptr);
  if (ctx->host_counters) (void)hipHostFree(ctx->host_counters);
  for (int i = 0; i < 2; ++i) {
    if (ctx->bounce[i]) (void)hipHostFree(ctx->bounce[i]);
    if (ctx->bounce_done[i]) (void)hipEventDestroy(ctx->bounce_done[i]);
  }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" int sgc_ctx_sync(sgc_ctx* ctx) {
  clear_error();
  if (!ctx) return fail(SGC_EINVAL, "null context");
  SGC_HIP(hipSetDevice(ctx->device));
  SGC_HIP(hipStreamSynchronize(ctx->stream));
  return SGC_OK;
}

extern "C" const char* sgc_ctx_device_name(const sgc_ctx* ctx) { return ctx ? ctx->name : ""; }

extern "C" int sgc_ctx_memory(sgc_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
  clear_error();
  if (!ctx) return fail(SGC_EINVAL, "null context");
  SGC_HIP(hipSetDevice(ctx->device));
  size_t f = 0, t = 0;
  SGC_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return SGC_OK;
}

extern "C" int sgc_ctx_set_similarity_kernel(sgc_ctx* ctx, int kind) {
  clear_error();
  if (!ctx) return fail(SGC_EINVAL, "null context");
  if (kind != SGC_SIM_KERNEL_TIERED && kind != SGC_SIM_KERNEL_WAVE_PER_EDGE)
    return fail(SGC_EINVAL, "unknown similarity kernel %d", kind);
  ctx->sim_impl = kind == SGC_SIM_KERNEL_TIERED ? kSimTiered : kSimWavePerEdge;
  return SGC_OK;
}

extern "C" int sgc_ctx_set_pruning(sgc_ctx* ctx, int enable) {
  clear_error();
  if (!ctx) return fail(SGC_EINVAL, "null context");
  ctx->prune = enable != 0;
  return SGC_OK;
}

namespace {
constexpr size_t kBounceBytes = 64u << 20;

// Host-to-device copy on the context stream.  A pageable source goes through
// the context's two page-locked buffers: the next chunk is copied in while
// the previous one is sent (the runtime's own staging of a pageable copy
// runs about a third slower).  The buffers are free again on return.
hipError_t upload_bytes(sgc_ctx* ctx, void* dst, const void* src, size_t bytes, bool src_pinned) {
  if (!bytes) return hipSuccess;
  if (!src_pinned && !ctx->bounce[0]) {
    bool ok = true;
    for (int i = 0; i < 2 && ok; ++i)
      ok = hipHostMalloc(&ctx->bounce[i], kBounceBytes) == hipSuccess &&
           hipEventCreateWithFlags(&ctx->bounce_done[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {  // the runtime stages the copy instead
      (void)hipGetLastError();
      for (int i = 0; i < 2; ++i) {
        if (ctx->bounce[i]) (void)hipHostFree(ctx->bounce[i]);
        if (ctx->bounce_done[i]) (void)hipEventDestroy(ctx->bounce_done[i]);
        ctx->bounce[i] = nullptr;
        ctx->bounce_done[i] = nullptr;
      }
    }
  }
  if (src_pinned || !ctx->bounce[0]) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  bool used[2] = {false, false};
  hipError_t e = hipSuccess;
  for (size_t off = 0, i = 0; off < bytes && e == hipSuccess; off += kBounceBytes, ++i) {
    const int b = (int)(i & 1);
    const size_t len = std::min(kBounceBytes, bytes - off);
    if (used[b]) e = hipEventSynchronize(ctx->bounce_done[b]);  // its last chunk is on the device
    if (e != hipSuccess) break;
    parallel_copy(ctx->bounce[b], (const char*)src + off, len);
    e = hipMemcpyAsync((char*)dst + off, ctx->bounce[b], len, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->bounce_done[b], ctx->stream);
    used[b] = true;
  }
  for (int b = 0; b < 2; ++b)
    if (used[b]) {
      const hipError_t w = hipEventSynchronize(ctx->bounce_done[b]);
      if (e == hipSuccess) e = w;
    }
  return e;
}
}  // namespace

extern "C" int sgc_dgraph_upload(sgc_ctx* ctx, const sgc_host_graph* g, sgc_dgraph** out) {
  clear_error();
  if (!ctx || !g || !out) return fail(SGC_EINVAL, "null argument");
  *out = nullptr;
  if (g->m >= (1ull << 32) || g->n >= (1ull << 31))
    return fail(SGC_ERANGE, "graph with n=%llu, m=%llu exceeds the in-memory path (m < 2^32); use sgc_scan_ooc",
                (unsigned long long)g->n, (unsigned long long)g->m);
  SGC_HIP(hipSetDevice(ctx->device));
  sgc_dgraph* d = new sgc_dgraph();
  d->ctx = ctx;
  d->n = (uint32_t)g->n;
  d->m = (uint32_t)g->m;
  std::vector<uint32_t> rp32(g->n + 1);
  uint64_t maxd = 0;
  narrow_row_ptr(g, rp32.data(), &maxd);
  d->max_degree = (uint32_t)maxd;
  auto cleanup = [d]() {
    if (d->mirror) (void)hipFree(d->mirror);
    if (d->ndeg) (void)hipFree(d->ndeg);
    if (d->woff) (void)hipFree(d->woff);
    if (d->row_ptr) (void)hipFree(d->row_ptr);
    if (d->adj) (void)hipFree(d->adj);
    if (d->src) (void)hipFree(d->src);
    delete d;
  };
  const size_t m_bytes = std::max<size_t>(g->m, 1) * 4;
  if (hipMalloc(&d->row_ptr, (g->n + 1) * 4) != hipSuccess || hipMalloc(&d->adj, m_bytes) != hipSuccess ||
      hipMalloc(&d->src, m_bytes) != hipSuccess) {
    (void)hipGetLastError();
    cleanup();
    return fail(SGC_ENOMEM, "device allocation for graph (n=%llu, m=%llu) failed", (unsigned long long)g->n,
                (unsigned long long)g->m);
  }
  d->row_begin = d->row_ptr;
  d->row_end = d->row_ptr + 1;
  // reverse-entry index: the similarity kernels write both entries of an
  // edge without searching for the second
  if (g->m && hipMalloc(&d->mirror, m_bytes) != hipSuccess) {
    (void)hipGetLastError();
    d->mirror = nullptr;  // the kernels find the reverse entry themselves
  }
  // neighbour degrees, streamed by the prep kernel on every pass
  if (g->m && hipMalloc(&d->ndeg, m_bytes) != hipSuccess) {
    (void)hipGetLastError();
    d->ndeg = nullptr;  // the prep kernel gathers them
  }
  // window-offset index when there is more than one window and it takes at
  // most four times the adjacency's bytes (it grows with n^2: R-MAT-26 at
  // edge factor 16 needs 17 GB, twice its adjacency, and a pass there runs
  // 2.5 times faster with it)
  const uint64_t woff_bytes = window_index_bytes(d->n, g->m);
  if (woff_bytes) {
    if (hipMalloc(&d->woff, woff_bytes) == hipSuccess) {
      d->wcount = window_count(d->n);
    } else {
      (void)hipGetLastError();
      d->woff = nullptr;  // the kernels find window boundaries while streaming instead
    }
  }
  hipError_t e = upload_bytes(ctx, d->row_ptr, rp32.data(), (g->n + 1) * 4, false);
  if (e == hipSuccess && g->m) e = upload_bytes(ctx, d->adj, g->adj, g->m * 4, g->pinned);
  if (e == hipSuccess) {
    launch_fill_src(d, ctx->stream);
    e = build_mirror(d, ctx->stream);
  }
  if (e == hipSuccess) {
    launch_neighbor_degrees(d, ctx->stream);
    launch_window_offsets(d, ctx->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    cleanup();
    return fail(SGC_EDEVICE, "graph upload failed: %s", hipGetErrorString(e));
  }
  *out = d;
  return SGC_OK;
}

extern "C" void sgc_dgraph_free(sgc_dgraph* d) {
  if (!d) return;
  (void)hipSetDevice(d->ctx->device);
  (void)hipStreamSynchronize(d->ctx->stream);
  if (d->mirror) (void)hipFree(d->mirror);
  if (d->ndeg) (void)hipFree(d->ndeg);
  if (d->woff) (void)hipFree(d->woff);
  if (d->row_ptr) (void)hipFree(d->row_ptr);
  if (d->adj) (void)hipFree(d->adj);
  if (d->src) (void)hipFree(d->src);
  delete d;
}

extern "C" int sgc_dgraph_get_tuning(const sgc_dgraph* d, sgc_tuning* t) {
  clear_error();
  if (!d || !t) return fail(SGC_EINVAL, "null argument");
  memset(t, 0, sizeof(*t));
  t->small_max_degree = d->tuning.small_max_degree;
  t->small_unit_work = d->tuning.small_unit_work;
  t->large_unit_work = d->tuning.large_unit_work;
  t->hash_tables = d->tuning.hash_tables;
  t->tiny_work = d->tuning.tiny_work;
  return SGC_OK;
}

extern "C" int sgc_dgraph_set_tuning(sgc_dgraph* d, const sgc_tuning* t) {
  clear_error();
  if (!d || !t) return fail(SGC_EINVAL, "null argument");
  if (t->small_max_degree < 1 || t->small_max_degree > 1024 || !(t->small_unit_work >= 64.0) ||
      !(t->large_unit_work >= 64.0) || t->hash_tables > 1 || t->tiny_work > 65536)
    return fail(SGC_EINVAL,
                "tuning out of range (small_max_degree 1..1024, unit work >= 64, hash_tables 0/1, tiny_work <= 65536)");
  d->tuning.small_max_degree = t->small_max_degree;
  d->tuning.small_unit_work = t->small_unit_work;
  d->tuning.large_unit_work = t->large_unit_work;
  d->tuning.hash_tables = t->hash_tables;
  d->tuning.tiny_work = t->tiny_work;
  return SGC_OK;
}

extern "C" uint64_t sgc_dgraph_n(const sgc_dgraph* d) { return d ? d->n : 0; }
extern "C" uint64_t sgc_dgraph_m(const sgc_dgraph* d) { return d ? d->m : 0; }

namespace sgc {
int check_unit_overflow(const uint64_t* c, uint64_t m) {
  if (!c[kCntOverflow]) return SGC_OK;
  return fail(SGC_EFORMAT,
              "the graph is not symmetric: %llu undecided entries are owned, more than the %llu work units a "
              "graph of %llu entries with every reverse entry can need",
              (unsigned long long)c[kCntOverflow], (unsigned long long)unit_capacity(m), (unsigned long long)m);
}
}  // namespace sgc

namespace {

// One plain pass over g with the arrays in *pb (carved by the caller).
// want_flags: the caller takes the similarity flags, so every edge is
// decided (no core-check pruning).
int run_pass(sgc_ctx* ctx, sgc_dgraph* g, double eps, int32_t mu, PassBuffers* pb, sgc_stats* st,
             bool want_flags) {
  SGC_HIP(hipSetDevice(ctx->device));
  // The reference's pruned core check (GPUSCAN++/GPUScan.cu:303) needs
  // mu >= 1 (for mu <= 0 no vertex is ever decided not core) and the
  // reverse-entry index of a resident graph.
  const bool prune = ctx->prune && mu >= 1 && !want_flags && g->mirror && ctx->sim_impl == kSimTiered;
  const ScanParams p{eps, mu, prune};
  hipStream_t s = ctx->stream;
  SGC_HIP(hipEventRecord(ctx->ev[0], s));
  launch_init(g, *pb, s);
  launch_prep(g, *pb, p, s);
  launch_plan(g, *pb, p, s);
  SGC_HIP(hipEventRecord(ctx->ev[1], s));
  launch_similarity(g, *pb, p, ctx->cu_count, ctx->sim_impl, ctx->ev[7], s);
  SGC_HIP(hipEventRecord(ctx->ev[2], s));
  launch_core(g, *pb, p, s);
  SGC_HIP(hipEventRecord(ctx->ev[3], s));
  launch_union(g, *pb, s);
  // deferred edges the clusters or borders need, decided now
  if (p.prune) launch_redo(g, *pb, p, ctx->cu_count, s);
  launch_flatten(g, *pb, s);
  SGC_HIP(hipEventRecord(ctx->ev[4], s));
  launch_border(g, *pb, s);
  SGC_HIP(hipEventRecord(ctx->ev[5], s));
  launch_classify(g, *pb, p, s);
  SGC_HIP(hipEventRecord(ctx->ev[6], s));
  SGC_HIP(hipGetLastError());
  SGC_HIP(hipMemcpyAsync(ctx->host_counters, pb->cnt, kCntSlots * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SGC_HIP(hipStreamSynchronize(s));
  if (int rc = check_unit_overflow(ctx->host_counters, g->m)) return rc;
  if (st) {
    memset(st, 0, sizeof(*st));
    st->ms_prep = elapsed(ctx->ev[0], ctx->ev[1]);
    st->ms_similarity = elapsed(ctx->ev[1], ctx->ev[2]);
    st->ms_similarity_large = g->m ? elapsed(ctx->ev[1], ctx->ev[7]) : 0.0;
    st->ms_core = elapsed(ctx->ev[2], ctx->ev[3]);
    st->ms_cluster = elapsed(ctx->ev[3], ctx->ev[4]);
    st->ms_border = elapsed(ctx->ev[4], ctx->ev[5]);
    st->ms_classify = elapsed(ctx->ev[5], ctx->ev[6]);
    st->ms_total = elapsed(ctx->ev[0], ctx->ev[6]);
    const uint64_t* c = ctx->host_counters;
    st->edges = g->m / 2;
    st->u_edges = c[kCntUEdges];
    st->similar_edges = c[kCntSimilar] / 2;
    st->cores = c[kCntCores];
    st->clusters = c[kCntClusters];
    st->borders = c[kCntBorders];
    st->hubs = c[kCntHubs];
    st->outliers = c[kCntOutliers];
    st->probe_elements = (double)c[kCntProbes];
    st->loaded_elements = (double)c[kCntLoaded];
    st->claims = (double)c[kCntClaims];
    st->redo_edges = c[kCntRedo];
    st->deferred_edges = c[kCntDeferred] - c[kCntRedo];
    st->pruned = prune ? 1 : 0;
#ifdef SGC_PHASE_TIMERS
    fprintf(stderr, "phase ticks (unit, window, stream, barrier, final) large: %llu %llu %llu %llu %llu small: "
            "%llu %llu %llu %llu %llu\n", (unsigned long long)c[kCntTimer0], (unsigned long long)c[kCntTimer0 + 1],
            (unsigned long long)c[kCntTimer0 + 2], (unsigned long long)c[kCntTimer0 + 3],
            (unsigned long long)c[kCntTimer0 + 4], (unsigned long long)c[kCntTimer0 + 5],
            (unsigned long long)c[kCntTimer0 + 6], (unsigned long long)c[kCntTimer0 + 7],
            (unsigned long long)c[kCntTimer0 + 8], (unsigned long long)c[kCntTimer0 + 9]);
#endif
  }
  return SGC_OK;
}

}  // namespace

namespace {

// Similarity-phase time of one pass (prep, plan and both tiers are run; the
// similarity kernels alone are timed).
int time_similarity(sgc_ctx* ctx, sgc_dgraph* g, double eps, PassBuffers* pb, double* ms) {
  int rc = carve(ctx, g, pb);
  if (rc) return rc;
  const ScanParams p{eps, 2};
  hipStream_t s = ctx->stream;
  launch_init(g, *pb, s);
  launch_prep(g, *pb, p, s);
  launch_plan(g, *pb, p, s);
  SGC_HIP(hipEventRecord(ctx->ev[1], s));
  launch_similarity(g, *pb, p, ctx->cu_count, kSimTiered, nullptr, s);
  SGC_HIP(hipEventRecord(ctx->ev[2], s));
  SGC_HIP(hipGetLastError());
  SGC_HIP(hipMemcpyAsync(ctx->host_counters, pb->cnt, kCntSlots * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SGC_HIP(hipStreamSynchronize(s));
  if (int rc = check_unit_overflow(ctx->host_counters, g->m)) return rc;
  *ms = elapsed(ctx->ev[1], ctx->ev[2]);
  return SGC_OK;
}

}  // namespace

extern "C" int sgc_dgraph_autotune(sgc_ctx* ctx, sgc_dgraph* g, double eps, sgc_tuning* best) {
  clear_error();
  if (!ctx || !g) return fail(SGC_EINVAL, "null argument");
  SGC_HIP(hipSetDevice(ctx->device));
  PassBuffers pb;
  uint32_t trials = 0;
  auto measure = [&](const SimTuning& t, double* ms) -> int {
    g->tuning = t;
    double a = 0, b = 0;
    int rc = time_similarity(ctx, g, eps, &pb, &a);
    if (!rc) rc = time_similarity(ctx, g, eps, &pb, &b);
    trials += 2;
    *ms = std::min(a, b);
    return rc;
  };
  const SimTuning defaults;
  double ms_default = 0;
  int rc = measure(defaults, &ms_default);
  if (rc) return rc;
  SimTuning cur = defaults;
  double cur_ms = ms_default;
  const double large_work[] = {2097152.0, 8388608.0, 33554432.0};
  const uint32_t small_deg[] = {128, 256, 512, 1024};
  const double small_work[] = {65536.0, 262144.0, 1048576.0};
  for (double v : large_work) {
    SimTuning t = cur;
    t.large_unit_work = v;
    double ms;
    if (v == cur.large_unit_work) continue;
    if ((rc = measure(t, &ms))) return rc;
    if (ms < cur_ms) cur = t, cur_ms = ms;
  }
  for (uint32_t v : small_deg) {
    SimTuning t = cur;
    t.small_max_degree = v;
    double ms;
    if (v == cur.small_max_degree) continue;
    if ((rc = measure(t, &ms))) return rc;
    if (ms < cur_ms) cur = t, cur_ms = ms;
  }
  for (double v : small_work) {
    SimTuning t = cur;
    t.small_unit_work = v;
    double ms;
    if (v == cur.small_unit_work) continue;
    if ((rc = measure(t, &ms))) return rc;
    if (ms < cur_ms) cur = t, cur_ms = ms;
  }
  {
    SimTuning t = cur;
    t.hash_tables = 1 - cur.hash_tables;
    double ms;
    if ((rc = measure(t, &ms))) return rc;
    if (ms < cur_ms) cur = t, cur_ms = ms;
  }
  const uint32_t tiny_work[] = {0, 1024, 2048, 4096};
  for (uint32_t v : tiny_work) {
    SimTuning t = cur;
    t.tiny_work = v;
    double ms;
    if (v == cur.tiny_work) continue;
    if ((rc = measure(t, &ms))) return rc;
    if (ms < cur_ms) cur = t, cur_ms = ms;
  }
  g->tuning = cur;
  if (best) {
    memset(best, 0, sizeof(*best));
    best->small_max_degree = cur.small_max_degree;
    best->small_unit_work = cur.small_unit_work;
    best->large_unit_work = cur.large_unit_work;
    best->hash_tables = cur.hash_tables;
    best->tiny_work = cur.tiny_work;
    best->ms_similarity = cur_ms;
    best->ms_default = ms_default;
    best->trials = trials;
  }
  return SGC_OK;
}

extern "C" int sgc_scan_device(sgc_ctx* ctx, sgc_dgraph* g, double eps, int32_t mu, int32_t* d_labels,
                               uint8_t* d_core, uint8_t* d_sim, sgc_stats* st) {
  clear_error();
  if (!ctx || !g) return fail(SGC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(SGC_EINVAL, "graph belongs to another context");
  SGC_HIP(hipSetDevice(ctx->device));
  PassBuffers pb;
  int rc = carve(ctx, g, &pb);
  if (!rc) rc = run_pass(ctx, g, eps, mu, &pb, st, d_sim != nullptr);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  if (d_labels) SGC_HIP(hipMemcpyAsync(d_labels, pb.label, (size_t)g->n * 4, hipMemcpyDeviceToDevice, s));
  if (d_core) SGC_HIP(hipMemcpyAsync(d_core, pb.core, g->n, hipMemcpyDeviceToDevice, s));
  if (d_sim && g->m) SGC_HIP(hipMemcpyAsync(d_sim, pb.flag, g->m, hipMemcpyDeviceToDevice, s));
  SGC_HIP(hipStreamSynchronize(s));
  return SGC_OK;
}

extern "C" int sgc_scan(sgc_ctx* ctx, sgc_dgraph* g, double eps, int32_t mu, int32_t* labels, uint8_t* core,
                        uint8_t* sim, sgc_stats* st) {
  clear_error();
  if (!ctx || !g) return fail(SGC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(SGC_EINVAL, "graph belongs to another context");
  SGC_HIP(hipSetDevice(ctx->device));
  PassBuffers pb;
  int rc = carve(ctx, g, &pb);
  if (!rc) rc = run_pass(ctx, g, eps, mu, &pb, st, sim != nullptr);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  if (labels) SGC_HIP(hipMemcpyAsync(labels, pb.label, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
  if (core) SGC_HIP(hipMemcpyAsync(core, pb.core, g->n, hipMemcpyDeviceToHost, s));
  if (sim && g->m) SGC_HIP(hipMemcpyAsync(sim, pb.flag, g->m, hipMemcpyDeviceToHost, s));
  SGC_HIP(hipStreamSynchronize(s));
  return SGC_OK;
}

extern "C" uint64_t sgc_noalloc_max_bytes(void) { return kScratchBytes; }

extern "C" uint64_t sgc_noalloc_bytes(const sgc_dgraph* g) { return g ? pass_layout(g).bytes : 0; }

// sgc_scan_device with the pass's work arrays in the static scratch of the
// context's GPU instead of the context workspace: no hipMalloc, hipFree or
// allocating library call on any path.  Calls on one GPU share its scratch
// (contexts included) and must not overlap; every array a pass reads before
// writing is cleared by the pass itself (k_init).
extern "C" int sgc_scan_device_noalloc(sgc_ctx* ctx, sgc_dgraph* g, double eps, int32_t mu, int32_t* d_labels,
                                       uint8_t* d_core, uint8_t* d_sim, sgc_stats* st) {
  clear_error();
  if (!ctx || !g) return fail(SGC_EINVAL, "null argument");
  if (g->ctx != ctx) return fail(SGC_EINVAL, "graph belongs to another context");
  if (!ctx->scratch) return fail(SGC_EDEVICE, "no static scratch on device %d", ctx->device);
  const PassLayout L = pass_layout(g);
  if (L.bytes > kScratchBytes)
    return fail(SGC_ERANGE,
                "a pass over n=%u, m=%u needs %zu bytes of work memory, more than the %zu bytes of static scratch "
                "behind sgc_scan_device_noalloc; use sgc_scan_device",
                g->n, g->m, L.bytes, kScratchBytes);
  SGC_HIP(hipSetDevice(ctx->device));
  PassBuffers pb;
  bind_pass((char*)ctx->scratch, L, g, (uint64_t*)ctx->counters.ptr, &pb);
  int rc = run_pass(ctx, g, eps, mu, &pb, st, d_sim != nullptr);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  if (d_labels) SGC_HIP(hipMemcpyAsync(d_labels, pb.label, (size_t)g->n * 4, hipMemcpyDeviceToDevice, s));
  if (d_core) SGC_HIP(hipMemcpyAsync(d_core, pb.core, g->n, hipMemcpyDeviceToDevice, s));
  if (d_sim && g->m) SGC_HIP(hipMemcpyAsync(d_sim, pb.flag, g->m, hipMemcpyDeviceToDevice, s));
  SGC_HIP(hipStreamSynchronize(s));
  return SGC_OK;
}

extern "C" int sgc_copy_bench(sgc_ctx* ctx, uint64_t bytes, double* ms) {
  clear_error();
  if (!ctx || bytes < 16) return fail(SGC_EINVAL, "bad argument");
  SGC_HIP(hipSetDevice(ctx->device));
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (a) (void)hipFree(a);
    return fail(SGC_ENOMEM, "device allocation of 2 x %llu bytes failed", (unsigned long long)bytes);
  }
  hipError_t e = hipMemsetAsync(a, 1, bytes, ctx->stream);
  if (e == hipSuccess) {
    launch_copy(a, b, bytes, ctx->stream);  // warm
    e = hipEventRecord(ctx->ev[0], ctx->stream);
  }
  if (e == hipSuccess) {
    launch_copy(a, b, bytes, ctx->stream);
    e = hipEventRecord(ctx->ev[1], ctx->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && ms) *ms = elapsed(ctx->ev[0], ctx->ev[1]);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(SGC_EDEVICE, "copy benchmark failed: %s", hipGetErrorString(e));
  return SGC_OK;
}

// ---------------------------------------------------------------------------
// Reference-compatible entry points
// ---------------------------------------------------------------------------

namespace {
std::mutex g_compat_mu;
sgc_ctx* g_compat_ctx = nullptr;

// The reference's error contract (GPUSCAN++/GPUScan.cu:24-32): a
// "GPUassert: <message> <file> <line>" line on stderr, then exit with a
// non-zero code.  The library's own entry points return status codes; only
// the two drop-ins below end the process.
[[noreturn]] void compat_abort(int code, int line) {
  fprintf(stderr, "GPUassert: %s %s %d\n", sgc_last_error(), __FILE__, line);
  fflush(stderr);
  exit(code ? code : 1);
}

#define SGC_COMPAT_CHECK(expr)                  \
  do {                                          \
    const int rc_ = (expr);                     \
    if (rc_ != SGC_OK) compat_abort(rc_, __LINE__); \
  } while (0)

// One plain pass over the caller's reference-layout graph on device 0;
// returns the malloc'd labels or ends the process.
int32_t* compat_scan(const sgc_graph_t* g, double eps, int32_t miu) {
  clear_error();
  if (!g) {
    fail(SGC_EINVAL, "graph_t is null");
    compat_abort(SGC_EINVAL, __LINE__);
  }
  if (g->n < 0 || g->m < 0 || g->n >= (1l << 31) || g->m >= (1l << 32)) {
    fail(SGC_EINVAL, "graph_t has n=%ld, m=%ld (need 0 <= n < 2^31, 0 <= m < 2^32)", g->n, g->m);
    compat_abort(SGC_EINVAL, __LINE__);
  }
  if (!g->num_edges || (g->m && !g->adj)) {
    fail(SGC_EINVAL, "graph_t has a null %s", !g->num_edges ? "num_edges" : "adj");
    compat_abort(SGC_EINVAL, __LINE__);
  }
  sgc_host_graph hg;
  hg.n = (uint64_t)g->n;
  hg.m = (uint64_t)g->m;
  std::vector<uint64_t> rp(hg.n + 1);
  for (uint64_t u = 0; u <= hg.n; ++u) rp[u] = g->num_edges[u];
  hg.row_ptr = rp.data();
  hg.adj = g->adj;
  // the device kernels trust the CSR: check the input contract
  // (GPUSCAN++/Graph.cpp:270-288 and ids < n) before anything runs
  int rc = validate_graph(&hg, false);
  sgc_dgraph* dg = nullptr;
  int32_t* out = nullptr;
  if (rc == SGC_OK) {
    std::lock_guard<std::mutex> lock(g_compat_mu);
    if (!g_compat_ctx) rc = sgc_ctx_create(0, &g_compat_ctx);
    if (rc == SGC_OK) rc = sgc_dgraph_upload(g_compat_ctx, &hg, &dg);
    if (rc == SGC_OK) {
      out = (int32_t*)malloc(std::max<size_t>(hg.n, 1) * sizeof(int32_t));
      if (!out) rc = fail(SGC_ENOMEM, "malloc of %llu labels failed", (unsigned long long)hg.n);
    }
    sgc_stats st;
    if (rc == SGC_OK) rc = sgc_scan(g_compat_ctx, dg, eps, miu, out, nullptr, nullptr, &st);
    if (dg) sgc_dgraph_free(dg);
    if (rc == SGC_OK) {
      // the reference's phase lines (GPUSCAN++/GPUScan.cu:890-989)
      fprintf(stderr, "\nPrep time: %.3lf s\n", st.ms_prep / 1000.0);
      fprintf(stderr, "\nCheck Core time: %.3lf s\n", (st.ms_similarity + st.ms_core) / 1000.0);
      fprintf(stderr, "core number: %llu\n", (unsigned long long)st.cores);
      fprintf(stderr, "\nFirst Link Phase time: %.3lf s\n", st.ms_cluster / 1000.0);
      fprintf(stderr, "Cluster non-core time: %.3lf s\n", st.ms_border / 1000.0);
      fprintf(stderr, "\nClassify Others time: %.3lf s\n", st.ms_classify / 1000.0);
    }
  }
  hg.row_ptr = nullptr;
  hg.adj = nullptr;
  if (rc != SGC_OK) {
    free(out);
    compat_abort(rc, __LINE__);
  }
  return out;
}
}  // namespace

extern "C" int32_t* sgc_GPUScan(sgc_graph_t* g, int32_t* Edge_u, int32_t* Edge_v, double eps, int32_t miu) {
  (void)Edge_u;
  (void)Edge_v;
  return compat_scan(g, eps, miu);
}

// C++ linkage, the reference's own symbol (GPUSCAN++/GPUScan.h:7).
int* GPUScan(graph_t* g, int* Edge_u, int* Edge_v, double eps, int miu) {
  static_assert(sizeof(graph_t) == sizeof(sgc_graph_t), "graph_t layout");
  static_assert(offsetof(graph_t, num_edges) == offsetof(sgc_graph_t, num_edges), "graph_t layout");
  static_assert(offsetof(graph_t, eid) == offsetof(sgc_graph_t, eid), "graph_t layout");
  (void)Edge_u;
  (void)Edge_v;
  return compat_scan(reinterpret_cast<const sgc_graph_t*>(g), eps, miu);
}
