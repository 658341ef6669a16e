// Out-of-core SCAN for graphs larger than a device budget (GPUSCAN++O:
// chips GPUSCAN++O/Graph.h:70-85, Graph.cpp:414-576; phase-major
// orchestration SCANoffload.cu:922-1464).
//
// Per-vertex state stays resident (the chip row table, similar-degree
// counts, union-find forest, labels, classify bounds, planner arrays: about
// 41 bytes per vertex); the rest of the budget holds one chip at a time.
//  * Similarity: rows are cut into P blocks, and consecutive blocks into
//    groups of k (4 where the budget allows).  A group's own chip decides the
//    undecided edges with both endpoints in the group; chip (group, j), for
//    every block j after the group, those with one endpoint in each (the
//    prep kernel's pair filter marks the other entries kFlagUSkip, or keeps
//    the group's flags from its earlier chips); every edge has both of its
//    rows in exactly one chip.  The resident prep, planner and similarity
//    kernels run on the chip unchanged through its row table (rows outside
//    the chip are never read).  The group stays on the device while j
//    advances, so the host sends about m * (1 + P / (2k)) entries rather
//    than a copy of every short row per owner.  Final flags return to a
//    whole-graph host flag array (a block's after its chip, a group's after
//    its last); the next block is staged and the previous chip's flags are
//    merged on the host while the device works.
//  * Core check, union-find, border and classify: chips of whole rows with
//    their flags, one pass per phase (each phase needs the previous one's
//    global result); when all rows fit in one chip it is uploaded once.
// Labels, core flags and similarity flags are bit-identical to the resident
// path's.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "device.h"
#include "device_util.h"
#include "ooc_host.h"

namespace sgc {
namespace {

// Rows r0 .. r0 + nr of a chip: off[] holds their offsets from the block's
// first entry, `shift` the block's first entry in the chip.
__global__ void __launch_bounds__(kBlock) k_rows_range(uint32_t r0, uint32_t nr, uint32_t shift,
                                                       const uint32_t* __restrict__ off, uint32_t* __restrict__ rb,
                                                       uint32_t* __restrict__ re) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < nr; i += stride) {
    rb[r0 + i] = shift + off[i];
    re[r0 + i] = shift + off[i + 1];
  }
}

// Per similarity chip: the counters, and the planner counts of the chip's
// rows [a0, a1) and [b0, b1) (the only rows its prep, planner and similarity
// kernels touch; the later phases clear every row again).
__global__ void __launch_bounds__(kBlock) k_chip_init(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1,
                                                      uint32_t* __restrict__ cnt_u, uint64_t* __restrict__ work_u,
                                                      uint64_t* __restrict__ cnt) {
  const uint64_t na = a1 - a0, total = na + (b1 - b0);
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += stride) {
    const uint64_t x = i < na ? a0 + i : b0 + (i - na);
    cnt_u[x] = 0;
    work_u[x] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < kCntSlots) cnt[threadIdx.x] = 0;
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

// Similarity chips pair a group of up to k consecutive row blocks (I
// entries at most) with one later row block (J entries at most), or hold a
// group alone.  Their device memory: adj, src and flags as [ j-buffer 0 |
// group | j-buffer 1 ] (I + 2J entries each; the row block sits right
// before the group in buffer 0 or right after it in buffer 1, so every chip
// is one contiguous entry range and the next chip's row block uploads while
// this one runs), and R packed edge records and R + 2 work units (16 B each)
// for one round of owners.  The group's flags stay on the device across its
// chips (a chip writes there only the entries it decides) and go back to the
// host once; a row block's flags go back after its chip, while the next chip
// computes in the other buffer.  A chip owns at most (I + J) / 2 undecided
// edges, decided in rounds of owners whose records start in a window of
// W = R - (largest degree), with R about half of that: two or three rounds.
// Per row-block entry a chip takes about 17 k + 26 bytes for groups of k
// blocks, and the host sends about m (1 + P / (2k)) entries for P row
// blocks, so for the same memory groups of 4 send about half of what pairs of
// single blocks send.
constexpr uint32_t kGroupChoices[] = {4, 2, 1};  // the largest the room allows
constexpr uint64_t kRowChipBytesPerEntry = 4 + 4 + 1;
constexpr uint64_t kMaxChipEntries = 0xFFFFFFF0ull;

struct ResidentLayout {
  size_t rb, re, sd, par, lab, lo, hi, core, cntu, work, cnt, scan, scan_bytes, bytes;
};

ResidentLayout resident_layout(uint32_t n) {
  Layout L;
  ResidentLayout r;
  r.rb = L.take((size_t)n * 4), r.re = L.take((size_t)n * 4);
  r.sd = L.take((size_t)n * 4), r.par = L.take((size_t)n * 4), r.lab = L.take((size_t)n * 4);
  r.lo = L.take((size_t)n * 4), r.hi = L.take((size_t)n * 4), r.core = L.take(n);
  r.cntu = L.take((size_t)n * 4), r.work = L.take((size_t)n * 8), r.cnt = L.take(kCntSlots * 8);
  r.scan_bytes = std::max<size_t>(scan_scratch_bytes(n), 1);
  r.scan = L.take(r.scan_bytes);
  r.bytes = L.bytes;
  return r;
}

struct SimChipArrays {
  int32_t* adj;   // I + 2J
  uint32_t* src;  // I + 2J
  uint8_t* flag;  // I + 2J
  uint4* uedge;   // R
  uint4* units;   // R + 2
};

SimChipArrays sim_arrays(char* base, uint64_t I, uint64_t J, uint64_t R, size_t* bytes) {
  Layout L;
  const uint64_t e = I + 2 * J;
  const size_t o_adj = L.take(e * 4), o_src = L.take(e * 4), o_flag = L.take(e),
               o_uedge = L.take(R * sizeof(uint4)), o_units = L.take((R + 2) * sizeof(uint4));
  if (bytes) *bytes = L.bytes;
  return SimChipArrays{(int32_t*)(base + o_adj), (uint32_t*)(base + o_src), (uint8_t*)(base + o_flag),
                       (uint4*)(base + o_uedge), (uint4*)(base + o_units)};
}

size_t sim_chip_bytes(uint64_t I, uint64_t J, uint64_t R) {
  size_t b = 0;
  sim_arrays(nullptr, I, J, R, &b);
  return b;
}

// Records per chip for chips owning at most C edges and rows of at most
// maxd entries: a window of at least C / 2 - maxd (and at least C / 4) plus
// the largest owner's count.
uint64_t record_capacity(uint64_t C, uint64_t maxd) {
  const uint64_t half = C / 2, w = half > maxd + C / 4 ? half - maxd : std::max<uint64_t>(C / 4, 1);
  return w + maxd;
}

// Row chip of the later phases: adj, src and flags of `cap` entries, then
// the row offsets of `rows` rows.
size_t row_chip_bytes(uint64_t cap, uint64_t rows) {
  Layout L;
  L.take(cap * 4), L.take(cap * 4), L.take(cap), L.take((rows + 1) * 4);
  return L.bytes;
}

// Row offsets of a group of rows_i rows and of the two j buffers of rows_j
// rows each, 256-byte aligned.
size_t offsets_slice(uint64_t rows) { return ((rows + 1) * 4 + 255) / 256 * 256; }
size_t offsets_bytes(uint64_t rows_i, uint64_t rows_j) { return offsets_slice(rows_i) + 2 * offsets_slice(rows_j); }

// Pinned copy of one row block, free again once its upload has run.
struct Staging {
  int32_t* adj = nullptr;
  uint32_t* off = nullptr;
  hipEvent_t free = nullptr;
  bool pending = false;  // an upload from it may still be in flight
};

struct OocRun {
  sgc_ctx* ctx = nullptr;
  const sgc_host_graph* g = nullptr;
  uint32_t n = 0;
  void* dev = nullptr;
  uint32_t *rb = nullptr, *re = nullptr, *off_i = nullptr, *off_j[2] = {nullptr, nullptr};
  PassBuffers pb{};
  char* chip = nullptr;
  Staging stage[3];  // block i, and one per j buffer
  // Side streams of the similarity phase (default priority): uploads and
  // flag read-back run there, overlapped with the chips' kernels on the
  // context stream.
  hipStream_t h2d = nullptr, d2h = nullptr;
  hipEvent_t up_i = nullptr, up_j[2] = {nullptr, nullptr};  // block uploads done
  hipEvent_t comp[2] = {nullptr, nullptr};                  // the chip of buffer b done on the device
  hipEvent_t back[2] = {nullptr, nullptr};                  // its row block's flags back on the host
  hipEvent_t back_i = nullptr;                              // the last group's flags back on the host
  uint8_t* h_flag[2] = {nullptr, nullptr};                  // a row block's flags (J)
  uint8_t* h_iflag = nullptr;                               // a group's flags (I)
  uint64_t* h_cnt[2] = {nullptr, nullptr};
  hipEvent_t t0[2] = {nullptr, nullptr}, tm[2] = {nullptr, nullptr}, t1[2] = {nullptr, nullptr};
  uint8_t* flag = nullptr;  // whole-graph flags, page-locked where the system allows
  bool flag_pageable = false;
  double bytes_h2d = 0, bytes_d2h = 0, ms_host = 0;

  ~OocRun() {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t st : {h2d, d2h})
      if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
      }
    for (Staging& st : stage) {
      if (st.adj) (void)hipHostFree(st.adj);
      if (st.off) (void)hipHostFree(st.off);
      if (st.free) (void)hipEventDestroy(st.free);
    }
    if (up_i) (void)hipEventDestroy(up_i);
    if (back_i) (void)hipEventDestroy(back_i);
    if (h_iflag) (void)hipHostFree(h_iflag);
    for (int i = 0; i < 2; ++i) {
      if (h_flag[i]) (void)hipHostFree(h_flag[i]);
      if (h_cnt[i]) (void)hipHostFree(h_cnt[i]);
      for (hipEvent_t e : {up_j[i], comp[i], back[i], t0[i], tm[i], t1[i]})
        if (e) (void)hipEventDestroy(e);
    }
    if (flag && flag_pageable) free(flag);
    else if (flag) (void)hipHostFree(flag);
    if (dev) (void)hipFree(dev);
  }
};

// One chip of the similarity phase: blocks i and j (j == i: one block).
// j_first: the chip's entries hold block j first (j-buffer 0), else block i.
struct ChipDesc {
  uint64_t r0i, r1i, r0j, r1j;
  uint64_t mi, mj;
  bool j_first;
};

int run_ooc(OocRun& R, double eps, int32_t mu, uint64_t budget, int32_t* labels, uint8_t* core_out, uint8_t* sim_out,
            sgc_stats* st, sgc_ooc_stats* oi) {
  sgc_ctx* ctx = R.ctx;
  const sgc_host_graph* g = R.g;
  const uint32_t n = R.n;
  hipStream_t s = ctx->stream;
  const double t_start = now_ms();

  // ---- resident layout and plans
  const ResidentLayout RLay = resident_layout(n);
  const size_t resident = RLay.bytes;
  if (budget <= resident + 8192)
    return fail(SGC_ENOMEM, "device budget %llu bytes is below the %llu resident bytes the graph's %u vertices need",
                (unsigned long long)budget, (unsigned long long)resident, n);
  const uint64_t room = budget - resident - 8192;
  double t0 = now_ms();
  std::vector<uint32_t> deg(n);
  ooc_degrees(g, deg.data());
  std::vector<uint64_t> blocks, row_bounds;
  uint64_t block_rows = 0, row_max = 0, row_rows = 0;
  uint64_t maxd = 1;
  for (uint32_t x = 0; x < n; ++x) maxd = std::max<uint64_t>(maxd, deg[x]);
  // Row blocks as large as the room allows once the records of the largest
  // degree and the row offsets are counted in (a block holds at most as many
  // rows as entries), in groups of the largest size whose chips still fit
  // with blocks of the longest row.
  int rc = SGC_OK;
  uint64_t I = 1, J = 1, RC = 1, rows_i = 0;  // group and block entries, records per round, group rows
  uint32_t group = 1;
  bool fits = false;
  for (const uint32_t k : kGroupChoices) {
    uint64_t cap = std::max<uint64_t>(std::min<uint64_t>(room / (17 * k + 26), kMaxChipEntries / (k + 2)), maxd);
    for (;;) {
      uint64_t block_max = 0;
      rc = ooc_plan_rows(g, cap, cap, &blocks, &block_max, &block_rows);
      if (rc) return rc;
      J = std::max<uint64_t>(block_max, 1);
      I = 1, rows_i = 0;
      for (size_t b = 0; b + 1 < blocks.size(); b += k) {
        const size_t e = std::min<size_t>(b + k, blocks.size() - 1);
        I = std::max<uint64_t>(I, g->row_ptr[blocks[e]] - g->row_ptr[blocks[b]]);
        rows_i = std::max<uint64_t>(rows_i, blocks[e] - blocks[b]);
      }
      RC = record_capacity((I + J) / 2 + 1, maxd);
      if (I + 2 * J <= kMaxChipEntries && sim_chip_bytes(I, J, RC) + offsets_bytes(rows_i, block_rows) <= room) {
        fits = true;
        break;
      }
      if (cap <= maxd) break;
      cap = std::max<uint64_t>(maxd, cap - cap / 8);
    }
    group = k;
    if (fits) break;
  }
  const size_t sim_need = sim_chip_bytes(I, J, RC) + offsets_bytes(rows_i, block_rows);
  // row chips for the later phases, with their own row offsets, in the same
  // memory
  uint64_t cap_row = std::min<uint64_t>(room / kRowChipBytesPerEntry, kMaxChipEntries), row_cap = n;
  for (;;) {
    rc = ooc_plan_rows(g, std::max<uint64_t>(cap_row, maxd), std::max<uint64_t>(row_cap, 1), &row_bounds, &row_max,
                       &row_rows);
    if (rc) return rc;
    if (row_chip_bytes(std::max<uint64_t>(row_max, 1), row_rows) <= room || cap_row <= maxd) break;
    cap_row = std::max<uint64_t>(maxd, cap_row - cap_row / 8);
    row_cap = std::min<uint64_t>(row_cap, cap_row);
  }
  const uint64_t W = RC - maxd;                          // window of one round of owners
  const uint64_t owned_max = (I + J) / 2 + 1;               // undecided edges one chip owns at most
  const uint32_t rounds = (uint32_t)((owned_max + W - 1) / W);
  const uint64_t sim_entries = I + J;                       // entries per chip at most
  const size_t chip_need = std::max(sim_need, row_chip_bytes(std::max<uint64_t>(row_max, 1), row_rows));
  R.ms_host += now_ms() - t0;
  if (resident + chip_need + 8192 > budget)
    return fail(SGC_ENOMEM, "device budget %llu bytes is below the %llu bytes this graph needs (sgc_ooc_min_budget)",
                (unsigned long long)budget, (unsigned long long)(resident + chip_need + 8192));

  // ---- device and pinned host memory
  if (hipMalloc(&R.dev, resident + chip_need) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SGC_ENOMEM, "device allocation of %llu bytes failed", (unsigned long long)(resident + chip_need));
  }
  char* base = (char*)R.dev;
  R.rb = (uint32_t*)(base + RLay.rb);
  R.re = (uint32_t*)(base + RLay.re);
  // the similarity chips' three row-offset arrays after their other arrays
  R.off_i = (uint32_t*)(base + resident + sim_chip_bytes(I, J, RC));
  R.off_j[0] = (uint32_t*)((char*)R.off_i + offsets_slice(rows_i));
  R.off_j[1] = (uint32_t*)((char*)R.off_j[0] + offsets_slice(block_rows));
  PassBuffers& pb = R.pb;
  pb.sd = (uint32_t*)(base + RLay.sd);
  pb.parent = (int32_t*)(base + RLay.par);
  pb.label = (int32_t*)(base + RLay.lab);
  pb.tiny = (uint32_t*)pb.label;
  pb.lo = (int32_t*)(base + RLay.lo);
  pb.hi = (int32_t*)(base + RLay.hi);
  pb.core = (uint8_t*)(base + RLay.core);
  pb.cnt_u = (uint32_t*)(base + RLay.cntu);
  pb.work_u = (uint64_t*)(base + RLay.work);
  pb.cnt = (uint64_t*)(base + RLay.cnt);
  // packed records of the similarity chips: first record and cursor of each
  // owner in lo's and sd's memory (free until the core phase)
  pb.rec_start = (uint32_t*)pb.lo;
  pb.rec_cur = pb.sd;
  pb.scan_tmp = base + RLay.scan;
  pb.scan_tmp_bytes = RLay.scan_bytes;
  R.chip = base + resident;

  // a page-locked graph is sent as it lies: only row offsets are staged.
  // Staging 0 holds a group (and the row chips' offsets), 1 and 2 a row
  // block each.
  for (int q = 0; q < 3; ++q) {
    Staging& sg = R.stage[q];
    const uint64_t entries = g->pinned ? 0 : std::max<uint64_t>(q ? J : I, 1);
    const uint64_t rows = std::max<uint64_t>(q ? block_rows : std::max(rows_i, row_rows), 1);
    if ((entries && hipHostMalloc((void**)&sg.adj, entries * 4) != hipSuccess) ||
        hipHostMalloc((void**)&sg.off, (rows + 1) * 4) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SGC_ENOMEM, "pinned host staging of %llu bytes failed",
                  (unsigned long long)(entries * 4 + (rows + 1) * 4));
    }
    SGC_HIP(hipEventCreateWithFlags(&sg.free, hipEventDisableTiming));
  }
  if (hipHostMalloc((void**)&R.h_iflag, std::max<uint64_t>(I, 1)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SGC_ENOMEM, "pinned host flag staging of %llu bytes failed", (unsigned long long)I);
  }
  SGC_HIP(hipEventCreateWithFlags(&R.back_i, hipEventDisableTiming));
  for (int i = 0; i < 2; ++i) {
    if (hipHostMalloc((void**)&R.h_flag[i], std::max<uint64_t>(J, 1)) != hipSuccess ||
        hipHostMalloc((void**)&R.h_cnt[i], kCntSlots * 8) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SGC_ENOMEM, "pinned host flag staging of %llu bytes failed", (unsigned long long)J);
    }
    SGC_HIP(hipEventCreateWithFlags(&R.up_j[i], hipEventDisableTiming));
    SGC_HIP(hipEventCreateWithFlags(&R.comp[i], hipEventDisableTiming));
    SGC_HIP(hipEventCreateWithFlags(&R.back[i], hipEventDisableTiming));
    SGC_HIP(hipEventCreate(&R.t0[i]));
    SGC_HIP(hipEventCreate(&R.tm[i]));
    SGC_HIP(hipEventCreate(&R.t1[i]));
  }
  SGC_HIP(hipEventCreateWithFlags(&R.up_i, hipEventDisableTiming));
  SGC_HIP(hipStreamCreateWithFlags(&R.h2d, hipStreamNonBlocking));
  SGC_HIP(hipStreamCreateWithFlags(&R.d2h, hipStreamNonBlocking));
  // the whole-graph flags only meet the device through the staged chip
  // flags and the row chips' uploads: ordinary memory serves when the system
  // will not lock m more bytes (the row-chip uploads then copy through the
  // runtime's own staging)
  if (hipHostMalloc((void**)&R.flag, std::max<uint64_t>(g->m, 1)) != hipSuccess) {
    (void)hipGetLastError();
    R.flag = nullptr;
    R.flag = (uint8_t*)malloc(std::max<uint64_t>(g->m, 1));
    if (!R.flag) return fail(SGC_ENOMEM, "host flag array of %llu bytes failed", (unsigned long long)g->m);
    R.flag_pageable = true;
  }

  const ScanParams p{eps, mu};
  uint64_t totals[kCntSlots] = {0};
  double ms_sim = 0, ms_sim_large = 0;

  // ---- similarity over chips: each group of blocks alone, then with every
  // later block.  Uploads on R.h2d, kernels on the context stream, flag
  // read-back on R.d2h, host merges in between.
  const size_t P = blocks.size() - 1;
  const SimChipArrays A = sim_arrays(R.chip, I, J, RC, nullptr);
  auto range_m = [&](size_t b0, size_t b1) { return g->row_ptr[blocks[b1]] - g->row_ptr[blocks[b0]]; };
  // rows of blocks [b0, b1) into pinned staging (offsets only for a
  // page-locked graph)
  auto stage_rows = [&](Staging& sg, size_t b0, size_t b1) -> int {
    if (sg.pending) SGC_HIP(hipEventSynchronize(sg.free));
    sg.pending = false;
    const double tb = now_ms();
    ooc_copy_rows(g, blocks[b0], blocks[b1], 0, g->pinned ? nullptr : sg.adj, sg.off);
    R.ms_host += now_ms() - tb;
    return SGC_OK;
  };
  // blocks [b0, b1) to chip entries [pos, pos + their entries), their
  // offsets (from their first entry) to off_dev; `done` is recorded on the
  // upload stream
  auto upload_rows = [&](Staging& sg, size_t b0, size_t b1, uint64_t pos, uint32_t* off_dev, hipEvent_t done) -> int {
    const uint64_t mb = range_m(b0, b1);
    const uint32_t nr = (uint32_t)(blocks[b1] - blocks[b0]);
    const int32_t* from = g->pinned ? g->adj + g->row_ptr[blocks[b0]] : sg.adj;
    if (mb) SGC_HIP(hipMemcpyAsync(A.adj + pos, from, mb * 4, hipMemcpyHostToDevice, R.h2d));
    SGC_HIP(hipMemcpyAsync(off_dev, sg.off, ((size_t)nr + 1) * 4, hipMemcpyHostToDevice, R.h2d));
    SGC_HIP(hipEventRecord(sg.free, R.h2d));
    sg.pending = true;
    SGC_HIP(hipEventRecord(done, R.h2d));
    R.bytes_h2d += (double)mb * 4 + (double)(nr + 1) * 4;
    return SGC_OK;
  };
  bool used[2] = {false, false};  // comp[b] / back[b] recorded
  ChipDesc prev{};
  int prev_b = 0;
  size_t chip_no = 0;
  // a chip's counters and its row block's flags, once they are back
  auto merge = [&](const ChipDesc& c, int b) -> int {
    SGC_HIP(hipEventSynchronize(R.back[b]));
    const double tb = now_ms();
    if (c.mj) ooc_merge_rows(g, c.r0j, c.r1j, R.h_flag[b], R.flag);
    for (int q = 0; q < kCntSlots; ++q) totals[q] += R.h_cnt[b][q];
    R.ms_host += now_ms() - tb;
    ms_sim += elapsed(R.t0[b], R.t1[b]);
    ms_sim_large += elapsed(R.t0[b], R.tm[b]);
    return SGC_OK;
  };
  // the last group's flags, once they are back
  bool group_back = false;  // back_i recorded and its flags not merged yet
  uint64_t back_r0 = 0, back_r1 = 0;
  auto merge_group = [&]() -> int {
    if (!group_back) return SGC_OK;
    SGC_HIP(hipEventSynchronize(R.back_i));
    const double tb = now_ms();
    ooc_merge_rows(g, back_r0, back_r1, R.h_iflag, R.flag);
    R.ms_host += now_ms() - tb;
    group_back = false;
    return SGC_OK;
  };
  const size_t total_chips = [&]() {
    size_t t = 0;
    for (size_t gi = 0; gi < P; gi += group) t += 1 + (P - std::min<size_t>(gi + group, P));
    return t;
  }();
  bool any_group = false;
  for (size_t gi = 0; g->m && gi < P; gi += group) {
    const size_t ge = std::min<size_t>(gi + group, P);
    ChipDesc c;
    c.r0i = blocks[gi], c.r1i = blocks[ge];
    c.mi = range_m(gi, ge);
    const uint32_t nri = (uint32_t)(c.r1i - c.r0i);
    // the group goes to chip entries [J, J + mi) once every chip of the
    // previous group has run
    if ((rc = stage_rows(R.stage[0], gi, ge))) return rc;
    for (int b = 0; b < 2; ++b)
      if (used[b]) SGC_HIP(hipStreamWaitEvent(R.h2d, R.comp[b], 0));
    if ((rc = upload_rows(R.stage[0], gi, ge, J, R.off_i, R.up_i))) return rc;
    SGC_HIP(hipStreamWaitEvent(s, R.up_i, 0));
    if (nri) k_fill_src_rows<<<grid_for(nri, kBlock / kWave), kBlock, 0, s>>>((uint32_t)c.r0i, nri, R.off_i, A.src + J);
    // the group's flags are rewritten once the previous group's are back
    if (any_group) SGC_HIP(hipStreamWaitEvent(s, R.back_i, 0));
    any_group = true;
    for (size_t j = ge - 1; j < P; ++j) {
      const bool alone = j + 1 == ge;  // the group's own chip
      const int b = (int)(chip_no & 1);
      uint64_t base = J;   // the chip's first entry in adj / src / flag
      uint64_t jpos = 0;   // its row block's first entry
      c.j_first = false;
      uint32_t nrj = 0;
      if (!alone) {
        c.r0j = blocks[j], c.r1j = blocks[j + 1];
        c.mj = range_m(j, j + 1);
        nrj = (uint32_t)(c.r1j - c.r0j);
        // block j right before the group (buffer 0) or right after it
        // (buffer 1), once the chip before last, which used this buffer, has
        // run
        jpos = b == 0 ? J - c.mj : J + c.mi;
        Staging& sj = R.stage[1 + b];
        if ((rc = stage_rows(sj, j, j + 1))) return rc;
        if (used[b]) SGC_HIP(hipStreamWaitEvent(R.h2d, R.comp[b], 0));
        if ((rc = upload_rows(sj, j, j + 1, jpos, R.off_j[b], R.up_j[b]))) return rc;
        SGC_HIP(hipStreamWaitEvent(s, R.up_j[b], 0));
        if (nrj)
          k_fill_src_rows<<<grid_for(nrj, kBlock / kWave), kBlock, 0, s>>>((uint32_t)c.r0j, nrj, R.off_j[b],
                                                                           A.src + jpos);
        if (b == 0) {
          base = jpos;
          c.j_first = true;
        }
      } else {
        c.r0j = c.r0i, c.r1j = c.r1i, c.mj = 0;
      }
      // buffer b's flags are free once the chip before last has been read back
      if (used[b]) SGC_HIP(hipStreamWaitEvent(s, R.back[b], 0));
      // rows of the group and the block in this chip's layout; rows outside
      // them are never read (the prep kernel reads degrees of entries with
      // one end in each only, every later kernel rows of the chip's owners
      // and of their edges' short rows)
      if (nri)
        k_rows_range<<<grid_for(nri, kBlock), kBlock, 0, s>>>((uint32_t)c.r0i, nri, c.j_first ? (uint32_t)c.mj : 0u,
                                                              R.off_i, R.rb, R.re);
      if (nrj)
        k_rows_range<<<grid_for(nrj, kBlock), kBlock, 0, s>>>((uint32_t)c.r0j, nrj, c.j_first ? 0u : (uint32_t)c.mi,
                                                              R.off_j[b], R.rb, R.re);
      const uint64_t m_chip = c.mi + c.mj;
      sgc_dgraph view;
      view.ctx = ctx;
      view.n = n;
      view.m = (uint32_t)m_chip;
      view.adj = A.adj + base;
      view.src = A.src + base;
      view.row_begin = R.rb;
      view.row_end = R.re;
      view.pair.lo0 = (uint32_t)c.r0i, view.pair.hi0 = (uint32_t)c.r1i;
      view.pair.lo1 = (uint32_t)c.r0j, view.pair.hi1 = (uint32_t)c.r1j;
      view.pair.on = true;
      view.pair.keep0 = !alone;  // the group's other entries were decided by its earlier chips
      PassBuffers cb = pb;
      cb.flag = A.flag + base;
      cb.uedge = A.uedge;
      cb.units = A.units;
      cb.ucap = (uint32_t)(RC + 2);
      {
        const uint32_t b0 = alone ? 0u : (uint32_t)c.r0j, b1 = alone ? 0u : (uint32_t)c.r1j;
        k_chip_init<<<grid_for((uint64_t)nri + (b1 - b0), kBlock), kBlock, 0, s>>>(
            (uint32_t)c.r0i, (uint32_t)c.r1i, b0, b1, cb.cnt_u, cb.work_u, cb.cnt);
      }
      launch_prep_packed(&view, cb, p, s);
      SGC_HIP(hipEventRecord(R.t0[b], s));
      launch_similarity_packed(&view, cb, p, ctx->cu_count, ctx->sim_impl, R.tm[b], s, (uint32_t)W, rounds);
      SGC_HIP(hipEventRecord(R.t1[b], s));
      SGC_HIP(hipGetLastError());
      // the counters before the next chip's init clears them
      SGC_HIP(hipMemcpyAsync(R.h_cnt[b], pb.cnt, kCntSlots * 8, hipMemcpyDeviceToHost, s));
      SGC_HIP(hipEventRecord(R.comp[b], s));
      SGC_HIP(hipStreamWaitEvent(R.d2h, R.comp[b], 0));
      if (c.mj) SGC_HIP(hipMemcpyAsync(R.h_flag[b], A.flag + jpos, c.mj, hipMemcpyDeviceToHost, R.d2h));
      SGC_HIP(hipEventRecord(R.back[b], R.d2h));
      used[b] = true;
      R.bytes_d2h += (double)c.mj;
      // while the device works: merge the previous chip, and the previous
      // group once this group's own chip is queued
      if (chip_no > 0 && (rc = merge(prev, prev_b))) return rc;
      if (alone && (rc = merge_group())) return rc;
      prev = c;
      prev_b = b;
      ++chip_no;
      if (chip_no % 32 == 0) progress("ooc: %zu of %zu similarity chips issued", chip_no, total_chips);
    }
    // the group's flags back to the host after its last chip (the d2h stream
    // already waits for that chip)
    if (c.mi) SGC_HIP(hipMemcpyAsync(R.h_iflag, A.flag + J, c.mi, hipMemcpyDeviceToHost, R.d2h));
    SGC_HIP(hipEventRecord(R.back_i, R.d2h));
    R.bytes_d2h += (double)c.mi;
    group_back = true;
    back_r0 = c.r0i, back_r1 = c.r1i;
  }
  if (chip_no && (rc = merge(prev, prev_b))) return rc;
  if ((rc = merge_group())) return rc;
  SGC_HIP(hipStreamSynchronize(R.h2d));
  SGC_HIP(hipStreamSynchronize(s));
  if ((rc = check_unit_overflow(totals, sim_entries))) return rc;
  const size_t nsim = chip_no;
  progress("ooc: similarity done over %zu chips of %zu blocks in groups of %u", nsim, P, group);

  // ---- core check, union-find, border, classify over row chips
  const double t_phase = now_ms();
  const size_t nrow = row_bounds.size() - 1;
  int32_t* c_adj = (int32_t*)R.chip;
  Layout RL;
  const size_t ro_adj = RL.take(std::max<uint64_t>(row_max, 1) * 4), ro_src = RL.take(std::max<uint64_t>(row_max, 1) * 4),
               ro_flag = RL.take(std::max<uint64_t>(row_max, 1)), ro_off = RL.take((row_rows + 1) * 4);
  (void)ro_adj;
  uint32_t* c_src = (uint32_t*)(R.chip + ro_src);
  uint8_t* c_flag = (uint8_t*)(R.chip + ro_flag);
  uint32_t* c_off = (uint32_t*)(R.chip + ro_off);
  sgc_dgraph whole;
  whole.ctx = ctx;
  whole.n = n;
  whole.m = 0;
  launch_init(&whole, pb, s);
  // row chip i on the device: rows, flags and entry rows; with_adj also the
  // neighbour ids (the core count reads flags and rows only)
  size_t loaded = (size_t)-1;
  bool loaded_adj = false;
  auto load_rows = [&](size_t i, sgc_dgraph* v, bool with_adj) -> int {
    const uint64_t r0 = row_bounds[i], r1 = row_bounds[i + 1];
    const uint64_t b0 = g->row_ptr[r0], mc = g->row_ptr[r1] - b0;
    v->ctx = ctx;
    v->n = n;
    v->m = (uint32_t)mc;
    v->adj = c_adj;
    v->src = c_src;
    if (loaded == i && (loaded_adj || !with_adj)) return SGC_OK;
    SGC_HIP(hipStreamSynchronize(s));  // staging and chip buffers are reused
    if (loaded != i) {
      const uint32_t nr = (uint32_t)(r1 - r0);
      uint32_t* ho = R.stage[0].off;
      for (uint32_t k = 0; k <= nr; ++k) ho[k] = (uint32_t)(g->row_ptr[r0 + k] - b0);
      SGC_HIP(hipMemcpyAsync(c_off, ho, ((size_t)nr + 1) * 4, hipMemcpyHostToDevice, s));
      if (mc) SGC_HIP(hipMemcpyAsync(c_flag, R.flag + b0, mc, hipMemcpyHostToDevice, s));
      R.bytes_h2d += (double)mc + (double)(nr + 1) * 4;
      if (nr) k_fill_src_rows<<<grid_for(nr, kBlock / kWave), kBlock, 0, s>>>((uint32_t)r0, nr, c_off, c_src);
      loaded = i;
      loaded_adj = false;
    }
    if (with_adj && !loaded_adj) {
      if (mc) SGC_HIP(hipMemcpyAsync(c_adj, g->adj + b0, mc * 4, hipMemcpyHostToDevice, s));
      R.bytes_h2d += (double)mc * 4;
      loaded_adj = true;
    }
    return SGC_OK;
  };
  PassBuffers rb = pb;
  rb.flag = c_flag;
  sgc_dgraph v;
  double ms_core = 0, ms_cluster = 0, ms_border = 0, ms_classify = 0;
  SGC_HIP(hipEventRecord(ctx->ev[3], s));
  for (size_t i = 0; i < nrow; ++i) {
    if ((rc = load_rows(i, &v, false))) return rc;
    launch_core_count(&v, rb, s);
  }
  launch_core_mark(&whole, rb, p, s);
  // without cores there is no core-core edge, no border and no label for
  // classify to see: the three passes over the row chips would return at
  // once, so their uploads are skipped (every vertex ends as an outlier, or
  // as the hub that mu <= 0 makes it, in finalize)
  SGC_HIP(hipMemcpyAsync(R.h_cnt[0], pb.cnt, kCntSlots * 8, hipMemcpyDeviceToHost, s));
  SGC_HIP(hipStreamSynchronize(s));
  const bool any_core = R.h_cnt[0][kCntMarked] != 0;
  progress("ooc: core check over %zu row chips, %llu cores", nrow, (unsigned long long)R.h_cnt[0][kCntMarked]);
  SGC_HIP(hipEventRecord(ctx->ev[4], s));
  for (size_t i = 0; any_core && i < nrow; ++i) {
    if ((rc = load_rows(nrow - 1 - i, &v, true))) return rc;  // last chip first: it is still loaded
    launch_union(&v, rb, s);
  }
  launch_flatten(&whole, rb, s);
  progress("ooc: union-find issued");
  SGC_HIP(hipEventRecord(ctx->ev[5], s));
  for (size_t i = 0; any_core && i < nrow; ++i) {
    if ((rc = load_rows(i, &v, true))) return rc;
    launch_border(&v, rb, s);
  }
  SGC_HIP(hipEventRecord(ctx->ev[6], s));
  for (size_t i = 0; any_core && i < nrow; ++i) {
    if ((rc = load_rows(nrow - 1 - i, &v, true))) return rc;
    launch_classify_edges(&v, rb, s, g->m);
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
    oi->row_blocks = (uint32_t)(blocks.size() - 1);
    oi->blocks_per_group = group;
  }
  return SGC_OK;
}

}  // namespace
}  // namespace sgc

extern "C" int sgc_ooc_min_budget(const sgc_host_graph* g, double eps, uint64_t* bytes) {
  using namespace sgc;
  (void)eps;  // block chips need the same room at every eps
  clear_error();
  if (!g || !bytes) return fail(SGC_EINVAL, "null argument");
  if (g->n >= (1ull << 31)) return fail(SGC_ERANGE, "n=%llu: vertex ids must fit int32", (unsigned long long)g->n);
  uint64_t maxd = 1;
  for (uint64_t x = 0; x < g->n; ++x) maxd = std::max<uint64_t>(maxd, g->row_ptr[x + 1] - g->row_ptr[x]);
  // a chip of two single blocks that each hold the longest row
  *bytes = resident_layout((uint32_t)g->n).bytes + 8192 +
           sim_chip_bytes(maxd, maxd, record_capacity(maxd + 1, maxd)) + offsets_bytes(maxd, maxd) + 4096;
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
