// Host orchestration of the gfx950 compute_dmsc pipeline and its C-ABI (include/dmsc_hip.h).
//
// Replaces the reference's GPU host layer (dmsc/csrc/dmsc_gpu.cpp, gpu/dmsc_cuda.hxx,
// gpu/workspace.hxx), which bounces between device kernels and host-side persistence / DAG
// assembly. Here every stage runs on the device; the host only launches kernels and reads back a
// handful of sizes (critical-cell counts, round termination flags, output sizes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "dmsc_hip.h"
#include "k_topo.hpp"

namespace dmsc {

void fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t st);
void scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, int64_t* d_total, Arena& ar, int slot, hipStream_t st);

// arena slots
enum : int {
  A_SCAN = 0, A_RS0, A_RS1,
  A_FSUCC, A_VSUCC, A_PJF, A_PJV, A_PLANES, A_PREMIN, A_PRESAD, A_PREMAX,
  A_C32, A_C64,
  A_MINIDS, A_MINVAL, A_SADIDS, A_SADVAL, A_MAXIDS, A_MAXVAL,
  A_CMAX, A_CMIN, A_LMAX, A_LMIN,
  A_KEYS, A_KEYS2, A_VALS2, A_ASC, A_ORDMIN, A_ORDMAX, A_REMOVED, A_POSMIN, A_POSMAX,
  A_SIDE0,  // per persistence side: 13 slots each
  A_SIDE_END = A_SIDE0 + 2 * 13,
  A_ASM0 = A_SIDE_END,  // assembly: surv, surv_pre, rows x2, rows_pre x2, sad_to_out, to_out x2
  A_ASM_END = A_ASM0 + 12,
  A_ARC0 = A_ASM_END,  // per arc side: 24 slots each
  A_ARC_END = A_ARC0 + 2 * 24,
};

enum SideSlot { P_PARG, P_PART, P_ALIVE, P_MINKEY, P_PAIR, P_PERS, P_CANCEL, P_RR, P_WL0, P_WL1, P_CNT, P_ROOT, P_TOOUT };
enum ArcSlot {
  G_BLEN, G_BOFF, G_ARENA, G_FLAG, G_PRE, G_CS0, G_CD0, G_KEYS, G_SLOT, G_KEYS2, G_SLOT2, G_CS, G_CD, G_DEADC,
  G_GPAR, G_LLEN, G_LC, G_LKD, G_DONE, G_WATCH, G_RR, G_WL0, G_WL1, G_CNT
};
enum ArcSlot2 { H_PJQ = 0, H_FLEN, H_PERPOS, H_POSPRE, H_SEGA, H_SEGB, H_RAW, H_NSEG, H_LONG };

}  // namespace dmsc

namespace dmsc {
thread_local KernelTimes* tl_times = nullptr;
}

using namespace dmsc;

struct dmsc_context {
  int device = 0;
  KernelTimes times;
  Arena ar;
  Arena ar2;  // arc emission scratch
  int64_t* pinned = nullptr;
  std::string err;
  // last-call record for dmsc_debug_fetch
  Dims dims{};
  int64_t nmin = 0, nsad = 0, nmax = 0;
  int rounds_persist = 0, rounds_geo[2] = {0, 0};
  std::vector<int32_t> pending_log;  // (round, pending min side, pending max side) at each readback
  bool dual = false;
  const int32_t *min_ids = nullptr, *sad_ids = nullptr, *max_ids = nullptr, *asc = nullptr;
  const int32_t *cmax = nullptr, *cmin = nullptr, *pair_min = nullptr, *pair_max = nullptr;
  const float *pers_min = nullptr, *pers_max = nullptr;
  const uint64_t *pjf = nullptr, *pjv = nullptr;
  // batch workers (own scratch + stream each)
  std::vector<std::unique_ptr<dmsc_context>> workers;
  hipStream_t stream = nullptr;
};

namespace {
int batch_workers() {
  const char* e = std::getenv("DMSC_BATCH_WORKERS");
  int n = e ? std::atoi(e) : 8;
  return n < 1 ? 1 : (n > 64 ? 64 : n);
}
}  // namespace

namespace {

struct Fail {
  dmsc_status st;
  std::string msg;
};

inline int slot_pts(bool faces, bool dual) { return (faces != dual) ? 1 : 2; }
inline int slot_e(bool faces, bool dual) { return (faces != dual) ? 4 : 5; }
inline int slot_p(bool faces, bool dual) { return (faces != dual) ? 6 : 7; }
inline int slot_pp(bool faces, bool dual) { return (faces != dual) ? 8 : 9; }
inline int slot_labels(bool faces, bool dual) { return (faces != dual) ? 11 : 15; }
inline int slot_arcs(bool faces, bool dual) { return (faces != dual) ? 12 : 16; }

void* alloc_out(dmsc_alloc_fn alloc, void* user, int field, dmsc_outputs* out, int slot, int ndim, int64_t d0, int64_t d1,
                int dtype) {
  int64_t dims[2] = {d0, d1};
  out->ndim[slot] = ndim;
  out->dims[slot][0] = d0;
  out->dims[slot][1] = ndim == 2 ? d1 : 0;
  out->dtype[slot] = dtype;
  void* p = alloc(user, field, slot, ndim, dims, dtype);
  int64_t n = ndim == 2 ? d0 * d1 : d0;
  if (!p && n > 0) throw Fail{DMSC_ERR_ALLOC, "output allocator returned NULL for slot " + std::to_string(slot)};
  out->data[slot] = p;
  return p;
}

void readback(dmsc_context* c, const void* dev, size_t bytes, hipStream_t st) {
  DMSC_CHECK(hipMemcpyAsync(c->pinned, dev, bytes, hipMemcpyDeviceToHost, st));
  DMSC_CHECK(hipStreamSynchronize(st));
}

struct SideBufs {
  int32_t *par_g, *par_t, *pair, *cancel, *rr, *wl0, *wl1, *cnt, *root, *to_out;
  uint8_t* alive;
  uint64_t* minkey;
  float* pers;
};

SideBufs side_bufs(Arena& ar, int side, int64_t next, int64_t ns) {
  const int b = A_SIDE0 + 13 * side;
  SideBufs s;
  s.par_g = ar.get<int32_t>(b + P_PARG, next);
  s.par_t = ar.get<int32_t>(b + P_PART, next);
  s.alive = ar.get<uint8_t>(b + P_ALIVE, next);
  s.minkey = ar.get<uint64_t>(b + P_MINKEY, next);
  s.pair = ar.get<int32_t>(b + P_PAIR, ns);
  s.pers = ar.get<float>(b + P_PERS, ns);
  s.cancel = ar.get<int32_t>(b + P_CANCEL, ns);
  s.rr = ar.get<int32_t>(b + P_RR, 2 * ns);
  s.wl0 = ar.get<int32_t>(b + P_WL0, ns);
  s.wl1 = ar.get<int32_t>(b + P_WL1, ns);
  s.cnt = ar.get<int32_t>(b + P_CNT, 2);
  s.root = ar.get<int32_t>(b + P_ROOT, next);
  s.to_out = ar.get<int32_t>(b + P_TOOUT, next);
  return s;
}

// Everything one traced arc side needs between the persistence rounds and emission.
struct ArcWork {
  bool on = false;
  ArcSide A{};
  GeoState G{};
  int64_t arena_total = 0, ncancel = 0, out_total = 0;
  int64_t* flen = nullptr;
  int64_t* pos_pre = nullptr;
  uint64_t* pjq = nullptr;
  int32_t* out_cells = nullptr;
  int32_t* out_off = nullptr;
  int32_t* out_arc_off = nullptr;
};

void run_one(dmsc_context* ctx, const float* f, int H, int W, float thr, uint32_t flags, hipStream_t st,
             dmsc_alloc_fn alloc, void* user, int field, dmsc_outputs* out) {
  if (!f) throw Fail{DMSC_ERR_INVALID_ARGUMENT, "Input tensor must be defined"};
  if (H <= 0 || W <= 0)
    throw Fail{DMSC_ERR_INVALID_ARGUMENT,
               "Spatial dimensions must be non-empty, got " + std::to_string(H) + "x" + std::to_string(W)};
  if ((int64_t)H + 1 > (int64_t)INT32_MAX / 4 / ((int64_t)W + 1))
    throw Fail{DMSC_ERR_TOO_LARGE,
               "Input is too large for 32-bit cell IDs: " + std::to_string(H) + "x" + std::to_string(W)};
  const bool dual = flags & DMSC_DUAL, grad = flags & DMSC_RETURN_GRADIENT;
  bool tmax = flags & DMSC_TRACE_MAX_ARCS, tmin = flags & DMSC_TRACE_MIN_ARCS;
  bool gmax = flags & DMSC_TRACE_MAX_GROUPS, gmin = flags & DMSC_TRACE_MIN_GROUPS;
  if (dual) {  // reference dmsc_gpu.cpp:71-74: internal "max" = faces, "min" = vertices
    std::swap(tmax, tmin);
    std::swap(gmax, gmin);
  }
  Arena& ar = ctx->ar;
  const Dims d = make_dims(H, W);
  std::memset(out, 0, sizeof(*out));
  out->shape[0] = H;
  out->shape[1] = W;
  out->ndim[0] = 1;
  out->dims[0][0] = 2;

  // outputs whose sizes are known up front
  int32_t* o_grad = (int32_t*)alloc_out(alloc, user, field, out,10, 1, grad ? 4 * d.S : 0, 0, DMSC_DTYPE_INT32);
  int32_t* o_flab = (int32_t*)alloc_out(alloc, user, field, out,slot_labels(true, dual), 2, gmax ? d.Ny : 0,
                                        gmax ? d.Nx : 0, DMSC_DTYPE_INT32);
  int32_t* o_vlab = (int32_t*)alloc_out(alloc, user, field, out,slot_labels(false, dual), 2, gmin ? H : 0, gmin ? W : 0,
                                        DMSC_DTYPE_INT32);

  // ---- gradient, critical bit planes, successor forests
  int32_t* fsucc = ar.get<int32_t>(A_FSUCC, d.S);
  int32_t* vsucc = ar.get<int32_t>(A_VSUCC, d.N);
  uint32_t* planes = ar.get<uint32_t>(A_PLANES, 4 * d.PW);
  launch_gradient(dual, f, d, grad ? o_grad : nullptr, fsucc, vsucc, planes, st);

  int32_t* c32 = ar.get<int32_t>(A_C32, 64);
  int64_t* c64 = ar.get<int64_t>(A_C64, 64);
  int32_t* d_err = c32 + 40;
  fill_i32(d_err, 0, 1, st);
  int32_t* pre_min = ar.get<int32_t>(A_PREMIN, d.PW);
  int32_t* pre_sad = ar.get<int32_t>(A_PRESAD, d.PW);
  int32_t* pre_max = ar.get<int32_t>(A_PREMAX, d.PW);
  scan_plane_popcounts(planes, nullptr, pre_min, d.PW, c32 + 0, ar, A_SCAN, st);
  scan_plane_popcounts(planes + d.PW, planes + 2 * d.PW, pre_sad, d.PW, c32 + 1, ar, A_SCAN, st);
  scan_plane_popcounts(planes + 3 * d.PW, nullptr, pre_max, d.PW, c32 + 2, ar, A_SCAN, st);
  readback(ctx, c32, 3 * sizeof(int32_t), st);
  const int32_t* hc = (const int32_t*)ctx->pinned;
  const int64_t nmin = hc[0], nsad = hc[1], nmax = hc[2];
  const RankView rv{planes, pre_min, pre_sad, pre_max, d.PW, d.WPR, d.Nx, d.W};

  int32_t* min_ids = ar.get<int32_t>(A_MINIDS, nmin);
  float* min_val = ar.get<float>(A_MINVAL, nmin);
  int32_t* sad_ids = ar.get<int32_t>(A_SADIDS, nsad);
  float* sad_val = ar.get<float>(A_SADVAL, nsad);
  int32_t* max_ids = ar.get<int32_t>(A_MAXIDS, nmax);
  float* max_val = ar.get<float>(A_MAXVAL, nmax);
  launch_compact_lists(dual, f, d, rv, min_ids, min_val, sad_ids, sad_val, max_ids, max_val, st);

  // ---- V-path forests: terminal extremum and depth of every face / vertex
  uint64_t* pjf = ar.get<uint64_t>(A_PJF, d.S);
  uint64_t* pjv = ar.get<uint64_t>(A_PJV, d.N);
  pointer_jump_init(fsucc, pjf, d.S, st);
  pointer_jump_init(vsucc, pjv, d.N, st);
  pointer_jump(pjf, d.S, d_err, st);
  pointer_jump(pjv, d.N, d_err, st);

  // ---- saddle records and sweep orders
  int32_t* cmax = ar.get<int32_t>(A_CMAX, 2 * nsad);
  int32_t* cmin = ar.get<int32_t>(A_CMIN, 2 * nsad);
  int32_t* lmax = ar.get<int32_t>(A_LMAX, 2 * nsad);
  int32_t* lmin = ar.get<int32_t>(A_LMIN, 2 * nsad);
  launch_saddle_records(d, rv, sad_ids, nsad, pjf, pjv, cmax, cmin, lmax, lmin, st);

  uint32_t* keys = ar.get<uint32_t>(A_KEYS, nsad);
  uint32_t* keys2 = ar.get<uint32_t>(A_KEYS2, nsad);
  int32_t* asc = ar.get<int32_t>(A_ASC, nsad);
  int32_t* vals2 = ar.get<int32_t>(A_VALS2, nsad);
  launch_sort_keys(nsad, sad_val, keys, st);
  iota_i32(asc, nsad, st);
  radix_sort_pairs(keys, asc, keys2, vals2, nsad, ar, A_RS0, st);
  int32_t* ord_min = ar.get<int32_t>(A_ORDMIN, nsad);
  int32_t* ord_max = ar.get<int32_t>(A_ORDMAX, nsad);
  // primal: min side ascending, max side descending; dual: the reverse (arcs_topology.hxx:178-181)
  int32_t* pos_min = ar.get<int32_t>(A_POSMIN, nsad);
  int32_t* pos_max = ar.get<int32_t>(A_POSMAX, nsad);
  if (!dual) launch_orders(nsad, asc, ord_min, ord_max, pos_min, pos_max, st);
  else launch_orders(nsad, asc, ord_max, ord_min, pos_max, pos_min, st);

  // ---- persistence rounds (both sides together)
  uint8_t* removed = ar.get<uint8_t>(A_REMOVED, nsad);
  fill_u8(removed, 0, nsad, st);
  SideBufs sb[2] = {side_bufs(ar, 0, nmin, nsad), side_bufs(ar, 1, nmax, nsad)};
  const bool record[2] = {tmin, tmax};
  PersistSide ps[2];
  for (int s = 0; s < 2; ++s) {
    const int64_t next = s ? nmax : nmin;
    iota_i32(sb[s].par_g, next, st);
    iota_i32(sb[s].par_t, next, st);
    fill_u8(sb[s].alive, 1, next, st);
    fill_u64(sb[s].minkey, ~0ull, next, st);
    fill_i32(sb[s].pair, -1, nsad, st);
    fill_i32(sb[s].cancel, -1, nsad, st);
    iota_i32(sb[s].wl0, nsad, st);
    fill_i32(sb[s].cnt, (int32_t)nsad, 1, st);
    fill_i32(sb[s].cnt + 1, 0, 1, st);
    PersistSide& p = ps[s];
    p.pos = s ? pos_max : pos_min;
    p.c = s ? cmax : cmin;
    p.sval = sad_val;
    p.ev = s ? max_val : min_val;
    p.eid = s ? max_ids : min_ids;
    p.par_g = sb[s].par_g;
    p.par_t = sb[s].par_t;
    p.alive = sb[s].alive;
    p.minkey = sb[s].minkey;
    p.pair = sb[s].pair;
    p.pers = sb[s].pers;
    p.cancel_d = record[s] ? sb[s].cancel : nullptr;
    p.rr = sb[s].rr;
    p.wl[0] = sb[s].wl0;
    p.wl[1] = sb[s].wl1;
    p.cnt = sb[s].cnt;
    p.err = d_err;
    p.maxside = s;
  }
  int round = 0;
  ctx->pending_log.clear();
  if (nsad > 0) {
    while (true) {
      const int batch = round < 4 ? 4 : 8;
      for (int b = 0; b < batch; ++b, ++round) launch_persist_round(dual, ps[0], ps[1], round, thr, removed, nsad, st);
      DMSC_CHECK(hipMemcpyAsync(ctx->pinned, sb[0].cnt + (round & 1), 4, hipMemcpyDeviceToHost, st));
      DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 4, sb[1].cnt + (round & 1), 4, hipMemcpyDeviceToHost, st));
      DMSC_CHECK(hipStreamSynchronize(st));
      const int32_t* pc = (const int32_t*)ctx->pinned;
      ctx->pending_log.insert(ctx->pending_log.end(), {round, pc[0], pc[1]});
      if (pc[0] == 0 && pc[1] == 0) break;
      if (round > 4 * nsad + 64) throw Fail{DMSC_ERR_INTERNAL, "persistence rounds did not converge"};
    }
  }
  launch_uf_flatten(sb[0].par_t, sb[0].root, nmin, st);
  launch_uf_flatten(sb[1].par_t, sb[1].root, nmax, st);

  // ---- assembly counts (cpu/workspace.hxx:107-168)
  int32_t* tmp_i32 = ar.get<int32_t>(A_ASM0 + 0, std::max(nmin, nmax));
  int32_t* surv = ar.get<int32_t>(A_ASM0 + 1, nsad);
  int32_t* surv_pre = ar.get<int32_t>(A_ASM0 + 2, nsad);
  int32_t* rows[2] = {ar.get<int32_t>(A_ASM0 + 3, nsad), ar.get<int32_t>(A_ASM0 + 4, nsad)};
  int32_t* rows_pre[2] = {ar.get<int32_t>(A_ASM0 + 5, nsad), ar.get<int32_t>(A_ASM0 + 6, nsad)};
  int32_t* sad_to_out = ar.get<int32_t>(A_ASM0 + 7, nsad);
  for (int s = 0; s < 2; ++s) {
    const int64_t next = s ? nmax : nmin;
    launch_u8_to_i32(sb[s].alive, tmp_i32, next, false, st);
    scan_exclusive<int32_t>(tmp_i32, sb[s].to_out, next, c32 + 3 + s, ar, A_SCAN, st);
    launch_mask_prefix(sb[s].alive, sb[s].to_out, next, st);
  }
  launch_side_counts(nsad, ord_min, removed, cmin, sb[0].root, sb[0].to_out, surv, rows[0], st);
  launch_side_counts(nsad, ord_max, removed, cmax, sb[1].root, sb[1].to_out, nullptr, rows[1], st);
  scan_exclusive<int32_t>(surv, surv_pre, nsad, c32 + 5, ar, A_SCAN, st);
  scan_exclusive<int32_t>(rows[0], rows_pre[0], nsad, c32 + 6, ar, A_SCAN, st);
  scan_exclusive<int32_t>(rows[1], rows_pre[1], nsad, c32 + 7, ar, A_SCAN, st);
  launch_sad_to_out(nsad, ord_min, removed, surv_pre, sad_to_out, st);

  // ---- arc geometry, per traced side: raw arcs and cancellation lists
  ArcWork aw[2];  // 0 = vertices (min arcs), 1 = faces (max arcs)
  const bool traced[2] = {tmin, tmax};
  for (int s = 0; s < 2; ++s) {
    if (!traced[s]) continue;
    ArcWork& w = aw[s];
    w.on = true;
    const int b = A_ARC0 + 24 * s;
    int32_t* blen = ar.get<int32_t>(b + G_BLEN, 2 * nsad);
    int64_t* boff = ar.get<int64_t>(b + G_BOFF, 2 * nsad);
    launch_arc_lengths(2 * nsad, s ? cmax : cmin, s ? lmax : lmin, blen, st);
    scan_i32_to_i64(blen, boff, 2 * nsad, c64 + 2 * s, ar, A_SCAN, st);
    int32_t* flag = ar.get<int32_t>(b + G_FLAG, nsad);
    int32_t* pre = ar.get<int32_t>(b + G_PRE, nsad);
    launch_cancel_flags(nsad, sb[s].cancel, flag, st);
    scan_exclusive<int32_t>(flag, pre, nsad, c32 + 8 + s, ar, A_SCAN, st);
    w.A.faces = s;
    w.A.succ = s ? fsucc : vsucc;
    w.A.Nx = d.Nx;
    w.A.W = W;
    w.A.H = H;
    w.A.ns = nsad;
    w.A.next = s ? nmax : nmin;
    w.A.sad_ids = sad_ids;
    w.A.tinit = s ? cmax : cmin;
    w.A.blen = blen;
    w.A.boff = boff;
  }
  readback(ctx, c32, 16 * sizeof(int32_t), st);
  int32_t h32[16];
  std::memcpy(h32, ctx->pinned, sizeof(h32));
  readback(ctx, c64, 4 * sizeof(int64_t), st);
  int64_t h64[4];
  std::memcpy(h64, ctx->pinned, sizeof(h64));
  const int64_t n_out_ext[2] = {h32[3], h32[4]};
  const int64_t n_sad_out = h32[5];
  const int64_t n_rows[2] = {h32[6], h32[7]};

  for (int s = 0; s < 2; ++s) {
    if (!aw[s].on) continue;
    ArcWork& w = aw[s];
    const int b = A_ARC0 + 24 * s;
    const int64_t next = s ? nmax : nmin;
    w.arena_total = h64[2 * s];
    w.ncancel = h32[8 + s];
    w.A.arena = ar.get<int32_t>(b + G_ARENA, w.arena_total);
    launch_raw_arcs(w.A, st);
    // cancellations sorted by (persistence, saddle id) (persistence_struct.hxx:15-19)
    const int64_t nc = w.ncancel;
    int32_t* cs0 = ar.get<int32_t>(b + G_CS0, nc);
    int32_t* cd0 = ar.get<int32_t>(b + G_CD0, nc);
    uint32_t* ck = ar.get<uint32_t>(b + G_KEYS, nc);
    int32_t* cslot = ar.get<int32_t>(b + G_SLOT, nc);
    launch_cancel_compact(nsad, sb[s].cancel, sb[s].pers, ar.get<int32_t>(b + G_PRE, nsad), cs0, cd0, ck, st);
    iota_i32(cslot, nc, st);
    radix_sort_pairs(ck, cslot, ar.get<uint32_t>(b + G_KEYS2, nc), ar.get<int32_t>(b + G_SLOT2, nc), nc, ar, A_RS0, st);
    GeoState& g = w.G;
    g.ncancel = nc;
    int32_t* ckey = ar.get<int32_t>(b + G_CS, nc);
    int32_t* deadc = ar.get<int32_t>(b + G_DEADC, next);
    fill_i32(deadc, -1, next, st);
    launch_cancel_finish(nc, cslot, cd0, ckey, deadc, st);
    g.cs = cs0;
    g.cd = cd0;
    g.key = ckey;
    g.deadc = deadc;
    g.gparent = ar.get<int32_t>(b + G_GPAR, next);
    g.linklen = ar.get<int32_t>(b + G_LLEN, next);
    g.linkc = ar.get<int32_t>(b + G_LC, next);
    g.linkkd = ar.get<int32_t>(b + G_LKD, next);
    g.done = ar.get<int32_t>(b + G_DONE, nc);
    g.watch = ar.get<uint64_t>(b + G_WATCH, next);
    g.rr = ar.get<int32_t>(b + G_RR, 4 * nc);
    g.wl[0] = ar.get<int32_t>(b + G_WL0, nc);
    g.wl[1] = ar.get<int32_t>(b + G_WL1, nc);
    g.cnt = ar.get<int32_t>(b + G_CNT, 2);
    g.err = d_err;
    iota_i32(g.gparent, next, st);
    fill_i32(g.linklen, 0, next, st);
    fill_i32(g.done, INT_MAX, nc, st);
    fill_u64(g.watch, ~0ull, next, st);
    iota_i32(g.wl[0], nc, st);
    fill_i32(g.cnt, (int32_t)nc, 1, st);
    fill_i32(g.cnt + 1, 0, 1, st);
  }
  // geometry rounds of both sides, interleaved, one readback per batch of rounds
  {
    bool live[2] = {aw[0].on && aw[0].ncancel > 0, aw[1].on && aw[1].ncancel > 0};
    int r = 0;
    ctx->rounds_geo[0] = ctx->rounds_geo[1] = 0;
    while (live[0] || live[1]) {
      const int batch = r < 4 ? 4 : 8;
      for (int k = 0; k < batch; ++k, ++r)
        for (int s = 0; s < 2; ++s)
          if (live[s]) launch_geo_round(aw[s].A, aw[s].G, r, aw[s].ncancel, st);
      for (int s = 0; s < 2; ++s)
        if (live[s]) DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 4 * s, aw[s].G.cnt + (r & 1), 4,
                                               hipMemcpyDeviceToHost, st));
      DMSC_CHECK(hipStreamSynchronize(st));
      for (int s = 0; s < 2; ++s)
        if (live[s]) {
          ctx->rounds_geo[s] = r;
          if (((const int32_t*)ctx->pinned)[s] == 0) live[s] = false;
          else if (r > 4 * aw[s].ncancel + 64) throw Fail{DMSC_ERR_INTERNAL, "geometry rounds did not converge"};
        }
    }
  }
  for (int s = 0; s < 2; ++s) {
    if (!aw[s].on) continue;
    ArcWork& w = aw[s];
    GeoState& g = w.G;
    const int64_t next = s ? nmax : nmin;
    // path lengths to the final roots and per-output-arc lengths (min order)
    Arena& a2 = ctx->ar2;
    const int hb = 16 * s;
    w.pjq = a2.get<uint64_t>(hb + H_PJQ, next + 1);
    launch_geo_q(g, next, w.pjq, st);
    w.flen = a2.get<int64_t>(hb + H_FLEN, 2 * nsad);
    int64_t* per_pos = a2.get<int64_t>(hb + H_PERPOS, nsad);
    w.pos_pre = a2.get<int64_t>(hb + H_POSPRE, nsad);
    launch_final_lengths(nsad, ord_min, removed, w.A, w.pjq, next, w.flen, per_pos, st);
    scan_exclusive<int64_t>(per_pos, w.pos_pre, nsad, c64 + 2 * s + 1, ar, A_SCAN, st);
  }
  readback(ctx, c64, 4 * sizeof(int64_t), st);
  std::memcpy(h64, ctx->pinned, sizeof(h64));
  for (int s = 0; s < 2; ++s) {
    if (!aw[s].on) continue;
    aw[s].out_total = h64[2 * s + 1];
    if (aw[s].out_total > INT32_MAX)
      throw Fail{DMSC_ERR_TOO_LARGE, "merged arc geometry exceeds 32-bit offsets"};
  }

  // ---- outputs (cpu/workspace.hxx:291-362 layout, dual swap by slot)
  int32_t* o_pts[2];
  for (int s = 0; s < 2; ++s)
    o_pts[s] = (int32_t*)alloc_out(alloc, user, field, out,slot_pts(s == 1, dual), 1, n_out_ext[s], 0, DMSC_DTYPE_INT32);
  int32_t* o_sad = (int32_t*)alloc_out(alloc, user, field, out,3, 1, n_sad_out, 0, DMSC_DTYPE_INT32);
  int32_t *o_e[2], *o_pp[2];
  float* o_p[2];
  for (int s = 0; s < 2; ++s) {
    o_e[s] = (int32_t*)alloc_out(alloc, user, field, out,slot_e(s == 1, dual), 2, n_rows[s], 2, DMSC_DTYPE_INT32);
    o_p[s] = (float*)alloc_out(alloc, user, field, out,slot_p(s == 1, dual), 1, n_sad_out, 0, DMSC_DTYPE_FLOAT32);
    o_pp[s] = (int32_t*)alloc_out(alloc, user, field, out,slot_pp(s == 1, dual), 1, n_sad_out, 0, DMSC_DTYPE_INT32);
  }
  for (int s = 0; s < 2; ++s) {
    const int sl = slot_arcs(s == 1, dual);
    ArcWork& w = aw[s];
    w.out_cells = (int32_t*)alloc_out(alloc, user, field, out,sl, 1, w.on ? w.out_total : 0, 0, DMSC_DTYPE_INT32);
    w.out_off = (int32_t*)alloc_out(alloc, user, field, out,sl + 1, 1, w.on ? n_sad_out + 1 : 0, 0, DMSC_DTYPE_INT32);
    w.out_arc_off = (int32_t*)alloc_out(alloc, user, field, out,sl + 2, 1, w.on ? n_sad_out : 0, 0, DMSC_DTYPE_INT32);
  }

  launch_emit_points(nmin, min_ids, sb[0].to_out, o_pts[0], st);
  launch_emit_points(nmax, max_ids, sb[1].to_out, o_pts[1], st);
  launch_emit_side(nsad, ord_min, removed, sad_to_out, cmin, sb[0].root, sb[0].to_out, sb[0].pers, sb[0].pair,
                   rows_pre[0], sad_ids, o_p[0], o_pp[0], o_e[0], o_sad, st);
  launch_emit_side(nsad, ord_max, removed, sad_to_out, cmax, sb[1].root, sb[1].to_out, sb[1].pers, sb[1].pair,
                   rows_pre[1], sad_ids, o_p[1], o_pp[1], o_e[1], nullptr, st);
  if (gmax) launch_labels_faces(d, rv, pjf, sb[1].root, sb[1].to_out, o_flab, st);
  if (gmin) launch_labels_verts(d, rv, pjv, sb[0].root, sb[0].to_out, o_vlab, st);

  for (int s = 0; s < 2; ++s) {
    ArcWork& w = aw[s];
    if (!w.on) continue;
    Arena& a2 = ctx->ar2;
    const int hb = 16 * s;
    const int64_t next = s ? nmax : nmin;
    const int64_t cap = w.out_total + 16;
    Seg* sa = a2.get<Seg>(hb + H_SEGA, cap);
    Seg* sbuf = a2.get<Seg>(hb + H_SEGB, cap);
    Seg* raw = a2.get<Seg>(hb + H_RAW, cap);
    int32_t* nseg = a2.get<int32_t>(hb + H_NSEG, 4);  // [0] list A, [1] list B, [2] raw
    fill_i32(nseg, 0, 4, st);
    launch_seed_segments(nsad, ord_min, removed, w.A, w.pjq, next, w.flen, w.pos_pre, sad_to_out, sa, nseg, raw,
                         nseg + 2, cap, d_err, w.out_off, w.out_arc_off, n_sad_out, c64 + 2 * s + 1, st);
    readback(ctx, nseg, 3 * sizeof(int32_t), st);
    int32_t n_in = ((const int32_t*)ctx->pinned)[0];
    Seg* in = sa;
    Seg* nxt = sbuf;
    int32_t* cnt_in = nseg;
    int32_t* cnt_out = nseg + 1;
    int guard = 0;
    while (n_in > 0) {
      if (n_in > cap) throw Fail{DMSC_ERR_INTERNAL, "arc segment list overflow"};
      fill_i32(cnt_out, 0, 1, st);
      launch_expand_segments(w.A, w.G, w.pjq, in, cnt_in, nxt, cnt_out, raw, nseg + 2, n_in, cap, st);
      readback(ctx, nseg, 3 * sizeof(int32_t), st);
      n_in = ((const int32_t*)ctx->pinned)[cnt_out - nseg];
      std::swap(in, nxt);
      std::swap(cnt_in, cnt_out);
      if (++guard > 1000000) throw Fail{DMSC_ERR_INTERNAL, "arc expansion did not terminate"};
    }
    const int32_t nraw = ((const int32_t*)ctx->pinned)[2];
    if (nraw > cap) throw Fail{DMSC_ERR_INTERNAL, "arc segment list overflow"};
    if (nraw > 0)
      launch_copy_raw(w.A, raw, nseg + 2, nraw, w.out_cells, w.out_total, d_err, a2.get<int32_t>(hb + H_LONG, nraw),
                      nseg + 3, st);
  }
  readback(ctx, d_err, sizeof(int32_t), st);
  if (ctx->times.on) ctx->times.collect();
  if (*(const int32_t*)ctx->pinned != 0)
    throw Fail{DMSC_ERR_INTERNAL, "device consistency check failed (flags " +
                                      std::to_string(*(const int32_t*)ctx->pinned) + ")"};

  // record for dmsc_debug_fetch
  ctx->dims = d;
  ctx->nmin = nmin;
  ctx->nsad = nsad;
  ctx->nmax = nmax;
  ctx->rounds_persist = round;
  ctx->dual = dual;
  ctx->min_ids = min_ids;
  ctx->sad_ids = sad_ids;
  ctx->max_ids = max_ids;
  ctx->asc = asc;
  ctx->cmax = cmax;
  ctx->cmin = cmin;
  ctx->pair_min = sb[0].pair;
  ctx->pair_max = sb[1].pair;
  ctx->pers_min = sb[0].pers;
  ctx->pers_max = sb[1].pers;
  ctx->pjf = pjf;
  ctx->pjv = pjv;
}

dmsc_status guarded(dmsc_context* ctx, const std::function<void()>& fn) {
  if (!ctx) return DMSC_ERR_INVALID_ARGUMENT;
  ctx->err.clear();
  tl_times = &ctx->times;
  try {
    DMSC_CHECK(hipSetDevice(ctx->device));
    fn();
    return DMSC_OK;
  } catch (const Fail& e) {
    ctx->err = e.msg;
    return e.st;
  } catch (const HipError& e) {
    ctx->err = std::string(e.what) + ": " + hipGetErrorString(e.err);
    return DMSC_ERR_HIP;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return DMSC_ERR_INTERNAL;
  }
}
}  // namespace

extern "C" {

dmsc_status dmsc_context_create(int device, dmsc_context** out) {
  if (!out) return DMSC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* c = new dmsc_context();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipHostMalloc((void**)&c->pinned, 4096) != hipSuccess) {
    delete c;
    return DMSC_ERR_HIP;
  }
  *out = c;
  return DMSC_OK;
}

void dmsc_context_destroy(dmsc_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto& w : ctx->workers) {
    w->ar.release();
    w->ar2.release();
    if (w->pinned) (void)hipHostFree(w->pinned);
    if (w->stream) (void)hipStreamDestroy(w->stream);
  }
  ctx->workers.clear();
  ctx->ar.release();
  ctx->ar2.release();
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  delete ctx;
}

dmsc_status dmsc_extract(dmsc_context* ctx, const float* d_field, int32_t H, int32_t W, float thr, uint32_t flags,
                         dmsc_stream_t stream, dmsc_alloc_fn alloc, void* user, dmsc_outputs* out) {
  if (!alloc || !out) return DMSC_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] { run_one(ctx, d_field, H, W, thr, flags, (hipStream_t)stream, alloc, user, 0, out); });
}

dmsc_status dmsc_extract_batch(dmsc_context* ctx, int32_t B, const float* const* d_fields, const int32_t* H,
                               const int32_t* W, const float* thr, uint32_t flags, dmsc_stream_t stream,
                               dmsc_alloc_fn alloc, void* user, dmsc_outputs* out) {
  if (!alloc || !out || B < 0 || (B > 0 && (!d_fields || !H || !W || !thr))) return DMSC_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    const int nw = std::min<int>(B, batch_workers());
    if (nw <= 1) {
      for (int b = 0; b < B; ++b)
        run_one(ctx, d_fields[b], H[b], W[b], thr[b], flags, (hipStream_t)stream, alloc, user, b, &out[b]);
      return;
    }
    // Worker contexts, each with its own scratch and its own (default-priority) HIP stream; fields
    // are taken from a shared counter, so small fields' latency-bound stages overlap on the device.
    while ((int)ctx->workers.size() < nw) {
      auto* w = new dmsc_context();
      w->device = ctx->device;
      DMSC_CHECK(hipHostMalloc((void**)&w->pinned, 4096));
      DMSC_CHECK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
      ctx->workers.emplace_back(w);
    }
    hipEvent_t ready;
    DMSC_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    DMSC_CHECK(hipEventRecord(ready, (hipStream_t)stream));  // inputs were produced on the caller's stream
    for (int k = 0; k < nw; ++k) DMSC_CHECK(hipStreamWaitEvent(ctx->workers[k]->stream, ready, 0));
    std::atomic<int> next{0};
    std::mutex mu;
    Fail first{DMSC_OK, ""};
    std::vector<std::thread> th;
    for (int k = 0; k < nw; ++k) {
      dmsc_context* w = ctx->workers[k].get();
      w->times.on = ctx->times.on;
      th.emplace_back([&, w] {
        tl_times = &w->times;
        if (hipSetDevice(w->device) != hipSuccess) return;
        for (int b; (b = next.fetch_add(1)) < B;) {
          try {
            run_one(w, d_fields[b], H[b], W[b], thr[b], flags, w->stream, alloc, user, b, &out[b]);
          } catch (const Fail& e) {
            std::lock_guard<std::mutex> lk(mu);
            if (first.st == DMSC_OK) first = e;
            next = B;
          } catch (const HipError& e) {
            std::lock_guard<std::mutex> lk(mu);
            if (first.st == DMSC_OK) first = Fail{DMSC_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.err)};
            next = B;
          } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(mu);
            if (first.st == DMSC_OK) first = Fail{DMSC_ERR_INTERNAL, e.what()};
            next = B;
          }
        }
      });
    }
    for (auto& t : th) t.join();
    (void)hipEventDestroy(ready);
    for (auto& w : ctx->workers) {
      for (auto& kv : w->times.acc) {
        ctx->times.add(kv.first.c_str(), 0.0);
        for (auto& mine : ctx->times.acc)
          if (mine.first == kv.first) {
            mine.second.first += kv.second.first;
            mine.second.second += kv.second.second - 1;
          }
      }
      w->times.acc.clear();
    }
    tl_times = &ctx->times;
    if (first.st != DMSC_OK) throw first;
  });
}

const char* dmsc_last_error(dmsc_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

dmsc_status dmsc_profile_enable(dmsc_context* ctx, int32_t on) {
  if (!ctx) return DMSC_ERR_INVALID_ARGUMENT;
  ctx->times.on = on != 0;
  ctx->times.acc.clear();
  return DMSC_OK;
}
int32_t dmsc_profile_count(dmsc_context* ctx) { return ctx ? (int32_t)ctx->times.acc.size() : 0; }
dmsc_status dmsc_profile_entry(dmsc_context* ctx, int32_t i, const char** name, double* total_ms, int64_t* launches) {
  if (!ctx || i < 0 || i >= (int32_t)ctx->times.acc.size()) return DMSC_ERR_INVALID_ARGUMENT;
  const auto& kv = ctx->times.acc[i];
  if (name) *name = kv.first.c_str();
  if (total_ms) *total_ms = kv.second.first;
  if (launches) *launches = kv.second.second;
  return DMSC_OK;
}

int64_t dmsc_debug_fetch(dmsc_context* ctx, const char* name, void* dst, int64_t cap) {
  if (!ctx || !name) return -1;
  std::string n(name);
  auto copy = [&](const void* src, int64_t count, size_t elem) -> int64_t {
    int64_t m = std::min(count, cap);
    if (m > 0 && src && dst) {
      if (hipMemcpy(dst, src, (size_t)m * elem, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    }
    return count;
  };
  if (n == "counts") {
    int32_t v[3] = {(int32_t)ctx->nmin, (int32_t)ctx->nsad, (int32_t)ctx->nmax};
    if (dst) std::memcpy(dst, v, sizeof(int32_t) * (size_t)std::min<int64_t>(3, cap));
    return 3;
  }
  if (n == "rounds") {
    int32_t v[3] = {ctx->rounds_persist, ctx->rounds_geo[0], ctx->rounds_geo[1]};
    if (dst) std::memcpy(dst, v, sizeof(int32_t) * (size_t)std::min<int64_t>(3, cap));
    return 3;
  }
  if (n == "persist_pending") {
    int64_t m = std::min<int64_t>((int64_t)ctx->pending_log.size(), cap);
    if (dst && m > 0) std::memcpy(dst, ctx->pending_log.data(), (size_t)m * 4);
    return (int64_t)ctx->pending_log.size();
  }
  if (n == "min_ids") return copy(ctx->min_ids, ctx->nmin, 4);
  if (n == "sad_ids") return copy(ctx->sad_ids, ctx->nsad, 4);
  if (n == "max_ids") return copy(ctx->max_ids, ctx->nmax, 4);
  if (n == "asc_order") return copy(ctx->asc, ctx->nsad, 4);
  if (n == "sad_cmax") return copy(ctx->cmax, 2 * ctx->nsad, 4);
  if (n == "sad_cmin") return copy(ctx->cmin, 2 * ctx->nsad, 4);
  if (n == "pair_min") return copy(ctx->pair_min, ctx->nsad, 4);
  if (n == "pair_max") return copy(ctx->pair_max, ctx->nsad, 4);
  if (n == "pers_min") return copy(ctx->pers_min, ctx->nsad, 4);
  if (n == "pers_max") return copy(ctx->pers_max, ctx->nsad, 4);
  if (n == "face_pj") return copy(ctx->pjf, ctx->dims.S, 8);
  if (n == "vert_pj") return copy(ctx->pjv, ctx->dims.N, 8);
  return -1;
}

}  // extern "C"
