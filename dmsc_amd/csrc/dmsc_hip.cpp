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
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "dmsc_hip.h"
#include "k_diff.hpp"
#include "k_topo.hpp"

namespace dmsc {

void fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t st);
void fill_f32_zero(float* p, int64_t n, hipStream_t st);
void scan_arc_offsets(const int32_t* blen, const uint8_t* removed, const int32_t* cancel, int32_t min_len, int64_t* out,
                      int64_t n, int64_t* d_total, Arena& ar, int slot, hipStream_t st);

// arena slots
enum : int {
  A_SCAN = 0, A_RS0, A_RS1,
  A_SCAN2, A_RS2, A_RS3, A_RSP2, A_RSP3,  // scratch of the per-side work run on the side stream
  A_FSUCC, A_VSUCC, A_PJF, A_PJV, A_PLANES, A_PREMIN, A_PRESAD, A_PREMAX,
  A_C32, A_C64,
  A_MINIDS, A_MINVAL, A_SADIDS, A_SADVAL, A_MAXIDS, A_MAXVAL,
  A_CMAX, A_CMIN, A_LMAX, A_LMIN, A_META, A_BASES32, A_SADFIELD, A_LOSSPART, A_SURV1, A_SURVR1,
  A_KEYS, A_KEYS2, A_VALS2, A_ASC, A_ORDMIN, A_ORDMAX, A_REMOVED, A_POSMIN, A_POSMAX, A_RSP0, A_RSP1,
  A_SIDE0,  // per persistence side: kSideSlots slots each
  A_SIDE_END = A_SIDE0 + 2 * 15,
  A_ASM0 = A_SIDE_END,  // assembly: counts, prefixes, output map, survivor lists, edge outputs (16 slots)
  A_ASM_END = A_ASM0 + 16,
  A_ARC0 = A_ASM_END,  // per arc side: kArcSlots slots each
  A_ARC_END = A_ARC0 + 2 * 21,
};

enum SideSlot { P_PARG, P_PART, P_ALIVE, P_MINKEY, P_PAIR, P_PERS, P_CANCEL, P_RR, P_WL, P_CNT, P_ROOT, P_TOOUT,
                P_CROOT, P_RKEY, P_WKEY, kSideSlots };
static_assert(A_SIDE_END - A_SIDE0 == 2 * kSideSlots, "side slot count");
enum ArcSlot {
  G_BLEN, G_BOFF, G_ARENA, G_FLAG, G_PRE, G_CS0, G_CD0, G_KEYS, G_CS, G_CD, G_DEADC,
  G_LINK, G_DONE, G_WATCH, G_RR, G_WL, G_CNT, G_JUMP, G_GLEN, G_TL, G_HEAD, kArcSlots
};
static_assert(A_ARC_END - A_ARC0 == 2 * kArcSlots, "arc slot count");
enum ArcSlot2 { H_PJQ = 0, H_FLEN, H_PERPOS, H_POSPRE, H_SEGA, H_SEGB, H_RAW, H_NSEG, H_CUMR, H_CUMC, H_ARCB, H_REC, H_XL, H_NCQ };

}  // namespace dmsc

namespace dmsc {
thread_local KernelTimes* tl_times = nullptr;
thread_local const Tunables* tl_tune = nullptr;
}

using namespace dmsc;

// ---- scratch of the allocation-free backward (dmsc_persistence_backward_noalloc)
// A fixed array in the code object: each GPU the library runs on holds its own copy (64 MiB) until
// the process ends. Host code never uses the array's name as an address: dmsc_context_create looks
// up each device's copy with hipGetSymbolAddress once and keeps that pointer here.
constexpr int64_t kNoAllocScratchFloats = int64_t(1) << 24;
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                      // guards the two tables below and the scratch's users
float* g_scratch_ptr[kMaxDevices] = {};       // device address of `scratch` on each device (null: not set up)
bool g_scratch_dirty[kMaxDevices] = {};       // true while the scratch may hold non-zero entries

// Once per device, with that device current: the device address of its `scratch`, zeroed (the
// backward kernels need it all zeros between calls).
dmsc_status scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return DMSC_OK;  // the allocation-free entry refuses such devices
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return DMSC_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DMSC_ERR_HIP;
  try {
    fill_f32_zero(static_cast<float*>(p), kNoAllocScratchFloats, nullptr);
  } catch (const HipError&) {
    return DMSC_ERR_HIP;
  }
  if (hipStreamSynchronize(nullptr) != hipSuccess) return DMSC_ERR_HIP;
  g_scratch_ptr[device] = static_cast<float*>(p);
  g_scratch_dirty[device] = false;
  return DMSC_OK;
}
}  // namespace

struct dmsc_context {
  // Every entry point holds this lock for its whole host-side run: calls on one context from several
  // threads take turns (the arena, pinned buffer, side stream and events are shared state).
  std::mutex mu;
  int device = 0;
  KernelTimes times;
  Tunables tune;
  hipStream_t cap_stream = nullptr;  // private stream that round batches are captured on
  hipStream_t side = nullptr;        // output emission, concurrent with the arc geometry
  hipStream_t late = nullptr;        // labels, concurrent with the last geometry rounds
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_mark = nullptr, ev_side = nullptr;
  hipEvent_t ev_late = nullptr, ev_late_join = nullptr;
  Arena ar;
  Arena ar2;  // arc emission scratch
  int64_t* pinned = nullptr;
  size_t pinned_cap = 0;
  std::map<std::thread::id, std::string> err;  // message of each thread's last failing call
  // last-call record for dmsc_debug_fetch
  int64_t S = 0, N = 0;
  int64_t nmin = 0, nsad = 0, nmax = 0;
  int rounds_persist = 0, rounds_geo[2] = {0, 0};
  int64_t arc_stats[4] = {0, 0, 0, 0};  // per traced side (vertices, faces): raw pieces copied, output cells
  std::vector<int32_t> pending_log;  // (round, pending min side, pending max side) at each readback
  std::vector<int32_t> geo_log;      // the same for the geometry rounds (first entry: round 0, cancellations)
  bool dual = false;
  const int32_t *min_ids = nullptr, *sad_ids = nullptr, *max_ids = nullptr, *asc = nullptr;
  const int32_t *cmax = nullptr, *cmin = nullptr, *pair_min = nullptr, *pair_max = nullptr;
  const float *pers_min = nullptr, *pers_max = nullptr;
  const uint64_t *pjf = nullptr, *pjv = nullptr;
  // persistence-loss backward scratch: grow-only, all zeros between calls unless bwd_dirty
  float* bwd_scratch = nullptr;
  int64_t bwd_cap = 0;  // floats
  bool bwd_dirty = false;
  hipEvent_t ev_bwd = nullptr;  // end of the last loss / backward work (orders calls on other streams)
  bool bwd_pending = false;
  // Further parts of a large batch (pipe_groups): each with its own context (scratch, pinned buffer,
  // streams) driven by a worker thread on its own stream, ordered after the caller's stream by ev_pin
  // and before it by its ev_pout.
  static constexpr int kMaxPeers = 7;
  dmsc_context* peer[kMaxPeers] = {};
  hipStream_t peer_st[kMaxPeers] = {};
  hipEvent_t ev_pin = nullptr, ev_pout[kMaxPeers] = {};
};

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
  int32_t *par_g, *par_t, *pair, *cancel, *rr, *wl, *cnt, *root, *to_out, *croot;
  uint32_t *rkey, *wkey;
  uint8_t* alive;
  uint64_t* minkey;
  float* pers;
};

SideBufs side_bufs(Arena& ar, int side, int64_t next, int64_t ns) {
  const int b = A_SIDE0 + kSideSlots * side;
  SideBufs s;
  s.par_g = ar.get<int32_t>(b + P_PARG, next);
  s.par_t = ar.get<int32_t>(b + P_PART, next);
  s.alive = ar.get<uint8_t>(b + P_ALIVE, next);
  s.minkey = ar.get<uint64_t>(b + P_MINKEY, next);
  s.pair = ar.get<int32_t>(b + P_PAIR, ns);
  s.pers = ar.get<float>(b + P_PERS, ns);
  s.cancel = ar.get<int32_t>(b + P_CANCEL, ns);
  s.rr = ar.get<int32_t>(b + P_RR, 2 * ns);
  s.wl = ar.get<int32_t>(b + P_WL, 2 * ns);
  s.cnt = ar.get<int32_t>(b + P_CNT, 2);
  s.root = ar.get<int32_t>(b + P_ROOT, next);
  s.to_out = ar.get<int32_t>(b + P_TOOUT, next);
  s.croot = ar.get<int32_t>(b + P_CROOT, 4 * ns);
  s.wkey = ar.get<uint32_t>(b + P_WKEY, 2 * ns);
  s.rkey = ar.get<uint32_t>(b + P_RKEY, ns);
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
  int4* rec = nullptr;  // per surviving saddle (output order): arc record, 32 bytes
  int64_t* arc_base = nullptr;  // [B+1] device
  uint64_t* pjq = nullptr;
  uint64_t* cum_r = nullptr;  // per extremum: raw / chain pieces of the links on its path to the root
  uint64_t* cum_c = nullptr;
  LinkRec* xl = nullptr;      // per extremum: its link record
  int4* ncq = nullptr;        // per extremum: raw / chain pieces and path length to the root, packed
  int32_t* out_cells = nullptr;
  int32_t* out_off = nullptr;
  int32_t* out_arc_off = nullptr;
  std::vector<int64_t> h_arc_base;
};

// A batch of rounds recorded once and replayed on the work stream; its kernels read the round number
// from device memory (set before each replay), so one graph serves every batch of the same size
// while the launch sizes recorded in it still fit the worklists.
struct CapturedRounds {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int batch = 0;
  int64_t items = 0;  // worklist size the launch grids were sized for
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~CapturedRounds() { reset(); }
  bool fits(int b, int64_t n) const { return exec && batch == b && n <= items && 4 * n >= items; }
  template <class F>
  void capture(dmsc_context* ctx, int b, int64_t n, F&& record) {
    reset();
    if (!ctx->cap_stream) DMSC_CHECK(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
    KernelTimes* saved = tl_times;
    tl_times = nullptr;  // no timing events inside a graph
    DMSC_CHECK(hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal));
    try {
      record(ctx->cap_stream);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(ctx->cap_stream, &g);
      if (g) (void)hipGraphDestroy(g);
      tl_times = saved;
      throw;
    }
    tl_times = saved;
    DMSC_CHECK(hipStreamEndCapture(ctx->cap_stream, &graph));
    DMSC_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    batch = b;
    items = n;
  }
};

// Rounds to launch before the next host check, after `done` rounds: rounds_first, then doubling
// up to rounds_max (a round past convergence finds an empty worklist and only costs its launch).
int round_batch(int done) {
  const Tunables& t = tune();
  int b = t.rounds_first;
  for (int sum = b; sum <= done && b < t.rounds_max; sum += b) b = std::min(2 * b, t.rounds_max);
  return b;
}
// The same, shortened once two readbacks show the pending count shrinking (pend after `done` rounds,
// prev_pend after prev_done): only the rounds that the last batch's per-round decay needs to bring it
// down to the tail threshold, after which the single-workgroup tail finishes. Rounds past the end
// would each still cost two launches.
int round_batch(int done, int64_t pend, int64_t prev_pend, int prev_done) {
  const Tunables& t = tune();
  int b = round_batch(done);
  if (t.tail_items > 0 && pend > t.tail_items && prev_pend > pend && done > prev_done) {
    const double q = std::pow((double)pend / (double)prev_pend, 1.0 / (double)(done - prev_done));
    if (q > 0.0 && q < 1.0) {
      const double need = std::log((double)t.tail_items / (double)pend) / std::log(q);
      b = std::max(1, std::min(b, (int)std::ceil(need)));
    }
  }
  return b;
}

// Work forked from the caller's stream onto the context's side stream (default priority) and joined
// back when the scope ends, on every path out of it, errors included: the caller's stream then
// orders everything after the side work (outputs written there are never released early).
struct StreamFork {
  dmsc_context* ctx;
  hipStream_t st;
  bool on;
  StreamFork(dmsc_context* c, hipStream_t s, bool enable) : ctx(c), st(s), on(enable) {
    if (!on) return;
    if (!ctx->side) DMSC_CHECK(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    if (!ctx->ev_fork) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    if (!ctx->ev_join) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    if (!ctx->ev_mark) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_mark, hipEventDisableTiming));
    if (!ctx->ev_side) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_side, hipEventDisableTiming));
    DMSC_CHECK(hipEventRecord(ctx->ev_fork, st));
    DMSC_CHECK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
  }
  hipStream_t stream() const { return on ? ctx->side : st; }
  // the side stream also waits for everything enqueued on `st` so far
  void follow() {
    if (!on) return;
    DMSC_CHECK(hipEventRecord(ctx->ev_fork, st));
    DMSC_CHECK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
  }
  // mark the side stream's current point; wait_mark() makes `st` wait for it
  void mark() {
    if (on) DMSC_CHECK(hipEventRecord(ctx->ev_mark, ctx->side));
  }
  void wait_mark() {
    if (on) DMSC_CHECK(hipStreamWaitEvent(st, ctx->ev_mark, 0));
  }
  // a second mark (per-side work run on the side stream): mark_side() records the side stream's
  // current point, wait_side() makes `st` wait for it
  void mark_side() {
    if (on) DMSC_CHECK(hipEventRecord(ctx->ev_side, ctx->side));
  }
  void wait_side() {
    if (on) DMSC_CHECK(hipStreamWaitEvent(st, ctx->ev_side, 0));
  }
  // A third stream (default priority) for work that waits for `st` at the point late_mark() was
  // called and is queued later, so that it runs beside whatever `st` runs by then.
  bool late_on = false;
  void late_mark() {
    if (!on) return;
    if (!ctx->ev_late) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_late, hipEventDisableTiming));
    DMSC_CHECK(hipEventRecord(ctx->ev_late, st));
  }
  hipStream_t late_stream() {
    if (!on) return st;
    if (!late_on) {
      if (!ctx->late) DMSC_CHECK(hipStreamCreateWithFlags(&ctx->late, hipStreamNonBlocking));
      if (!ctx->ev_late_join) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_late_join, hipEventDisableTiming));
      DMSC_CHECK(hipStreamWaitEvent(ctx->late, ctx->ev_late, 0));
      late_on = true;
    }
    return ctx->late;
  }
  void join() {
    if (!on) return;
    on = false;
    DMSC_CHECK(hipEventRecord(ctx->ev_join, ctx->side));
    DMSC_CHECK(hipStreamWaitEvent(st, ctx->ev_join, 0));
    if (late_on) {
      late_on = false;
      DMSC_CHECK(hipEventRecord(ctx->ev_late_join, ctx->late));
      DMSC_CHECK(hipStreamWaitEvent(st, ctx->ev_late_join, 0));
    }
  }
  ~StreamFork() {
    if (!on) return;
    if (hipEventRecord(ctx->ev_join, ctx->side) != hipSuccess || hipStreamWaitEvent(st, ctx->ev_join, 0) != hipSuccess)
      (void)hipStreamSynchronize(ctx->side);
    if (late_on && (hipEventRecord(ctx->ev_late_join, ctx->late) != hipSuccess ||
                    hipStreamWaitEvent(st, ctx->ev_late_join, 0) != hipSuccess))
      (void)hipStreamSynchronize(ctx->late);
  }
};


// One field of a call, with the caller's index (the allocator's `field` argument).
struct FieldIn {
  const float* f;
  int H, W;
  float thr;
  int index;
};

void validate_field(const FieldIn& in) {
  if (!in.f) throw Fail{DMSC_ERR_INVALID_ARGUMENT, "Input tensor must be defined"};
  if (in.H <= 0 || in.W <= 0)
    throw Fail{DMSC_ERR_INVALID_ARGUMENT,
               "Spatial dimensions must be non-empty, got " + std::to_string(in.H) + "x" + std::to_string(in.W)};
  if ((int64_t)in.H + 1 > (int64_t)INT32_MAX / 4 / ((int64_t)in.W + 1))
    throw Fail{DMSC_ERR_TOO_LARGE,
               "Input is too large for 32-bit cell IDs: " + std::to_string(in.H) + "x" + std::to_string(in.W)};
}

void ensure_pinned(dmsc_context* c, size_t bytes) {
  if (bytes <= c->pinned_cap) return;
  if (c->pinned) DMSC_CHECK(hipHostFree(c->pinned));
  c->pinned = nullptr;
  c->pinned_cap = 0;
  size_t cap = std::max<size_t>(bytes + bytes / 2, 4096);
  DMSC_CHECK(hipHostMalloc((void**)&c->pinned, cap));
  c->pinned_cap = cap;
}

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Output allocation for a group of fields. A one-field group gets the field's own arrays in their
// final shapes; a larger group gets one buffer per slot (allocator called with field = -1) holding
// the fields' arrays back to back in field order, and each field's output points into it.
struct OutAlloc {
  dmsc_alloc_fn alloc;
  void* user;
  const FieldIn* fin;
  dmsc_outputs* outs;
  int B;
  // numel(f) and shape of slot k for field f
  void* slot(int k, int ndim, const std::function<std::pair<int64_t, int64_t>(int)>& shape, int dtype) {
    if (B == 1) {
      auto sh = shape(0);
      return alloc_out(alloc, user, fin[0].index, &outs[fin[0].index], k, ndim, sh.first, sh.second, dtype);
    }
    std::vector<int64_t> numel(B);
    int64_t total = 0;
    for (int f = 0; f < B; ++f) {
      auto sh = shape(f);
      numel[f] = ndim == 2 ? sh.first * sh.second : sh.first;
      dmsc_outputs* o = &outs[fin[f].index];
      o->ndim[k] = ndim;
      o->dims[k][0] = sh.first;
      o->dims[k][1] = ndim == 2 ? sh.second : 0;
      o->dtype[k] = dtype;
      total += numel[f];
    }
    int64_t dims[2] = {total, 0};
    char* p = (char*)alloc(user, -1, k, 1, dims, dtype);
    if (!p && total > 0) throw Fail{DMSC_ERR_ALLOC, "output allocator returned NULL for slot " + std::to_string(k)};
    int64_t off = 0;
    for (int f = 0; f < B; ++f) {
      outs[fin[f].index].data[k] = numel[f] > 0 ? p + 4 * off : nullptr;
      off += numel[f];
    }
    return p;
  }
};

// All fields of a group go through the same launches: per-cell arrays are the concatenation of the
// fields' arrays, critical-cell lists are concatenated in field order (ids stay field-local), and
// the persistence and geometry rounds run for every field at once (a saddle only ever meets
// extrema of its own field, so fields never interact).
void run_group(dmsc_context* ctx, const FieldIn* fin, int B, uint32_t flags, hipStream_t st, dmsc_alloc_fn alloc,
               void* user, dmsc_outputs* outs) {
  const bool dual = flags & DMSC_DUAL, grad = flags & DMSC_RETURN_GRADIENT;
  bool tmax = flags & DMSC_TRACE_MAX_ARCS, tmin = flags & DMSC_TRACE_MIN_ARCS;
  bool gmax = flags & DMSC_TRACE_MAX_GROUPS, gmin = flags & DMSC_TRACE_MIN_GROUPS;
  if (dual) {  // reference dmsc_gpu.cpp:71-74: internal "max" = faces, "min" = vertices
    std::swap(tmax, tmin);
    std::swap(gmax, gmin);
  }
  Arena& ar = ctx->ar;

  // ---- field table, bases and tiles
  std::vector<FieldDesc> fd(B);
  std::vector<Dims> dims(B);
  std::vector<int64_t> slot_b(B + 1), pix_b(B + 1), word_b(B + 1), tile_b(B + 1);
  std::vector<float> thr(B);
  int64_t S = 0, N = 0, PW = 0, T = 0;
  for (int f = 0; f < B; ++f) {
    const Dims d = make_dims(fin[f].H, fin[f].W);
    dims[f] = d;
    fd[f] = FieldDesc{fin[f].f, d.H, d.W, d.Nx, d.Ny, d.WPR, fin[f].thr, N, S, PW};
    slot_b[f] = S;
    pix_b[f] = N;
    word_b[f] = PW;
    thr[f] = fin[f].thr;
    tile_b[f] = T;
    T += gradient_tiles(d.H, d.W, ctx->tune.grad_rows);  // gradient tiles, row-major within the field
    S += d.S;
    N += d.N;
    PW += d.PW;
  }
  slot_b[B] = S;
  pix_b[B] = N;
  word_b[B] = PW;
  tile_b[B] = T;
  if (S > (int64_t)INT32_MAX - 1) throw Fail{DMSC_ERR_TOO_LARGE, "batch group exceeds 32-bit cell indices"};

  const size_t o_fd = 0, o_sb = align_up(o_fd + sizeof(FieldDesc) * B, 16),
               o_pb = align_up(o_sb + 8 * (B + 1), 16), o_wb = align_up(o_pb + 8 * (B + 1), 16),
               o_thr = align_up(o_wb + 8 * (B + 1), 16), o_tl = align_up(o_thr + 4 * B, 16),
               meta_bytes = align_up(o_tl + 8 * (B + 1), 16);
  ensure_pinned(ctx, std::max<size_t>(meta_bytes, 64 * (size_t)(B + 1) + 1024));
  {
    char* h = (char*)ctx->pinned;
    std::memcpy(h + o_fd, fd.data(), sizeof(FieldDesc) * B);
    std::memcpy(h + o_sb, slot_b.data(), 8 * (B + 1));
    std::memcpy(h + o_pb, pix_b.data(), 8 * (B + 1));
    std::memcpy(h + o_wb, word_b.data(), 8 * (B + 1));
    std::memcpy(h + o_thr, thr.data(), 4 * B);
    std::memcpy(h + o_tl, tile_b.data(), 8 * (B + 1));
  }
  char* meta = ar.get<char>(A_META, meta_bytes);
  DMSC_CHECK(hipMemcpyAsync(meta, ctx->pinned, meta_bytes, hipMemcpyHostToDevice, st));
  Batch bt{(const FieldDesc*)(meta + o_fd), (const int64_t*)(meta + o_sb), (const int64_t*)(meta + o_pb),
           (const int64_t*)(meta + o_wb), B};
  const float* d_thr = (const float*)(meta + o_thr);
  const int64_t* d_tile_b = (const int64_t*)(meta + o_tl);

  for (int f = 0; f < B; ++f) {
    dmsc_outputs* o = &outs[fin[f].index];
    std::memset(o, 0, sizeof(*o));
    o->shape[0] = fin[f].H;
    o->shape[1] = fin[f].W;
    o->ndim[0] = 1;
    o->dims[0][0] = 2;
  }
  OutAlloc oa{alloc, user, fin, outs, B};

  // outputs whose sizes are known up front
  int32_t* o_grad = (int32_t*)oa.slot(10, 1, [&](int f) { return std::make_pair(grad ? 4 * dims[f].S : 0, (int64_t)0); },
                                      DMSC_DTYPE_INT32);
  int32_t* o_flab = (int32_t*)oa.slot(
      slot_labels(true, dual), 2,
      [&](int f) { return gmax ? std::make_pair((int64_t)dims[f].Ny, (int64_t)dims[f].Nx) : std::make_pair(0L, 0L); },
      DMSC_DTYPE_INT32);
  int32_t* o_vlab = (int32_t*)oa.slot(
      slot_labels(false, dual), 2,
      [&](int f) { return gmin ? std::make_pair((int64_t)dims[f].H, (int64_t)dims[f].W) : std::make_pair(0L, 0L); },
      DMSC_DTYPE_INT32);

  // ---- gradient, critical bit planes, successor forests
  // successor arrays only for the traced sides (the raw-arc walks are their only readers)
  int32_t* fsucc = tmax ? ar.get<int32_t>(A_FSUCC, S) : nullptr;
  int32_t* vsucc = tmin ? ar.get<int32_t>(A_VSUCC, N) : nullptr;
  uint32_t* planes = ar.get<uint32_t>(A_PLANES, 4 * PW);
  uint64_t* pjf = ar.get<uint64_t>(A_PJF, S);
  uint64_t* pjv = ar.get<uint64_t>(A_PJV, N);
  launch_gradient(dual, ctx->tune.grad_rows, bt, d_tile_b, T, grad ? o_grad : nullptr, fsucc, vsucc, planes, PW, pjf,
                  pjv, st);

  int32_t* c32 = ar.get<int32_t>(A_C32, 64);
  int64_t* c64 = ar.get<int64_t>(A_C64, 64);
  int32_t* d_err = c32 + 40;
  fill_i32(d_err, 0, 1, st);
  int32_t* pre_min = ar.get<int32_t>(A_PREMIN, PW);
  int32_t* pre_sad = ar.get<int32_t>(A_PRESAD, PW);
  int32_t* pre_max = ar.get<int32_t>(A_PREMAX, PW);
  scan_plane_popcounts(planes, nullptr, pre_min, PW, c32 + 0, ar, A_SCAN, st);
  scan_plane_popcounts(planes + PW, planes + 2 * PW, pre_sad, PW, c32 + 1, ar, A_SCAN, st);
  scan_plane_popcounts(planes + 3 * PW, nullptr, pre_max, PW, c32 + 2, ar, A_SCAN, st);
  // per-field list bases: b32 rows 0 min, 1 saddles, 2 max, 3-4 output extrema (min, max side),
  // 5 output saddles, 6-7 edge rows (min, max side); each row B+1 entries
  int32_t* b32 = ar.get<int32_t>(A_BASES32, 8 * (size_t)(B + 1));
  auto row32 = [&](int r) { return b32 + (size_t)r * (B + 1); };
  bases_at<int32_t, int64_t>(pre_min, bt.word_b, PW, c32 + 0, B, row32(0), st);
  bases_at<int32_t, int64_t>(pre_sad, bt.word_b, PW, c32 + 1, B, row32(1), st);
  bases_at<int32_t, int64_t>(pre_max, bt.word_b, PW, c32 + 2, B, row32(2), st);
  readback(ctx, c32, 3 * sizeof(int32_t), st);
  const int32_t* hc = (const int32_t*)ctx->pinned;
  const int64_t nmin = hc[0], nsad = hc[1], nmax = hc[2];
  const RankView rv{planes, pre_min, pre_sad, pre_max, PW, bt};
  const int32_t* sad_b = row32(1);

  int32_t* min_ids = ar.get<int32_t>(A_MINIDS, nmin);
  uint64_t* min_key = ar.get<uint64_t>(A_MINVAL, nmin);
  int32_t* sad_ids = ar.get<int32_t>(A_SADIDS, nsad);
  float* sad_val = ar.get<float>(A_SADVAL, nsad);
  int32_t* sad_field = B > 1 ? ar.get<int32_t>(A_SADFIELD, nsad) : nullptr;
  int32_t* max_ids = ar.get<int32_t>(A_MAXIDS, nmax);
  uint64_t* max_key = ar.get<uint64_t>(A_MAXVAL, nmax);
  launch_compact_lists(dual, rv, min_ids, min_key, sad_ids, sad_val, sad_field, max_ids, max_key, st);

  // ---- V-path forests: terminal extremum and depth of every face / vertex (from the tile-resolved state)
  pointer_jump(pjf, S, d_err, st);
  pointer_jump(pjv, N, d_err, st);

  // ---- saddle records and sweep orders
  int32_t* cmax = ar.get<int32_t>(A_CMAX, 2 * nsad);
  int32_t* cmin = ar.get<int32_t>(A_CMIN, 2 * nsad);
  int32_t* lmax = ar.get<int32_t>(A_LMAX, 2 * nsad);
  int32_t* lmin = ar.get<int32_t>(A_LMIN, 2 * nsad);
  launch_saddle_records(rv, sad_ids, sad_field, nsad, pjf, pjv, cmax, cmin, lmax, lmin, st);

  uint32_t* keys = ar.get<uint32_t>(A_KEYS, nsad);
  uint32_t* keys2 = ar.get<uint32_t>(A_KEYS2, nsad);
  int32_t* asc = ar.get<int32_t>(A_ASC, nsad);
  int32_t* vals2 = ar.get<int32_t>(A_VALS2, nsad);
  launch_sort_keys(nsad, sad_val, keys, st);
  radix_sort_pairs(keys, asc, keys2, vals2, nsad, ar, A_RS0, A_RSP0, st, true);  // (value key, saddle index)
  if (B > 1) {  // stable second pass on the field: (field, value, id) order
    launch_gather_field(nsad, asc, B, sad_b, keys, st);
    int bits = 0;
    while ((1 << bits) < B) ++bits;
    radix_sort_pairs_bits(keys, asc, keys2, vals2, nsad, bits, ar, A_RS0, A_RSP0, st);
  }
  // primal: min side ascending, max side descending; dual: the reverse (arcs_topology.hxx:178-181).
  // The ascending side's order is `asc` itself.
  // positions in the ascending order only: the descending side mirrors them within the field's range
  int32_t* const ord_min = dual ? ar.get<int32_t>(A_ORDMIN, nsad) : asc;
  int32_t* const ord_max = dual ? asc : ar.get<int32_t>(A_ORDMAX, nsad);
  int32_t* pos_asc = ar.get<int32_t>(A_POSMIN, nsad);
  launch_orders(nsad, asc, B, sad_b, nullptr, dual ? ord_min : ord_max, pos_asc, nullptr, st);

  // ---- raw arc lengths of the traced sides. The raw arcs need no persistence result: they are
  // traced on the side stream while the persistence rounds run on `st`.
  ArcWork aw[2];  // 0 = vertices (min arcs), 1 = faces (max arcs)
  const bool traced[2] = {tmin, tmax};
  for (int s = 0; s < 2; ++s) {
    if (!traced[s]) continue;
    ArcWork& w = aw[s];
    w.on = true;
    const int b = A_ARC0 + kArcSlots * s;
    int32_t* blen = ar.get<int32_t>(b + G_BLEN, 2 * nsad);
    int64_t* boff = ar.get<int64_t>(b + G_BOFF, 2 * nsad);
    int4* tl = ar.get<int4>(b + G_TL, nsad);
    launch_arc_lengths(nsad, s ? cmax : cmin, s ? lmax : lmin, blen, tl, st);
    w.A.faces = s;
    w.A.succ = s ? fsucc : vsucc;
    w.A.fd = bt.fd;
    w.A.sfield = sad_field;
    w.A.ns = nsad;
    w.A.next = s ? nmax : nmin;
    w.A.sad_ids = sad_ids;
    w.A.tinit = s ? cmax : cmin;
    w.A.blen = blen;
    w.A.tl = tl;
    w.A.boff = boff;
    w.A.head = ar.get<int4>(b + G_HEAD, 2 * nsad);
  }
  StreamFork fork(ctx, st, tmin || tmax);

  // ---- persistence rounds (both sides together)
  uint8_t* removed = ar.get<uint8_t>(A_REMOVED, nsad);
  SideBufs sb[2] = {side_bufs(ar, 0, nmin, nsad), side_bufs(ar, 1, nmax, nsad)};
  const bool record[2] = {tmin, tmax};
  PersistSide ps[2];
  for (int s = 0; s < 2; ++s) {
    PersistSide& p = ps[s];
    p.pos = pos_asc;
    p.sad_b = sad_b;
    p.pos_rev = (s == 1) != dual;  // primal: the max side descends; dual: the min side
    p.c = s ? cmax : cmin;
    p.sval = sad_val;
    p.ekey = s ? max_key : min_key;
    p.par_g = sb[s].par_g;
    p.par_t = sb[s].par_t;
    p.alive = sb[s].alive;
    p.minkey = sb[s].minkey;
    p.pair = sb[s].pair;
    p.pair_other = sb[1 - s].pair;
    p.pers = sb[s].pers;
    p.cancel_d = record[s] ? sb[s].cancel : nullptr;
    p.rr = sb[s].rr;
    p.croot = sb[s].croot;
    p.wkey = sb[s].wkey;
    p.rkey = sb[s].rkey;
    p.wl = sb[s].wl;
    p.wls = nsad;
    p.cnt = sb[s].cnt;
    p.err = d_err;
    p.thr = d_thr;
    p.sfield = sad_field;
    p.maxside = s;
  }
  launch_persist_init(ps[0], ps[1], nmin, nmax, nsad, sb[0].cancel, sb[1].cancel, removed, st);
  int round = 0;
  ctx->pending_log.clear();
  std::fill(std::begin(ctx->arc_stats), std::end(ctx->arc_stats), int64_t(0));
  int32_t* d_round = c32 + 50;  // round number for graph replays
  CapturedRounds pgraph, ggraph;
  if (nsad > 0) {
    int64_t pend = 2 * nsad;  // pending items of both sides at the last readback (launch sizes)
    int64_t prev_pend = -1;   // ... and at the readback before, prev_round rounds in
    int prev_round = 0;
    while (true) {
      const int round_at_launch = round;  // pend was read back after this many rounds
      bool tail = false;
      {
        KScope sweep_("persist_sweep", st, true);  // the rounds of this batch, back to back on `st`
        // few pending saddles: one workgroup runs every remaining round (at most one round per item)
        tail = ctx->tune.tail_items > 0 && pend <= ctx->tune.tail_items;
        if (tail) {
          launch_persist_tail(dual, ps[0], ps[1], round, (int)pend + 2, removed, c32 + 48, st);
        } else {
          const int batch = round_batch(round, pend, prev_pend, prev_round);
          if (ctx->tune.graph_rounds) {
            if (!pgraph.fits(batch, pend))
              pgraph.capture(ctx, batch, pend, [&](hipStream_t cs) {
                for (int b = 0; b < batch; ++b) {
                  launch_persist_round(dual, ps[0], ps[1], 0, removed, pend, cs, d_round);
                  launch_round_inc(d_round, cs);
                }
              });
            fill_i32(d_round, round, 1, st);
            KScope ks_("persist_round", st);
            DMSC_CHECK(hipGraphLaunch(pgraph.exec, st));
            round += batch;
          } else {
            for (int b = 0; b < batch; ++b, ++round) launch_persist_round(dual, ps[0], ps[1], round, removed, pend, st);
          }
        }
      }
      if (tail) {
        DMSC_CHECK(hipMemcpyAsync(ctx->pinned, c32 + 48, 4, hipMemcpyDeviceToHost, st));
        DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 8, sb[0].cnt, 8, hipMemcpyDeviceToHost, st));
        DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 16, sb[1].cnt, 8, hipMemcpyDeviceToHost, st));
        DMSC_CHECK(hipStreamSynchronize(st));
        const int32_t* h = (const int32_t*)ctx->pinned;
        round = h[0];
        const int32_t left[2] = {h[2 + (round & 1)], h[4 + (round & 1)]};
        std::memcpy(ctx->pinned, left, 8);
      } else {
        DMSC_CHECK(hipMemcpyAsync(ctx->pinned, sb[0].cnt + (round & 1), 4, hipMemcpyDeviceToHost, st));
        DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 4, sb[1].cnt + (round & 1), 4, hipMemcpyDeviceToHost, st));
        DMSC_CHECK(hipStreamSynchronize(st));
      }
      const int32_t* pc = (const int32_t*)ctx->pinned;
      ctx->pending_log.insert(ctx->pending_log.end(), {round, pc[0], pc[1]});
      if (pc[0] == 0 && pc[1] == 0) break;
      prev_pend = round > 0 ? pend : -1;
      prev_round = round_at_launch;
      pend = (int64_t)pc[0] + pc[1];
      if (round > 4 * nsad + 64) throw Fail{DMSC_ERR_INTERNAL, "persistence rounds did not converge"};
    }
  }

  // ---- cancellation lists of the traced sides, ahead of the assembly: their counts come back first,
  // so the cancellation order (a sort, the merge keys, the first geometry worklist) is queued on the
  // side stream while the assembly below runs on `st`
  for (int s = 0; s < 2; ++s) {
    if (!aw[s].on) continue;
    const int b = A_ARC0 + kArcSlots * s;
    const int64_t next = s ? nmax : nmin;
    int32_t* flag = ar.get<int32_t>(b + G_FLAG, nsad);
    int32_t* pre = ar.get<int32_t>(b + G_PRE, nsad);
    launch_cancel_flags(nsad, sb[s].cancel, flag, st);
    scan_exclusive<int32_t>(flag, pre, nsad, c32 + 8 + s, ar, A_SCAN, st);
    GeoState& g = aw[s].G;
    int32_t* cs0 = ar.get<int32_t>(b + G_CS0, nsad);
    int32_t* cd0 = ar.get<int32_t>(b + G_CD0, nsad);
    uint32_t* ck = ar.get<uint32_t>(b + G_KEYS, nsad);
    launch_cancel_compact(nsad, sb[s].cancel, sb[s].pers, pre, cs0, cd0, ck, st);
    g.cs = cs0;
    g.cd = cd0;
    // arena offsets of the arcs an output can use (the surviving saddles' and this side's
    // cancellations') that are longer than kHeadCells (the shorter ones live in their head records)
    aw[s].A.removed = removed;
    aw[s].A.cancel = sb[s].cancel;
    scan_arc_offsets(aw[s].A.blen, removed, sb[s].cancel, kHeadCells + 1, ar.get<int64_t>(b + G_BOFF, 2 * nsad),
                     2 * nsad, c64 + 2 * s, ar, A_SCAN, st);

    g.link = ar.get<int4>(b + G_LINK, next);
    g.watch = ar.get<uint64_t>(b + G_WATCH, next);
    g.gk = ar.get<uint64_t>(b + G_JUMP, 2 * next);  // (forest entry, kill word) per extremum
    launch_geo_init(g, next, st);  // every extremum its own root, no killer, no watch
  }
  constexpr size_t kCancelPin = 64;  // pinned bytes 64..71: the two cancellation counts
  constexpr size_t kArenaPin = 72;   // pinned bytes 72..103: c64[0..3], the arena sizes at c64[0] and c64[2]
  if (fork.on) {
    DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + kCancelPin, c32 + 8, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + kArenaPin, c64, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    fork.follow();  // the side stream waits for the compacted lists (and, with them, the counts)
  }

  // ---- assembly counts (cpu/workspace.hxx:107-168)
  int32_t* surv = ar.get<int32_t>(A_ASM0 + 1, nsad);
  int32_t* surv_pre = ar.get<int32_t>(A_ASM0 + 2, nsad);
  int32_t* rows[2] = {ar.get<int32_t>(A_ASM0 + 3, nsad), ar.get<int32_t>(A_ASM0 + 4, nsad)};
  int32_t* rows_pre[2] = {ar.get<int32_t>(A_ASM0 + 5, nsad), ar.get<int32_t>(A_ASM0 + 6, nsad)};
  int32_t* out_list_mx = ar.get<int32_t>(A_ASM0 + 7, nsad);  // per max-order survivor: its output index
  for (int s = 0; s < 2; ++s) {
    const int64_t next = s ? nmax : nmin;
    scan_u8_flags(sb[s].alive, sb[s].to_out, next, c32 + 3 + s, ar, A_SCAN, st);
    // first output extremum of each field (sampled before dead extrema are masked)
    bases_at<int32_t, int32_t>(sb[s].to_out, row32(s ? 2 : 0), next, c32 + 3 + s, B, row32(3 + s), st);
    launch_root_out(sb[s].par_t, sb[s].alive, sb[s].to_out, sb[s].root, next, st);  // masks to_out too
  }
  int32_t* surv_mx = ar.get<int32_t>(A_ASM0 + 8, nsad);
  int32_t* surv_mx_pre = ar.get<int32_t>(A_ASM0 + 9, nsad);
  int2* edges[2] = {ar.get<int2>(A_ASM0 + 12, nsad), ar.get<int2>(A_ASM0 + 13, nsad)};  // by order position
  int2* edge_list[2] = {ar.get<int2>(A_ASM0 + 14, nsad), ar.get<int2>(A_ASM0 + 15, nsad)};  // by survivor rank
  launch_side_counts(nsad, ord_min, removed, cmin, sb[0].root, surv, rows[0], edges[0], st);
  launch_side_counts(nsad, ord_max, removed, cmax, sb[1].root, surv_mx, rows[1], edges[1], st);
  scan_exclusive<int32_t>(surv, surv_pre, nsad, c32 + 5, ar, A_SCAN, st);
  scan_exclusive<int32_t>(surv_mx, surv_mx_pre, nsad, c32 + 10, ar, A_SCAN, st);
  scan_exclusive<int32_t>(rows[0], rows_pre[0], nsad, c32 + 6, ar, A_SCAN, st);
  scan_exclusive<int32_t>(rows[1], rows_pre[1], nsad, c32 + 7, ar, A_SCAN, st);
  // orders are grouped by field, so field f owns positions [sad_b[f], sad_b[f+1]) of each order
  bases_at<int32_t, int32_t>(surv_pre, sad_b, nsad, c32 + 5, B, row32(5), st);
  bases_at<int32_t, int32_t>(rows_pre[0], sad_b, nsad, c32 + 6, B, row32(6), st);
  bases_at<int32_t, int32_t>(rows_pre[1], sad_b, nsad, c32 + 7, B, row32(7), st);
  // survivor lists: min order (position = output saddle index) and max order, with their edge rows
  int32_t* surv_list[2] = {ar.get<int32_t>(A_ASM0 + 10, nsad), ar.get<int32_t>(A_SURV1, nsad)};
  int32_t* surv_rows[2] = {ar.get<int32_t>(A_ASM0 + 11, nsad), ar.get<int32_t>(A_SURVR1, nsad)};
  launch_survivors(nsad, ord_min, surv, surv_pre, rows_pre[0], edges[0], surv_list[0], surv_rows[0], edge_list[0],
                   nullptr, sad_b, B, nullptr, st);
  launch_survivors(nsad, ord_max, surv_mx, surv_mx_pre, rows_pre[1], edges[1], surv_list[1], surv_rows[1], edge_list[1],
                   surv_pre, sad_b, B, out_list_mx, st);

  DMSC_CHECK(hipMemcpyAsync(ctx->pinned, c32, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 128, b32, 8 * sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost, st));
  // Per traced side, on the side stream while the assembly runs: cancellations ordered by (persistence,
  // saddle id) (persistence_struct.hxx:15-19), as the merge key of each (position in that order; the
  // lists stay in saddle-id order, which keeps them spatially coherent), and the first geometry
  // worklist. Fields never meet, so one batch-wide stable sort orders every field's cancellations.
  if (fork.on) {
    DMSC_CHECK(hipEventSynchronize(ctx->ev_fork));  // the counts copied before the assembly
    int32_t hc[2];
    std::memcpy(hc, (const char*)ctx->pinned + kCancelPin, sizeof(hc));
    const hipStream_t cst = fork.stream();
    for (int s = 0; s < 2; ++s) {
      ArcWork& w = aw[s];
      if (!w.on) continue;
      const int b = A_ARC0 + kArcSlots * s;
      const int64_t nc = hc[s];
      w.ncancel = nc;
      GeoState& g = w.G;
      g.ncancel = nc;
      int32_t* ckey = ar.get<int32_t>(b + G_CS, nc);
      g.key = ckey;
      g.rr = ar.get<int32_t>(b + G_RR, 4 * nc);
      g.wl = ar.get<int4>(b + G_WL, 2 * nc);
      g.wls = nc;
      g.glen = ar.get<int2>(b + G_GLEN, 2 * nc);
      g.cnt = ar.get<int32_t>(b + G_CNT, 2);
      g.err = d_err;
      radix_rank(ar.get<uint32_t>(b + G_KEYS, nsad), ckey, nc, ar, A_RS2, A_RSP2, cst);
      launch_geo_items(nc, g.cs, g.cd, ckey, w.A.tinit, g.list(0), g.gk, g.cnt, cst);  // list 0, its count, killers
    }
    fork.mark_side();  // the geometry rounds wait for this point, not for the emission queued below
    // The raw arcs go on the same stream after that point: they run beside the assembly and the
    // geometry rounds, and only the expansion waits for them.
    int64_t tot[4];
    std::memcpy(tot, (const char*)ctx->pinned + kArenaPin, sizeof(tot));
    for (int s = 0; s < 2; ++s) {
      if (!aw[s].on) continue;
      ArcWork& w = aw[s];
      w.arena_total = tot[2 * s];
      w.A.arena = ar.get<int32_t>(A_ARC0 + kArcSlots * s + G_ARENA, w.arena_total);
      launch_raw_arcs(w.A, cst);
    }
    fork.mark();  // the expansion waits for the raw arcs only
  }
  DMSC_CHECK(hipStreamSynchronize(st));
  int32_t h32[16];
  std::memcpy(h32, ctx->pinned, sizeof(h32));
  std::vector<int32_t> hb(8 * (size_t)(B + 1));
  std::memcpy(hb.data(), (char*)ctx->pinned + 128, sizeof(int32_t) * hb.size());
  auto hrow = [&](int r, int f) -> int64_t { return hb[(size_t)r * (B + 1) + f]; };
  const int64_t n_sad_total = h32[5];  // surviving saddles of all fields

  // ---- outputs that do not depend on the arc geometry (cpu/workspace.hxx:291-362 layout, dual swap by
  // slot). Their kernels only read persistence and assembly results, which nothing below writes, so
  // they run on the side stream while the arc geometry (latency-bound rounds) runs on `st`.
  auto n_out_ext = [&](int s, int f) { return hrow(3 + s, f + 1) - hrow(3 + s, f); };
  auto n_sad_out = [&](int f) { return hrow(5, f + 1) - hrow(5, f); };
  auto n_rows = [&](int s, int f) { return hrow(6 + s, f + 1) - hrow(6 + s, f); };
  auto one = [](int64_t n) { return std::make_pair(n, (int64_t)0); };
  int32_t* o_pts[2];
  for (int s = 0; s < 2; ++s)
    o_pts[s] = (int32_t*)oa.slot(slot_pts(s == 1, dual), 1, [&](int f) { return one(n_out_ext(s, f)); },
                                 DMSC_DTYPE_INT32);
  int32_t* o_sad = (int32_t*)oa.slot(3, 1, [&](int f) { return one(n_sad_out(f)); }, DMSC_DTYPE_INT32);
  int32_t *o_e[2], *o_pp[2];
  float* o_p[2];
  for (int s = 0; s < 2; ++s) {
    o_e[s] = (int32_t*)oa.slot(slot_e(s == 1, dual), 2, [&](int f) { return std::make_pair(n_rows(s, f), (int64_t)2); },
                               DMSC_DTYPE_INT32);
    o_p[s] = (float*)oa.slot(slot_p(s == 1, dual), 1, [&](int f) { return one(n_sad_out(f)); }, DMSC_DTYPE_FLOAT32);
    o_pp[s] = (int32_t*)oa.slot(slot_pp(s == 1, dual), 1, [&](int f) { return one(n_sad_out(f)); }, DMSC_DTYPE_INT32);
  }
  // Per-side arc work between the assembly and the expansion runs side 0 on `st` and, when both sides are
  // traced, side 1 on the side stream (with its own scan and sort scratch): the cancellation order here,
  // the path lengths after the geometry rounds.
  const bool split = fork.on && aw[0].on && aw[1].on;
  auto side_stream = [&](int s) { return split && s == 1 ? fork.stream() : st; };
  fork.follow();  // the side stream also waits for the persistence and assembly results
  {
    const hipStream_t es = fork.stream();
    launch_emit_points(nmin, min_ids, sb[0].to_out, o_pts[0], es);
    launch_emit_points(nmax, max_ids, sb[1].to_out, o_pts[1], es);
    const EmitBases eb[2] = {{sad_field, row32(5), row32(3)}, {sad_field, row32(5), row32(4)}};
    launch_emit_side(n_sad_total, surv_list[0], nullptr, nullptr, edge_list[0], sb[0].root, sb[0].pers,
                     sb[0].pair, surv_rows[0], sad_ids, eb[0], o_p[0], o_pp[0], o_e[0], o_sad, es);
    launch_emit_side(n_sad_total, surv_list[1], nullptr, out_list_mx, edge_list[1], sb[1].root, sb[1].pers,
                     sb[1].pair, surv_rows[1], sad_ids, eb[1], o_p[1], o_pp[1], o_e[1], nullptr, es);
  }
  // The labels read only persistence and assembly results. They are bandwidth-bound, so they wait
  // for the geometry rounds to fall to a sixteenth of their cancellations (the latency-bound rounds)
  // and then run on a stream of their own beside them.
  fork.late_mark();
  bool labels_queued = false;
  auto queue_labels = [&](hipStream_t ls) {
    if (labels_queued) return;
    labels_queued = true;
    if (gmax) launch_labels_faces(rv, S, pjf, sb[1].root, row32(4), o_flab, ls);
    if (gmin) launch_labels_verts(rv, N, pjv, sb[0].root, row32(3), o_vlab, ls);
  };
  fork.wait_side();
  // geometry rounds of both sides, interleaved, one readback per batch of rounds
  {
    bool live[2] = {aw[0].on && aw[0].ncancel > 0, aw[1].on && aw[1].ncancel > 0};
    int64_t pend[2] = {aw[0].ncancel, aw[1].ncancel};  // pending at the last readback (launch sizes)
    int r = 0;
    ctx->rounds_geo[0] = ctx->rounds_geo[1] = 0;
    ctx->geo_log.assign({0, (int32_t)aw[0].ncancel, (int32_t)aw[1].ncancel});
    int ggraph_mask = -1;
    int64_t gprev = -1;  // pending of the live sides at the readback before the last, gprev_r rounds in
    int gprev_r = 0;
    while (live[0] || live[1]) {
      const ArcSide as[2] = {aw[0].A, aw[1].A};
      const GeoState gs[2] = {aw[0].G, aw[1].G};
      const int mask = (live[0] ? 1 : 0) | (live[1] ? 2 : 0);
      const int64_t gpend = (live[0] ? pend[0] : 0) + (live[1] ? pend[1] : 0);
      const bool gtail = ctx->tune.tail_items > 0 && gpend <= ctx->tune.tail_items;
      if (gtail) {
        launch_geo_tail(as, gs, mask, r, (int)gpend + 64, c32 + 49, st);
        DMSC_CHECK(hipMemcpyAsync(ctx->pinned, c32 + 49, 4, hipMemcpyDeviceToHost, st));
        for (int s = 0; s < 2; ++s)
          if (live[s])
            DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 8 + 8 * s, aw[s].G.cnt, 8, hipMemcpyDeviceToHost, st));
        DMSC_CHECK(hipStreamSynchronize(st));
        const int32_t* h = (const int32_t*)ctx->pinned;
        r = h[0];
        const int32_t left[2] = {live[0] ? h[2 + (r & 1)] : 0, live[1] ? h[4 + (r & 1)] : 0};
        std::memcpy(ctx->pinned, left, 8);
      } else {
        const int batch = round_batch(r, gpend, gprev, gprev_r);
        gprev = gpend;
        gprev_r = r;
        const int64_t items = pend[0] + pend[1];
        if (ctx->tune.graph_rounds) {
          if (!ggraph.fits(batch, items) || ggraph_mask != mask)
            ggraph.capture(ctx, batch, items, [&](hipStream_t cs) {
              for (int k = 0; k < batch; ++k) {
                launch_geo_round2(as, gs, mask, 0, items, cs, d_round);
                launch_round_inc(d_round, cs);
              }
            });
          ggraph_mask = mask;
          fill_i32(d_round, r, 1, st);
          KScope ks_("geo_round", st);
          DMSC_CHECK(hipGraphLaunch(ggraph.exec, st));
          r += batch;
        } else {
          for (int k = 0; k < batch; ++k, ++r) launch_geo_round2(as, gs, mask, r, items, st);
        }
        for (int s = 0; s < 2; ++s)
          if (live[s]) DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 4 * s, aw[s].G.cnt + (r & 1), 4,
                                                 hipMemcpyDeviceToHost, st));
        DMSC_CHECK(hipStreamSynchronize(st));
      }
      const int32_t* hp = (const int32_t*)ctx->pinned;  // pending per side
      ctx->geo_log.insert(ctx->geo_log.end(), {r, live[0] ? hp[0] : 0, live[1] ? hp[1] : 0});
      if (16 * ((live[0] ? (int64_t)hp[0] : 0) + (live[1] ? (int64_t)hp[1] : 0)) <= aw[0].ncancel + aw[1].ncancel)
        queue_labels(fork.late_stream());
      for (int s = 0; s < 2; ++s)
        if (live[s]) {
          ctx->rounds_geo[s] = r;
          pend[s] = hp[s];
          if (pend[s] == 0) live[s] = false;
          else if (r > 4 * aw[s].ncancel + 64) throw Fail{DMSC_ERR_INTERNAL, "geometry rounds did not converge"};
        }
    }
  }
  queue_labels(fork.late_stream());  // when no geometry round ran
  if (split) fork.follow();  // the side stream waits for the geometry rounds
  for (int s = 0; s < 2; ++s) {
    if (!aw[s].on) continue;
    ArcWork& w = aw[s];
    GeoState& g = w.G;
    const int64_t next = s ? nmax : nmin;
    // path lengths to the final roots and per-output-arc lengths (min order)
    const hipStream_t cst = side_stream(s);
    Arena& a2 = ctx->ar2;
    const int hb2 = 16 * s;
    w.pjq = a2.get<uint64_t>(hb2 + H_PJQ, next + 1);
    launch_geo_q(g, next, w.pjq, cst);
    w.cum_r = a2.get<uint64_t>(hb2 + H_CUMR, next + 1);
    w.cum_c = a2.get<uint64_t>(hb2 + H_CUMC, next + 1);
    w.xl = a2.get<LinkRec>(hb2 + H_XL, next + 1);
    w.ncq = a2.get<int4>(hb2 + H_NCQ, next + 1);
    launch_link_counts(w.A, g, w.pjq, next, w.cum_r, w.cum_c, w.xl, w.ncq, cst);
    w.flen = a2.get<int64_t>(hb2 + H_FLEN, 2 * nsad);
    int64_t* per_pos = a2.get<int64_t>(hb2 + H_PERPOS, nsad);
    w.pos_pre = a2.get<int64_t>(hb2 + H_POSPRE, nsad);
    // over the surviving saddles in output order (position = output saddle index)
    w.rec = a2.get<int4>(hb2 + H_REC, 2 * nsad);  // 32 bytes per surviving saddle
    launch_final_lengths(n_sad_total, surv_list[0], w.A, w.pjq, w.rec, w.flen, per_pos, cst);
    scan_exclusive<int64_t>(per_pos, w.pos_pre, n_sad_total, c64 + 2 * s + 1, ar, cst == st ? A_SCAN : A_SCAN2, cst);
    w.arc_base = a2.get<int64_t>(hb2 + H_ARCB, B + 1);
    bases_at<int64_t, int32_t>(w.pos_pre, row32(5), n_sad_total, c64 + 2 * s + 1, B, w.arc_base, cst);
    DMSC_CHECK(hipMemcpyAsync((char*)ctx->pinned + 8 * (size_t)(B + 1) * s, w.arc_base, 8 * (size_t)(B + 1),
                              hipMemcpyDeviceToHost, cst));
  }
  if (aw[0].on || aw[1].on) {
    DMSC_CHECK(hipStreamSynchronize(st));
    if (split) DMSC_CHECK(hipStreamSynchronize(fork.stream()));
    for (int s = 0; s < 2; ++s) {
      if (!aw[s].on) continue;
      ArcWork& w = aw[s];
      w.h_arc_base.resize(B + 1);
      std::memcpy(w.h_arc_base.data(), (char*)ctx->pinned + 8 * (size_t)(B + 1) * s, 8 * (size_t)(B + 1));
      w.out_total = w.h_arc_base[B];
      for (int f = 0; f < B; ++f)
        if (w.h_arc_base[f + 1] - w.h_arc_base[f] > INT32_MAX)
          throw Fail{DMSC_ERR_TOO_LARGE, "merged arc geometry exceeds 32-bit offsets"};
      if (w.out_total > INT32_MAX - 64)
        throw Fail{DMSC_ERR_TOO_LARGE, "merged arc geometry of the batch exceeds 32-bit segment counts"};
    }
  }

  // ---- arc outputs (cpu/workspace.hxx:291-362 layout, dual swap by slot)
  for (int s = 0; s < 2; ++s) {
    const int sl = slot_arcs(s == 1, dual);
    ArcWork& w = aw[s];
    w.out_cells = (int32_t*)oa.slot(
        sl, 1, [&](int f) { return one(w.on ? w.h_arc_base[f + 1] - w.h_arc_base[f] : 0); }, DMSC_DTYPE_INT32);
    w.out_off = (int32_t*)oa.slot(sl + 1, 1, [&](int f) { return one(w.on ? n_sad_out(f) + 1 : 0); }, DMSC_DTYPE_INT32);
    w.out_arc_off = (int32_t*)oa.slot(sl + 2, 1, [&](int f) { return one(w.on ? n_sad_out(f) : 0); }, DMSC_DTYPE_INT32);
  }

  fork.wait_mark();  // raw arcs traced
  // Expansion of both sides in lockstep: side 0 on `st`, side 1 on the side stream when both are
  // traced (the side stream first waits for the geometry results), one host check per batch of levels.
  const bool two = aw[0].on && aw[1].on;
  if (two) fork.follow();
  struct Expansion {
    hipStream_t es = nullptr;
    Seg* lists[2] = {nullptr, nullptr};
    Seg* raw = nullptr;
    int32_t* nseg = nullptr;
    int64_t cap = 0;
    int level = 0;
    int32_t n_in = 0;
  } ex[2];
  auto pin32 = [&](int s) { return (int32_t*)((char*)ctx->pinned + 64 * s); };
  for (int s = 0; s < 2; ++s) {
    ArcWork& w = aw[s];
    if (!w.on) continue;
    Expansion& e = ex[s];
    e.es = (two && s == 1) ? fork.stream() : st;
    Arena& a2 = ctx->ar2;
    const int hb2 = 16 * s;
    const int64_t next = s ? nmax : nmin;
    e.cap = w.out_total + 16;
    e.lists[0] = a2.get<Seg>(hb2 + H_SEGA, e.cap);
    e.lists[1] = a2.get<Seg>(hb2 + H_SEGB, e.cap);
    e.raw = a2.get<Seg>(hb2 + H_RAW, e.cap);
    // chain lists alternate between two buffers; their counters rotate over three (level L reads
    // nseg[L%3], appends to nseg[(L+1)%3] and clears nseg[(L+2)%3]); nseg[3] counts raw pieces
    e.nseg = a2.get<int32_t>(hb2 + H_NSEG, 4);
    fill_i32(e.nseg, 0, 4, e.es);
    launch_seed_segments(n_sad_total, surv_list[0], w.A, w.rec, next, w.flen, w.pos_pre, e.lists[0], e.nseg, e.raw,
                         e.nseg + 3, e.cap, d_err, w.out_off, w.out_arc_off, w.arc_base, row32(5), B, e.es);
    DMSC_CHECK(hipMemcpyAsync(pin32(s), e.nseg, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, e.es));
  }
  // a side's raw pieces are copied as soon as its own expansion has ended (on its stream), so the
  // copy of the shorter side overlaps the other side's remaining levels
  bool copied[2] = {false, false};
  auto start_copy = [&](int s) {
    ArcWork& w = aw[s];
    Expansion& e = ex[s];
    copied[s] = true;
    const int32_t nraw = pin32(s)[3];
    ctx->arc_stats[2 * s] = nraw;
    ctx->arc_stats[2 * s + 1] = w.out_total;
    if (nraw > e.cap) throw Fail{DMSC_ERR_INTERNAL, "arc segment list overflow"};
    if (nraw > 0)
      launch_copy_raw(w.A, e.raw, e.nseg + 3, std::min<int64_t>(nraw, e.cap), w.out_cells, w.out_total, d_err, e.es);
  };
  for (int s = 0; s < 2; ++s)
    if (aw[s].on) {
      DMSC_CHECK(hipStreamSynchronize(ex[s].es));
      ex[s].n_in = pin32(s)[0];
      if (ex[s].n_in <= 0) start_copy(s);
    }
  while ((aw[0].on && ex[0].n_in > 0) || (aw[1].on && ex[1].n_in > 0)) {
    bool ran[2] = {false, false};
    for (int s = 0; s < 2; ++s) {
      ArcWork& w = aw[s];
      Expansion& e = ex[s];
      if (!w.on || e.n_in <= 0) continue;
      if (e.n_in > e.cap) throw Fail{DMSC_ERR_INTERNAL, "arc segment list overflow"};
      // a batch of levels per host check; a level past the last one finds no chains and only clears
      const int batch = e.level < 4 ? std::min(2, ctx->tune.expand_levels) : ctx->tune.expand_levels;
      for (int k = 0; k < batch; ++k, ++e.level)
        launch_expand_segments(w.A, w.G, w.ncq, w.xl, e.lists[e.level & 1], e.nseg + e.level % 3,
                               e.lists[(e.level + 1) & 1], e.nseg + (e.level + 1) % 3, e.nseg + (e.level + 2) % 3, e.raw,
                               e.nseg + 3, std::max<int64_t>(e.n_in, 2048 * 256), e.cap, e.es);
      DMSC_CHECK(hipMemcpyAsync(pin32(s), e.nseg, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, e.es));
      ran[s] = true;
      if (e.level > 1000000) throw Fail{DMSC_ERR_INTERNAL, "arc expansion did not terminate"};
    }
    for (int s = 0; s < 2; ++s)
      if (ran[s]) {
        DMSC_CHECK(hipStreamSynchronize(ex[s].es));
        ex[s].n_in = pin32(s)[ex[s].level % 3];
        if (ex[s].n_in <= 0) start_copy(s);
      }
  }
  for (int s = 0; s < 2; ++s)
    if (aw[s].on && !copied[s]) start_copy(s);
  fork.join();
  readback(ctx, d_err, sizeof(int32_t), st);
  if (ctx->times.level) ctx->times.collect();
  if (*(const int32_t*)ctx->pinned != 0)
    throw Fail{DMSC_ERR_INTERNAL, "device consistency check failed (flags " +
                                      std::to_string(*(const int32_t*)ctx->pinned) + ")"};

  // record for dmsc_debug_fetch (batch-wide arrays; for one field, that field's)
  ctx->S = S;
  ctx->N = N;
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

// Splits a call's fields into groups whose concatenated cell arrays stay within 32-bit indices and
// a bounded scratch footprint, and runs each group through the batch-wide pipeline.
constexpr int64_t kGroupSlots = int64_t(1) << 28;
constexpr int kGroupFields = 1 << 16;

// slots per group: DMSC_GROUP_SLOTS when set (smaller groups, e.g. to bound scratch memory)
int64_t group_slots() {
  const char* e = std::getenv("DMSC_GROUP_SLOTS");
  if (!e || !*e) return kGroupSlots;
  const long long v = std::atoll(e);
  return v < 1 ? 1 : std::min<int64_t>(v, kGroupSlots);
}

int64_t field_slots(const FieldIn& f) { return ((int64_t)f.H + 1) * ((int64_t)f.W + 1); }

void run_range(dmsc_context* ctx, const FieldIn* fields, size_t n, uint32_t flags, hipStream_t st,
               dmsc_alloc_fn alloc, void* user, dmsc_outputs* outs) {
  const int64_t cap = group_slots();
  size_t i = 0;
  while (i < n) {
    size_t j = i;
    int64_t slots = 0;
    while (j < n && (int)(j - i) < kGroupFields) {
      const int64_t s = field_slots(fields[j]);
      if (j > i && slots + s > cap) break;
      slots += s;
      ++j;
    }
    run_group(ctx, fields + i, (int)(j - i), flags, st, alloc, user, outs);
    i = j;
  }
}

// The output allocator of a pipelined batch: the parts' threads call it one at a time.
struct SerialAlloc {
  dmsc_alloc_fn fn;
  void* user;
  std::mutex mu;
};
void* serial_alloc(void* u, int32_t field, int32_t slot, int32_t ndim, const int64_t* dims, int32_t dtype) {
  auto* a = static_cast<SerialAlloc*>(u);
  std::lock_guard<std::mutex> lock(a->mu);
  return a->fn(a->user, field, slot, ndim, dims, dtype);
}

// Peer context i, its stream and its ordering events (created on first use, kept with ctx).
dmsc_context* peer_of(dmsc_context* ctx, int i) {
  if (!ctx->peer[i]) {
    dmsc_context* p = nullptr;
    if (dmsc_context_create(ctx->device, &p) != DMSC_OK)
      throw Fail{DMSC_ERR_HIP, "could not create the context of a part of the batch"};
    ctx->peer[i] = p;
  }
  if (!ctx->peer_st[i]) DMSC_CHECK(hipStreamCreateWithFlags(&ctx->peer_st[i], hipStreamNonBlocking));
  if (!ctx->ev_pin) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_pin, hipEventDisableTiming));
  if (!ctx->ev_pout[i]) DMSC_CHECK(hipEventCreateWithFlags(&ctx->ev_pout[i], hipEventDisableTiming));
  return ctx->peer[i];
}

// A large batch as parts of about equal slots run side by side: the first on the calling thread and
// stream, each other one on a worker thread with a peer context and stream. Each part is a pipeline
// of latency-bound round loops and bandwidth-bound passes with host checks in between; run together,
// one part's round tails and host waits overlap the others' passes (+6 % on the 512-field batch with
// two parts). Results do not depend on the split: fields never meet. Part g holds the fields
// [bounds[g], bounds[g + 1]).
void run_pipelined(dmsc_context* ctx, const std::vector<FieldIn>& fields, const std::vector<size_t>& bounds,
                   uint32_t flags, hipStream_t st, dmsc_alloc_fn alloc, void* user, dmsc_outputs* outs) {
  const int G = (int)bounds.size() - 1;
  dmsc_context* peers[dmsc_context::kMaxPeers] = {};
  for (int i = 0; i + 1 < G; ++i) peers[i] = peer_of(ctx, i);
  SerialAlloc sa{alloc, user, {}};
  DMSC_CHECK(hipEventRecord(ctx->ev_pin, st));
  for (int i = 0; i + 1 < G; ++i) {
    peers[i]->tune = ctx->tune;  // every part runs with the caller's scheduling parameters
    peers[i]->times.level = ctx->times.level;
    DMSC_CHECK(hipStreamWaitEvent(ctx->peer_st[i], ctx->ev_pin, 0));  // the peer's inputs are ready
  }
  std::exception_ptr merr, perr[dmsc_context::kMaxPeers];
  std::vector<std::thread> workers;
  for (int i = 0; i + 1 < G; ++i)
    workers.emplace_back([&, i] {
      try {
        DMSC_CHECK(hipSetDevice(ctx->device));
        tl_times = &peers[i]->times;
        tl_tune = &peers[i]->tune;
        run_range(peers[i], fields.data() + bounds[i + 1], bounds[i + 2] - bounds[i + 1], flags, ctx->peer_st[i],
                  serial_alloc, &sa, outs);
      } catch (...) {
        perr[i] = std::current_exception();
      }
    });
  try {
    run_range(ctx, fields.data(), bounds[1], flags, st, serial_alloc, &sa, outs);
  } catch (...) {
    merr = std::current_exception();
  }
  for (std::thread& w : workers) w.join();
  // the caller's stream orders everything after the peers' work (their outputs, and their scratch reuse)
  bool joined = true;
  for (int i = 0; i + 1 < G; ++i) {
    const bool ok = hipEventRecord(ctx->ev_pout[i], ctx->peer_st[i]) == hipSuccess &&
                    hipStreamWaitEvent(st, ctx->ev_pout[i], 0) == hipSuccess;
    if (!ok) (void)hipStreamSynchronize(ctx->peer_st[i]);
    joined = joined && ok;
  }
  // per-stage times and the debug counts cover every part (the debug arrays are the first part's)
  for (int i = 0; i + 1 < G; ++i) {
    dmsc_context* peer = peers[i];
    for (const auto& kv : peer->times.acc) {
      bool found = false;
      for (auto& mine : ctx->times.acc)
        if (mine.first == kv.first) {
          mine.second.first += kv.second.first;
          mine.second.second += kv.second.second;
          found = true;
        }
      if (!found) ctx->times.acc.push_back(kv);
    }
    peer->times.acc.clear();
  }
  if (merr) std::rethrow_exception(merr);
  for (int i = 0; i + 1 < G; ++i)
    if (perr[i]) std::rethrow_exception(perr[i]);
  for (int i = 0; i + 1 < G; ++i) {
    const dmsc_context* peer = peers[i];
    ctx->nmin += peer->nmin;
    ctx->nsad += peer->nsad;
    ctx->nmax += peer->nmax;
    ctx->S += peer->S;
    ctx->N += peer->N;
    for (int k = 0; k < 4; ++k) ctx->arc_stats[k] += peer->arc_stats[k];
    ctx->rounds_persist = std::max(ctx->rounds_persist, peer->rounds_persist);
    for (int s = 0; s < 2; ++s) ctx->rounds_geo[s] = std::max(ctx->rounds_geo[s], peer->rounds_geo[s]);
  }
  if (!joined) throw Fail{DMSC_ERR_HIP, "could not order the caller's stream after the other parts of the batch"};
}

void run_fields(dmsc_context* ctx, const std::vector<FieldIn>& fields, uint32_t flags, hipStream_t st,
                dmsc_alloc_fn alloc, void* user, dmsc_outputs* outs) {
  for (const FieldIn& f : fields) validate_field(f);
  int64_t total = 0;
  for (const FieldIn& f : fields) total += field_slots(f);
  // parts: pipe_groups, at most one per field and at least pipe_min_slots slots each
  const size_t n = fields.size();
  const int64_t by_size = ctx->tune.pipe_min_slots > 0 ? total / ctx->tune.pipe_min_slots : INT64_MAX;
  const int G = (int)std::min<int64_t>(std::min<int64_t>(ctx->tune.pipe_groups, (int64_t)n), by_size);
  if (G >= 2) {
    // part g ends at the first field where the prefix slots reach g / G of the total, each part
    // keeping at least one field
    std::vector<size_t> bounds{0};
    size_t j = 0;
    int64_t pre = 0;
    for (int g = 1; g < G; ++g) {
      while (j < n - (size_t)(G - g) && (j == bounds.back() || (int64_t)G * pre < (int64_t)g * total))
        pre += field_slots(fields[j++]);
      bounds.push_back(j);
    }
    bounds.push_back(n);
    run_pipelined(ctx, fields, bounds, flags, st, alloc, user, outs);
    return;
  }
  run_range(ctx, fields.data(), n, flags, st, alloc, user, outs);
}

// Persistence-loss fields in chunks of at most kBwdChunk fields and cap scratch floats, handed to
// fn(const BwdChunk&) one at a time (a field that needs more than cap is an error). fields / H / W /
// gscale / grad may be null (loss only: no field data is read). Builds each chunk on the stack.
template <class Fn>
void for_each_bwd_chunk(int B, const float* const* fields, const int32_t* H, const int32_t* W, const dmsc_pairs* pairs,
                        const float* gscale, float* const* grad, int64_t cap, Fn&& fn) {
  BwdChunk cur{};
  int64_t used = 0, blocks = 0;
  auto flush = [&]() {
    if (cur.n == 0) return;
    bwd_assign_blocks(cur);
    fn(static_cast<const BwdChunk&>(cur));
    const int64_t next_base = cur.loss_base + cur.n;
    cur = BwdChunk{};
    cur.loss_base = next_base;
    used = blocks = 0;
  };
  for (int b = 0; b < B; ++b) {
    const dmsc_pairs& p = pairs[b];
    if (p.n_sad < 0 || p.n_min < 0 || p.n_max < 0 ||
        (p.n_sad > 0 && (!p.sad_pts || !p.p_min || !p.ppairs_min || !p.p_max || !p.ppairs_max)) ||
        (p.n_min > 0 && !p.min_pts) || (p.n_max > 0 && !p.max_pts) || p.n_sad > INT32_MAX || p.n_min > INT32_MAX ||
        p.n_max > INT32_MAX)
      throw Fail{DMSC_ERR_INVALID_ARGUMENT, "persistence pairs of field " + std::to_string(b) + " are incomplete"};
    BwdField F{};
    F.sad_pts = p.sad_pts;
    F.p_min = p.p_min;
    F.pp_min = p.ppairs_min;
    F.min_pts = p.min_pts;
    F.p_max = p.p_max;
    F.pp_max = p.ppairs_max;
    F.max_pts = p.max_pts;
    F.n_sad = (int32_t)p.n_sad;
    F.n_min = (int32_t)p.n_min;
    F.n_max = (int32_t)p.n_max;
    int64_t fc = 0, fb = 0;
    if (fields) {
      validate_field(FieldIn{fields[b], H[b], W[b], 0.f, b});
      if (!grad[b]) throw Fail{DMSC_ERR_INVALID_ARGUMENT, "gradient output must be defined"};
      F.f = fields[b];
      F.grad = grad[b];
      F.gscale = gscale + b;
      F.H = H[b];
      F.W = W[b];
      fc = bwd_scratch_floats(H[b], W[b]);
      fb = bwd_blocks_for(H[b], W[b], p.n_sad);
      if (fc > cap)
        throw Fail{DMSC_ERR_TOO_LARGE, "field " + std::to_string(b) + " (" + std::to_string(H[b]) + "x" +
                                           std::to_string(W[b]) + ") needs " + std::to_string(fc) +
                                           " scratch floats; at most " + std::to_string(cap) + " fit"};
    }
    if (cur.n == kBwdChunk || used + fc > cap || blocks + fb > INT32_MAX) flush();
    cur.fld[cur.n++] = F;
    used += fc;
    blocks += fb;
  }
  flush();
}

// Checks every field and returns the most scratch one chunk needs (nothing is launched).
int64_t bwd_scratch_needed(int B, const float* const* fields, const int32_t* H, const int32_t* W,
                           const dmsc_pairs* pairs, const float* gscale, float* const* grad, int64_t cap) {
  int64_t need = 0;
  for_each_bwd_chunk(B, fields, H, W, pairs, gscale, grad, cap,
                     [&](const BwdChunk& c) { need = std::max(need, c.scratch); });
  return need;
}

// Launches the backward of every chunk on `st` with `scratch` (cap floats): clears the whole
// scratch first when it may hold leftovers, and orders the work after the previous loss / backward
// of this context (possibly on another stream). The fields are checked before anything is queued.
void run_backward(dmsc_context* ctx, int B, const float* const* fields, const int32_t* H, const int32_t* W,
                  const dmsc_pairs* pairs, const float* gscale, float* const* grad, uint32_t flags, float max_pers,
                  float* scratch, int64_t cap, bool& dirty, hipStream_t st) {
  (void)bwd_scratch_needed(B, fields, H, W, pairs, gscale, grad, cap);
  if (ctx->bwd_pending) DMSC_CHECK(hipStreamWaitEvent(st, ctx->ev_bwd, 0));
  if (dirty) fill_f32_zero(scratch, cap, st);
  dirty = true;  // until every chunk's gather (which re-zeroes the scratch) is queued
  for_each_bwd_chunk(B, fields, H, W, pairs, gscale, grad, cap, [&](const BwdChunk& c) {
    launch_persistence_backward_chunk((flags & DMSC_DUAL) != 0, c, max_pers, scratch, st);
  });
  dirty = false;
  DMSC_CHECK(hipEventRecord(ctx->ev_bwd, st));
  ctx->bwd_pending = true;
}

// Runs one entry point under the context lock; exceptions become a status plus the calling
// thread's error message. The HIP device of the calling thread is restored afterwards.
dmsc_status guarded(dmsc_context* ctx, const std::function<void()>& fn) {
  if (!ctx) return DMSC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  std::string& err = ctx->err[std::this_thread::get_id()];
  err.clear();
  tl_times = &ctx->times;
  tl_tune = &ctx->tune;
  int prev_dev = -1;
  (void)hipGetDevice(&prev_dev);
  dmsc_status st = DMSC_OK;
  try {
    DMSC_CHECK(hipSetDevice(ctx->device));
    fn();
  } catch (const Fail& e) {
    err = e.msg;
    st = e.st;
  } catch (const HipError& e) {
    err = std::string(e.what) + ": " + hipGetErrorString(e.err);
    st = DMSC_ERR_HIP;
  } catch (const std::exception& e) {
    err = e.what();
    st = DMSC_ERR_INTERNAL;
  }
  if (prev_dev >= 0 && prev_dev != ctx->device) (void)hipSetDevice(prev_dev);
  return st;
}
}  // namespace

extern "C" {

dmsc_status dmsc_context_create(int device, dmsc_context** out) {
  if (!out) return DMSC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int prev_dev = -1;
  (void)hipGetDevice(&prev_dev);
  auto* c = new dmsc_context();
  c->device = device;
  dmsc_status st = DMSC_OK;
  if (hipSetDevice(device) != hipSuccess || hipHostMalloc((void**)&c->pinned, 4096) != hipSuccess) {
    st = DMSC_ERR_HIP;
  } else {
    c->pinned_cap = 4096;
    if (hipEventCreateWithFlags(&c->ev_bwd, hipEventDisableTiming) != hipSuccess) st = DMSC_ERR_HIP;
    if (st == DMSC_OK) st = scratch_setup(device);
  }
  if (prev_dev >= 0 && prev_dev != device) (void)hipSetDevice(prev_dev);
  if (st != DMSC_OK) {
    if (c->ev_bwd) (void)hipEventDestroy(c->ev_bwd);
    if (c->pinned) (void)hipHostFree(c->pinned);
    delete c;
    return st;
  }
  *out = c;
  return DMSC_OK;
}

void dmsc_context_destroy(dmsc_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (int i = 0; i < dmsc_context::kMaxPeers; ++i) {
    if (ctx->peer[i]) dmsc_context_destroy(ctx->peer[i]);
    if (ctx->peer_st[i]) (void)hipStreamDestroy(ctx->peer_st[i]);
    if (ctx->ev_pout[i]) (void)hipEventDestroy(ctx->ev_pout[i]);
  }
  if (ctx->ev_pin) (void)hipEventDestroy(ctx->ev_pin);
  ctx->ar.release();
  ctx->ar2.release();
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->late) (void)hipStreamDestroy(ctx->late);
  if (ctx->ev_late) (void)hipEventDestroy(ctx->ev_late);
  if (ctx->ev_late_join) (void)hipEventDestroy(ctx->ev_late_join);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->ev_mark) (void)hipEventDestroy(ctx->ev_mark);
  if (ctx->ev_side) (void)hipEventDestroy(ctx->ev_side);
  if (ctx->ev_bwd) (void)hipEventDestroy(ctx->ev_bwd);
  if (ctx->bwd_scratch) (void)hipFree(ctx->bwd_scratch);
  delete ctx;
}

dmsc_status dmsc_extract(dmsc_context* ctx, const float* d_field, int32_t H, int32_t W, float thr, uint32_t flags,
                         dmsc_stream_t stream, dmsc_alloc_fn alloc, void* user, dmsc_outputs* out) {
  if (!alloc || !out) return DMSC_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    run_fields(ctx, {FieldIn{d_field, H, W, thr, 0}}, flags, (hipStream_t)stream, alloc, user, out);
  });
}

dmsc_status dmsc_extract_batch(dmsc_context* ctx, int32_t B, const float* const* d_fields, const int32_t* H,
                               const int32_t* W, const float* thr, uint32_t flags, dmsc_stream_t stream,
                               dmsc_alloc_fn alloc, void* user, dmsc_outputs* out) {
  if (!alloc || !out || B < 0 || (B > 0 && (!d_fields || !H || !W || !thr))) return DMSC_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    std::vector<FieldIn> fields((size_t)B);
    for (int b = 0; b < B; ++b) fields[b] = FieldIn{d_fields[b], H[b], W[b], thr[b], b};
    run_fields(ctx, fields, flags, (hipStream_t)stream, alloc, user, out);
  });
}

dmsc_status dmsc_persistence_backward(dmsc_context* ctx, int32_t B, const float* const* d_fields, const int32_t* H,
                                      const int32_t* W, uint32_t flags, const dmsc_pairs* pairs,
                                      float max_persistence, const float* d_grad_scale, float* const* d_grad,
                                      dmsc_stream_t stream) {
  if (B < 0 || (B > 0 && (!d_fields || !H || !W || !pairs || !d_grad_scale || !d_grad))) return DMSC_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    const hipStream_t st = (hipStream_t)stream;
    const int64_t need = bwd_scratch_needed(B, d_fields, H, W, pairs, d_grad_scale, d_grad, INT64_MAX);
    if (need == 0) return;
    if (ctx->bwd_cap < need) {  // grow: the old buffer may still be in use by the last call's stream
      if (ctx->bwd_pending) DMSC_CHECK(hipEventSynchronize(ctx->ev_bwd));
      if (ctx->bwd_scratch) DMSC_CHECK(hipFree(ctx->bwd_scratch));
      ctx->bwd_scratch = nullptr;
      ctx->bwd_cap = 0;
      const int64_t cap = need + need / 4 + 1024;
      DMSC_CHECK(hipMalloc((void**)&ctx->bwd_scratch, (size_t)cap * sizeof(float)));
      ctx->bwd_cap = cap;
      ctx->bwd_dirty = true;
    }
    run_backward(ctx, B, d_fields, H, W, pairs, d_grad_scale, d_grad, flags, max_persistence, ctx->bwd_scratch,
                 ctx->bwd_cap, ctx->bwd_dirty, st);
    if (ctx->times.level) {
      DMSC_CHECK(hipStreamSynchronize(st));
      ctx->times.collect();
    }
  });
}

dmsc_status dmsc_persistence_backward_noalloc(dmsc_context* ctx, int32_t B, const float* const* d_fields,
                                              const int32_t* H, const int32_t* W, uint32_t flags,
                                              const dmsc_pairs* pairs, float max_persistence,
                                              const float* d_grad_scale, float* const* d_grad, dmsc_stream_t stream) {
  if (B < 0 || (B > 0 && (!d_fields || !H || !W || !pairs || !d_grad_scale || !d_grad))) return DMSC_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
      throw Fail{DMSC_ERR_INVALID_ARGUMENT, "the allocation-free backward supports devices 0.." +
                                                std::to_string(kMaxDevices - 1)};
    // no timing events here: their pool may create events
    KernelTimes* saved = tl_times;
    tl_times = nullptr;
    struct Restore {
      KernelTimes* t;
      ~Restore() { tl_times = t; }
    } restore{saved};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    float* s = g_scratch_ptr[ctx->device];
    if (!s) throw Fail{DMSC_ERR_HIP, "the allocation-free scratch is not set up on this device"};
    run_backward(ctx, B, d_fields, H, W, pairs, d_grad_scale, d_grad, flags, max_persistence, s,
                 kNoAllocScratchFloats, g_scratch_dirty[ctx->device], (hipStream_t)stream);
  });
}

dmsc_status dmsc_persistence_loss(dmsc_context* ctx, int32_t B, const dmsc_pairs* pairs, float max_persistence,
                                  float* d_loss, dmsc_stream_t stream) {
  if (B < 0 || (B > 0 && (!pairs || !d_loss))) return DMSC_ERR_INVALID_ARGUMENT;
  return guarded(ctx, [&] {
    (void)bwd_scratch_needed(B, nullptr, nullptr, nullptr, pairs, nullptr, nullptr, INT64_MAX);  // checks
    // partial sums of one chunk at a time; chunks on the same stream reuse the buffer in order
    const hipStream_t st = (hipStream_t)stream;
    double* part = ctx->ar.get<double>(A_LOSSPART, loss_partials(kBwdChunk));
    if (ctx->bwd_pending) DMSC_CHECK(hipStreamWaitEvent(st, ctx->ev_bwd, 0));
    for_each_bwd_chunk(B, nullptr, nullptr, nullptr, pairs, nullptr, nullptr, INT64_MAX,
                       [&](const BwdChunk& c) { launch_persistence_loss_chunk(c, max_persistence, part, d_loss, st); });
    DMSC_CHECK(hipEventRecord(ctx->ev_bwd, st));
    ctx->bwd_pending = true;
  });
}

namespace {
int* param_slot(dmsc_context* ctx, const std::string& n, int64_t& lo, int64_t& hi) {
  Tunables& t = ctx->tune;
  if (n == "round_grid") { lo = 64; hi = 65536; return &t.round_grid; }
  if (n == "small_list") { lo = 0; hi = 1 << 30; return &t.small_list; }
  if (n == "rounds_first") { lo = 1; hi = 4096; return &t.rounds_first; }
  if (n == "rounds_max") { lo = 1; hi = 4096; return &t.rounds_max; }
  if (n == "expand_levels") { lo = 1; hi = 4096; return &t.expand_levels; }
  if (n == "graph_rounds") { lo = 0; hi = 1; return &t.graph_rounds; }
  if (n == "tail_items") { lo = 0; hi = 1 << 16; return &t.tail_items; }
  if (n == "grad_rows") { lo = 8; hi = 32; return &t.grad_rows; }
  if (n == "side_rounds") { lo = 0; hi = 4096; return &t.side_rounds; }
  if (n == "pipe_groups") { lo = 1; hi = 1 + dmsc_context::kMaxPeers; return &t.pipe_groups; }
  if (n == "pipe_min_slots") { lo = 0; hi = INT32_MAX; return &t.pipe_min_slots; }
  return nullptr;
}
}  // namespace

dmsc_status dmsc_set_param(dmsc_context* ctx, const char* name, int64_t value) {
  if (!ctx || !name) return DMSC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  int64_t lo = 0, hi = 0;
  int* p = param_slot(ctx, name, lo, hi);
  if (!p || value < lo || value > hi) return DMSC_ERR_INVALID_ARGUMENT;
  if (p == &ctx->tune.grad_rows && value != 8 && value != 16 && value != 32) return DMSC_ERR_INVALID_ARGUMENT;
  *p = (int)value;
  return DMSC_OK;
}

dmsc_status dmsc_get_param(dmsc_context* ctx, const char* name, int64_t* value) {
  if (!ctx || !name || !value) return DMSC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  int64_t lo = 0, hi = 0;
  int* p = param_slot(ctx, name, lo, hi);
  if (!p) return DMSC_ERR_INVALID_ARGUMENT;
  *value = *p;
  return DMSC_OK;
}

const char* dmsc_last_error(dmsc_context* ctx) {
  if (!ctx) return "null context";
  std::lock_guard<std::mutex> lock(ctx->mu);
  return ctx->err[std::this_thread::get_id()].c_str();  // the entry stays until this thread's next call
}

dmsc_status dmsc_profile_enable(dmsc_context* ctx, int32_t on) {
  if (!ctx) return DMSC_ERR_INVALID_ARGUMENT;
  if (on < 0 || on > 2) return DMSC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->times.level = on;
  ctx->times.acc.clear();
  return DMSC_OK;
}
int32_t dmsc_profile_count(dmsc_context* ctx) {
  if (!ctx) return 0;
  std::lock_guard<std::mutex> lock(ctx->mu);
  return (int32_t)ctx->times.acc.size();
}
dmsc_status dmsc_profile_entry(dmsc_context* ctx, int32_t i, const char** name, double* total_ms, int64_t* launches) {
  if (!ctx) return DMSC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (i < 0 || i >= (int32_t)ctx->times.acc.size()) return DMSC_ERR_INVALID_ARGUMENT;
  const auto& kv = ctx->times.acc[i];
  if (name) *name = kv.first.c_str();
  if (total_ms) *total_ms = kv.second.first;
  if (launches) *launches = kv.second.second;
  return DMSC_OK;
}

int64_t dmsc_debug_fetch(dmsc_context* ctx, const char* name, void* dst, int64_t cap) {
  if (!ctx || !name) return -1;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return -1;
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
  if (n == "arc_stats") {
    const int64_t m = std::min<int64_t>(4, cap);
    if (dst && m > 0) std::memcpy(dst, ctx->arc_stats, (size_t)m * sizeof(int64_t));
    return 4;
  }
  if (n == "geo_pending") {
    int64_t m = std::min<int64_t>((int64_t)ctx->geo_log.size(), cap);
    if (dst && m > 0) std::memcpy(dst, ctx->geo_log.data(), (size_t)m * 4);
    return (int64_t)ctx->geo_log.size();
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
  if (n == "face_pj") return copy(ctx->pjf, ctx->S, 8);
  if (n == "vert_pj") return copy(ctx->pjv, ctx->N, 8);
  return -1;
}

}  // extern "C"
