// Host-side plumbing shared by the kernel translation units: device scratch arena, launch
// helpers and the launcher declarations of every kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

#include "dmsc_common.hpp"

namespace dmsc {

struct HipError {
  hipError_t err;
  const char* what;
};

#define DMSC_CHECK(call)                                \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) throw ::dmsc::HipError{e_, #call}; \
  } while (0)

#define DMSC_LAUNCH_CHECK() DMSC_CHECK(hipGetLastError())

// Grow-only device scratch, one slot per named buffer. Buffers keep their contents between calls
// only by accident: every stage initialises what it reads.
class Arena {
 public:
  ~Arena() { release(); }
  void* get(int slot, size_t bytes) {
    if ((size_t)slot >= slots_.size()) slots_.resize(slot + 1);
    Slot& s = slots_[slot];
    if (bytes == 0) bytes = 256;
    if (s.cap < bytes) {
      if (s.p) DMSC_CHECK(hipFree(s.p));
      s.p = nullptr;
      size_t cap = bytes + bytes / 4 + 4096;
      DMSC_CHECK(hipMalloc(&s.p, cap));
      s.cap = cap;
    }
    return s.p;
  }
  template <class T>
  T* get(int slot, size_t count) {
    return static_cast<T*>(get(slot, count * sizeof(T)));
  }
  void release() {
    for (auto& s : slots_)
      if (s.p) (void)hipFree(s.p);
    slots_.clear();
  }

 private:
  struct Slot {
    void* p = nullptr;
    size_t cap = 0;
  };
  std::vector<Slot> slots_;
};

// Device-side error flags (bit set by kernels that hit a loop cap or a capacity bound).
enum : int32_t { ERR_PJ_CAP = 1, ERR_UF_CAP = 2, ERR_SEG_CAP = 4, ERR_GEO_CAP = 8 };
constexpr int kLoopCap = 1 << 22;
__device__ __forceinline__ void report(int32_t* err, int32_t code) {
  if (err) atomicOr(err, code);
}

// Optional per-kernel timing with HIP events on the launch stream (enabled per context; used by
// bench.py to measure the dominant kernel's average duration live). Events come from a reusable pool.
struct KernelTimes {
  int level = 0;  // 0: off, 1: key scopes only (a few events per call), 2: every scope
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<std::string, std::pair<double, int64_t>>> acc;  // name -> (total ms, launches)
  hipEvent_t take() {
    if (pool.empty()) {
      hipEvent_t e;
      DMSC_CHECK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  void add(const char* name, double ms) {
    for (auto& kv : acc)
      if (kv.first == name) {
        kv.second.first += ms;
        kv.second.second += 1;
        return;
      }
    acc.push_back({name, {ms, 1}});
  }
  // call after the stream has been synchronised
  void collect() {
    for (auto& r : pending) {
      float ms = 0.f;
      DMSC_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
      add(r.name, ms);
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
};
extern thread_local KernelTimes* tl_times;

// Scheduling parameters of a context (set through dmsc_set_param; results never depend on them).
struct Tunables {
  int round_grid = 2048;     // most blocks a round kernel (persistence, geometry) is launched with
  int small_list = 1 << 16;  // worklists up to this size are processed one item per thread
  int rounds_first = 4;      // rounds launched before the first host check of a round loop
  int rounds_max = 16;       // most rounds between host checks (the batch doubles up to this)
  int expand_levels = 8;     // arc expansion levels between host checks
  int graph_rounds = 0;      // 1: each batch of rounds is a captured hipGraph, replayed while it fits
  int tail_items = 2048;     // pending items at a host check up to which one workgroup runs every
                             // remaining round in one launch (0: never)
  int grad_rows = 16;        // slot rows of a gradient tile (8, 16 or 32; tiles are 64 slots wide)
  int side_rounds = 1;       // first persistence rounds run side by side (vertex side, then face side,
                             // which drops what the vertex side just merged); later rounds run both
                             // sides in each launch
  int pipe_groups = 4;       // a large batch runs as this many parts side by side ...
  int pipe_min_slots = 1 << 23;  // ... of at least this many slots each
};
extern thread_local const Tunables* tl_tune;
inline const Tunables& tune() {
  static const Tunables defaults;
  return tl_tune ? *tl_tune : defaults;
}

// Events around a stretch of launches on one stream. Timing events are stream markers that cost a
// few microseconds of GPU time each, so only `key` scopes record at level 1.
struct KScope {
  KernelTimes* t;
  const char* name;
  hipStream_t st;
  hipEvent_t a{};
  KScope(const char* n, hipStream_t s, bool key = false)
      : t(tl_times && tl_times->level >= (key ? 1 : 2) ? tl_times : nullptr), name(n), st(s) {
    if (t) {
      a = t->take();
      DMSC_CHECK(hipEventRecord(a, st));
    }
  }
  ~KScope() noexcept(false) {
    if (t) {
      hipEvent_t b = t->take();
      DMSC_CHECK(hipEventRecord(b, st));
      t->pending.push_back({name, a, b});
    }
  }
};

// 64-bit atomic min of a value that only ever decreases. With `check`, the current value is read
// first (a plain load: it may come from a cache) and the atomic is skipped when it could not lower
// it: the atomics of many saddles posting to one hot root then shrink to the few that still lower
// it. The read is conservative: a stale value is never below the current one.
__device__ __forceinline__ void post_min(uint64_t* a, uint64_t key, bool check) {
  if (check && __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) <= key) return;
  atomicMin((unsigned long long*)a, (unsigned long long)key);
}

// The (up to) two root posts of one round item: atomic min of `key` into *a and *b (null: none).
struct RootPosts {
  uint64_t* a = nullptr;
  uint64_t* b = nullptr;
  uint64_t key = 0;
};

// Atomic min posts of a whole wave, combined first: a lane whose address equals the previous
// lane's takes the smaller key (segmented min over each run of equal addresses, by doubling), and
// only the last lane of a run issues the atomic. Pays off when neighbouring lanes often post to
// one address. Every lane of the wave must call it; lanes with nothing to post pass a null address.
// `check`: the posting lane reads first and skips an atomic that could not lower the value.
__device__ __forceinline__ void wave_post_min(uint64_t* addr, uint64_t key, bool check) {
  const int lane = threadIdx.x & 63;
  const unsigned long long a = (unsigned long long)addr;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long pa = __shfl_up(a, o, 64);
    const unsigned long long pk = __shfl_up((unsigned long long)key, o, 64);
    if (lane >= o && pa == a && pk < key) key = pk;
  }
  const unsigned long long na = __shfl_down(a, 1, 64);
  if (addr && (lane == 63 || na != a)) post_min(addr, key, check);
}

// Barrier for kernels whose phases hand data to each other through global memory inside one
// workgroup: agent-scope release before and acquire after, so the next phase reads every store and
// atomic of the previous one (the acquire drops lines the CU's L1 may still hold).
__device__ __forceinline__ void phase_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Threads of the one-workgroup kernels that run the last rounds of a round loop.
constexpr int kTailThreads = 1024;

// Inclusive prefix sum over the 64 lanes of a wave (every lane must take part).
__device__ __forceinline__ int32_t wave_inclusive_sum(int32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Block-aggregated reservation for 256-thread blocks: every thread asks for n[l] >= 0 consecutive
// slots of list l; one atomicAdd per list and block (same-address atomics from every wave would
// serialise on one L2 channel). Returns the thread's first slot per list; slots follow thread
// order within the block. Must be reached by every thread of the block.
template <int NL>
__device__ __forceinline__ void block_reserve(int32_t* const (&counter)[NL], const int32_t (&n)[NL],
                                              int32_t (&first)[NL]) {
  __shared__ int32_t s_tot[NL][4];
  __shared__ int32_t s_base[NL];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t incl[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    incl[l] = wave_inclusive_sum(n[l]);
    if (lane == 63) s_tot[l][wid] = incl[l];
  }
  __syncthreads();
  if (threadIdx.x < NL) {
    int32_t* cp = counter[0];
#pragma unroll
    for (int k = 1; k < NL; ++k)
      if ((int)threadIdx.x == k) cp = counter[k];
    const int l = threadIdx.x;
    const int32_t t = s_tot[l][0] + s_tot[l][1] + s_tot[l][2] + s_tot[l][3];
    s_base[l] = t > 0 ? atomicAdd(cp, t) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    int32_t b = s_base[l];
    for (int w = 0; w < wid; ++w) b += s_tot[l][w];
    first[l] = b + incl[l] - n[l];
  }
  __syncthreads();
}

// Block-uniform grid-stride loop over chunks of `per_block` items (every thread of a block runs the
// same iterations; use with block_reserve).
#define DMSC_BLOCK_LOOP(b, n, per_block)                                                       \
  for (int64_t b = (int64_t)blockIdx.x * (per_block); b < (n); b += (int64_t)gridDim.x * (per_block))

// Wave-uniform grid-stride loop: every lane of a wave runs the same iterations (for wave-wide
// shuffles over a group of 64 consecutive items).
#define DMSC_WAVE_LOOP(j, n)                                                                        \
  for (int64_t j##_base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); j##_base < (n);   \
       j##_base += (int64_t)gridDim.x * blockDim.x)

inline int grid_for(int64_t n, int block, int max_blocks = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

// Geometry of one field.
struct Dims {
  int H, W, Nx, Ny;
  int64_t S, N;  // slots (H+1)(W+1), pixels H*W
  int WPR;       // 32-slot words per slot row
  int64_t PW;    // words per bit plane
};
inline Dims make_dims(int H, int W) {
  Dims d;
  d.H = H;
  d.W = W;
  d.Nx = W + 1;
  d.Ny = H + 1;
  d.S = (int64_t)d.Nx * d.Ny;
  d.N = (int64_t)H * W;
  d.WPR = (d.Nx + 31) / 32;
  d.PW = (int64_t)d.Ny * d.WPR;
  return d;
}

// ------------------------------------------------------------------ batches of fields
// All fields of a call are processed by the same launches. Every per-cell array is the concatenation
// of the fields' arrays; a field's slice starts at its base offset. Cell ids stay field-local.
struct FieldDesc {
  const float* f;  // H x W input
  int H, W, Nx, Ny, WPR;
  float thr;
  int64_t pix_base, slot_base, word_base;
};
struct TileRef {
  int field, tx, ty;
};
struct Batch {
  const FieldDesc* fd;    // [B]
  const int64_t* slot_b;  // [B+1] slot bases
  const int64_t* pix_b;   // [B+1] pixel bases
  const int64_t* word_b;  // [B+1] bit-plane word bases
  int B;
};
// index of the field whose [base[f], base[f+1]) range holds idx
__device__ __forceinline__ int field_of(const int64_t* __restrict__ base, int B, int64_t idx) {
  int lo = 0, hi = B - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// field of idx given the field `from` of a smaller index of the same range (base[from] <= idx): a few
// steps forward. Kernels whose block (or wave) covers consecutive indices search once for the first
// index of the block (a uniform search) and step from there per thread.
template <class T>
__device__ __forceinline__ int field_from(const T* __restrict__ base, int B, int from, int64_t idx) {
  int f = from;
  while (f + 1 < B && base[f + 1] <= idx) ++f;
  return f;
}

// same over int32 bases (list positions)
__device__ __forceinline__ int field_of64(const int32_t* __restrict__ base, int B, int64_t idx) {
  int lo = 0, hi = B - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Critical-cell rank structure over the batch: bit planes (type t at planes + t*PW, PW = total
// words) plus exclusive popcount prefixes per word for minima (type 0), saddles (types 1+2) and
// maxima (type 3). Ranks are global list indices (fields' lists are concatenated in field order).
struct RankView {
  const uint32_t* planes;
  const int32_t* pre_min;
  const int32_t* pre_sad;
  const int32_t* pre_max;
  int64_t PW;
  Batch bt;
};

// ------------------------------------------------------------------ launchers (k_*.hip)
// primitives
template <class T>
void scan_exclusive(const T* in, T* out, int64_t n, T* d_total, Arena& ar, int slot, hipStream_t st);
// exclusive scan of 0/1 flags stored as bytes (nonzero = 1)
void scan_u8_flags(const uint8_t* in, int32_t* out, int64_t n, int32_t* d_total, Arena& ar, int slot, hipStream_t st);
void scan_plane_popcounts(const uint32_t* plane_a, const uint32_t* plane_b, int32_t* out, int64_t n, int32_t* d_total,
                          Arena& ar, int slot, hipStream_t st);
// iota_vals: sort the pairs (keys[i], i); vals is then only written
void radix_sort_pairs(uint32_t* keys, int32_t* vals, uint32_t* keys_tmp, int32_t* vals_tmp, int64_t n, Arena& ar,
                      int slot, int pslot, hipStream_t st, bool iota_vals = false);
// rank[i] = position of element i in the stable ascending order of the 32-bit keys (keys only read)
void radix_rank(const uint32_t* keys, int32_t* rank, int64_t n, Arena& ar, int slot, int pslot, hipStream_t st);
void pointer_jump(uint64_t* pj, int64_t n, int32_t* err, hipStream_t st);
void pointer_jump_init(const int32_t* succ, uint64_t* pj, int64_t n, hipStream_t st);
void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st);
void fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t st);
void iota_i32(int32_t* p, int64_t n, hipStream_t st);
template <class T, class S>
void bases_at(const T* pre, const S* starts, int64_t n, const T* total, int B, T* out, hipStream_t st);

// grid stages (over all fields of a batch)
// also writes the initial pointer-jumping state of both V-path forests (pjf over slots, pjv over pixels)
// tile_b: [B+1] first gradient tile of each field (64 x rows slots per tile, row-major within a field;
// rows is 8, 16 or 32); pw, fsucc and vsucc may be null (not written)
int gradient_tiles(int H, int W, int rows);
void launch_gradient(bool dual, int rows, const Batch& bt, const int64_t* tile_b, int64_t ntiles, int32_t* pw,
                     int32_t* fsucc, int32_t* vsucc, uint32_t* planes, int64_t PW, uint64_t* pjf, uint64_t* pjv,
                     hipStream_t st);
// min_key / max_key: pack_key(value, cell id) per extremum; sad_field: field of each saddle (null: not
// wanted)
void launch_compact_lists(bool dual, const RankView& rv, int32_t* min_ids, uint64_t* min_key,
                          int32_t* sad_ids, float* sad_val, int32_t* sad_field, int32_t* max_ids, uint64_t* max_key,
                          hipStream_t st);
void launch_saddle_records(const RankView& rv, const int32_t* sad_ids, const int32_t* sad_field, int64_t nsad,
                           const uint64_t* pj_face, const uint64_t* pj_vert, int32_t* cmax, int32_t* cmin,
                           int32_t* lmax, int32_t* lmin, hipStream_t st);
// labels: root_out[x] = output index of extremum x's thresholded root (-1: none); out_base[f] = output
// index of field f's first extremum of that side
void launch_labels_faces(const RankView& rv, int64_t S, const uint64_t* pj_face, const int32_t* root_out,
                         const int32_t* out_base, int32_t* out, hipStream_t st);
void launch_labels_verts(const RankView& rv, int64_t N, const uint64_t* pj_vert, const int32_t* root_out,
                         const int32_t* out_base, int32_t* out, hipStream_t st);
void radix_sort_pairs_bits(uint32_t* keys, int32_t* vals, uint32_t* keys_tmp, int32_t* vals_tmp, int64_t n, int bits,
                           Arena& ar, int slot, int pslot, hipStream_t st, bool iota_vals = false);

}  // namespace dmsc
