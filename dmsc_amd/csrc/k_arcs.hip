// Arc geometry: raw V-path arcs, the arc-merging contraction of cancelled saddles, and emission of
// the merged ridge / valley arcs.
//
// Reference (cpu/arcs_simplification.hxx:310-504): cancellations sorted by (persistence, saddle id)
// are applied in order to a union-find over the extrema of one side; a cancellation whose dead
// extremum d is the current root of one of its saddle's two arc targets (and the roots differ)
// links d to the other root, with the link path
//     L(d) = reverse(arc_dead ++ path(T_dead -> d)) ++ arc_other ++ path(T_other -> root_other),
// where path(x -> r) concatenates the link paths along x's parent chain. A surviving arc is
// arc ++ path(T -> final root). (The reference stores these concatenations as a DAG; here the
// parent forest plus per-link metadata is the whole representation.)
//
// Rounds: a pending cancellation c is decided when
//  (A) neither current root R_k of its targets can still be linked by a pending cancellation that
//      precedes c (the one whose dead is R_k is done, absent, or not earlier than c), and
//  (B) if c links d, no pending cancellation that precedes c currently has d as a root (so links
//      made out of order never sit below an undecided earlier cancellation).
// Under (A) and (B) every decision and link path equals the sequential one; the smallest pending
// cancellation always satisfies both.

#include "k_topo.hpp"

namespace dmsc {

__device__ __forceinline__ int32_t sat_add(int32_t a, int32_t b) {
  int64_t s = (int64_t)a + (int64_t)b;
  return s > 0x7FFFFFFF ? 0x7FFFFFFF : (int32_t)s;
}

// ------------------------------------------------------------------ raw arcs
__global__ void k_arc_lengths(int64_t ns, const int32_t* __restrict__ c, const int32_t* __restrict__ steps,
                              int32_t* __restrict__ blen, int4* __restrict__ tl) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < ns; s += (int64_t)gridDim.x * blockDim.x) {
    const int2 cc = *reinterpret_cast<const int2*>(c + 2 * s);
    const int2 st = *reinterpret_cast<const int2*>(steps + 2 * s);
    const int2 l = make_int2(cc.x >= 0 ? st.x + 1 : 0, cc.y >= 0 ? st.y + 1 : 0);
    *reinterpret_cast<int2*>(blen + 2 * s) = l;
    tl[s] = make_int4(cc.x, cc.y, l.x, l.y);
  }
}
void launch_arc_lengths(int64_t ns, const int32_t* c, const int32_t* steps, int32_t* blen, int4* tl, hipStream_t st) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_arc_lengths, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, c, steps, blen, tl);
  DMSC_LAUNCH_CHECK();
}

// start cell (field-local forest index) of arc k of saddle s; -1 if absent
__device__ __forceinline__ int arc_start(const ArcSide& A, const FieldDesc& fd, int s, int k) {
  const int id = A.sad_ids[s], q = id >> 2, t = id & 3;
  const int Nx = fd.Nx, W = fd.W;
  if (A.faces) return k == 0 ? q : (t == 1 ? q + Nx : q + 1);
  const int y = q / Nx, x = q - y * Nx;
  int v[2], n = 0;
  if (t == 1) {
    if (x - 1 >= 0) v[n++] = y * W + x - 1;
    if (x < W) v[n++] = y * W + x;
  } else {
    if (y - 1 >= 0) v[n++] = (y - 1) * W + x;
    if (y < fd.H) v[n++] = y * W + x;
  }
  return k < n ? v[k] : -1;
}
// cell id of a field-local forest node
__device__ __forceinline__ int32_t cell_of(const ArcSide& A, const FieldDesc& fd, int node) {
  if (A.faces) return 4 * node + 3;
  const int y = node / fd.W, x = node - y * fd.W;
  return 4 * (y * fd.Nx + x);
}

// The arcs that an output can use (their saddle survives or this side cancels it): walk the successor
// forest from the start cell to the extremum. An arc of at most kHeadCells cells goes to its head
// record (one 16-byte store, indexed by arc), a longer one to the arena. A wave takes kRawChunk x 64
// consecutive arcs: it writes the short ones at once and queues the long ones in LDS, then walks the
// queue with every lane taking the next queued arc as soon as its own ends, so that one long arc no
// longer holds 63 idle lanes (the walk is one dependent load per cell).
constexpr int kRawChunk = 8, kRawWaves = 4;
__global__ __launch_bounds__(64 * kRawWaves) void k_raw_arcs(ArcSide A) {
  __shared__ int2 q[kRawWaves][64 * kRawChunk];  // per wave: (arc offset in the chunk, length) of its long arcs
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1;
  const int64_t n = 2 * A.ns, chunk = 64 * kRawChunk;
  for (int64_t c0 = ((int64_t)blockIdx.x * kRawWaves + wv) * chunk; c0 < n;
       c0 += (int64_t)gridDim.x * kRawWaves * chunk) {
    int nq = 0;
    for (int u = 0; u < kRawChunk; ++u) {
      const int64_t a = c0 + u * 64 + lane;
      int len = 0;
      if (a < n) {
        len = A.blen[a];
        if (len > 0 && A.removed[a >> 1] && A.cancel[a >> 1] < 0) len = 0;
      }
      if (len > 0 && len <= kHeadCells) {
        const int s = (int)(a >> 1);
        const FieldDesc& fd = A.fd[A.sfield ? A.sfield[s] : 0];
        const int base = (int)(A.faces ? fd.slot_base : fd.pix_base);
        int cur = base + arc_start(A, fd, s, (int)(a & 1));
        int32_t h[kHeadCells] = {-1, -1, -1, -1};
#pragma unroll
        for (int t = 0; t < kHeadCells; ++t)
          if (t < len) {
            h[t] = cell_of(A, fd, cur - base);
            if (t + 1 < len) cur = A.succ[cur];
          }
        A.head[a] = make_int4(h[0], h[1], h[2], h[3]);
      }
      const bool lng = len > kHeadCells;
      const uint64_t m = __ballot(lng);
      if (lng) q[wv][nq + __popcll(m & below)] = make_int2(u * 64 + lane, len);
      nq += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    // queued arcs: rem cells left to write at o, the walk at cur (field-local cell ids from base)
    int rem = 0, cur = 0, base = 0, W = 1, Nx = 1;
    int32_t* o = nullptr;
    for (int head = 0;;) {
      const uint64_t idle = __ballot(rem == 0);
      if (head >= nq && idle == ~0ull) break;  // wave-uniform
      if (rem == 0) {
        const int r = head + __popcll(idle & below);
        if (r < nq) {
          const int2 e = q[wv][r];
          const int64_t a = c0 + e.x;
          const int s = (int)(a >> 1);
          const FieldDesc& fd = A.fd[A.sfield ? A.sfield[s] : 0];
          base = (int)(A.faces ? fd.slot_base : fd.pix_base);
          W = fd.W;
          Nx = fd.Nx;
          cur = base + arc_start(A, fd, s, (int)(a & 1));
          o = A.arena + A.boff[a];
          rem = e.y;
        }
      }
      head += __popcll(idle);
      if (rem > 0) {
        const int node = cur - base;
        int32_t cell;
        if (A.faces) {
          cell = 4 * node + 3;
        } else {
          const int y = node / W;
          cell = 4 * (y * Nx + node - y * W);
        }
        *o++ = cell;
        if (--rem > 0) cur = A.succ[cur];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}
void launch_raw_arcs(const ArcSide& a, hipStream_t st) {
  KScope ks_("raw_arcs", st);
  if (a.ns == 0) return;
  const int64_t chunk = 64 * kRawChunk, waves = (2 * a.ns + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_raw_arcs, dim3(grid_for(waves, kRawWaves, 16384)), dim3(64 * kRawWaves), 0, st, a);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ cancellation order
__global__ void k_cancel_flags(int64_t ns, const int32_t* __restrict__ cancel_d, int32_t* __restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ns; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = cancel_d[i] >= 0 ? 1 : 0;
}
void launch_cancel_flags(int64_t ns, const int32_t* cancel_d, int32_t* flag, hipStream_t st) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_cancel_flags, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, cancel_d, flag);
  DMSC_LAUNCH_CHECK();
}
// compaction in saddle (= cell id) order, so a stable sort on persistence yields (pers, saddle id)
__global__ void k_cancel_compact(int64_t ns, const int32_t* __restrict__ cancel_d, const float* __restrict__ pers,
                                 const int32_t* __restrict__ pre, int32_t* __restrict__ cs, int32_t* __restrict__ cd,
                                 uint32_t* __restrict__ keys) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ns; i += (int64_t)gridDim.x * blockDim.x) {
    const int d = cancel_d[i];
    if (d < 0) continue;
    const int j = pre[i];
    cs[j] = (int)i;
    cd[j] = d;
    keys[j] = sortable_key(pers[i]);
  }
}
void launch_cancel_compact(int64_t ns, const int32_t* cancel_d, const float* pers, const int32_t* pre, int32_t* cs,
                           int32_t* cd, uint32_t* keys, hipStream_t st) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_cancel_compact, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, cancel_d, pers, pre, cs, cd,
                     keys);
  DMSC_LAUNCH_CHECK();
}
// ------------------------------------------------------------------ geometry rounds
// initial worklist: every cancellation (saddle-id order) with its key and its two arc targets; and
// kill word of d = (pending, key of the cancellation whose dead extremum d is)
__global__ void k_geo_items(int64_t nc, const int32_t* __restrict__ cs, const int32_t* __restrict__ cd,
                            const int32_t* __restrict__ key, const int32_t* __restrict__ tinit, int4* __restrict__ wl,
                            uint64_t* __restrict__ gk, int32_t* __restrict__ cnt) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt[0] = (int32_t)nc;
    cnt[1] = 0;
  }
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const int s = cs[c];
    const int2 t = *reinterpret_cast<const int2*>(tinit + 2 * s);
    const int k = key[c];
    wl[c] = make_int4((int)c, k, t.x, t.y);
    gk[2 * (int64_t)cd[c] + 1] = kKillPending | (uint32_t)k;
  }
}
void launch_geo_items(int64_t nc, const int32_t* cs, const int32_t* cd, const int32_t* key, const int32_t* tinit,
                      int4* wl, uint64_t* gk, int32_t* cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_geo_items, dim3(grid_for(std::max<int64_t>(nc, 1), 256, 16384)), dim3(256), 0, st, nc, cs, cd,
                     key, tinit, wl, gk, cnt);
  DMSC_LAUNCH_CHECK();
}

// Roots in the geometry forest and the summed link lengths on the way, by path halving over the
// forest entries gk[2 e]: (ancestor J, length to J) pairs. The forest only grows at roots (a link is only ever
// added from a root), so a shortcut stays valid once written; finds run in their own kernel (or
// phase), so concurrent shortcuts are all consistent 64-bit pairs, and an older pair read from a
// cache is still valid: the accesses are single 64-bit loads and stores without a coherence scope.
constexpr uint32_t kVoid = 0xFFFFFFFFu;
// x[k] <- root of x[k] (-1: the void), acc[k] <- the length to it, for the two arc targets of one
// item, walked in lockstep (each step issues both walks' loads before using either). Find rounds
// only compress paths, so the walks are independent whatever their interleaving.
__device__ __forceinline__ void geo_root2(const GeoState& g, int (&x)[2], int32_t (&acc)[2]) {
  acc[0] = acc[1] = 0;
  bool done[2] = {x[0] < 0, x[1] < 0};
  for (int it = 0; !(done[0] && done[1]); ++it) {
    if (it >= kLoopCap) {
      report(g.err, ERR_GEO_CAP);
      break;
    }
    uint64_t cur[2] = {0, 0}, nx[2] = {0, 0};
    uint32_t J[2] = {0, 0};
    int32_t L[2] = {0, 0};
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (!done[k]) cur[k] = __hip_atomic_load(&g.gk[2 * (int64_t)x[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (done[k]) continue;
      J[k] = (uint32_t)cur[k];
      L[k] = (int32_t)(cur[k] >> 32);
      if (J[k] == (uint32_t)x[k]) {
        done[k] = true;
      } else if (J[k] == kVoid) {
        acc[k] = sat_add(acc[k], L[k]);
        x[k] = -1;
        done[k] = true;
      }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (!done[k]) nx[k] = __hip_atomic_load(&g.gk[2 * (int64_t)J[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (done[k]) continue;
      const uint32_t JJ = (uint32_t)nx[k];
      if (JJ == J[k]) {  // J is the root
        acc[k] = sat_add(acc[k], L[k]);
        x[k] = (int)J[k];
        done[k] = true;
        continue;
      }
      const int32_t L2 = sat_add(L[k], (int32_t)(nx[k] >> 32));
      __hip_atomic_store(&g.gk[2 * (int64_t)x[k]], (uint64_t)JJ | ((uint64_t)(uint32_t)L2 << 32), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WAVEFRONT);
      acc[k] = sat_add(acc[k], L2);
      x[k] = JJ == kVoid ? -1 : (int)JJ;
      if (x[k] < 0) done[k] = true;
    }
  }
}

// watch[R] is only ever read by R's killer (condition B of its decision), so cancellation kc posts on
// root R only while R's killer is pending and does not come before kc (an earlier killer is not
// affected by a larger key, and a decided or absent killer reads nothing).
__device__ __forceinline__ bool watch_needed(uint64_t kill_word, int kc) {
  return (kill_word >> 32) == 0xFFFFFFFFull && (uint32_t)kill_word != 0xFFFFFFFFu && (int)(uint32_t)kill_word >= kc;
}

// R (kill word k) stays a root until after cancellation `kc` (key): the cancellation that could link
// it is done, absent, or comes later in (persistence, saddle id) order. Evaluated in the find: a
// killer decided during this round's apply carries this round's number, so the answer there is the
// same as here.
__device__ __forceinline__ bool root_settled(uint64_t k, int kc, int round) {
  return (uint32_t)k == 0xFFFFFFFFu || (uint32_t)(k >> 32) < (uint32_t)round || (int)(uint32_t)k >= kc;
}

// worklist item: (cancellation, key, arc target 0, arc target 1); results go to rr at the item's
// list position: (R0, R1, L0, L1) with bit 31 of a length set when that root is not settled. In
// round 0 the forest has no links yet: every target is its own root at length 0. Returns the watch
// posts on the two roots.
__device__ __forceinline__ RootPosts geo_find_one(const ArcSide& A, const GeoState& g, const int4 it, int64_t j, int cur,
                                                  uint64_t tag, bool first) {
  int R0 = it.z, R1 = it.w;
  int32_t L0 = 0, L1 = 0;
  if (!first) {  // from round 1 on the item holds the roots it last waited with, lens(cur) the lengths to them
    int x[2] = {it.z, it.w};
    int32_t acc[2];
    geo_root2(g, x, acc);
    R0 = x[0];
    R1 = x[1];
    const int2 base = g.lens(cur)[j];
    L0 = sat_add(base.x, acc[0]);
    L1 = sat_add(base.y, acc[1]);
  }
  const int round = (int)(0xFFFFFFFFu - (uint32_t)(tag >> 32));
  const uint64_t k0 = R0 >= 0 ? g.gk[2 * (int64_t)R0 + 1] : ~0ull, k1 = R1 >= 0 ? g.gk[2 * (int64_t)R1 + 1] : ~0ull;
  const int32_t f0 = root_settled(k0, it.y, round) ? 0 : INT32_MIN, f1 = root_settled(k1, it.y, round) ? 0 : INT32_MIN;
  *reinterpret_cast<int4*>(g.rr + 4 * j) = make_int4(R0, R1, L0 | f0, L1 | f1);
  RootPosts p;
  p.key = tag | (uint32_t)it.y;
  if (R0 >= 0 && watch_needed(k0, it.y)) p.a = &g.watch[R0];
  if (R1 >= 0 && watch_needed(k1, it.y)) p.b = &g.watch[R1];
  return p;
}

// Worklists keep saddle-id order: neighbouring lanes often post to the same extremum, so a wave
// combines its posts first; from round 1 on the posting lane also reads before its atomic.
__device__ __forceinline__ void geo_post(const RootPosts& p, bool first) {
  wave_post_min(p.a, p.key, !first);
  wave_post_min(p.b, p.key, !first);
}

__device__ __forceinline__ void geo_find_phase(const ArcSide& A, const GeoState& g, int round) {
  const int cur = round & 1;
  const int n = g.cnt[cur];
  if (blockIdx.x == 0 && threadIdx.x == 0) g.cnt[cur ^ 1] = 0;
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  DMSC_WAVE_LOOP(jw, (int64_t)n) {
    const int64_t j = jw_base + (threadIdx.x & 63);
    RootPosts p;
    if (j < n) p = geo_find_one(A, g, g.list(cur)[j], j, cur, tag, round == 0);
    geo_post(p, round == 0);
  }
}

// Both traced sides of one round in one launch: side 0's items first, then side 1's.
__global__ void k_geo_find2(ArcSide A0, GeoState G0, ArcSide A1, GeoState G1, int live, int round,
                            const int32_t* __restrict__ d_round) {
  if (d_round) round = *d_round;  // graph replays take the round number from the device
  const int cur = round & 1;
  const int n0 = (live & 1) ? G0.cnt[cur] : 0, n1 = (live & 2) ? G1.cnt[cur] : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (live & 1) G0.cnt[cur ^ 1] = 0;
    if (live & 2) G1.cnt[cur ^ 1] = 0;
  }
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  DMSC_WAVE_LOOP(jw, (int64_t)n0 + n1) {
    const int64_t j = jw_base + (threadIdx.x & 63);
    RootPosts p;
    if (j < (int64_t)n0 + n1) {
      if (j >= n0) p = geo_find_one(A1, G1, G1.list(cur)[j - n0], j - n0, cur, tag, round == 0);
      else p = geo_find_one(A0, G0, G0.list(cur)[j], j, cur, tag, round == 0);
    }
    geo_post(p, round == 0);
  }
}
__global__ void k_geo_find(ArcSide A, GeoState g, int round) { geo_find_phase(A, g, round); }

// Cancellation c (item it, round results r = (R0, R1, L0, L1) with both roots settled, so the
// lengths carry no flag; dead extremum d) is decided: record it, and link d when it is one of c's
// roots (kd).
__device__ __forceinline__ void geo_decide(const ArcSide& A, const GeoState& g, const int4 it, const int4 r, int d,
                                           int kd, int round) {
  const int c = it.x, kc = it.y;
  g.gk[2 * (int64_t)d + 1] = ((uint64_t)(uint32_t)round << 32) | (uint32_t)kc;  // c is decided (d is c's dead extremum)
  if (kd >= 0) {
    const int s = g.cs[c], ko = 1 - kd;
    const int Ro = kd == 0 ? r.y : r.x;
    const int32_t Ld = kd == 0 ? r.z : r.w, Lo = kd == 0 ? r.w : r.z;
    const int32_t ll = sat_add(sat_add(Ld, A.blen[2 * s + kd]), sat_add(A.blen[2 * s + ko], Lo));
    g.link[d] = make_int4(Ro, ll, s, kd);
    g.gk[2 * (int64_t)d] = (uint64_t)(Ro < 0 ? kVoid : (uint32_t)Ro) | ((uint64_t)(uint32_t)ll << 32);
  }
}

// Waiting item `it` (round results r) takes slot o of list k with the roots it found and the
// lengths to them, so that the next find resumes there.
__device__ __forceinline__ void geo_park(const GeoState& g, int k, int o, const int4 it, const int4 r) {
  g.list(k)[o] = make_int4(it.x, it.y, r.x, r.y);
  g.lens(k)[o] = make_int2(r.z & INT32_MAX, r.w & INT32_MAX);  // without the settled flags
}

// which of c's current roots is its dead extremum d (-1: neither)
__device__ __forceinline__ int geo_kd(const int4 r, int d) {
  if (r.x == d && r.x != r.y) return 0;
  if (r.y == d && r.x != r.y) return 1;
  return -1;
}

// Decides one pending cancellation (index c, saddle-id order); returns true when it must wait.
__device__ __forceinline__ bool geo_apply_one(const ArcSide& A, const GeoState& g, const int4 it, int64_t j, int round,
                                              uint64_t tag) {
  const int4 r = *reinterpret_cast<const int4*>(g.rr + 4 * j);
  if (r.z < 0 || r.w < 0) return true;  // a root is not settled (flags of the find)
  const int d = g.cd[it.x];
  const int kc = it.y;
  const int kd = geo_kd(r, d);
  if (kd >= 0 && g.watch[d] != (tag | (uint32_t)kc)) return true;
  geo_decide(A, g, it, r, d, kd, round);
  return false;
}

// Small worklists (the tail of the rounds) are latency-bound: one item per thread there.
constexpr int kGeoItems = 4, kGeoChunk = 256 * kGeoItems, kGeoSmall = 1 << 16;

template <int ITEMS>
__device__ __forceinline__ void geo_apply_chunks(const ArcSide& A, const GeoState& g, int round, int cur, int n,
                                                 uint64_t tag) {
  DMSC_BLOCK_LOOP(start, (int64_t)n, 256 * ITEMS) {
    const int64_t j0 = start + (int64_t)threadIdx.x * ITEMS;
    int4 c[ITEMS];
    int32_t nwait = 0;
    uint32_t waitmask = 0;
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      c[u] = make_int4(0, 0, 0, 0);
      if (j0 + u < n) {
        c[u] = g.list(cur)[j0 + u];
        if (geo_apply_one(A, g, c[u], j0 + u, round, tag)) {
          waitmask |= 1u << u;
          ++nwait;
        }
      }
    }
    int32_t* const cnt[1] = {&g.cnt[cur ^ 1]};
    const int32_t req[1] = {nwait};
    int32_t first[1];
    block_reserve<1>(cnt, req, first);
    int32_t o = first[0];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u)
      if (waitmask >> u & 1) {
        geo_park(g, cur ^ 1, o++, c[u], *reinterpret_cast<const int4*>(g.rr + 4 * (j0 + u)));
      }
  }
}

__device__ __forceinline__ void geo_apply_phase(const ArcSide& A, const GeoState& g, int round) {
  const int cur = round & 1;
  const int n = g.cnt[cur];
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  if (n <= kGeoSmall) geo_apply_chunks<1>(A, g, round, cur, n, tag);
  else geo_apply_chunks<kGeoItems>(A, g, round, cur, n, tag);
}
__global__ __launch_bounds__(256) void k_geo_apply(ArcSide A, GeoState g, int round) { geo_apply_phase(A, g, round); }

// Both traced sides of one round in one launch; chunks of side 0 come first, so every chunk (and
// block) has one side.
// ITEMS consecutive items per thread, their loads issued stage by stage (items and round results;
// dead extrema and the killers of both roots; the watch words) so that the latencies overlap.
template <int ITEMS>
__device__ __forceinline__ void geo_apply_chunk(const ArcSide& A, const GeoState& g, int round, int cur, int64_t start,
                                                int n, uint64_t tag) {
  const int64_t j0 = start + (int64_t)threadIdx.x * ITEMS;
  const int4* const in = g.list(cur);
  int4 c[ITEMS], r[ITEMS];
  int d[ITEMS];
  bool ready[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    const bool v = j0 + u < n;
    c[u] = v ? in[j0 + u] : make_int4(-1, 0, 0, 0);
    r[u] = v ? *reinterpret_cast<const int4*>(g.rr + 4 * (j0 + u)) : make_int4(-1, -1, 0, 0);
  }
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    d[u] = -1;
    ready[u] = c[u].x >= 0 && r[u].z >= 0 && r[u].w >= 0;  // both roots settled (flags of the find)
    if (ready[u]) d[u] = g.cd[c[u].x];  // the dead extremum matters only then
  }
  int kd[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    kd[u] = ready[u] ? geo_kd(r[u], d[u]) : -1;
    if (ready[u] && kd[u] >= 0) ready[u] = g.watch[d[u]] == (tag | (uint32_t)c[u].y);
  }
  int32_t nwait = 0;
  uint32_t waitmask = 0;
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if (c[u].x < 0) continue;
    if (!ready[u]) {
      waitmask |= 1u << u;
      ++nwait;
    } else {
      geo_decide(A, g, c[u], r[u], d[u], kd[u], round);
    }
  }
  int32_t* const cnt[1] = {&g.cnt[cur ^ 1]};
  const int32_t req[1] = {nwait};
  int32_t first[1];
  block_reserve<1>(cnt, req, first);
  int32_t o = first[0];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u)
    if (waitmask >> u & 1) geo_park(g, cur ^ 1, o++, c[u], r[u]);
}
template <int ITEMS>
__device__ __forceinline__ void geo_apply_both(const ArcSide& A0, const GeoState& G0, const ArcSide& A1,
                                               const GeoState& G1, int round, int cur, int n0, int n1, uint64_t tag) {
  constexpr int CH = 256 * ITEMS;
  const int64_t nc0 = ceil_div(n0, CH), nc = nc0 + ceil_div(n1, CH);
  for (int64_t ch = blockIdx.x; ch < nc; ch += gridDim.x) {
    if (ch >= nc0) geo_apply_chunk<ITEMS>(A1, G1, round, cur, (ch - nc0) * CH, n1, tag);
    else geo_apply_chunk<ITEMS>(A0, G0, round, cur, ch * CH, n0, tag);
  }
}
__global__ __launch_bounds__(256) void k_geo_apply2(ArcSide A0, GeoState G0, ArcSide A1, GeoState G1, int live,
                                                    int round, const int32_t* __restrict__ d_round, int small_list) {
  if (d_round) round = *d_round;
  const int cur = round & 1;
  const int n0 = (live & 1) ? G0.cnt[cur] : 0, n1 = (live & 2) ? G1.cnt[cur] : 0;
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  if ((int64_t)n0 + n1 <= small_list) geo_apply_both<1>(A0, G0, A1, G1, round, cur, n0, n1, tag);
  else geo_apply_both<kGeoItems>(A0, G0, A1, G1, round, cur, n0, n1, tag);
}

void launch_geo_round2(const ArcSide* a, const GeoState* g, int live, int round, int64_t max_items, hipStream_t st,
                       const int32_t* d_round) {
  KScope ks_("geo_round", st);
  const Tunables& tu = tune();
  hipLaunchKernelGGL(k_geo_find2, dim3(grid_for(max_items, 256, tu.round_grid)), dim3(256), 0, st, a[0], g[0], a[1],
                     g[1], live, round, d_round);
  DMSC_LAUNCH_CHECK();
  const int per = max_items <= tu.small_list ? 256 : kGeoChunk;
  hipLaunchKernelGGL(k_geo_apply2, dim3(grid_for(max_items, per, tu.round_grid)), dim3(256), 0, st, a[0], g[0], a[1],
                     g[1], live, round, d_round, tu.small_list);
  DMSC_LAUNCH_CHECK();
}

// one item of the tail kernel's apply phase; waiting items take next-list slots from an LDS counter
__device__ __forceinline__ void geo_tail_apply(const ArcSide& A, const GeoState& g, int j, int cur, int r, uint64_t tag,
                                               int32_t* s_next) {
  const int4 it = g.list(cur)[j];
  if (geo_apply_one(A, g, it, j, r, tag))
    geo_park(g, cur ^ 1, atomicAdd(s_next, 1), it, *reinterpret_cast<const int4*>(g.rr + 4 * j));
}

// The last geometry rounds of both traced sides in one workgroup (same structure as the
// persistence tail kernel): the pending counts end in cnt[r & 1], the round reached in *round_out.
__global__ __launch_bounds__(kTailThreads) void k_geo_tail(ArcSide A0, GeoState G0, ArcSide A1, GeoState G1, int live,
                                                           int round0, int max_rounds,
                                                           int32_t* __restrict__ round_out) {
  __shared__ int32_t s_next[2];
  int n0 = (live & 1) ? __hip_atomic_load(&G0.cnt[round0 & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  int n1 = (live & 2) ? __hip_atomic_load(&G1.cnt[round0 & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  int r = round0;
  for (; r < round0 + max_rounds && n0 + n1 > 0; ++r) {
    const int cur = r & 1;
    const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)r) << 32;
    if (threadIdx.x == 0) s_next[0] = s_next[1] = 0;
    for (int j = threadIdx.x; j < n0 + n1; j += kTailThreads) {
      const RootPosts p = j >= n0 ? geo_find_one(A1, G1, G1.list(cur)[j - n0], j - n0, cur, tag, r == 0)
                                  : geo_find_one(A0, G0, G0.list(cur)[j], j, cur, tag, r == 0);
      if (p.a) post_min(p.a, p.key, r > 0);
      if (p.b) post_min(p.b, p.key, r > 0);
    }
    phase_barrier();
    for (int j = threadIdx.x; j < n0 + n1; j += kTailThreads) {
      if (j >= n0) geo_tail_apply(A1, G1, j - n0, cur, r, tag, &s_next[1]);
      else geo_tail_apply(A0, G0, j, cur, r, tag, &s_next[0]);
    }
    phase_barrier();
    n0 = s_next[0];
    n1 = s_next[1];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (live & 1) G0.cnt[r & 1] = n0;
    if (live & 2) G1.cnt[r & 1] = n1;
    *round_out = r;
  }
}

void launch_geo_tail(const ArcSide* a, const GeoState* g, int live, int round0, int max_rounds, int32_t* round_out,
                     hipStream_t st) {
  KScope ks_("geo_round", st);
  hipLaunchKernelGGL(k_geo_tail, dim3(1), dim3(kTailThreads), 0, st, a[0], g[0], a[1], g[1], live, round0, max_rounds,
                     round_out);
  DMSC_LAUNCH_CHECK();
}

void launch_geo_round(const ArcSide& a, const GeoState& g, int round, int64_t max_items, hipStream_t st) {
  KScope ks_("geo_round", st);
  const int grid = grid_for(max_items, 256, 2048);
  hipLaunchKernelGGL(k_geo_find, dim3(grid), dim3(256), 0, st, a, g, round);
  DMSC_LAUNCH_CHECK();
  const int per = max_items <= kGeoSmall ? 256 : kGeoChunk;
  hipLaunchKernelGGL(k_geo_apply, dim3(grid_for(max_items, per, 2048)), dim3(256), 0, st, a, g, round);
  DMSC_LAUNCH_CHECK();
}

__global__ void k_geo_init(int4* __restrict__ link, uint64_t* __restrict__ gk,
                           uint64_t* __restrict__ watch, int64_t next) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < next; e += (int64_t)gridDim.x * blockDim.x) {
    link[e] = make_int4((int)e, 0, 0, 0);
    gk[2 * e] = (uint64_t)(uint32_t)e;
    gk[2 * e + 1] = ~0ull;
    watch[e] = ~0ull;
  }
}
void launch_geo_init(const GeoState& g, int64_t next, hipStream_t st) {
  if (next == 0) return;
  hipLaunchKernelGGL(k_geo_init, dim3(grid_for(next, 256, 16384)), dim3(256), 0, st, g.link, g.gk, g.watch,
                     next);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ path lengths to the final root
// pjq[e] = (final root or `next` for the void, length of path(e -> root)), by pointer jumping with
// saturating sums over the geometry forest. It starts from the rounds' own shortcuts: gk[2 e] is an
// ancestor of e (its parent when nothing compressed it) with the saturated length to it.
__global__ void k_geo_q_init(GeoState g, int64_t next, uint64_t* __restrict__ pj) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= next; e += (int64_t)gridDim.x * blockDim.x) {
    uint32_t J = (uint32_t)e, D = 0;
    if (e < next) {
      const uint64_t gj = g.gk[2 * e];
      const uint32_t a = (uint32_t)gj;
      if (a != (uint32_t)e) {
        J = a == kVoid ? (uint32_t)next : a;
        D = (uint32_t)(gj >> 32);
      }
    }
    pj[e] = (uint64_t)J | ((uint64_t)D << 32);
  }
}
// Saturating pointer jumping (same access rules as k_pj: unscoped single 64-bit accesses).
__global__ void k_pj_sat(uint64_t* __restrict__ pj, int64_t n, int32_t* __restrict__ err) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t cur = __hip_atomic_load(&pj[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    uint32_t J = (uint32_t)cur;
    int32_t D = (int32_t)(cur >> 32);
    if (J == (uint32_t)i) continue;
    for (int it = 0;; ++it) {
      if (it >= kLoopCap) {  // only a cyclic forest can get here
        report(err, ERR_GEO_CAP);
        break;
      }
      const uint64_t nx = __hip_atomic_load(&pj[J], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      const uint32_t JJ = (uint32_t)nx;
      if (JJ == J) break;
      D = sat_add(D, (int32_t)(nx >> 32));
      J = JJ;
      __hip_atomic_store(&pj[i], (uint64_t)J | ((uint64_t)(uint32_t)D << 32), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  }
}
void launch_geo_q(const GeoState& g, int64_t next, uint64_t* pjq, hipStream_t st) {
  KScope ks_("geo_q", st);
  hipLaunchKernelGGL(k_geo_q_init, dim3(grid_for(next + 1, 256, 16384)), dim3(256), 0, st, g, next, pjq);
  DMSC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pj_sat, dim3(grid_for(next + 1, 256, 16384)), dim3(256), 0, st, pjq, next + 1, g.err);
  DMSC_LAUNCH_CHECK();
}

__device__ __forceinline__ int32_t qlen(const uint64_t* pjq, int x) { return x >= 0 ? (int32_t)(pjq[x] >> 32) : 0; }

// ------------------------------------------------------------------ output arcs
// Per surviving saddle, in output order, what the output stages read about its two arcs, in one
// 32-byte record: targets, raw lengths, path lengths from the targets to their final roots and those
// roots. The lengths pass gathers them once by saddle id; the seeds pass then reads them in order.
struct alignas(32) ArcRec {
  int32_t T[2];     // arc targets (-1: no arc)
  int32_t L[2];     // raw lengths
  int32_t q[2];     // path length from the target to its final root (0 without an arc)
  int32_t root[2];  // the target's final root (pjq entry; `next` for the void)
};
static_assert(sizeof(ArcRec) == 32, "ArcRec is 32 bytes");

__global__ void k_final_lengths(int64_t ns, const int32_t* __restrict__ ord, ArcSide A,
                                const uint64_t* __restrict__ pjq, ArcRec* __restrict__ rec,
                                int64_t* __restrict__ flen, int64_t* __restrict__ per_pos) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ns; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = ord[k];
    const int4 tl = A.tl[i];
    const int2 T = make_int2(tl.x, tl.y), L = make_int2(tl.z, tl.w);
    const uint64_t p0 = T.x >= 0 ? pjq[T.x] : 0, p1 = T.y >= 0 ? pjq[T.y] : 0;
    ArcRec r;
    r.T[0] = T.x;
    r.T[1] = T.y;
    r.L[0] = L.x;
    r.L[1] = L.y;
    r.q[0] = (int32_t)(p0 >> 32);
    r.q[1] = (int32_t)(p1 >> 32);
    r.root[0] = (int32_t)(uint32_t)p0;
    r.root[1] = (int32_t)(uint32_t)p1;
    rec[k] = r;
    const int64_t l0 = T.x >= 0 ? (int64_t)L.x + r.q[0] : 0, l1 = T.y >= 0 ? (int64_t)L.y + r.q[1] : 0;
    flen[2 * k] = l0;
    flen[2 * k + 1] = l1;
    per_pos[k] = l0 + l1;
  }
}
void launch_final_lengths(int64_t ns, const int32_t* ord, const ArcSide& a, const uint64_t* pjq, void* rec,
                          int64_t* flen, int64_t* per_pos, hipStream_t st) {
  KScope ks_("final_lengths", st);
  if (ns == 0) return;
  hipLaunchKernelGGL(k_final_lengths, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, ord, a, pjq,
                     static_cast<ArcRec*>(rec), flen, per_pos);
  DMSC_LAUNCH_CHECK();
}

__device__ __forceinline__ int64_t seg_od(int64_t off, int fwd) { return off << 1 | (fwd ? 1 : 0); }

// Seeds: per surviving saddle arc, its raw arc then CHAIN(target -> final root); also the field-local
// CSR offsets.
__global__ __launch_bounds__(256) void k_seed_segments(int64_t ns, const int32_t* __restrict__ ord,
                                ArcSide A, const ArcRec* __restrict__ rec, int64_t next,
                                const int64_t* __restrict__ flen, const int64_t* __restrict__ pos_pre,
                                Seg* __restrict__ chains,
                                int32_t* __restrict__ nchain, Seg* __restrict__ raw, int32_t* __restrict__ nraw,
                                int64_t cap, int32_t* __restrict__ g_err, int32_t* __restrict__ out_off,
                                int32_t* __restrict__ out_arc_off, const int64_t* __restrict__ arc_base) {
  DMSC_BLOCK_LOOP(kb, ns, 256) {
    const int64_t kk0 = kb + threadIdx.x;
    bool has_raw[2] = {false, false}, has_chain[2] = {false, false};
    Seg rs[2], cs[2];
    if (kk0 < ns) {
      // position kk0 of the survivor list is output saddle kk0
      const int i = ord[kk0];
      const int fi = A.sfield ? A.sfield[i] : 0;
      const int64_t base = pos_pre[kk0];
      const int64_t local = base - arc_base[fi];  // field-local output offset
      const int64_t f0 = flen[2 * kk0];
      out_off[kk0 + fi] = (int32_t)local;
      out_arc_off[kk0] = (int32_t)(local + f0);
      const ArcRec r = rec[kk0];
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        const int a = 2 * i + kk, T = r.T[kk];
        if (T < 0) continue;
        const int64_t off = base + (kk ? f0 : 0);
        has_raw[kk] = true;
        rs[kk] = Seg{a, r.L[kk], seg_od(off, 1)};
        if (r.q[kk] > 0) {
          const uint32_t root = (uint32_t)r.root[kk];
          has_chain[kk] = true;
          cs[kk] = Seg{T, root == (uint32_t)next ? -1 : (int)root, seg_od(off + r.L[kk], 1)};
        }
      }
    }
    int32_t* const cnt[2] = {nraw, nchain};
    const int32_t req[2] = {(has_raw[0] ? 1 : 0) + (has_raw[1] ? 1 : 0), (has_chain[0] ? 1 : 0) + (has_chain[1] ? 1 : 0)};
    int32_t first[2];
    block_reserve<2>(cnt, req, first);
    if (first[0] + req[0] > cap || first[1] + req[1] > cap) {
      if (req[0] + req[1] > 0) report(g_err, ERR_SEG_CAP);
      continue;
    }
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      if (has_raw[kk]) raw[first[0]++] = rs[kk];
      if (has_chain[kk]) chains[first[1]++] = cs[kk];
    }
  }
}
// the closing entry of each field's offsets: its total merged length
__global__ void k_set_last_offsets(int32_t* __restrict__ out_off, const int64_t* __restrict__ arc_base,
                                   const int32_t* __restrict__ sad_out, int B) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < B; f += gridDim.x * blockDim.x)
    out_off[sad_out[f + 1] + f] = (int32_t)(arc_base[f + 1] - arc_base[f]);
}

// ------------------------------------------------------------------ link expansion
// The link of extremum d (made by cancellation c, saddle s, dead arc kd, other arc ko) expands to
//   reverse(CHAIN(T_kd -> d)) ++ reverse(raw kd) ++ raw ko ++ CHAIN(T_ko -> R_o)
// with the chains present when their length is positive and raw ko when the arc exists.
struct LinkParts {
  int s, kd, ko, Td, To, Ro;
  int32_t Alen, Bk, Bo, Clen;
};
// lk = link[d] of a linked extremum d
__device__ __forceinline__ LinkParts link_parts(const ArcSide& A, const uint64_t* __restrict__ pjq, int d,
                                                const int4 lk) {
  LinkParts L;
  L.kd = lk.w;
  L.ko = 1 - L.kd;
  L.s = lk.z;
  const int4 tl = A.tl[L.s];
  L.Td = L.kd ? tl.y : tl.x;
  L.To = L.kd ? tl.x : tl.y;
  L.Ro = lk.x;
  L.Alen = qlen(pjq, L.Td) - qlen(pjq, d);
  L.Bk = L.kd ? tl.w : tl.z;
  L.Bo = L.kd ? tl.z : tl.w;
  L.Clen = L.To >= 0 ? qlen(pjq, L.To) - qlen(pjq, L.Ro) : 0;
  return L;
}

// cum_r / cum_c: (parent or the void node `next`, raw / chain pieces of the link) for linked
// extrema, (self, 0) for roots; pointer jumping then turns them into path sums.
// The link record of every extremum e < next (the expansion then reads one record per link instead of
// the link, the saddle's targets and four path lengths).
__global__ void k_link_counts(ArcSide A, GeoState g, const uint64_t* __restrict__ pjq, int64_t next,
                              uint64_t* __restrict__ cum_r, uint64_t* __restrict__ cum_c, LinkRec* __restrict__ xl) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= next; e += (int64_t)gridDim.x * blockDim.x) {
    uint32_t J = (uint32_t)e, nr = 0, nc = 0;
    if (e < next) {
      const int4 lk = g.link[e];
      LinkRec R{(int)e, -1, -1, 0, 0, 0, 0, 0};
      if (lk.x != (int)e) {
        J = lk.x < 0 ? (uint32_t)next : (uint32_t)lk.x;
        const LinkParts L = link_parts(A, pjq, (int)e, lk);
        nr = 1 + (L.Bo > 0 ? 1 : 0);
        nc = (L.Alen > 0 ? 1 : 0) + (L.Clen > 0 ? 1 : 0);
        R = LinkRec{lk.x, L.Td, L.To, 2 * L.s + L.kd, L.Alen, L.Bk, L.Bo, L.Clen};
      }
      xl[e] = R;
    }
    cum_r[e] = (uint64_t)J | ((uint64_t)nr << 32);
    cum_c[e] = (uint64_t)J | ((uint64_t)nc << 32);
  }
}
__global__ void k_pack_counts(const uint64_t* __restrict__ pjq, const uint64_t* __restrict__ cum_r,
                              const uint64_t* __restrict__ cum_c, int64_t n, int4* __restrict__ ncq) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    ncq[e] = make_int4((int32_t)(cum_r[e] >> 32), (int32_t)(cum_c[e] >> 32), (int32_t)(pjq[e] >> 32), 0);
}
void launch_link_counts(const ArcSide& a, const GeoState& g, const uint64_t* pjq, int64_t next, uint64_t* cum_r,
                        uint64_t* cum_c, LinkRec* xl, int4* ncq, hipStream_t st) {
  KScope ks_("link_counts", st);
  hipLaunchKernelGGL(k_link_counts, dim3(grid_for(next + 1, 256, 16384)), dim3(256), 0, st, a, g, pjq, next, cum_r,
                     cum_c, xl);
  DMSC_LAUNCH_CHECK();
  pointer_jump(cum_r, next + 1, g.err, st);
  pointer_jump(cum_c, next + 1, g.err, st);
  hipLaunchKernelGGL(k_pack_counts, dim3(grid_for(next + 1, 256, 16384)), dim3(256), 0, st, pjq, cum_r, cum_c, next + 1,
                     ncq);
  DMSC_LAUNCH_CHECK();
}

// One expansion level. Each chain (a -> b) reserves exactly its links' pieces (path-sum
// differences of the raw / chain piece counts, read with its length from the ends' packed records)
// with one atomic per wave, then walks its links writing them.
__global__ __launch_bounds__(256) void k_expand(ArcSide A, GeoState g, const int4* __restrict__ ncq,
                         const LinkRec* __restrict__ xl,
                         const Seg* __restrict__ in, const int32_t* __restrict__ nin,
                         Seg* __restrict__ out, int32_t* __restrict__ nout, int32_t* __restrict__ nreset,
                         Seg* __restrict__ raw, int32_t* __restrict__ nraw, int64_t cap) {
  // levels are launched back to back without host checks: the input count comes from the device
  // (clamped to the list capacity; an overflow was already reported) and this level clears the
  // counter the next level will append to (the one the previous level read)
  const int n = (int)(*nin < cap ? *nin : cap);
  if (blockIdx.x == 0 && threadIdx.x == 0) *nreset = 0;
  DMSC_BLOCK_LOOP(jb, (int64_t)n, 256) {
    const int64_t jj = jb + threadIdx.x;
    Seg sg{0, 0, 0};
    int32_t nr = 0, nc = 0, Lc = 0;
    if (jj < n) {
      sg = in[jj];
      const int4 ca = ncq[sg.a], cb = sg.b >= 0 ? ncq[sg.b] : make_int4(0, 0, 0, 0);
      nr = ca.x - cb.x;
      nc = ca.y - cb.y;
      Lc = ca.z - cb.z;
    }
    int32_t* const cnt[2] = {nraw, nout};
    const int32_t req[2] = {nr, nc};
    int32_t first[2];
    block_reserve<2>(cnt, req, first);
    int32_t pr = first[0], pc = first[1];
    if (jj >= n) continue;
    if ((int64_t)pr + nr > cap || (int64_t)pc + nc > cap) {
      report(g.err, ERR_SEG_CAP);
      continue;
    }
    const int b = sg.b;
    const int fwd = (int)(sg.od & 1);
    const int64_t off = sg.od >> 1;
    const int pr_end = pr + nr, pc_end = pc + nc;
    int y = sg.a;
    int64_t acc = 0;
    for (int it = 0; y >= 0 && y != b; ++it) {
      const LinkRec L = xl[y];
      if (L.next == y) break;
      if (it >= kLoopCap) {
        report(g.err, ERR_GEO_CAP);
        break;
      }
      // the link's length (the sum cannot overflow here: a saturated length on any expanded path
      // makes the merged arcs' total exceed 32 bits, which the host refuses before expanding)
      const int32_t len = L.Alen + L.Bk + L.Bo + L.Clen;
      const int need_r = 1 + (L.Bo > 0 ? 1 : 0), need_c = (L.Alen > 0 ? 1 : 0) + (L.Clen > 0 ? 1 : 0);
      if (pr + need_r > pr_end || pc + need_c > pc_end) {  // counts disagree with the walk: never overrun
        report(g.err, ERR_SEG_CAP);
        break;
      }
      int64_t o = fwd ? off + acc : off + Lc - acc - len;
      if (fwd) {
        if (L.Alen > 0) out[pc++] = Seg{L.Td, y, seg_od(o, 0)};
        o += L.Alen;
        raw[pr++] = Seg{L.arc, L.Bk, seg_od(o, 0)};
        o += L.Bk;
        if (L.Bo > 0) raw[pr++] = Seg{L.arc ^ 1, L.Bo, seg_od(o, 1)};
        o += L.Bo;
        if (L.Clen > 0) out[pc++] = Seg{L.To, L.next, seg_od(o, 1)};
      } else {
        if (L.Clen > 0) out[pc++] = Seg{L.To, L.next, seg_od(o, 0)};
        o += L.Clen;
        if (L.Bo > 0) raw[pr++] = Seg{L.arc ^ 1, L.Bo, seg_od(o, 0)};
        o += L.Bo;
        raw[pr++] = Seg{L.arc, L.Bk, seg_od(o, 1)};
        o += L.Bk;
        if (L.Alen > 0) out[pc++] = Seg{L.Td, y, seg_od(o, 1)};
      }
      acc += len;
      y = L.next;
    }
  }
}
void launch_expand_segments(const ArcSide& a, const GeoState& g, const int4* ncq, const LinkRec* xl, const Seg* in,
                            const int32_t* nin, Seg* out, int32_t* nout, int32_t* nreset, Seg* raw, int32_t* nraw,
                            int64_t max_in, int64_t cap, hipStream_t st) {
  KScope ks_("expand_segments", st);
  hipLaunchKernelGGL(k_expand, dim3(grid_for(max_in, 256, 8192)), dim3(256), 0, st, a, g, ncq, xl, in, nin, out, nout,
                     nreset, raw, nraw, cap);
  DMSC_LAUNCH_CHECK();
}

// Raw pieces -> output cells. A piece of at most kHeadCells cells is written by its own lane from the
// arc's head record. A wave lays the cells of its other short pieces end to end (prefix sum of
// lengths) and copies that run 64 cells at a time from the arena, so reads and writes are contiguous
// within each piece; pieces of kLongPiece cells or more are then copied one at a time by the whole
// wave, without the per-cell piece search.
constexpr int32_t kLongPiece = 128;
__global__ __launch_bounds__(256) void k_copy_raw(ArcSide A, const Seg* __restrict__ raw,
                                                  const int32_t* __restrict__ nraw, int64_t max_n,
                                                  int32_t* __restrict__ out, int64_t out_total,
                                                  int32_t* __restrict__ err) {
  const int n = (int)(*nraw < max_n ? *nraw : max_n);
  const int lane = threadIdx.x & 63;
  DMSC_WAVE_LOOP(j, (int64_t)n) {
    const int64_t jj = j_base + lane;
    int32_t len = 0;
    int64_t dst = 0;
    int fwd = 1, arc = 0;
    if (jj < n) {
      const Seg sg = raw[jj];
      len = sg.b;  // raw pieces carry their length
      dst = sg.od >> 1;
      fwd = (int)(sg.od & 1);
      arc = sg.a;
      if (dst < 0 || dst + len > out_total) {  // inconsistent offsets: never write out of bounds
        report(err, ERR_SEG_CAP);
        len = 0;
      }
    }
    const bool hd = len <= kHeadCells;
    int64_t src = 0;
    if (hd) {
      if (len > 0) {
        const int4 h4 = A.head[arc];
        const int32_t h[kHeadCells] = {h4.x, h4.y, h4.z, h4.w};
#pragma unroll
        for (int t = 0; t < kHeadCells; ++t)
          if (t < len) out[dst + (fwd ? t : len - 1 - t)] = h[t];
      }
    } else {
      src = A.boff[arc];
    }
    // reversed pieces read src + len - 1 - t: encode as a negative start
    const int64_t srcv = fwd ? src : ~(src + len - 1);
    // long pieces are copied one at a time by the whole wave (below); the others are packed
    const bool lng = len >= kLongPiece;
    const int32_t slen = (lng || hd) ? 0 : len;
    const int32_t incl = wave_inclusive_sum(slen);
    const int32_t start = incl - slen;
    const int32_t total = __shfl(incl, 63, 64);
    for (int32_t base = 0; base < total; base += 64) {
      const int32_t idx = base + lane;
      int k = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        const int32_t sc = __shfl(start, k + step, 64);
        if (sc <= idx) k += step;
      }
      const int32_t t = idx - __shfl(start, k, 64);
      const int64_t d = __shfl(dst, k, 64);
      const int64_t sv = __shfl(srcv, k, 64);
      if (idx < total) out[d + t] = A.arena[sv >= 0 ? sv + t : ~sv - t];
    }
    for (uint64_t m = __ballot(lng); m; m &= m - 1) {  // wave-uniform
      const int l = __builtin_ctzll(m);
      const int32_t L = __shfl(len, l, 64);
      const int64_t d = __shfl(dst, l, 64), sv = __shfl(srcv, l, 64);
      const int32_t* __restrict__ a = A.arena;
      int32_t t = lane;
      for (; t + 192 < L; t += 256) {  // four loads in flight per lane
        const int32_t v0 = a[sv >= 0 ? sv + t : ~sv - t];
        const int32_t v1 = a[sv >= 0 ? sv + t + 64 : ~sv - t - 64];
        const int32_t v2 = a[sv >= 0 ? sv + t + 128 : ~sv - t - 128];
        const int32_t v3 = a[sv >= 0 ? sv + t + 192 : ~sv - t - 192];
        out[d + t] = v0;
        out[d + t + 64] = v1;
        out[d + t + 128] = v2;
        out[d + t + 192] = v3;
      }
      for (; t < L; t += 64) out[d + t] = a[sv >= 0 ? sv + t : ~sv - t];
    }
  }
}

void launch_copy_raw(const ArcSide& a, const Seg* raw, const int32_t* nraw, int64_t max_raw, int32_t* out,
                     int64_t out_total, int32_t* err, hipStream_t st) {
  KScope ks_("copy_raw", st);
  hipLaunchKernelGGL(k_copy_raw, dim3(grid_for(max_raw, 256, 16384)), dim3(256), 0, st, a, raw, nraw, max_raw, out, out_total,
                     err);
  DMSC_LAUNCH_CHECK();
}

void launch_seed_segments(int64_t ns, const int32_t* ord, const ArcSide& a, const void* rec, int64_t next,
                          const int64_t* flen, const int64_t* pos_pre, Seg* chains, int32_t* nchain, Seg* raw, int32_t* nraw, int64_t cap,
                          int32_t* err, int32_t* out_offsets, int32_t* out_arc_offsets, const int64_t* arc_base,
                          const int32_t* sad_out, int B, hipStream_t st) {
  KScope ks_("seed_segments", st);
  if (ns > 0) {
    hipLaunchKernelGGL(k_seed_segments, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, ord, a,
                       static_cast<const ArcRec*>(rec), next, flen, pos_pre, chains, nchain, raw, nraw, cap, err, out_offsets, out_arc_offsets,
                       arc_base);
    DMSC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_set_last_offsets, dim3(grid_for(B, 256, 64)), dim3(256), 0, st, out_offsets, arc_base, sad_out, B);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
