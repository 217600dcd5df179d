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
__global__ void k_arc_lengths(int64_t n2, const int32_t* __restrict__ c, const int32_t* __restrict__ steps,
                              int32_t* __restrict__ blen) {
  for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < n2; a += (int64_t)gridDim.x * blockDim.x)
    blen[a] = c[a] >= 0 ? steps[a] + 1 : 0;
}
void launch_arc_lengths(int64_t n2, const int32_t* c, const int32_t* steps, int32_t* blen, hipStream_t st) {
  if (n2 == 0) return;
  hipLaunchKernelGGL(k_arc_lengths, dim3(grid_for(n2, 256, 16384)), dim3(256), 0, st, n2, c, steps, blen);
  DMSC_LAUNCH_CHECK();
}

// start cell (forest index) of arc k of saddle s; -1 if absent
__device__ __forceinline__ int arc_start(const ArcSide& A, int s, int k) {
  const int id = A.sad_ids[s], q = id >> 2, t = id & 3;
  if (A.faces) return k == 0 ? q : (t == 1 ? q + A.Nx : q + 1);
  const int y = q / A.Nx, x = q - y * A.Nx;
  int v[2], n = 0;
  if (t == 1) {
    if (x - 1 >= 0) v[n++] = y * A.W + x - 1;
    if (x < A.W) v[n++] = y * A.W + x;
  } else {
    if (y - 1 >= 0) v[n++] = (y - 1) * A.W + x;
    if (y < A.H) v[n++] = y * A.W + x;
  }
  return k < n ? v[k] : -1;
}
__device__ __forceinline__ int32_t cell_of(const ArcSide& A, int node) {
  if (A.faces) return 4 * node + 3;
  const int y = node / A.W, x = node - y * A.W;
  return 4 * (y * A.Nx + x);
}

// one thread per arc: walk the successor forest from the start cell to the extremum
__global__ void k_raw_arcs(ArcSide A) {
  for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < 2 * A.ns; a += (int64_t)gridDim.x * blockDim.x) {
    const int len = A.blen[a];
    if (len == 0) continue;
    int cur = arc_start(A, (int)(a >> 1), (int)(a & 1));
    int32_t* o = A.arena + A.boff[a];
    for (int j = 0; j < len; ++j) {
      o[j] = cell_of(A, cur);
      cur = A.succ[cur];
    }
  }
}
void launch_raw_arcs(const ArcSide& a, hipStream_t st) {
  KScope ks_("raw_arcs", st);
  if (a.ns == 0) return;
  hipLaunchKernelGGL(k_raw_arcs, dim3(grid_for(2 * a.ns, 256, 16384)), dim3(256), 0, st, a);
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
// Cancellations stay in saddle-id order (index j, spatially coherent); their sorted position is the
// merge key: key[j] = position of j in (persistence, saddle id) order; deadc[d] = j.
__global__ void k_cancel_finish(int64_t nc, const int32_t* __restrict__ slot, const int32_t* __restrict__ cd,
                                int32_t* __restrict__ key, int32_t* __restrict__ deadc) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const int j = slot[c];
    key[j] = (int)c;
    deadc[cd[j]] = j;
  }
}
void launch_cancel_finish(int64_t nc, const int32_t* slot, const int32_t* cd, int32_t* key, int32_t* deadc,
                          hipStream_t st) {
  if (nc == 0) return;
  hipLaunchKernelGGL(k_cancel_finish, dim3(grid_for(nc, 256, 16384)), dim3(256), 0, st, nc, slot, cd, key, deadc);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ geometry rounds
__device__ __forceinline__ void geo_root(const GeoState& g, int x, int& root, int32_t& len) {
  int32_t acc = 0;
  for (int it = 0; x >= 0; ++it) {
    if (it >= kLoopCap) {
      report(g.err, ERR_GEO_CAP);
      break;
    }
    const int p = g.gparent[x];
    if (p == x) break;
    acc = sat_add(acc, g.linklen[x]);
    x = p;
  }
  root = x;
  len = acc;
}

__global__ void k_geo_find(ArcSide A, GeoState g, int round) {
  const int cur = round & 1;
  const int n = g.cnt[cur];
  if (blockIdx.x == 0 && threadIdx.x == 0) g.cnt[cur ^ 1] = 0;
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int c = g.wl[cur][j];
    const int s = g.cs[c];
    int R0, R1;
    int32_t L0, L1;
    geo_root(g, A.tinit[2 * s], R0, L0);
    geo_root(g, A.tinit[2 * s + 1], R1, L1);
    g.rr[4 * c] = R0;
    g.rr[4 * c + 1] = R1;
    g.rr[4 * c + 2] = L0;
    g.rr[4 * c + 3] = L1;
    const uint64_t key = tag | (uint32_t)g.key[c];
    if (R0 >= 0) atomicMin((unsigned long long*)&g.watch[R0], (unsigned long long)key);
    if (R1 >= 0) atomicMin((unsigned long long*)&g.watch[R1], (unsigned long long)key);
  }
}

// R stays a root until after cancellation `kc` (key): the cancellation that could link it is done,
// absent, or comes later in (persistence, saddle id) order.
__device__ __forceinline__ bool root_settled(const GeoState& g, int R, int kc, int round) {
  if (R < 0) return true;
  const int dk = g.deadc[R];
  return dk < 0 || g.done[dk] < round || g.key[dk] >= kc;
}

// Decides one pending cancellation (index c, saddle-id order); returns true when it must wait.
__device__ __forceinline__ bool geo_apply_one(const ArcSide& A, const GeoState& g, int c, int round, uint64_t tag) {
  const int R0 = g.rr[4 * c], R1 = g.rr[4 * c + 1];
  const int32_t L0 = g.rr[4 * c + 2], L1 = g.rr[4 * c + 3];
  const int d = g.cd[c];
  const int kc = g.key[c];
  int kd = -1;
  if (R0 == d && R0 != R1) kd = 0;
  else if (R1 == d && R0 != R1) kd = 1;
  bool ready = root_settled(g, R0, kc, round) && root_settled(g, R1, kc, round);
  if (ready && kd >= 0) ready = g.watch[d] == (tag | (uint32_t)kc);
  if (!ready) return true;
  g.done[c] = round;
  if (kd >= 0) {
    const int s = g.cs[c], ko = 1 - kd;
    const int Ro = kd == 0 ? R1 : R0;
    const int32_t Ld = kd == 0 ? L0 : L1, Lo = kd == 0 ? L1 : L0;
    g.gparent[d] = Ro;
    g.linklen[d] = sat_add(sat_add(Ld, A.blen[2 * s + kd]), sat_add(A.blen[2 * s + ko], Lo));
    g.linkc[d] = c;
    g.linkkd[d] = kd;
  }
  return false;
}

__global__ void k_geo_apply(ArcSide A, GeoState g, int round) {
  const int cur = round & 1;
  const int n = g.cnt[cur];
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  const int lane = threadIdx.x & 63;
  DMSC_WAVE_LOOP(j, (int64_t)n) {
    const int64_t jj = j_base + lane;
    bool wait = false;
    int c = 0;
    if (jj < n) {
      c = g.wl[cur][jj];
      wait = geo_apply_one(A, g, c, round, tag);
    }
    const int pos = wave_append(&g.cnt[cur ^ 1], wait);
    if (wait) g.wl[cur ^ 1][pos] = c;
  }
}

void launch_geo_round(const ArcSide& a, const GeoState& g, int round, int64_t max_items, hipStream_t st) {
  KScope ks_("geo_round", st);
  const int grid = grid_for(max_items, 256, 2048);
  hipLaunchKernelGGL(k_geo_find, dim3(grid), dim3(256), 0, st, a, g, round);
  DMSC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_geo_apply, dim3(grid), dim3(256), 0, st, a, g, round);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ path lengths to the final root
// pjq[e] = (final root or `next` for the void, length of path(e -> root)), by pointer jumping with
// saturating sums over the geometry forest.
__global__ void k_geo_q_init(GeoState g, int64_t next, uint64_t* __restrict__ pj) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= next; e += (int64_t)gridDim.x * blockDim.x) {
    uint32_t J = (uint32_t)e, D = 0;
    if (e < next) {
      const int p = g.gparent[e];
      if (p != (int)e) {
        J = p < 0 ? (uint32_t)next : (uint32_t)p;
        D = (uint32_t)g.linklen[e];
      }
    }
    pj[e] = (uint64_t)J | ((uint64_t)D << 32);
  }
}
__global__ void k_pj_sat(uint64_t* __restrict__ pj, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t cur = __hip_atomic_load(&pj[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t J = (uint32_t)cur;
    int32_t D = (int32_t)(cur >> 32);
    if (J == (uint32_t)i) continue;
    while (true) {
      const uint64_t nx = __hip_atomic_load(&pj[J], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t JJ = (uint32_t)nx;
      if (JJ == J) break;
      D = sat_add(D, (int32_t)(nx >> 32));
      J = JJ;
      __hip_atomic_store(&pj[i], (uint64_t)J | ((uint64_t)(uint32_t)D << 32), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}
void launch_geo_q(const GeoState& g, int64_t next, uint64_t* pjq, hipStream_t st) {
  KScope ks_("geo_q", st);
  hipLaunchKernelGGL(k_geo_q_init, dim3(grid_for(next + 1, 256, 16384)), dim3(256), 0, st, g, next, pjq);
  DMSC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pj_sat, dim3(grid_for(next + 1, 256, 16384)), dim3(256), 0, st, pjq, next + 1);
  DMSC_LAUNCH_CHECK();
}

__device__ __forceinline__ int32_t qlen(const uint64_t* pjq, int x) { return x >= 0 ? (int32_t)(pjq[x] >> 32) : 0; }

// ------------------------------------------------------------------ output arcs
__global__ void k_final_lengths(int64_t ns, const int32_t* __restrict__ ord, const uint8_t* __restrict__ removed,
                                ArcSide A, const uint64_t* __restrict__ pjq, int64_t* __restrict__ flen,
                                int64_t* __restrict__ per_pos) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ns; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = ord[k];
    int64_t l[2] = {0, 0};
    if (!removed[i]) {
      for (int kk = 0; kk < 2; ++kk) {
        const int a = 2 * i + kk, T = A.tinit[a];
        if (T >= 0) l[kk] = (int64_t)A.blen[a] + qlen(pjq, T);
      }
    }
    flen[2 * k] = l[0];
    flen[2 * k + 1] = l[1];
    per_pos[k] = l[0] + l[1];
  }
}
void launch_final_lengths(int64_t ns, const int32_t* ord, const uint8_t* removed, const ArcSide& a, const uint64_t* pjq,
                          int64_t /*next*/, int64_t* flen, int64_t* per_pos, hipStream_t st) {
  KScope ks_("final_lengths", st);
  if (ns == 0) return;
  hipLaunchKernelGGL(k_final_lengths, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, ord, removed, a, pjq, flen,
                     per_pos);
  DMSC_LAUNCH_CHECK();
}

__device__ __forceinline__ void push_seg(Seg* list, int32_t* n, int64_t cap, int32_t* err, int kind, int a, int b,
                                         int dir, int64_t off) {
  const int pos = atomicAdd(n, 1);
  if (pos >= cap) {
    report(err, ERR_SEG_CAP);
    return;
  }
  list[pos] = Seg{kind, a, b, dir, off};
}

// Seeds: per surviving saddle arc, RAW(arc) then CHAIN(target -> final root); also the CSR offsets.
__global__ void k_seed_segments(int64_t ns, const int32_t* __restrict__ ord, const uint8_t* __restrict__ removed,
                                ArcSide A, const uint64_t* __restrict__ pjq, int64_t next,
                                const int64_t* __restrict__ flen, const int64_t* __restrict__ pos_pre,
                                const int32_t* __restrict__ sad_to_out, Seg* __restrict__ segs,
                                int32_t* __restrict__ nseg, Seg* __restrict__ raw, int32_t* __restrict__ nraw,
                                int64_t cap, int32_t* __restrict__ g_err, int32_t* __restrict__ out_off,
                                int32_t* __restrict__ out_arc_off) {
  const int lane = threadIdx.x & 63;
  DMSC_WAVE_LOOP(k, ns) {
    const int64_t kk0 = k_base + lane;
    bool has_raw[2] = {false, false}, has_chain[2] = {false, false};
    Seg rs[2], cs[2];
    if (kk0 < ns) {
      const int i = ord[kk0];
      if (!removed[i]) {
        const int o = sad_to_out[i];
        const int64_t base = pos_pre[kk0];
        out_off[o] = (int32_t)base;
        out_arc_off[o] = (int32_t)(base + flen[2 * kk0]);
        for (int kk = 0; kk < 2; ++kk) {
          const int a = 2 * i + kk, T = A.tinit[a];
          if (T < 0) continue;
          const int64_t off = base + (kk ? flen[2 * kk0] : 0);
          has_raw[kk] = true;
          rs[kk] = Seg{0, a, 0, 1, off};
          if (qlen(pjq, T) > 0) {
            const uint32_t root = (uint32_t)pjq[T];
            has_chain[kk] = true;
            cs[kk] = Seg{1, T, root == (uint32_t)next ? -1 : (int)root, 1, off + A.blen[a]};
          }
        }
      }
    }
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const int pr = wave_append(nraw, has_raw[kk]);
      if (has_raw[kk]) {
        if (pr < cap) raw[pr] = rs[kk];
        else report(g_err, ERR_SEG_CAP);
      }
      const int pc = wave_append(nseg, has_chain[kk]);
      if (has_chain[kk]) {
        if (pc < cap) segs[pc] = cs[kk];
        else report(g_err, ERR_SEG_CAP);
      }
    }
  }
}
__global__ void k_set_last_offset(int32_t* out_off, int64_t n_out, const int64_t* total) {
  out_off[n_out] = (int32_t)*total;
}

// One expansion round: CHAIN -> LINK segments; LINK -> two RAW and up to two CHAIN segments.
__global__ void k_expand(ArcSide A, GeoState g, const uint64_t* __restrict__ pjq, const Seg* __restrict__ in,
                         const int32_t* __restrict__ nin, Seg* __restrict__ out, int32_t* __restrict__ nout,
                         Seg* __restrict__ raw, int32_t* __restrict__ nraw, int64_t cap) {
  const int n = *nin;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const Seg sg = in[j];
    if (sg.kind == 1) {
      const int b = sg.b;
      const int32_t Lc = qlen(pjq, sg.a) - qlen(pjq, b);
      int y = sg.a;
      int64_t acc = 0;
      for (int it = 0; y >= 0 && y != b && g.gparent[y] != y; ++it) {
        if (it >= kLoopCap) {
          report(g.err, ERR_GEO_CAP);
          break;
        }
        const int32_t len = g.linklen[y];
        const int64_t o = sg.dir ? sg.off + acc : sg.off + Lc - acc - len;
        push_seg(out, nout, cap, g.err, 2, y, 0, sg.dir, o);
        acc += len;
        y = g.gparent[y];
      }
    } else {
      const int d = sg.a;
      const int c = g.linkc[d], kd = g.linkkd[d], ko = 1 - kd;
      const int s = g.cs[c];
      const int Td = A.tinit[2 * s + kd], To = A.tinit[2 * s + ko];
      const int Ro = g.gparent[d];
      const int32_t Alen = qlen(pjq, Td) - qlen(pjq, d);
      const int32_t Bk = A.blen[2 * s + kd], Bo = A.blen[2 * s + ko];
      const int32_t Clen = To >= 0 ? qlen(pjq, To) - qlen(pjq, Ro) : 0;
      if (sg.dir) {
        int64_t o = sg.off;
        if (Alen > 0) push_seg(out, nout, cap, g.err, 1, Td, d, 0, o);
        o += Alen;
        push_seg(raw, nraw, cap, g.err, 0, 2 * s + kd, 0, 0, o);
        o += Bk;
        if (Bo > 0) push_seg(raw, nraw, cap, g.err, 0, 2 * s + ko, 0, 1, o);
        o += Bo;
        if (Clen > 0) push_seg(out, nout, cap, g.err, 1, To, Ro, 1, o);
      } else {
        int64_t o = sg.off;
        if (Clen > 0) push_seg(out, nout, cap, g.err, 1, To, Ro, 0, o);
        o += Clen;
        if (Bo > 0) push_seg(raw, nraw, cap, g.err, 0, 2 * s + ko, 0, 0, o);
        o += Bo;
        push_seg(raw, nraw, cap, g.err, 0, 2 * s + kd, 0, 1, o);
        o += Bk;
        if (Alen > 0) push_seg(out, nout, cap, g.err, 1, Td, d, 1, o);
      }
    }
  }
}
void launch_expand_segments(const ArcSide& a, const GeoState& g, const uint64_t* pjq, const Seg* in, const int32_t* nin,
                            Seg* out, int32_t* nout, Seg* raw, int32_t* nraw, int64_t max_in, int64_t cap,
                            hipStream_t st) {
  KScope ks_("expand_segments", st);
  hipLaunchKernelGGL(k_expand, dim3(grid_for(max_in, 256, 4096)), dim3(256), 0, st, a, g, pjq, in, nin, out, nout, raw,
                     nraw, cap);
  DMSC_LAUNCH_CHECK();
}

// RAW segments: copy the raw arc (forward or reversed) into the output. Short arcs (the common case)
// are copied by one thread each; arcs longer than kLongArc are deferred to a wave-per-arc pass.
constexpr int kLongArc = 48;

__global__ void k_copy_raw_short(ArcSide A, const Seg* __restrict__ raw, const int32_t* __restrict__ nraw,
                                 int32_t* __restrict__ out, int64_t out_total, int32_t* __restrict__ err,
                                 int32_t* __restrict__ longlist, int32_t* __restrict__ nlong) {
  const int n = *nraw;
  const int lane = threadIdx.x & 63;
  DMSC_WAVE_LOOP(j, (int64_t)n) {
    const int64_t jj = j_base + lane;
    bool is_long = false;
    if (jj < n) {
      const Seg sg = raw[jj];
      const int len = A.blen[sg.a];
      if (sg.off < 0 || sg.off + len > out_total) {  // inconsistent offsets: never write out of bounds
        report(err, ERR_SEG_CAP);
      } else if (len > kLongArc) {
        is_long = true;
      } else {
        const int32_t* src = A.arena + A.boff[sg.a];
        int32_t* dst = out + sg.off;
        if (sg.dir)
          for (int t = 0; t < len; ++t) dst[t] = src[t];
        else
          for (int t = 0; t < len; ++t) dst[t] = src[len - 1 - t];
      }
    }
    const int p = wave_append(nlong, is_long);
    if (is_long) longlist[p] = (int)jj;
  }
}

__global__ void k_copy_raw_long(ArcSide A, const Seg* __restrict__ raw, const int32_t* __restrict__ longlist,
                                const int32_t* __restrict__ nlong, int32_t* __restrict__ out) {
  const int n = *nlong;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t j = wave; j < n; j += nwaves) {
    const Seg sg = raw[longlist[j]];
    const int len = A.blen[sg.a];
    const int32_t* src = A.arena + A.boff[sg.a];
    int32_t* dst = out + sg.off;
    if (sg.dir)
      for (int t = lane; t < len; t += 64) dst[t] = src[t];
    else
      for (int t = lane; t < len; t += 64) dst[t] = src[len - 1 - t];
  }
}

void launch_copy_raw(const ArcSide& a, const Seg* raw, const int32_t* nraw, int64_t max_raw, int32_t* out,
                     int64_t out_total, int32_t* err, int32_t* longlist, int32_t* nlong, hipStream_t st) {
  KScope ks_("copy_raw", st);
  fill_i32(nlong, 0, 1, st);
  hipLaunchKernelGGL(k_copy_raw_short, dim3(grid_for(max_raw, 256, 16384)), dim3(256), 0, st, a, raw, nraw, out,
                     out_total, err, longlist, nlong);
  DMSC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_copy_raw_long, dim3(2048), dim3(256), 0, st, a, raw, longlist, nlong, out);
  DMSC_LAUNCH_CHECK();
}

void launch_seed_segments(int64_t ns, const int32_t* ord, const uint8_t* removed, const ArcSide& a,
                          const uint64_t* pjq, int64_t next, const int64_t* flen, const int64_t* pos_pre,
                          const int32_t* sad_to_out, Seg* segs, int32_t* nseg, Seg* raw, int32_t* nraw, int64_t cap,
                          int32_t* err, int32_t* out_offsets, int32_t* out_arc_offsets, int64_t n_out,
                          const int64_t* d_total, hipStream_t st) {
  if (ns > 0) {
    hipLaunchKernelGGL(k_seed_segments, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, ord, removed, a, pjq, next,
                       flen, pos_pre, sad_to_out, segs, nseg, raw, nraw, cap, err, out_offsets, out_arc_offsets);
    DMSC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_set_last_offset, dim3(1), dim3(1), 0, st, out_offsets, n_out, d_total);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
