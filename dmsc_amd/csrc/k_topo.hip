// Persistence (elder rule with a void root, thresholded union-find, cancellation records) as
// rounds of independent saddle applications, and the assembly of the graph outputs.
//
// Sequential reference: persistence.hxx:50-163 visits each side's saddles in order and merges the
// younger of the two extremum components into the older one. Here every round:
//   A) each pending saddle finds its two current component roots (the root of a component is its
//      oldest extremum); equal roots => a cycle (pair -1, persistence -1). Otherwise it decides
//      which root dies (the younger) and posts its order position into minkey[root] of BOTH roots
//      (64-bit atomicMin, tagged by round);
//   B) a saddle is applied when it holds minkey of the root that dies at it. No earlier pending
//      saddle touches that component, and the other component can only grow older before this
//      saddle, so the dying extremum, the persistence and the thresholded merge are exactly those
//      of the sequential sweep. Dying roots are distinct within a round.
// Posting to both roots is necessary: an earlier pending saddle may see an incomplete component
// on its other side (a merge it depends on is still pending) and so misjudge which root it kills;
// it must still block every later saddle that touches either of its current roots.
// The smallest pending saddle always holds its dying root, so every round makes progress.
//
// Both sides run together, and a saddle that merges on one side is a cycle on the other. With the
// void vertex added, the complex is a 2-sphere: the minimum graph (minima + void, saddles joining the
// ends of their two vertex V-paths) and the maximum graph (maxima, saddles joining the ends of their
// two face V-paths) are dual graphs on it, and the two sweeps visit the saddles in opposite orders.
// The merging saddles of an ascending union-find sweep form the minimum spanning forest of a graph,
// those of the descending sweep of its dual the maximum spanning forest, and in a plane graph and
// its dual these are complements (a spanning tree of one is the complement of a spanning tree of the
// other). So a pending saddle whose pair on the other side is already set resolves as a cycle here
// (pair -1, persistence -1) without a root walk or a post: in the find, and in the apply before a
// waiting saddle is appended to the next worklist. Round 0 runs side 0 first and then side 1, which
// drops every saddle side 0 merged in it. Most cycle saddles wait many rounds for the merges that
// close them; dropping them early halves the visits of a sweep, and a cycle is a no-op of the
// sequential sweep, so the decisions of the other saddles are unchanged.

#include "internal.hpp"
#include "k_topo.hpp"

namespace dmsc {

__device__ __forceinline__ int uf_find(int32_t* __restrict__ p, int x, int32_t* err = nullptr) {
  if (x < 0) return -1;
  for (int it = 0;; ++it) {
    if (it >= kLoopCap) {  // only a cyclic parent array can get here
      report(err, ERR_UF_CAP);
      return -1;
    }
    const int px = p[x];
    if (px == x) return x;
    if (px < 0) return -1;
    const int ppx = p[px];
    if (ppx == px) return px;
    p[x] = ppx;  // path halving: always an ancestor, so concurrent finds stay valid
    if (ppx < 0) return -1;
    x = ppx;
  }
}

// uf_find of two elements in lockstep: each step issues both walks' loads before using either.
// r[k] <- root of x[k] (-1: none). Halving writes only ever store ancestors, so interleaving the
// two walks changes no result.
__device__ __forceinline__ void uf_find2(int32_t* __restrict__ p, const int (&x0)[2], int (&r)[2],
                                         int32_t* err = nullptr) {
  int x[2] = {x0[0], x0[1]};
  bool done[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    r[k] = -1;
    done[k] = x[k] < 0;
  }
  for (int it = 0; !(done[0] && done[1]); ++it) {
    if (it >= kLoopCap) {  // only a cyclic parent array can get here
      report(err, ERR_UF_CAP);
#pragma unroll
      for (int k = 0; k < 2; ++k)
        if (!done[k]) r[k] = -1;
      return;
    }
    int px[2] = {0, 0}, ppx[2] = {0, 0};
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (!done[k]) px[k] = p[x[k]];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (done[k]) continue;
      if (px[k] == x[k]) {
        r[k] = x[k];
        done[k] = true;
      } else if (px[k] < 0) {
        done[k] = true;
      }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (!done[k]) ppx[k] = p[px[k]];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (done[k]) continue;
      if (ppx[k] == px[k]) {
        r[k] = px[k];
        done[k] = true;
        continue;
      }
      p[x[k]] = ppx[k];
      if (ppx[k] < 0) done[k] = true;
      else x[k] = ppx[k];
    }
  }
}

// rr[2i] = dying root (-2: cycle); rr[2i+1] = (surviving root + 1) << 1 | (dying root came from c[2i+1])
// Round 0 starts from the saddle's extrema (cc = c[i]). Later rounds start from the decision the
// saddle waited with (cc, in rr's format): while that dying root is still a root it still dies (the
// other component can only have been merged into an older root, or the void), so only the surviving
// root is looked up again and no keys are compared; otherwise both roots are found again. ps: the
// saddle's sweep position. Returns the two minkey posts (null: none).
template <bool DUAL>
__device__ __forceinline__ RootPosts persist_find_decide(const PersistSide& s, const int i, int64_t j, uint64_t tag,
                                                         bool first, const int2 cc, uint32_t ps) {
  RootPosts p;
  int ra, rb;
  bool b_dies = false, decided = false;
  if (first) {  // no merges yet, every extremum is its own root
    ra = cc.x < 0 ? -1 : cc.x;
    rb = cc.y < 0 ? -1 : cc.y;
  } else {
    const int d0 = cc.x, s0 = (cc.y >> 1) - 1;
    const bool b0 = cc.y & 1;
    int r2[2];
    uf_find2(s.par_g, {d0, s0}, r2, s.err);
    const int rd = r2[0], rs = r2[1];
    ra = b0 ? rs : rd;
    rb = b0 ? rd : rs;
    if (rd == d0 && rs != rd) {
      b_dies = b0;
      decided = true;
    }
  }
  if (!decided) {
    if (ra == rb) {
      s.pair[i] = -1;
      s.pers[i] = -1.0f;
      s.rr[2 * j] = -2;
      return p;
    }
    if (ra == -1) {
      b_dies = true;
    } else if (rb == -1) {
      b_dies = false;
    } else {
      // min side: value_greater -> r1 dies (persistence.hxx:73-80); max side: value_less (:130-137)
      const uint64_t ka = s.ekey[ra], kb = s.ekey[rb];
      const bool a_dies = s.maxside ? higher<DUAL>(key_value(kb), key_id(kb), key_value(ka), key_id(ka))
                                    : higher<DUAL>(key_value(ka), key_id(ka), key_value(kb), key_id(kb));
      b_dies = !a_dies;
    }
  }
  const int dying = b_dies ? rb : ra, surv = b_dies ? ra : rb;
  *reinterpret_cast<int2*>(s.rr + 2 * j) = make_int2(dying, ((surv + 1) << 1) | (b_dies ? 1 : 0));
  s.rkey[j] = ps;
  p.key = tag | ps;
  p.a = &s.minkey[dying];
  if (surv >= 0) p.b = &s.minkey[surv];
  return p;
}
// the saddle's position in this side's sweep order (its merge key)
__device__ __forceinline__ uint32_t sweep_pos(const PersistSide& s, int i) {
  const int32_t p = s.pos[i];
  if (!s.pos_rev) return (uint32_t)p;
  int64_t lo = 0, hi = s.wls;  // one field: every saddle
  if (s.sfield) {
    const int f = s.sfield[i];
    lo = s.sad_b[f];
    hi = s.sad_b[f + 1];
  }
  return (uint32_t)(lo + hi - 1 - p);
}

// saddle i (list position j) merged on the other side, so it is a cycle on this one
__device__ __forceinline__ void persist_prune(const PersistSide& s, const int i, int64_t j) {
  s.pers[i] = -1.0f;  // pair[i] is still -1 from the start state
  s.rr[2 * j] = -2;
}

template <bool DUAL>
__device__ __forceinline__ RootPosts persist_find_one(const PersistSide& s, const int i, int64_t j, int cur, uint64_t tag,
                                                      bool first) {
  if (first) return persist_find_decide<DUAL>(s, i, j, tag, first, *reinterpret_cast<const int2*>(s.c + 2 * i),
                                              sweep_pos(s, i));
  if (s.pair_other[i] >= 0) {
    persist_prune(s, i, j);
    return RootPosts{};
  }
  return persist_find_decide<DUAL>(s, i, j, tag, first, s.roots(cur)[j], s.keys(cur)[j]);
}

// One side's list, ITEMS items per lane: a wave takes 64 * ITEMS consecutive items, item u of a lane
// at base + 64 u + lane, and issues each stage's loads (list entries; extrema or cached roots and
// sweep positions) for all its items before the next, so that their latencies overlap. Worklists
// keep saddle-id order, so the lanes of a wave hold nearby saddles that often share an extremum:
// their posts are combined per wave first; from round 1 on the posting lane also reads before its
// atomic (hot roots of big components).
template <bool DUAL, int ITEMS>
__device__ __forceinline__ void persist_find_list(const PersistSide& s, int cur, int64_t n, uint64_t tag, bool first,
                                                  bool prune) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int32_t* const in = s.list(cur);
  for (int64_t base = wave * 64 * ITEMS; base < n; base += nwaves * 64 * ITEMS) {
    int i[ITEMS];
    int2 cc[ITEMS];
    uint32_t ps[ITEMS];
    bool pruned[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      const int64_t j = base + 64 * u + lane;
      i[u] = j < n ? (first ? (int)j : in[j]) : -1;  // list 0 holds every saddle in id order
    }
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      cc[u] = make_int2(-1, -1);
      ps[u] = 0;
      pruned[u] = false;
      if (i[u] >= 0) {
        if (first) {  // nothing has merged yet on this side (the other side's round 0 may have run)
          cc[u] = *reinterpret_cast<const int2*>(s.c + 2 * i[u]);
          ps[u] = sweep_pos(s, i[u]);
          pruned[u] = prune && s.pair_other[i[u]] >= 0;
        } else {  // by list position: written where the apply appended the waiting saddle
          const int64_t j = base + 64 * u + lane;
          cc[u] = s.roots(cur)[j];
          ps[u] = s.keys(cur)[j];
          pruned[u] = s.pair_other[i[u]] >= 0;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      RootPosts p;
      if (i[u] >= 0) {
        if (pruned[u]) persist_prune(s, i[u], base + 64 * u + lane);
        else p = persist_find_decide<DUAL>(s, i[u], base + 64 * u + lane, tag, first, cc[u], ps[u]);
      }
      wave_post_min(p.a, p.key, !first);
      wave_post_min(p.b, p.key, !first);
    }
  }
}

// sides: bit 0 / bit 1 = side 0 / side 1 take part in this launch; bit 2 = the round is round 0 of
// side 1 run after side 0's round 0, so side 1 already drops the saddles side 0 merged
template <bool DUAL>
__device__ __forceinline__ void persist_find_phase(const PersistSide& s0, const PersistSide& s1, int round, int sides) {
  const int cur = round & 1;
  const int n0 = (sides & 1) ? s0.cnt[cur] : 0, n1 = (sides & 2) ? s1.cnt[cur] : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (sides & 1) s0.cnt[cur ^ 1] = 0;
    if (sides & 2) s1.cnt[cur ^ 1] = 0;
  }
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  const bool first = round == 0, prune = !first || (sides & 4);
  // the round's decisions go to rr at the item's list position
  if ((int64_t)n0 + n1 > (1 << 16)) {
    persist_find_list<DUAL, 4>(s0, cur, n0, tag, first, prune);
    persist_find_list<DUAL, 4>(s1, cur, n1, tag, first, prune);
  } else {
    persist_find_list<DUAL, 1>(s0, cur, n0, tag, first, prune);
    persist_find_list<DUAL, 1>(s1, cur, n1, tag, first, prune);
  }
}
template <bool DUAL>
__global__ void k_persist_find(PersistSide s0, PersistSide s1, int round, const int32_t* __restrict__ d_round,
                               int sides) {
  if (d_round) round = *d_round;  // graph replays take the round number from the device
  persist_find_phase<DUAL>(s0, s1, round, sides);
}

// Applies saddle i, which holds minkey of its dying root (r = its rr pair; sv its value, dv the dying
// extremum's, thr its field's threshold).
template <bool DUAL>
__device__ __forceinline__ void persist_apply_values(const PersistSide& s, const int i, const int2 r, float sv, float dv,
                                                     float thr, uint8_t* __restrict__ removed) {
  const int dying = r.x;
  const int surv = (r.y >> 1) - 1;
  const bool b_dies = r.y & 1;
  float p;
  if (s.maxside) p = DUAL ? sv - dv : dv - sv;
  else p = DUAL ? dv - sv : sv - dv;
  s.pers[i] = p;
  s.pair[i] = dying;
  s.par_g[dying] = surv;
  if (p <= thr) {
    s.alive[dying] = 0;
    removed[i] = 1;
    const int2 cc = *reinterpret_cast<const int2*>(s.c + 2 * i);
    const int a = cc.x, b = cc.y;
    int r2[2];
    uf_find2(s.par_t, {b_dies ? b : a, b_dies ? a : b}, r2, s.err);
    const int d = r2[0], t = r2[1];
    if (d != t && d != -1) s.par_t[d] = t;
    if (s.cancel_d) s.cancel_d[i] = d;
  }
}
template <bool DUAL>
__device__ __forceinline__ void persist_apply_winner(const PersistSide& s, const int i, const int2 r,
                                                     uint8_t* __restrict__ removed) {
  persist_apply_values<DUAL>(s, i, r, s.sval[i], key_value(s.ekey[r.x]), s.thr[s.sfield ? s.sfield[i] : 0], removed);
}

// saddle i, waiting with decision r and sweep position ps, takes slot o of list k: the next find
// resumes from its roots
__device__ __forceinline__ void persist_park(const PersistSide& s, int k, int o, const int i, const int2 r,
                                             uint32_t ps) {
  s.list(k)[o] = i;
  s.roots(k)[o] = r;  // the decision it waits with
  s.keys(k)[o] = ps;
}

// Decides one pending saddle i; returns true when it must wait for a later round.
template <bool DUAL>
__device__ __forceinline__ bool persist_apply_one(const PersistSide& s, const int i, int64_t j, uint64_t tag,
                                                  uint8_t* __restrict__ removed) {
  const int2 r = *reinterpret_cast<const int2*>(s.rr + 2 * j);
  if (r.x == -2) return false;
  if (s.minkey[r.x] != (tag | s.rkey[j])) return true;
  persist_apply_winner<DUAL>(s, i, r, removed);
  return false;
}

// One chunk of a side's worklist: kApplyItems consecutive items per thread; waiting items are
// appended to the next worklist with one block-wide reservation (order preserved within the chunk).
// The loads of the thread's items are issued stage by stage (list entries and decisions, then the
// holder tests) so that their latencies overlap. Small worklists (the tail of the sweep) are
// latency-bound: one item per thread there.
constexpr int kApplyItems = 4, kApplyChunk = 256 * kApplyItems;

template <bool DUAL, int ITEMS>
__device__ __forceinline__ void persist_apply_chunk(const PersistSide& s, int cur, int64_t start, int n, uint64_t tag,
                                                    uint8_t* __restrict__ removed, bool prune) {
  const int64_t j0 = start + (int64_t)threadIdx.x * ITEMS;
  const int32_t* const in = s.list(cur);
  int k[ITEMS];
  int2 r[ITEMS];
  uint64_t mk[ITEMS];
  uint32_t ps[ITEMS];
  if (ITEMS == 4 && j0 + 4 <= n) {
    // a whole run: 16-byte loads (rr and rkey are indexed by list position, so j0 % 4 == 0 aligns
    // them; the list itself sits at an offset of the worklist buffer)
    const int4 kv = (reinterpret_cast<uintptr_t>(in + j0) & 15) == 0
                        ? *reinterpret_cast<const int4*>(in + j0)
                        : make_int4(in[j0], in[j0 + 1], in[j0 + 2], in[j0 + 3]);
    const int4 ra = *reinterpret_cast<const int4*>(s.rr + 2 * j0);
    const int4 rb = *reinterpret_cast<const int4*>(s.rr + 2 * j0 + 4);
    const uint4 pv = *reinterpret_cast<const uint4*>(s.rkey + j0);
    const int kk[4] = {kv.x, kv.y, kv.z, kv.w};
    const int2 rv[4] = {make_int2(ra.x, ra.y), make_int2(ra.z, ra.w), make_int2(rb.x, rb.y), make_int2(rb.z, rb.w)};
    const uint32_t pp[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      k[u] = kk[u & 3];
      r[u] = rv[u & 3];
      ps[u] = pp[u & 3];
    }
  } else {
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      const bool v = j0 + u < n;
      k[u] = v ? in[j0 + u] : 0;
      r[u] = v ? *reinterpret_cast<const int2*>(s.rr + 2 * (j0 + u)) : make_int2(-3, 0);  // -3: past the list
      ps[u] = v ? s.rkey[j0 + u] : 0u;
    }
  }
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    mk[u] = 0;
    if (r[u].x >= 0) mk[u] = s.minkey[r[u].x];
  }
  uint32_t waitmask = 0, winmask = 0;
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if (r[u].x < 0) continue;  // a cycle (decided in the find) or past the list
    if (mk[u] != (tag | ps[u])) waitmask |= 1u << u;
    else winmask |= 1u << u;
  }
  // a waiting saddle that the other side has merged by now is a cycle here: it leaves the worklist
  // (a merge made elsewhere in this launch may not be seen yet; the next find catches that one)
  int32_t po[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) po[u] = prune && (waitmask >> u & 1) ? s.pair_other[k[u]] : -1;
  int32_t nwait = 0;
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    if (po[u] >= 0) {
      s.pers[k[u]] = -1.0f;
      waitmask &= ~(1u << u);
    }
    nwait += waitmask >> u & 1;
  }
  // winners: their value loads for all items first, then the decisions
  float sv[ITEMS], dv[ITEMS], th[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u) {
    sv[u] = dv[u] = th[u] = 0.f;
    if (winmask >> u & 1) {
      sv[u] = s.sval[k[u]];
      dv[u] = key_value(s.ekey[r[u].x]);
      th[u] = s.thr[s.sfield ? s.sfield[k[u]] : 0];
    }
  }
#pragma unroll
  for (int u = 0; u < ITEMS; ++u)
    if (winmask >> u & 1) persist_apply_values<DUAL>(s, k[u], r[u], sv[u], dv[u], th[u], removed);
  int32_t* const cnt[1] = {&s.cnt[cur ^ 1]};
  const int32_t req[1] = {nwait};
  int32_t first[1];
  block_reserve<1>(cnt, req, first);
  int32_t o = first[0];
#pragma unroll
  for (int u = 0; u < ITEMS; ++u)
    if (waitmask >> u & 1) persist_park(s, cur ^ 1, o++, k[u], r[u], ps[u]);
}

// Chunks of side 0 come first, then those of side 1, so every chunk (and block) has one side.
template <bool DUAL>
__device__ __forceinline__ void persist_apply_phase(const PersistSide& s0, const PersistSide& s1, int round,
                                                    uint8_t* __restrict__ removed, int small_list, int sides) {
  const int cur = round & 1;
  const int n0 = (sides & 1) ? s0.cnt[cur] : 0, n1 = (sides & 2) ? s1.cnt[cur] : 0;
  const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)round) << 32;
  const bool prune = round > 0 || (sides & 4);  // side 0's round 0 runs before anything merged on side 1
  if ((int64_t)n0 + n1 <= small_list) {
    const int64_t nc0 = ceil_div(n0, 256), nc = nc0 + ceil_div(n1, 256);
    for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
      if (c >= nc0) persist_apply_chunk<DUAL, 1>(s1, cur, (c - nc0) * 256, n1, tag, removed, prune);
      else persist_apply_chunk<DUAL, 1>(s0, cur, c * 256, n0, tag, removed, prune);
    }
    return;
  }
  const int64_t nc0 = ceil_div(n0, kApplyChunk), nc = nc0 + ceil_div(n1, kApplyChunk);
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    if (c >= nc0) persist_apply_chunk<DUAL, kApplyItems>(s1, cur, (c - nc0) * kApplyChunk, n1, tag, removed, prune);
    else persist_apply_chunk<DUAL, kApplyItems>(s0, cur, c * kApplyChunk, n0, tag, removed, prune);
  }
}
template <bool DUAL>
__global__ __launch_bounds__(256) void k_persist_apply(PersistSide s0, PersistSide s1, int round,
                                                       const int32_t* __restrict__ d_round,
                                                       uint8_t* __restrict__ removed, int small_list, int sides) {
  if (d_round) round = *d_round;
  persist_apply_phase<DUAL>(s0, s1, round, removed, small_list, sides);
}

// Start state of both sides' rounds in one launch: every extremum its own root (both union-finds),
// alive, no post; every saddle unpaired, not cancelled, not removed, in list 0.
__global__ void k_persist_init(PersistSide s0, PersistSide s1, int64_t n0, int64_t n1, int64_t ns,
                               int32_t* __restrict__ cancel0, int32_t* __restrict__ cancel1,
                               uint8_t* __restrict__ removed) {
  const int64_t m = max(max(n0, n1), ns);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
    if (e < n0) {
      s0.par_g[e] = (int32_t)e;
      s0.par_t[e] = (int32_t)e;
      s0.alive[e] = 1;
      s0.minkey[e] = ~0ull;
    }
    if (e < n1) {
      s1.par_g[e] = (int32_t)e;
      s1.par_t[e] = (int32_t)e;
      s1.alive[e] = 1;
      s1.minkey[e] = ~0ull;
    }
    if (e < ns) {
      s0.pair[e] = -1;
      s1.pair[e] = -1;
      cancel0[e] = -1;
      cancel1[e] = -1;
      s0.list(0)[e] = (int32_t)e;
      s1.list(0)[e] = (int32_t)e;
      removed[e] = 0;
    }
    if (e == 0) {
      s0.cnt[0] = (int32_t)ns;
      s0.cnt[1] = 0;
      s1.cnt[0] = (int32_t)ns;
      s1.cnt[1] = 0;
    }
  }
}
void launch_persist_init(const PersistSide& s0, const PersistSide& s1, int64_t n0, int64_t n1, int64_t ns,
                         int32_t* cancel0, int32_t* cancel1, uint8_t* removed, hipStream_t st) {
  const int64_t m = std::max(std::max(n0, n1), std::max<int64_t>(ns, 1));
  hipLaunchKernelGGL(k_persist_init, dim3(grid_for(m, 256, 16384)), dim3(256), 0, st, s0, s1, n0, n1, ns, cancel0,
                     cancel1, removed);
  DMSC_LAUNCH_CHECK();
}

// one item of the tail kernel's apply phase; waiting saddles take next-list slots from an LDS counter
template <bool DUAL>
__device__ __forceinline__ void persist_tail_apply(const PersistSide& s, int j, int cur, uint64_t tag,
                                                   uint8_t* __restrict__ removed, int32_t* s_next) {
  const int i = s.list(cur)[j];
  if (persist_apply_one<DUAL>(s, i, j, tag, removed))
    persist_park(s, cur ^ 1, atomicAdd(s_next, 1), i, *reinterpret_cast<const int2*>(s.rr + 2 * j), s.rkey[j]);
}

// The tail of the sweep in one workgroup: once few saddles are pending, a round would cost two
// launches for a few hundred items. One block runs find, barrier, apply, barrier per round until both
// worklists drain or max_rounds have run, then leaves the pending counts in cnt[r & 1] and the round
// reached in *round_out. Waiting saddles take next-list slots from an LDS counter (the list order
// affects no decision: every decision compares posted keys). Loop bounds are block-uniform (LDS).
template <bool DUAL>
__global__ __launch_bounds__(kTailThreads) void k_persist_tail(PersistSide s0, PersistSide s1, int round0,
                                                               int max_rounds, uint8_t* __restrict__ removed,
                                                               int32_t* __restrict__ round_out) {
  __shared__ int32_t s_next[2];
  int n0 = __hip_atomic_load(&s0.cnt[round0 & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int n1 = __hip_atomic_load(&s1.cnt[round0 & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int r = round0;
  for (; r < round0 + max_rounds && n0 + n1 > 0; ++r) {
    const int cur = r & 1;
    const uint64_t tag = (uint64_t)(0xFFFFFFFFu - (uint32_t)r) << 32;
    if (threadIdx.x == 0) s_next[0] = s_next[1] = 0;
    for (int j = threadIdx.x; j < n0 + n1; j += kTailThreads) {
      const RootPosts p = j >= n0 ? persist_find_one<DUAL>(s1, s1.list(cur)[j - n0], j - n0, cur, tag, r == 0)
                                  : persist_find_one<DUAL>(s0, s0.list(cur)[j], j, cur, tag, r == 0);
      if (p.a) post_min(p.a, p.key, r > 0);
      if (p.b) post_min(p.b, p.key, r > 0);
    }
    phase_barrier();
    for (int j = threadIdx.x; j < n0 + n1; j += kTailThreads) {
      if (j >= n0) persist_tail_apply<DUAL>(s1, j - n0, cur, tag, removed, &s_next[1]);
      else persist_tail_apply<DUAL>(s0, j, cur, tag, removed, &s_next[0]);
    }
    phase_barrier();
    n0 = s_next[0];
    n1 = s_next[1];
    __syncthreads();  // every thread has read the counts before the next round clears them
  }
  if (threadIdx.x == 0) {
    s0.cnt[r & 1] = n0;
    s1.cnt[r & 1] = n1;
    *round_out = r;
  }
}

void launch_persist_tail(bool dual, const PersistSide& s0, const PersistSide& s1, int round0, int max_rounds,
                         uint8_t* removed, int32_t* round_out, hipStream_t st) {
  KScope ks_("persist_round", st);
  if (dual)
    hipLaunchKernelGGL(k_persist_tail<true>, dim3(1), dim3(kTailThreads), 0, st, s0, s1, round0, max_rounds, removed,
                       round_out);
  else
    hipLaunchKernelGGL(k_persist_tail<false>, dim3(1), dim3(kTailThreads), 0, st, s0, s1, round0, max_rounds, removed,
                       round_out);
  DMSC_LAUNCH_CHECK();
}

__global__ void k_round_inc(int32_t* d_round) { *d_round += 1; }
void launch_round_inc(int32_t* d_round, hipStream_t st) {
  hipLaunchKernelGGL(k_round_inc, dim3(1), dim3(1), 0, st, d_round);
  DMSC_LAUNCH_CHECK();
}

// one find and one apply launch over the sides in `sides` (persist_find_phase's mask)
static void persist_round_launches(bool dual, const PersistSide& s0, const PersistSide& s1, int round,
                                   uint8_t* removed, int64_t max_items, hipStream_t st, const int32_t* d_round,
                                   int sides) {
  const Tunables& tu = tune();
  const int g = grid_for(max_items, 256, tu.round_grid);
  if (dual) hipLaunchKernelGGL(k_persist_find<true>, dim3(g), dim3(256), 0, st, s0, s1, round, d_round, sides);
  else hipLaunchKernelGGL(k_persist_find<false>, dim3(g), dim3(256), 0, st, s0, s1, round, d_round, sides);
  DMSC_LAUNCH_CHECK();
  const int ga = grid_for(max_items, max_items <= tu.small_list ? 256 : kApplyChunk, tu.round_grid);
  if (dual)
    hipLaunchKernelGGL(k_persist_apply<true>, dim3(ga), dim3(256), 0, st, s0, s1, round, d_round, removed,
                       tu.small_list, sides);
  else
    hipLaunchKernelGGL(k_persist_apply<false>, dim3(ga), dim3(256), 0, st, s0, s1, round, d_round, removed,
                       tu.small_list, sides);
  DMSC_LAUNCH_CHECK();
}

// Round 0 visits every saddle on both sides. Run side 0 (the vertex side, which has the more extrema
// and so the more merges) first: side 1's round 0 then drops every saddle side 0 merged, with one
// coalesced read and two writes instead of a visit. The first tune().side_rounds rounds run side by
// side this way (the next round of side 0 drops what side 1 merged, and so on); later rounds run
// both sides in each launch, half the launches for the short worklists of the tail.
void launch_persist_round(bool dual, const PersistSide& s0, const PersistSide& s1, int round, uint8_t* removed,
                          int64_t max_items, hipStream_t st, const int32_t* d_round) {
  KScope ks_("persist_round", st);
  if (round < tune().side_rounds && !d_round) {
    persist_round_launches(dual, s0, s1, round, removed, max_items, st, nullptr, 1);
    persist_round_launches(dual, s0, s1, round, removed, max_items, st, nullptr, 2 | 4);
    return;
  }
  persist_round_launches(dual, s0, s1, round, removed, max_items, st, d_round, 3);
}

// ------------------------------------------------------------------ assembly (cpu/workspace.hxx:86-365)
// root_out[x] <- output index of x's thresholded root (the root of x in par_t, as the reference's
// flattening finds it, cell_groups.hxx:128-146; -1: no root or the root is not output). Output indices of live extrema are distinct, so two extrema share a root output
// exactly when they share a live root; every consumer below compares these instead of the roots.
// The same pass masks the prefix into to_out (-1 for dead extrema, in place): a live root's entry
// reads the same before and after its own thread masks, and a dead one is never read.
__global__ void k_root_out(int32_t* __restrict__ par, const uint8_t* __restrict__ alive, int32_t* to_out,
                           int32_t* __restrict__ root_out, int64_t n) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const int r = uf_find(par, (int)x);
    root_out[x] = (r >= 0 && alive[r]) ? to_out[r] : -1;
    if (!alive[x]) to_out[x] = -1;
  }
}
void launch_root_out(int32_t* par, const uint8_t* alive, int32_t* to_out, int32_t* root_out, int64_t n,
                     hipStream_t st) {
  KScope ks_("root_out", st);
  if (n == 0) return;
  hipLaunchKernelGGL(k_root_out, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, par, alive, to_out, root_out, n);
  DMSC_LAUNCH_CHECK();
}

// Per position k of one side's order (sequential writes; the per-saddle reads are gathers: visiting
// saddles in id order and scattering the counts measured slower except on heavy-tie noise): survival
// (min order), edge rows, and the rows' extremum outputs edges[k] = (first, second) (-1: no row; the
// second is dropped when it repeats the first).
__global__ void k_side_counts(int64_t ns, const int32_t* __restrict__ ord, const uint8_t* __restrict__ removed,
                              const int32_t* __restrict__ c, const int32_t* __restrict__ root_out,
                              int32_t* __restrict__ surv, int32_t* __restrict__ rows, int2* __restrict__ edges) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ns; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = ord[k];
    const bool alive = !removed[i];
    if (surv) surv[k] = alive ? 1 : 0;
    int o1 = -1, o2 = -1;
    if (alive) {
      const int2 ab = *reinterpret_cast<const int2*>(c + 2 * i);
      o1 = ab.x >= 0 ? root_out[ab.x] : -1;
      o2 = ab.y >= 0 ? root_out[ab.y] : -1;
      if (o2 == o1) o2 = -1;
    }
    rows[k] = (o1 >= 0 ? 1 : 0) + (o2 >= 0 ? 1 : 0);
    edges[k] = make_int2(o1, o2);
  }
}
void launch_side_counts(int64_t ns, const int32_t* ord, const uint8_t* removed, const int32_t* c,
                        const int32_t* root_out, int32_t* surv, int32_t* rows, int2* edges, hipStream_t st) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_side_counts, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, ord, removed, c, root_out,
                     surv, rows, edges);
  DMSC_LAUNCH_CHECK();
}

// Surviving saddles of one side's order as a list: list[r] = saddle, row_list[r] = its first edge row,
// r = its rank among the survivors in that order (for the min order r is the output saddle index).
// out_list (the max order only): out_list[r] = the saddle's output index. Within a field the max order
// is the min order reversed, so position k mirrors position lo + hi - 1 - k of the min order, whose
// survivor prefix min_pre holds that index (survival does not depend on the order).
__global__ void k_survivors(int64_t ns, const int32_t* __restrict__ ord, const int32_t* __restrict__ surv,
                            const int32_t* __restrict__ surv_pre, const int32_t* __restrict__ row_pre,
                            const int2* __restrict__ edges, int32_t* __restrict__ list,
                            int32_t* __restrict__ row_list, int2* __restrict__ edge_list,
                            const int32_t* __restrict__ min_pre, const int32_t* __restrict__ sad_b, int B,
                            int32_t* __restrict__ out_list) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ns; k += (int64_t)gridDim.x * blockDim.x) {
    if (!surv[k]) continue;
    const int r = surv_pre[k];
    list[r] = ord[k];
    row_list[r] = row_pre[k];
    edge_list[r] = edges[k];
    if (out_list) {
      int64_t lo = 0, hi = ns;
      if (B > 1) {
        const int f = field_of64(sad_b, B, k);
        lo = sad_b[f];
        hi = sad_b[f + 1];
      }
      out_list[r] = min_pre[lo + hi - 1 - k];
    }
  }
}
void launch_survivors(int64_t ns, const int32_t* ord, const int32_t* surv, const int32_t* surv_pre,
                      const int32_t* row_pre, const int2* edges, int32_t* list, int32_t* row_list, int2* edge_list,
                      const int32_t* min_pre, const int32_t* sad_b, int B, int32_t* out_list, hipStream_t st) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_survivors, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, ord, surv, surv_pre, row_pre,
                     edges, list, row_list, edge_list, min_pre, sad_b, B, out_list);
  DMSC_LAUNCH_CHECK();
}

// extremum lists -> output points
__global__ void k_emit_points(int64_t n, const int32_t* __restrict__ ids, const int32_t* __restrict__ to_out,
                              int32_t* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int o = to_out[e];
    if (o >= 0) out[o] = ids[e];
  }
}
void launch_emit_points(int64_t n, const int32_t* ids, const int32_t* to_out, int32_t* out, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_emit_points, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, n, ids, to_out, out);
  DMSC_LAUNCH_CHECK();
}

// One side's per-saddle outputs, visiting that side's order: persistence, paired extremum (through
// the thresholded root), and edge rows (saddle output index, extremum output index).
__global__ void k_emit_side(int64_t ns, const int32_t* __restrict__ ord, const uint8_t* __restrict__ removed,
                            const int32_t* __restrict__ out_list, const int2* __restrict__ edge_list,
                            const int32_t* __restrict__ root_out, const float* __restrict__ pers,
                            const int32_t* __restrict__ pair,
                            const int32_t* __restrict__ row_pre, const int32_t* __restrict__ sad_ids, EmitBases eb,
                            float* __restrict__ out_p, int32_t* __restrict__ out_pp, int32_t* __restrict__ out_e,
                            int32_t* __restrict__ out_sad) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ns; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = ord[k];
    if (removed && removed[i]) continue;
    int sbase = 0, ebase = 0;
    if (eb.sfield) {
      const int fi = eb.sfield[i];
      sbase = eb.sad_out[fi];
      ebase = eb.ext_out[fi];
    }
    const int so = out_list ? out_list[k] : (int)k;  // no list: ord is the output order
    if (out_sad) out_sad[so] = sad_ids[i];
    out_p[so] = pers[i];
    const int pe = pair[i];
    const int po = pe >= 0 ? root_out[pe] : -1;
    out_pp[so] = po >= 0 ? po - ebase : -1;
    const int2 o = edge_list[k];
    int row = row_pre[k];
    if (o.x >= 0) {
      out_e[2 * (int64_t)row] = so - sbase;
      out_e[2 * (int64_t)row + 1] = o.x - ebase;
      ++row;
    }
    if (o.y >= 0) {
      out_e[2 * (int64_t)row] = so - sbase;
      out_e[2 * (int64_t)row + 1] = o.y - ebase;
    }
  }
}
void launch_emit_side(int64_t ns, const int32_t* ord, const uint8_t* removed, const int32_t* out_list,
                      const int2* edge_list, const int32_t* root_out, const float* pers, const int32_t* pair,
                      const int32_t* row_pre, const int32_t* sad_ids, const EmitBases& eb, float* out_p, int32_t* out_pp,
                      int32_t* out_e, int32_t* out_sad, hipStream_t st) {
  KScope ks_("emit_side", st);
  if (ns == 0) return;
  hipLaunchKernelGGL(k_emit_side, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, ord, removed, out_list, edge_list,
                     root_out, pers, pair, row_pre, sad_ids, eb, out_p, out_pp, out_e, out_sad);
  DMSC_LAUNCH_CHECK();
}

// Both sweep orders (ascending / descending within each field) and their inverses (saddle index ->
// position). `asc` is grouped by field, ascending (value, id) inside each group.
// The orders are grouped by field, so position k belongs to the field whose position range holds
// it: a binary search over the B + 1 field bases (cached) instead of a gather by saddle.
// Consecutive positions of a block share a field (orders are grouped by field): the block searches
// the field bases once for its first position and each thread steps forward from there.
__global__ void k_orders(int64_t ns, const int32_t* __restrict__ asc, int B, const int32_t* __restrict__ sad_b,
                         int32_t* __restrict__ fwd, int32_t* __restrict__ rev, int32_t* __restrict__ pos_fwd,
                         int32_t* __restrict__ pos_rev) {
  for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x; k0 < ns; k0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = k0 + threadIdx.x;
    const int fb = B > 1 ? field_of64(sad_b, B, k0) : 0;
    if (k >= ns) continue;
    const int v = asc[k];
    int64_t lo = 0, hi = ns;
    if (B > 1) {
      const int fi = field_from(sad_b, B, fb, k);
      lo = sad_b[fi];
      hi = sad_b[fi + 1];
    }
    const int64_t r = lo + hi - 1 - k;
    if (fwd) fwd[k] = v;
    rev[r] = v;
    pos_fwd[v] = (int)k;
    if (pos_rev) pos_rev[v] = (int)r;
  }
}
void launch_orders(int64_t ns, const int32_t* asc, int B, const int32_t* sad_b, int32_t* fwd, int32_t* rev,
                   int32_t* pos_fwd, int32_t* pos_rev, hipStream_t st) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_orders, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, st, ns, asc, B, sad_b, fwd, rev,
                     pos_fwd, pos_rev);
  DMSC_LAUNCH_CHECK();
}

// keys[k] = field of saddle idx[k] (the second, field-major pass of the saddle sort). Saddles are
// numbered field by field, so that is a binary search over the B + 1 saddle bases (in LDS when they
// fit) rather than a gather from a per-saddle array.
constexpr int kLdsBases = 4096;
__global__ void k_gather_field(int64_t n, const int32_t* __restrict__ idx, int B, const int32_t* __restrict__ sad_b,
                               uint32_t* __restrict__ keys) {
  __shared__ int32_t lb[kLdsBases];
  const bool in_lds = B < kLdsBases;
  if (in_lds) {
    for (int i = threadIdx.x; i <= B; i += blockDim.x) lb[i] = sad_b[i];
    __syncthreads();
  }
  const int32_t* const base = in_lds ? lb : sad_b;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    keys[k] = (uint32_t)field_of64(base, B, idx[k]);
}
void launch_gather_field(int64_t n, const int32_t* idx, int B, const int32_t* sad_b, uint32_t* keys, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_field, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, idx, B, sad_b, keys);
  DMSC_LAUNCH_CHECK();
}

__global__ void k_sort_keys(int64_t n, const float* __restrict__ v, uint32_t* __restrict__ keys) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    keys[i] = sortable_key(v[i]);
}
void launch_sort_keys(int64_t n, const float* v, uint32_t* keys, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_sort_keys, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, n, v, keys);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
