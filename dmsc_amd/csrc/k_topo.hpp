// Launchers for persistence, assembly and arc geometry stages.
#pragma once

#include "internal.hpp"

namespace dmsc {

// One side (vertex-minima or face-maxima) of the persistence sweep.
struct PersistSide {
  const int32_t* pos;   // saddle index -> position in the ascending-value order
  const int32_t* sad_b; // [B+1] first position of each field in the orders
  int pos_rev;          // this side sweeps in descending order: its position is the mirror of pos within
                        // the saddle's field range (the merge key)
  const int32_t* c;     // per saddle: c[2i], c[2i+1] = extremum indices of its two V-paths (-1 = none)
  const float* sval;    // per saddle value
  const uint64_t* ekey; // per extremum pack_key(value, cell id)
  int32_t* par_g;       // global union-find parents (-1 = joined the void)
  int32_t* par_t;       // thresholded union-find parents
  uint8_t* alive;       // per extremum
  uint64_t* minkey;     // per extremum: (round tag << 32 | order position), atomic min
  int32_t* pair;        // per saddle: dying extremum index, -1 for a cycle
  const int32_t* pair_other;  // the other side's pair array: a saddle that merged there (>= 0) is a
                              // cycle on this side
  float* pers;          // per saddle persistence (-1 for a cycle)
  int32_t* cancel_d;    // per saddle thresholded dead extremum (null when arcs are not traced)
  int32_t* rr;          // per worklist position: dying root (-2 = cycle), (surviving root + 1) << 1 | side bit
  int32_t* croot;       // per position of each worklist (2 ints, list k at croot + 2 k wls): the decision the
                        // saddle waited with, in rr's format (the next find resumes there)
  uint32_t* wkey;       // per position of each worklist (list k at wkey + k wls): the saddle's sweep position
  uint32_t* rkey;       // per worklist position: the saddle's sweep position (its merge key), from the find
  int32_t* wl;          // ping-pong worklists of pending saddle indices: list k at wl + k * wls
  int64_t wls;
  // (one buffer and an offset rather than an array of two pointers: a run-time index into a struct
  // member array makes the compiler keep the kernel argument in scratch memory)
  __host__ __device__ int32_t* list(int k) const { return wl + (int64_t)k * wls; }
  __host__ __device__ int2* roots(int k) const { return reinterpret_cast<int2*>(croot) + (int64_t)k * wls; }
  __host__ __device__ uint32_t* keys(int k) const { return wkey + (int64_t)k * wls; }
  int32_t* cnt;         // [2] worklist sizes (device)
  int32_t* err;         // device error flags
  const float* thr;     // per field persistence threshold
  const int32_t* sfield;  // per saddle field index (null: a single field)
  int maxside;
};

// start state of both sides' persistence rounds (one launch)
void launch_persist_init(const PersistSide& s0, const PersistSide& s1, int64_t n0, int64_t n1, int64_t ns,
                         int32_t* cancel0, int32_t* cancel1, uint8_t* removed, hipStream_t st);
// d_round (optional): the kernels read the round number from there instead of `round` (graph replay)
void launch_persist_round(bool dual, const PersistSide& s0, const PersistSide& s1, int round, uint8_t* removed,
                          int64_t max_items, hipStream_t st, const int32_t* d_round = nullptr);
void launch_round_inc(int32_t* d_round, hipStream_t st);
// Every remaining round in one single-workgroup launch (for short worklists); pending counts are
// left in cnt[r & 1] of each side and the round reached in *round_out (device).
void launch_persist_tail(bool dual, const PersistSide& s0, const PersistSide& s1, int round0, int max_rounds,
                         uint8_t* removed, int32_t* round_out, hipStream_t st);

// to_out: exclusive prefix of the alive flags, masked to -1 for dead extrema by this pass;
// root_out[x] <- output index of the thresholded root of extremum x (its root in par), -1 when none
void launch_root_out(int32_t* par, const uint8_t* alive, int32_t* to_out, int32_t* root_out, int64_t n,
                     hipStream_t st);
// per position: survival, edge rows and their extremum outputs (edges: int2 per position)
void launch_side_counts(int64_t ns, const int32_t* ord, const uint8_t* removed, const int32_t* c,
                        const int32_t* root_out, int32_t* surv, int32_t* rows, int2* edges, hipStream_t st);
// Survivors of one order as lists by survivor rank (saddle, first edge row, edge outputs); for the max
// order also out_list[r] = output index, read from the min order's survivor prefix min_pre at the
// mirrored position (sad_b: [B+1] field bases of the orders). min_pre, out_list: null for the min order.
void launch_survivors(int64_t ns, const int32_t* ord, const int32_t* surv, const int32_t* surv_pre,
                      const int32_t* row_pre, const int2* edges, int32_t* list, int32_t* row_list, int2* edge_list,
                      const int32_t* min_pre, const int32_t* sad_b, int B, int32_t* out_list, hipStream_t st);
void launch_emit_points(int64_t n, const int32_t* ids, const int32_t* to_out, int32_t* out, hipStream_t st);
// ord / removed: a side order over all saddles with their removal flags, or a survivor list with
// removed == null (row_pre indexed like ord). Output indices are batch-wide positions; index values written into the outputs are made
// field-local with the per-field bases (sfield null: a single field, bases unused).
struct EmitBases {
  const int32_t* sfield;    // per saddle field
  const int32_t* sad_out;   // [B+1] first output saddle of each field
  const int32_t* ext_out;   // [B+1] first output extremum (this side) of each field
};
// out_list[k]: output index of the saddle at position k of ord (nullptr: position k is output saddle k)
// edge_list[k]: the edge rows' extremum outputs of the saddle at position k (from the side counts)
void launch_emit_side(int64_t ns, const int32_t* ord, const uint8_t* removed, const int32_t* out_list,
                      const int2* edge_list, const int32_t* root_out, const float* pers, const int32_t* pair,
                      const int32_t* row_pre, const int32_t* sad_ids, const EmitBases& eb, float* out_p, int32_t* out_pp,
                      int32_t* out_e, int32_t* out_sad, hipStream_t st);
// Sweep orders from the ascending order `asc`, grouped by field (sad_b: [B+1] saddle bases):
// fwd = asc, rev = asc reversed within each field's range, and their inverses.
void launch_orders(int64_t ns, const int32_t* asc, int B, const int32_t* sad_b, int32_t* fwd, int32_t* rev,
                   int32_t* pos_fwd, int32_t* pos_rev, hipStream_t st);
// keys[k] = field of saddle idx[k] (saddles are numbered field by field: sad_b holds the B + 1 bases)
void launch_gather_field(int64_t n, const int32_t* idx, int B, const int32_t* sad_b, uint32_t* keys, hipStream_t st);
void launch_sort_keys(int64_t n, const float* v, uint32_t* keys, hipStream_t st);

// ------------------------------------------------------------------ arc geometry (k_arcs.hip)
// One traced side of the arc geometry (faces for ridges/"max", vertices for valleys/"min").
struct ArcSide {
  int faces;               // 1: face V-paths (slot-indexed successors), 0: vertex V-paths (pixel-indexed)
  const int32_t* succ;     // successor array of the forest (global slot / pixel indices)
  const FieldDesc* fd;     // per field geometry and bases
  const int32_t* sfield;   // per saddle field (null: a single field)
  int64_t ns;              // saddles
  int64_t next;            // extrema of this side
  const int32_t* sad_ids;
  const int32_t* tinit;    // per arc (2s+k): initial target extremum, -1 = no arc
  const int32_t* blen;     // per arc: raw length in cells (steps + 1), 0 = no arc
  const int4* tl;          // per saddle: both arcs' targets and raw lengths (tinit[2s], tinit[2s+1], blen[2s],
                           // blen[2s+1]) in one 16-byte record for the stages that gather by saddle
  const uint8_t* removed;  // per saddle: cancelled (on either side)
  const int32_t* cancel;   // per saddle: this side's dead extremum when this side cancels it, else -1
                           // (an arc is traced when its saddle survives or this side cancels it: no
                           // output reads the arcs of the other side's cancellations)
  const int64_t* boff;     // per traced arc longer than kHeadCells: offset in the raw arena
  int32_t* arena;          // raw arcs longer than kHeadCells, cell ids from the saddle outwards
  int4* head;              // per arc of at most kHeadCells cells: its cells (the arena holds the longer ones)
};
constexpr int kHeadCells = 4;

// per saddle: blen[2s..2s+1] and tl[s] (targets c[2s..2s+1] with those lengths)
void launch_arc_lengths(int64_t ns, const int32_t* c, const int32_t* steps, int32_t* blen, int4* tl, hipStream_t st);
void launch_raw_arcs(const ArcSide& a, hipStream_t st);

struct GeoState {
  // cancellations indexed in saddle-id order; key = position in (persistence, saddle id) order
  int64_t ncancel;
  const int32_t* cs;       // cancel -> saddle index
  const int32_t* cd;       // cancel -> dead extremum
  const int32_t* key;      // cancel -> sorted position
  int4* link;              // per extremum, one record: (parent in the geometry forest: self for a root,
                           // -1 for the void; length of its link path (saturating); saddle of the
                           // cancellation that linked it; which arc (0/1) of that saddle led to it)
  uint64_t* gk;            // per extremum e, one 16-byte record: gk[2e] = path-compressed view of the parents
                           // (ancestor or kVoid, link length to it); gk[2e+1] = kill word: (round its killing
                           // cancellation was decided, or all ones while pending) << 32 | that cancellation's
                           // key (all ones: no killer). A root walk's last read brings the root's kill word.
  uint64_t* watch;         // per extremum: (tag << 32 | cancel) atomic min
  int32_t* rr;             // per worklist position: R0, R1, len0, len1
  int4* wl;                // worklists (cancellation, key, arc target 0, arc target 1): list k at
  int64_t wls;             // wl + k * wls
  __host__ __device__ int4* list(int k) const { return wl + (int64_t)k * wls; }
  int2* glen;              // per position of each worklist (list k at glen + k wls): path lengths from the
                           // cancellation's arc targets to the two roots its item holds from round 1 on
  __host__ __device__ int2* lens(int k) const { return glen + (int64_t)k * wls; }
  int32_t* cnt;
  int32_t* err;
};
// link[e] = (e, 0, 0, 0), forest entry (e, 0): every extremum its own root; kill and watch words all ones
void launch_geo_init(const GeoState& g, int64_t next, hipStream_t st);
void launch_geo_round(const ArcSide& a, const GeoState& g, int round, int64_t max_items, hipStream_t st);
// one round of both traced sides (bit s of `live`) in one find and one apply launch
void launch_geo_round2(const ArcSide* a, const GeoState* g, int live, int round, int64_t max_items, hipStream_t st,
                       const int32_t* d_round = nullptr);
// Geometry counterpart of launch_persist_tail (both traced sides, `live` mask).
void launch_geo_tail(const ArcSide* a, const GeoState* g, int live, int round0, int max_rounds, int32_t* round_out,
                     hipStream_t st);
void launch_cancel_flags(int64_t ns, const int32_t* cancel_d, int32_t* flag, hipStream_t st);
void launch_cancel_compact(int64_t ns, const int32_t* cancel_d, const float* pers, const int32_t* pre,
                           int32_t* cs, int32_t* cd, uint32_t* keys, hipStream_t st);
constexpr uint64_t kKillPending = 0xFFFFFFFFull << 32;
// list 0 (every cancellation with its key and arc targets, saddle-id order), its count, and the
// kill word of every dead extremum (gk: the records of GeoState::gk)
void launch_geo_items(int64_t nc, const int32_t* cs, const int32_t* cd, const int32_t* key, const int32_t* tinit,
                      int4* wl, uint64_t* gk, int32_t* cnt, hipStream_t st);
void launch_geo_q(const GeoState& g, int64_t next, uint64_t* pjq, hipStream_t st);

// Output arc lengths for surviving saddles in min order (ord: survivor list, position = output saddle);
// rec receives one 32-byte record per position (targets, raw lengths, path lengths to the final roots,
// those roots) for the seeds pass. Then segment expansion.
void launch_final_lengths(int64_t ns, const int32_t* ord, const ArcSide& a, const uint64_t* pjq, void* rec,
                          int64_t* flen, int64_t* per_pos, hipStream_t st);
// A piece of a merged arc: a raw arc (a = arc index, b = its length) or a chain of links (from
// extremum a up to its ancestor b, -1 = the void). od = output offset << 1 | forward bit.
struct Seg {
  int32_t a, b;
  int64_t od;
};
// Everything the expansion needs about the link of one extremum d, gathered once (32 bytes, one
// aligned half cache line): the link expands to reverse(chain(Td -> d)) ++ reverse(raw arc) ++
// raw arc ^ 1 ++ chain(To -> next) with those pieces' lengths Alen, Bk, Bo, Clen.
struct alignas(32) LinkRec {
  int32_t next;      // parent (-1: the void); d itself for a root
  int32_t Td, To;    // the arc targets of the linking saddle on d's side and on the other side
  int32_t arc;       // arc index 2 s + kd of the saddle's arc on d's side (the other one is arc ^ 1)
  int32_t Alen, Bk, Bo, Clen;
};
// Per extremum, after the geometry rounds: cumulative counts of the raw / chain pieces that the
// links on its path to the final root expand into (pointer-jumped (root, count) pairs), and its
// link record; then ncq[e] = (raw pieces, chain pieces, path length to the final root, 0), the three
// path sums a chain's two ends are read for, in one 16-byte record.
void launch_link_counts(const ArcSide& a, const GeoState& g, const uint64_t* pjq, int64_t next, uint64_t* cum_r,
                        uint64_t* cum_c, LinkRec* xl, int4* ncq, hipStream_t st);
// arc_base: [B+1] first output cell of each field; sad_out: [B+1] first output saddle of each field.
// A field's offsets array has one more entry than its saddles, so field f's entries start at
// sad_out[f] + f in the concatenated offsets output.
void launch_seed_segments(int64_t ns, const int32_t* ord, const ArcSide& a, const void* rec, int64_t next,
                          const int64_t* flen, const int64_t* pos_pre, Seg* chains, int32_t* nchain, Seg* raw, int32_t* nraw, int64_t cap,
                          int32_t* err, int32_t* out_offsets, int32_t* out_arc_offsets, const int64_t* arc_base,
                          const int32_t* sad_out, int B, hipStream_t st);
// One level of expansion: every chain becomes the raw pieces and sub-chains of its links.
// n = min(*nin, cap) chains are read; *nreset is cleared (the counter the following level appends to).
void launch_expand_segments(const ArcSide& a, const GeoState& g, const int4* ncq, const LinkRec* xl, const Seg* in,
                            const int32_t* nin, Seg* out,
                            int32_t* nout, int32_t* nreset, Seg* raw, int32_t* nraw, int64_t max_in, int64_t cap,
                            hipStream_t st);
void launch_copy_raw(const ArcSide& a, const Seg* raw, const int32_t* nraw, int64_t max_raw, int32_t* out,
                     int64_t out_total, int32_t* err, hipStream_t st);

}  // namespace dmsc
