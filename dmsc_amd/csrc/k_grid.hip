// Grid-shaped stages: discrete gradient + critical bit planes (one LDS-tiled stencil pass),
// order-preserving critical lists, saddle records and basin/peak labels.
#include <type_traits>

#include "internal.hpp"

namespace dmsc {

__constant__ constexpr PairLut c_lut = make_pair_lut();

// ------------------------------------------------------------------ K1: gradient
constexpr int GT_X = 64;

// In-tile V-path state of phase 3, one 32-bit word per node: the node it points to (local index),
// an exit flag (that node's successor leaves the tile; its global id is kept in the tile's exit
// table) and the steps to get there. A node is final when it points to itself (a root or an
// invalid slot) or to an exit node. Both forests' words of a node share one 64-bit LDS entry.
constexpr uint32_t kPjNodeBits = 11;  // tiles of up to 2048 slots
constexpr uint32_t kPjNodeMask = (1u << kPjNodeBits) - 1u, kPjExit = 1u << kPjNodeBits;
constexpr int kPjStepShift = kPjNodeBits + 1;
__device__ __forceinline__ uint32_t pj_word(int node, bool exit, uint32_t steps) {
  return (uint32_t)node | (exit ? kPjExit : 0u) | steps << kPjStepShift;
}
__device__ __forceinline__ int pj_node(uint32_t w) { return (int)(w & kPjNodeMask); }
__device__ __forceinline__ bool pj_final(uint32_t w, int self) { return (w & kPjExit) || pj_node(w) == self; }
// w points to node n whose word is w2: w now points where w2 does, steps added
__device__ __forceinline__ uint32_t pj_join(uint32_t w, uint32_t w2) {
  return (w2 & (kPjNodeMask | kPjExit)) | ((w >> kPjStepShift) + (w2 >> kPjStepShift)) << kPjStepShift;
}

constexpr int log2_exact(int n) { return n <= 1 ? 0 : 1 + log2_exact(n / 2); }

// Bijection of block ids onto [0, n): the blocks of XCD k (ids k, k + 8, ...) get the k-th run of
// consecutive ids, in order.
__device__ __forceinline__ int xcd_run(int b, int n) {
  const int q = n >> 3, r = n & 7, k = b & 7, j = b >> 3;
  return k * q + (k < r ? k : r) + j;
}

// Lower star of one vertex as a 32-bit summary: edge mask (bits 0-3, L_edges order), face mask
// (bits 4-7, L_faces order), paired edge estar (bits 8-10, 4 = none) and the face-queue result
// (bits 11-18, 2 bits per face: 0 unpaired, 1 + which of its two edges). A vertex outside the image
// has an empty star.
constexpr uint32_t kEmptyStar = 4u << 8;
__device__ __forceinline__ uint32_t star_word(const LowerStar& ls) {
  const uint32_t idx = (uint32_t)(ls.emask | ls.fmask << 4 | ls.estar << 8);
  return idx | (uint32_t)c_lut.v[idx] << 11;
}
// Partner of edge e of the star of vertex slot qo: the vertex itself when e is its paired edge, the
// face the queue paired with e, or -1 (critical).
__device__ __forceinline__ int star_edge_partner(uint32_t w, int e, int qo, int Nx) {
  int partner = ((w >> 8) & 7) == (uint32_t)e ? 4 * qo : -1;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cj = (int)(w >> (11 + 2 * j)) & 3;
    if (cj && face_edge(j, cj - 1) == e) partner = 4 * (qo + (j & 1) + (j >> 1) * Nx) + 3;
  }
  return partner;
}
// Partner of vertex qo itself: its paired edge (the summary's estar is < 4).
__device__ __forceinline__ int star_edge_partner_vertex(uint32_t w, int qo, int Nx) {
  const int e = (int)(w >> 8) & 7;
  return 4 * (qo + (e == 1 ? 1 : 0) + (e == 3 ? Nx : 0)) + (e < 2 ? 1 : 2);
}
// Partner of face j of the star of vertex slot qo: the edge the queue paired it with, or -1.
__device__ __forceinline__ int star_face_partner(uint32_t w, int j, int qo, int Nx) {
  const int cj = (int)(w >> (11 + 2 * j)) & 3;
  if (!cj) return -1;
  const int e = face_edge(j, cj - 1);
  return 4 * (qo + (e == 1 ? 1 : 0) + (e == 3 ? Nx : 0)) + (e < 2 ? 1 : 2);
}

// One GT_X x TY tile of slots per 256-thread block (a wave per slot row). Phase 1 evaluates the lower
// star of every vertex that can own a cell of the tile ((TY+1) x (GT_X+1) vertices, from pixels
// staged in LDS with a 2-pixel halo) and keeps one summary word per vertex; tiles whose vertices'
// neighbourhoods all lie inside the image (almost all of them) take a path with every existence
// test folded away. Phase 2 gives each slot the partners of its four cells from the summaries of the
// vertices that can own them (every valid cell lies in the lower star of exactly one vertex: its
// vertex, its edges' and face's highest vertex among the slot's vertex and its left, upper and
// upper-left neighbours) and writes them as one 16-byte store, the face / vertex V-path successors,
// and the four critical bit planes via wave ballots. Phase 3 follows every face and vertex V-path
// inside the tile by pointer doubling in LDS (log2 of the tile's slots at most) and writes the
// initial pointer-jumping state: (in-tile root or first node outside the tile, steps to it). The
// global pointer jumping then moves one tile crossing per step instead of one cell.
template <bool DUAL, int TY>
__global__ __launch_bounds__(256) void k_gradient(Batch bt, const int64_t* __restrict__ tile_b,
                                                  int32_t* __restrict__ pw_all, int32_t* __restrict__ fsucc_all,
                                                  int32_t* __restrict__ vsucc_all, uint32_t* __restrict__ planes,
                                                  int64_t PW, uint64_t* __restrict__ pjf_all,
                                                  uint64_t* __restrict__ pjv_all) {
  constexpr int GP_W = GT_X + 3, GP_H = TY + 3, TN = TY * GT_X;
  constexpr int SV_W = GT_X + 1, SV_H = TY + 1;
  static_assert(TN <= (1 << kPjNodeBits), "tile too large for the packed phase-3 words");
  // phases 1-2: staged pixels and star summaries; phase 3 reuses that space as its second buffer
  constexpr int kStageBytes = (GP_H * GP_W + SV_H * SV_W) * 4;
  constexpr int kBufBytes = TN * 8;
  __shared__ __align__(16) unsigned char stage[kStageBytes > kBufBytes ? kStageBytes : kBufBytes];
  __shared__ uint64_t pjA[TN];      // phase 3: (face word, vertex word) per node
  __shared__ int32_t exits[2][TN];  // global successor of each exit node (faces, vertices)
  float (*pix)[GP_W] = reinterpret_cast<float (*)[GP_W]>(stage);
  uint32_t (*sv)[SV_W] = reinterpret_cast<uint32_t (*)[SV_W]>(stage + GP_H * GP_W * 4);

  // Workgroups go to the 8 XCDs in turn (block b to XCD b mod 8); each XCD takes one contiguous run
  // of tiles instead, so the tile rows' partial cache lines at tile edges (rows of Nx = W + 1 slots
  // are not line-aligned) are completed in the same L2 by the neighbouring tile.
  const int tile = xcd_run(blockIdx.x, gridDim.x);
  const int tfield = field_of(tile_b, bt.B, tile);
  const FieldDesc fd = bt.fd[tfield];
  const int64_t tlocal = tile - tile_b[tfield];
  const int ntx = (fd.Nx + GT_X - 1) / GT_X;
  const TileRef tr{tfield, (int)(tlocal % ntx), (int)(tlocal / ntx)};
  const float* __restrict__ f = fd.f;
  const int H = fd.H, W = fd.W, WPR = fd.WPR;
  const int Nx = W + 1, Ny = H + 1;
  const int x0 = tr.tx * GT_X, y0 = tr.ty * TY;
  const int sbase = (int)fd.slot_base, pbase = (int)fd.pix_base;  // batch totals fit int32 (host check)
  int32_t* __restrict__ pw = pw_all ? pw_all + 4 * fd.slot_base : nullptr;
  int32_t* __restrict__ fsucc = fsucc_all ? fsucc_all + fd.slot_base : nullptr;  // null: not wanted
  int32_t* __restrict__ vsucc = vsucc_all ? vsucc_all + fd.pix_base : nullptr;
  const int64_t wbase = fd.word_base;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  // pixels (y0-2 .. y0+TY, x0-2 .. x0+GT_X): the GP_H x GP_W staging area as one flat range over the
  // block's threads, every load issued before any is stored
  {
    constexpr int NP = GP_H * GP_W, KP = (NP + 255) / 256;
    float v[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int i = tid + 256 * k, r = i / GP_W, c = i - r * GP_W;
      const int py = y0 - 2 + r, px = x0 - 2 + c;
      v[k] = (i < NP && py >= 0 && py < H && px >= 0 && px < W) ? f[(int64_t)py * W + px] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < KP; ++k)
      if (tid + 256 * k < NP) (&pix[0][0])[tid + 256 * k] = v[k];
  }
  __syncthreads();

  // phase 1: star summary of vertex (y0-1+ly, x0-1+lx), a wave per vertex row (lane = column) and the
  // last column after
  const bool interior = y0 >= 2 && x0 >= 2 && y0 + TY <= H - 1 && x0 + GT_X <= W - 1;
  auto star_at = [&](int ly, int lx, auto interior_tag) -> uint32_t {
    constexpr bool IN = decltype(interior_tag)::value;
    const int vy = y0 - 1 + ly, vx = x0 - 1 + lx;
    if (!IN && (vy < 0 || vy >= H || vx < 0 || vx >= W)) return kEmptyStar;
    float nb[3][3];
    bool ex[3][3];
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int py = vy + dy, px = vx + dx;
        ex[dy + 1][dx + 1] = IN || (py >= 0 && py < H && px >= 0 && px < W);
        nb[dy + 1][dx + 1] = pix[ly + dy + 1][lx + dx + 1];
      }
    return star_word(lower_star<DUAL>(nb, ex));
  };
  // the SV_H x SV_W summaries as one flat range over the block's threads (a wave per row would leave
  // the last row and the last column to one wave)
  auto phase1 = [&](auto interior_tag) {
    constexpr int NV = SV_H * SV_W;
    for (int i = tid; i < NV; i += 256) {
      const int ly = i / SV_W, lx = i - ly * SV_W;
      sv[ly][lx] = star_at(ly, lx, interior_tag);
    }
  };
  if (interior) phase1(std::true_type{});
  else phase1(std::false_type{});
  __syncthreads();

  for (int r = wv; r < TY; r += 4) {
    const int y = y0 + r, x = x0 + lane;
    const bool in = (y < Ny) && (x < Nx);
    int c0 = -1, c1 = -1, c2 = -1, c3 = -1;
    if (in) {
      const int q = y * Nx + x;
      // summaries of the vertices (y, x), (y, x-1), (y-1, x), (y-1, x-1)
      const uint32_t w00 = sv[r + 1][lane + 1], wl = sv[r + 1][lane], wu = sv[r][lane + 1], wul = sv[r][lane];
      c0 = ((w00 >> 8) & 7) < 4 ? star_edge_partner_vertex(w00, q, Nx) : -1;
      // H-edge (y, x): the left edge of (y, x) or the right edge of (y, x-1)
      if (w00 & 1) c1 = star_edge_partner(w00, 0, q, Nx);
      else if (wl & 2) c1 = star_edge_partner(wl, 1, q - 1, Nx);
      // V-edge (y, x): the up edge of (y, x) or the down edge of (y-1, x)
      if (w00 & 4) c2 = star_edge_partner(w00, 2, q, Nx);
      else if (wu & 8) c2 = star_edge_partner(wu, 3, q - Nx, Nx);
      // face (y, x): UL of (y, x), UR of (y, x-1), DL of (y-1, x) or DR of (y-1, x-1)
      if (w00 & 16) c3 = star_face_partner(w00, 0, q, Nx);
      else if (wl & 32) c3 = star_face_partner(wl, 1, q - 1, Nx);
      else if (wu & 64) c3 = star_face_partner(wu, 2, q - Nx, Nx);
      else if (wul & 128) c3 = star_face_partner(wul, 3, q - Nx - 1, Nx);
    }
    const bool k0 = in && c0 < 0 && y < H && x < W;
    const bool k1 = in && c1 < 0 && y < H;
    const bool k2 = in && c2 < 0 && x < W;
    const bool k3 = in && c3 < 0;
    const int e = r * GT_X + lane;
    uint32_t fw = pj_word(e, false, 0), vw = pj_word(e, false, 0);  // itself: a root or an invalid slot
    if (in) {
      const int q = y * Nx + x;
      if (pw) *reinterpret_cast<int4*>(pw + 4 * (int64_t)q) = make_int4(c0, c1, c2, c3);
      int fs = q, dy = 0, dx = 0;
      if (c3 >= 0) {
        if (c3 == 4 * (q - Nx) + 1) { fs = q - Nx; dy = -1; }
        else if (c3 == 4 * q + 1) { fs = q + Nx; dy = 1; }
        else if (c3 == 4 * (q - 1) + 2) { fs = q - 1; dx = -1; }
        else { fs = q + 1; dx = 1; }
        const int rr = r + dy, ll = lane + dx;
        if (rr >= 0 && rr < TY && ll >= 0 && ll < GT_X) {
          fw = pj_word(rr * GT_X + ll, false, 1);
        } else {
          fw = pj_word(e, true, 1);
          exits[0][e] = sbase + fs;
        }
      }
      if (fsucc) fsucc[q] = sbase + fs;
      if (y < H && x < W) {
        const int p = y * W + x;
        int vs = p;
        dy = 0;
        dx = 0;
        if (c0 >= 0) {
          if (c0 == 4 * q + 1) { vs = p - 1; dx = -1; }
          else if (c0 == 4 * (q + 1) + 1) { vs = p + 1; dx = 1; }
          else if (c0 == 4 * q + 2) { vs = p - W; dy = -1; }
          else { vs = p + W; dy = 1; }
          const int rr = r + dy, ll = lane + dx;
          if (rr >= 0 && rr < TY && ll >= 0 && ll < GT_X) {
            vw = pj_word(rr * GT_X + ll, false, 1);
          } else {
            vw = pj_word(e, true, 1);
            exits[1][e] = pbase + vs;
          }
        }
        if (vsucc) vsucc[p] = pbase + vs;
      }
    }
    pjA[e] = (uint64_t)fw | (uint64_t)vw << 32;
    const uint64_t b0 = __ballot(k0), b1 = __ballot(k1), b2 = __ballot(k2), b3 = __ballot(k3);
    if (y < Ny && lane < 8) {
      const int t = lane >> 1, half = lane & 1;
      const uint64_t bb = t == 0 ? b0 : t == 1 ? b1 : t == 2 ? b2 : b3;
      const int wi = (x0 >> 5) + half;
      if (wi < WPR) planes[t * PW + wbase + (int64_t)y * WPR + wi] = (uint32_t)(bb >> (32 * half));
    }
  }
  __syncthreads();

  // phase 3: pointer doubling inside the tile, both forests per 64-bit word (double-buffered through
  // the staging space)
  uint64_t* A = pjA;
  uint64_t* Bf = reinterpret_cast<uint64_t*>(stage);
#pragma unroll 1
  for (int it = 0; it < log2_exact(TN); ++it) {
    int moved = 0;  // a pass in which no word can advance further leaves every chain resolved
#pragma unroll
    for (int k = 0; k < TN / 256; ++k) {
      const int e = tid + 256 * k;
      const uint64_t w = A[e];
      uint32_t fwd = (uint32_t)w, vwd = (uint32_t)(w >> 32);
      if (!pj_final(fwd, e)) {
        const int n = pj_node(fwd);
        const uint32_t w2 = (uint32_t)A[n];
        moved |= !pj_final(w2, n);
        fwd = pj_join(fwd, w2);
      }
      if (!pj_final(vwd, e)) {
        const int n = pj_node(vwd);
        const uint32_t w2 = (uint32_t)(A[n] >> 32);
        moved |= !pj_final(w2, n);
        vwd = pj_join(vwd, w2);
      }
      Bf[e] = (uint64_t)fwd | (uint64_t)vwd << 32;
    }
    const int any = __syncthreads_or(moved);
    uint64_t* t = A;
    A = Bf;
    Bf = t;
    if (!any) break;
  }
  uint64_t* __restrict__ pjf = pjf_all + fd.slot_base;
  uint64_t* __restrict__ pjv = pjv_all + fd.pix_base;
  for (int r = wv; r < TY; r += 4) {
    const int e = r * GT_X + lane;
    const int y = y0 + r, x = x0 + lane;
    if (y >= Ny || x >= Nx) continue;
    const int q = y * Nx + x;
    const uint64_t w = A[e];
    uint32_t wd = (uint32_t)w;
    int n = pj_node(wd);
    uint32_t target = (wd & kPjExit) ? (uint32_t)exits[0][n] : (uint32_t)(sbase + (y0 + n / GT_X) * Nx + x0 + n % GT_X);
    pjf[q] = (uint64_t)target | (uint64_t)(wd >> kPjStepShift) << 32;
    if (y < H && x < W) {
      wd = (uint32_t)(w >> 32);
      n = pj_node(wd);
      target = (wd & kPjExit) ? (uint32_t)exits[1][n] : (uint32_t)(pbase + (y0 + n / GT_X) * W + x0 + n % GT_X);
      pjv[y * W + x] = (uint64_t)target | (uint64_t)(wd >> kPjStepShift) << 32;
    }
  }
}

int gradient_tiles(int H, int W, int rows) { return ceil_div(H + 1, rows) * ceil_div(W + 1, GT_X); }

template <bool DUAL>
static void launch_gradient_rows(int rows, dim3 grid, hipStream_t st, const Batch& bt, const int64_t* tile_b,
                                 int32_t* pw, int32_t* fsucc, int32_t* vsucc, uint32_t* planes, int64_t PW,
                                 uint64_t* pjf, uint64_t* pjv) {
  if (rows == 8)
    hipLaunchKernelGGL((k_gradient<DUAL, 8>), grid, dim3(256), 0, st, bt, tile_b, pw, fsucc, vsucc, planes, PW, pjf, pjv);
  else if (rows == 32)
    hipLaunchKernelGGL((k_gradient<DUAL, 32>), grid, dim3(256), 0, st, bt, tile_b, pw, fsucc, vsucc, planes, PW, pjf, pjv);
  else
    hipLaunchKernelGGL((k_gradient<DUAL, 16>), grid, dim3(256), 0, st, bt, tile_b, pw, fsucc, vsucc, planes, PW, pjf, pjv);
}

void launch_gradient(bool dual, int rows, const Batch& bt, const int64_t* tile_b, int64_t ntiles, int32_t* pw,
                     int32_t* fsucc, int32_t* vsucc, uint32_t* planes, int64_t PW, uint64_t* pjf, uint64_t* pjv,
                     hipStream_t st) {
  KScope ks_("gradient", st, true);
  if (ntiles == 0) return;
  const dim3 grid((unsigned)ntiles);
  if (dual) launch_gradient_rows<true>(rows, grid, st, bt, tile_b, pw, fsucc, vsucc, planes, PW, pjf, pjv);
  else launch_gradient_rows<false>(rows, grid, st, bt, tile_b, pw, fsucc, vsucc, planes, PW, pjf, pjv);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ ranks
// index of a critical cell within its (batch-wide) list: the reference's fast_crit_map,
// gradient.hxx:198-231, as a popcount over the field's bit plane plus the word prefix
__device__ __forceinline__ int rank_in_plane(const uint32_t* plane, const int32_t* pre, int64_t row_word, int x) {
  const int64_t w = row_word + (x >> 5);
  const uint32_t below = (1u << (x & 31)) - 1u;
  return pre[w] + __popc(plane[w] & below);
}
// q: global slot index of a face of field f
__device__ __forceinline__ int rank_max_slot(const RankView& rv, int f, int q) {
  const FieldDesc& fd = rv.bt.fd[f];
  const int ql = q - (int)fd.slot_base;
  const int y = ql / fd.Nx, x = ql - y * fd.Nx;
  return rank_in_plane(rv.planes + 3 * rv.PW, rv.pre_max, fd.word_base + (int64_t)y * fd.WPR, x);
}
// p: global pixel index of a vertex of field f
__device__ __forceinline__ int rank_min_pixel(const RankView& rv, int f, int p) {
  const FieldDesc& fd = rv.bt.fd[f];
  const int pl = p - (int)fd.pix_base;
  const int y = pl / fd.W, x = pl - y * fd.W;
  return rank_in_plane(rv.planes, rv.pre_min, fd.word_base + (int64_t)y * fd.WPR, x);
}

// ------------------------------------------------------------------ critical lists
// Set bits in id order (the reference's single-thread order, critical_points.hxx:14-65), with the
// extremum key / saddle value. A wave takes 64 consecutive plane words: each lane counts its word's
// entries, a wave prefix sum places them, the lanes write the entries into an LDS stage as 16-bit
// (word, bit, type) codes beside a table of each word's row, first column and field, and the wave
// then writes the chunk's part of each list with consecutive lanes on consecutive entries (one lane
// per word would scatter 4-byte stores over 64 cache lines). No entry needs a division or a field
// search: its row, column and field come from its word's table entry.
constexpr int kCompactChunk = 64;  // words per wave
constexpr int kCompactWaves = 2;   // waves per block (the stage holds 2 saddles per slot: 8 KB a wave)

template <bool DUAL>
__global__ __launch_bounds__(64 * kCompactWaves) void k_compact_lists(RankView rv, int32_t* __restrict__ min_ids,
                                                                    uint64_t* __restrict__ min_key,
                                                                    int32_t* __restrict__ sad_ids, float* __restrict__ sad_val,
                                                                    int32_t* __restrict__ sad_field,
                                                                    int32_t* __restrict__ max_ids,
                                                                    uint64_t* __restrict__ max_key) {
  // code = word in chunk << 7 | bit << 2 | cell type
  __shared__ uint16_t stage[kCompactWaves][kCompactChunk * 32 * 2];  // up to 2 saddles per slot
  __shared__ int4 winfo[kCompactWaves][kCompactChunk];               // (row, first column, field, -)
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int B = rv.bt.B;
  const int64_t nchunks = (rv.PW + kCompactChunk - 1) / kCompactChunk;
  // block-uniform loop: every wave runs the same iterations (the block synchronises in between)
  for (int64_t cb = (int64_t)blockIdx.x * kCompactWaves; cb < nchunks; cb += (int64_t)gridDim.x * kCompactWaves) {
    const int64_t chunk = cb + wid;
    const int64_t w = chunk * kCompactChunk + lane;
    const bool valid = chunk < nchunks && w < rv.PW;
    uint32_t m[4] = {0, 0, 0, 0};
    if (valid) {
      const int fi = field_of(rv.bt.word_b, B, w);
      const FieldDesc& fd = rv.bt.fd[fi];
      const int64_t lw = w - fd.word_base;
      const int yrow = (int)(lw / fd.WPR);
      winfo[wid][lane] = make_int4(yrow, (int)(lw - (int64_t)yrow * fd.WPR) * 32, fi, 0);
#pragma unroll
      for (int t = 0; t < 4; ++t) m[t] = rv.planes[t * rv.PW + w];
    }
    const int64_t w0 = chunk * kCompactChunk;
#pragma unroll
    for (int list = 0; list < 3; ++list) {
      // entries of this lane's word, in id order
      const uint32_t ma = list == 0 ? m[0] : list == 1 ? m[1] : m[3];
      const uint32_t mb = list == 1 ? m[2] : 0u;
      const int cnt = __popc(ma) + __popc(mb);
      const int incl = wave_inclusive_sum(cnt);
      const int total = __shfl(incl, 63, 64);
      int o = incl - cnt;
      uint32_t mm = ma | mb;
      while (mm) {
        const int bit = __builtin_ctz(mm);
        mm &= mm - 1;
        const int c = lane << 7 | bit << 2;
        if (list == 0) stage[wid][o++] = (uint16_t)c;
        else if (list == 2) stage[wid][o++] = (uint16_t)(c | 3);
        else {
          if (ma >> bit & 1) stage[wid][o++] = (uint16_t)(c | 1);
          if (mb >> bit & 1) stage[wid][o++] = (uint16_t)(c | 2);
        }
      }
      __syncthreads();
      if (chunk < nchunks && total > 0) {
        const int32_t* pre = list == 0 ? rv.pre_min : list == 1 ? rv.pre_sad : rv.pre_max;
        const int64_t base = pre[w0];
        for (int t = lane; t < total; t += 64) {
          const int c = stage[wid][t];
          const int4 wi = winfo[wid][c >> 7];
          const int ty = wi.x, tx = wi.y + ((c >> 2) & 31), type = c & 3;
          const FieldDesc& fd = rv.bt.fd[wi.z];
          const int id = 4 * (ty * fd.Nx + tx) + type;
          const int64_t oi = base + t;
          if (list == 0) {
            min_ids[oi] = id;
            min_key[oi] = pack_key(fd.f[(int64_t)ty * fd.W + tx], id);
          } else if (list == 2) {
            max_ids[oi] = id;
            max_key[oi] = pack_key(fd.f[top_pixel<DUAL>(3, ty, tx, fd.H, fd.W, fd.f)], id);
          } else {
            sad_ids[oi] = id;
            sad_val[oi] = fd.f[top_pixel<DUAL>(type, ty, tx, fd.H, fd.W, fd.f)];
            if (sad_field) sad_field[oi] = wi.z;
          }
        }
      }
      __syncthreads();
    }
  }
}

void launch_compact_lists(bool dual, const RankView& rv, int32_t* min_ids, uint64_t* min_key, int32_t* sad_ids,
                          float* sad_val, int32_t* sad_field, int32_t* max_ids, uint64_t* max_key, hipStream_t st) {
  KScope ks_("compact_lists", st);
  if (rv.PW == 0) return;
  const int g = grid_for((rv.PW + kCompactChunk - 1) / kCompactChunk, kCompactWaves, 16384);
  if (dual)
    hipLaunchKernelGGL(k_compact_lists<true>, dim3(g), dim3(64 * kCompactWaves), 0, st, rv, min_ids, min_key, sad_ids,
                       sad_val, sad_field, max_ids, max_key);
  else
    hipLaunchKernelGGL(k_compact_lists<false>, dim3(g), dim3(64 * kCompactWaves), 0, st, rv, min_ids, min_key, sad_ids,
                       sad_val, sad_field, max_ids, max_key);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ saddle records
// Per saddle (list order): the extremum reached by each of its four V-paths and the step count
// (reference arcs_topology.hxx:67-172). Arc k of a side is its k-th existing start cell; a missing
// start (border edge, vertex side only) gives target -1 and length 0. Forest indices are global
// (field base + local slot / pixel); extremum indices are batch-wide list ranks.
__global__ void k_saddle_records(RankView rv, const int32_t* __restrict__ sad_ids, const int32_t* __restrict__ sad_field,
                                 int64_t nsad, const uint64_t* __restrict__ pj_face,
                                 const uint64_t* __restrict__ pj_vert, int32_t* __restrict__ cmax,
                                 int32_t* __restrict__ cmin, int32_t* __restrict__ lmax, int32_t* __restrict__ lmin) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nsad; i += (int64_t)gridDim.x * blockDim.x) {
    const int fi = sad_field ? sad_field[i] : 0;
    const FieldDesc& fd = rv.bt.fd[fi];
    const int Nx = fd.Nx, W = fd.W, H = fd.H;
    const int id = sad_ids[i], q = id >> 2, t = id & 3;
    const int y = q / Nx, x = q - y * Nx;
    const int sb = (int)fd.slot_base, pb = (int)fd.pix_base;
    const int f1 = sb + q, f2 = sb + ((t == 1) ? q + Nx : q + 1);
    const uint64_t a = pj_face[f1], b = pj_face[f2];
    cmax[2 * i] = rank_max_slot(rv, fi, (int)(uint32_t)a);
    cmax[2 * i + 1] = rank_max_slot(rv, fi, (int)(uint32_t)b);
    lmax[2 * i] = (int)(a >> 32);
    lmax[2 * i + 1] = (int)(b >> 32);
    int v[2], nv = 0;
    if (t == 1) {
      if (x - 1 >= 0) v[nv++] = y * W + x - 1;
      if (x < W) v[nv++] = y * W + x;
    } else {
      if (y - 1 >= 0) v[nv++] = (y - 1) * W + x;
      if (y < H) v[nv++] = y * W + x;
    }
    for (int k = 0; k < 2; ++k) {
      if (k < nv) {
        const uint64_t c = pj_vert[pb + v[k]];
        cmin[2 * i + k] = rank_min_pixel(rv, fi, (int)(uint32_t)c);
        lmin[2 * i + k] = (int)(c >> 32);
      } else {
        cmin[2 * i + k] = -1;
        lmin[2 * i + k] = 0;
      }
    }
  }
}

void launch_saddle_records(const RankView& rv, const int32_t* sad_ids, const int32_t* sad_field, int64_t nsad,
                           const uint64_t* pj_face, const uint64_t* pj_vert, int32_t* cmax, int32_t* cmin,
                           int32_t* lmax, int32_t* lmin, hipStream_t st) {
  KScope ks_("saddle_records", st);
  if (nsad == 0) return;
  hipLaunchKernelGGL(k_saddle_records, dim3(grid_for(nsad, 256, 16384)), dim3(256), 0, st, rv, sad_ids, sad_field, nsad,
                     pj_face, pj_vert, cmax, cmin, lmax, lmin);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ labels
// label(cell) = field-local output index of the thresholded-UF root of its V-path terminal, -1 for
// the void or a cancelled root (reference cell_groups.hxx:96-229). The output is the concatenation
// of the fields' label arrays (slot / pixel order).
// o: output index of the extremum's thresholded root (-1: none)
__device__ __forceinline__ int32_t label_of(int32_t o, int32_t base) { return o >= 0 ? o - base : -1; }
// Each thread takes kLabelUnroll cells of its block's chunk per iteration, one per block-wide stride,
// and issues every cell's loads before the dependent ones (successor entry, then the rank words, then
// the root's output index), so each wave keeps several of those chains in flight.
constexpr int kLabelUnroll = 4;
template <bool FACES>
__global__ __launch_bounds__(256) void k_labels(int64_t n, RankView rv, const uint64_t* __restrict__ pj,
                                                const int32_t* __restrict__ root_out,
                                                const int32_t* __restrict__ out_base, int32_t* __restrict__ out) {
  const int64_t* const base = FACES ? rv.bt.slot_b : rv.bt.pix_b;
  constexpr int U = kLabelUnroll;
  for (int64_t c0 = (int64_t)blockIdx.x * (256 * U); c0 < n; c0 += (int64_t)gridDim.x * (256 * U)) {
    const int fb = field_of(base, rv.bt.B, c0);  // the chunk's first cell: one uniform search
    int fi[U];
    uint64_t w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = c0 + u * 256 + threadIdx.x;
      fi[u] = q < n ? field_from(base, rv.bt.B, fb, q) : fb;
      w[u] = q < n ? pj[q] : 0;
    }
    int32_t rk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = (int)(uint32_t)w[u];
      rk[u] = c0 + u * 256 + threadIdx.x < n ? (FACES ? rank_max_slot(rv, fi[u], t) : rank_min_pixel(rv, fi[u], t)) : 0;
    }
    int32_t o[U];
#pragma unroll
    for (int u = 0; u < U; ++u) o[u] = root_out[rk[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = c0 + u * 256 + threadIdx.x;
      if (q < n) out[q] = label_of(o[u], out_base[fi[u]]);
    }
  }
}
void launch_labels_faces(const RankView& rv, int64_t S, const uint64_t* pj_face, const int32_t* root_out,
                         const int32_t* out_base, int32_t* out, hipStream_t st) {
  KScope ks_("labels_faces", st);
  if (S == 0) return;
  hipLaunchKernelGGL(k_labels<true>, dim3(grid_for((S + kLabelUnroll - 1) / kLabelUnroll, 256, 16384)), dim3(256), 0, st,
                     S, rv, pj_face, root_out, out_base, out);
  DMSC_LAUNCH_CHECK();
}
void launch_labels_verts(const RankView& rv, int64_t N, const uint64_t* pj_vert, const int32_t* root_out,
                         const int32_t* out_base, int32_t* out, hipStream_t st) {
  KScope ks_("labels_verts", st);
  if (N == 0) return;
  hipLaunchKernelGGL(k_labels<false>, dim3(grid_for((N + kLabelUnroll - 1) / kLabelUnroll, 256, 16384)), dim3(256), 0,
                     st, N, rv, pj_vert, root_out, out_base, out);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
