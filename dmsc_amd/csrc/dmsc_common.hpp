// Shared definitions for the gfx950 compute_dmsc kernels: cell codec, the (value, id) total order,
// and the lower-star face/edge pairing table.
//
// Cell complex (reference dmsc/csrc/cell_complex.hxx:13-51): slot q = y*Nx + x with Nx = W+1 over an
// (H+1) x (W+1) grid; cell id = 4*q + type, type 0 vertex (pixel, valid y<H,x<W), 1 H-edge (valid
// y<H), 2 V-edge (valid x<W), 3 face (always valid). Invalid padding cells are never critical and
// stay -1 in the gradient.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DMSC_HD __host__ __device__ __forceinline__

namespace dmsc {

// ------------------------------------------------------------------ (value, pixel id) order
// std::tie semantics of the reference (cpu/cell_compare.hxx:9-31): when neither a<b nor b<a (equal
// values, -0 == +0, or a NaN operand) the pixel id decides. `higher(a, ia, b, ib)` is
// value_greater<DUAL>: in the dual complex every comparison is reversed.
template <bool DUAL>
DMSC_HD bool higher(float a, int ia, float b, int ib) {
  if (!DUAL) return (b < a) || (!(a < b) && ib < ia);
  return (a < b) || (!(b < a) && ia < ib);
}

// Same relation when the id order of the two pixels is known statically (a_id_larger = id(a) > id(b)).
template <bool DUAL>
DMSC_HD bool higher_known(float a, float b, bool a_id_larger) {
  if (!DUAL) return a_id_larger ? !(a < b) : (b < a);
  return a_id_larger ? (a < b) : !(b < a);
}

// float -> uint32 whose unsigned order equals the float order; -0 is folded onto +0 (the reference
// compares them equal) and every NaN maps to one key above +inf.
DMSC_HD uint32_t sortable_key(float f) {
  if (f != f) return 0xFFFFFFFFu;
  if (f == 0.0f) f = 0.0f;
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// ------------------------------------------------------------------ value of a cell
// Pixel index of the "highest" vertex of cell (t, y, x) in the reference's candidate order
// (cpu/cell_compare.hxx:33-63); the cell value is the field value there.
template <bool DUAL>
__device__ __forceinline__ int top_pixel(int t, int y, int x, int H, int W, const float* __restrict__ f) {
  int best = -1;
  float bv = 0.f;
  auto consider = [&](int vy, int vx) {
    int id = vy * W + vx;
    float v = f[id];
    if (best == -1 || higher<DUAL>(v, id, bv, best)) {
      best = id;
      bv = v;
    }
  };
  if (t == 0) {
    consider(y, x);
  } else if (t == 1) {
    if (x - 1 >= 0) consider(y, x - 1);
    if (x < W) consider(y, x);
  } else if (t == 2) {
    if (y - 1 >= 0) consider(y - 1, x);
    if (y < H) consider(y, x);
  } else {
    if (y - 1 >= 0 && x - 1 >= 0) consider(y - 1, x - 1);
    if (y - 1 >= 0 && x < W) consider(y - 1, x);
    if (y < H && x - 1 >= 0) consider(y, x - 1);
    if (y < H && x < W) consider(y, x);
  }
  return best;
}

// An extremum's (value, cell id) in one 64-bit word (value bits low, id high): one load per compare.
DMSC_HD uint64_t pack_key(float v, int32_t id) {
  return (uint64_t)__builtin_bit_cast(uint32_t, v) | ((uint64_t)(uint32_t)id << 32);
}
DMSC_HD float key_value(uint64_t k) { return __builtin_bit_cast(float, (uint32_t)k); }
DMSC_HD int32_t key_id(uint64_t k) { return (int32_t)(k >> 32); }

// ------------------------------------------------------------------ lower-star pairing table
// Lower star of vertex v = (y, x) (reference cpu/gradient.hxx:55-74):
//   edges in the reference's L_edges order: 0 = H(y,x) "left", 1 = H(y,x+1) "right",
//                                           2 = V(y,x) "up",   3 = V(y+1,x) "down";
//   faces in L_faces order:  0 = F(y,x) "UL", 1 = F(y,x+1) "UR", 2 = F(y+1,x) "DL", 3 = F(y+1,x+1) "DR".
// Every incident edge and face of a valid vertex is a valid cell, so membership is the only
// variable. A face in the lower star always has both of its v-incident edges in it.
// Face j's v-incident edges, (first, second):
constexpr DMSC_HD int face_edge(int j, int which) {
  // UL: left, up   UR: up, right   DL: left, down   DR: right, down  (2-bit entries, packed)
  return ((which ? 0xF6 : 0x48) >> (2 * j)) & 3;
}
// The two faces bounding edge e, in the order the reference pushes them (gradient.hxx:174-181):
// H edge (ey,ex): F(ey,ex), F(ey+1,ex);  V edge (ey,ex): F(ey,ex), F(ey,ex+1).
constexpr DMSC_HD int edge_face(int e, int which) {
  // left: UL,DL  right: UR,DR  up: UL,UR  down: DL,DR  (2-bit entries, packed)
  return ((which ? 0xDE : 0x84) >> (2 * e)) & 3;
}

// Result of the reference FIFO face queue (gradient.hxx:115-191) for one lower star, as a byte:
// bits [2j, 2j+1] = 0 if face j stays unpaired, 1 if paired with face_edge(j,0), 2 with face_edge(j,1).
// index = edge_mask | face_mask << 4 | estar << 8, estar in 0..3 (edge paired with v) or 4 (none).
constexpr int kLutSize = 5 * 256;

struct PairLut {
  uint8_t v[kLutSize];
};

constexpr uint8_t lower_star_queue(int emask, int fmask, int estar) {
  int epaired[4] = {0, 0, 0, 0};
  int fpaired[4] = {0, 0, 0, 0};  // 0 = free, else 1 + which edge slot (0/1)
  if (estar < 4) epaired[estar] = 1;
  int q[16] = {};
  int qn = 0, head = 0;
  for (int j = 0; j < 4; ++j)
    if (fmask >> j & 1) q[qn++] = j;
  while (head < qn) {
    int f = q[head++];
    if (fpaired[f]) continue;
    int nfree = 0, freew = -1;
    for (int w = 0; w < 2; ++w) {
      int e = face_edge(f, w);
      if ((emask >> e & 1) && !epaired[e]) {
        ++nfree;
        freew = w;
      }
    }
    if (nfree != 1) continue;
    int e = face_edge(f, freew);
    fpaired[f] = 1 + freew;
    epaired[e] = 1;
    for (int w = 0; w < 2; ++w) {
      int nf = edge_face(e, w);
      if (nf == f || !(fmask >> nf & 1) || fpaired[nf]) continue;
      bool queued = false;
      for (int k = head; k < qn; ++k) queued = queued || (q[k] == nf);
      if (!queued && qn < 16) q[qn++] = nf;
    }
  }
  uint8_t out = 0;
  for (int j = 0; j < 4; ++j) out |= (uint8_t)(fpaired[j] << (2 * j));
  return out;
}

constexpr PairLut make_pair_lut() {
  PairLut t{};
  for (int es = 0; es < 5; ++es)
    for (int fm = 0; fm < 16; ++fm)
      for (int em = 0; em < 16; ++em) t.v[em | fm << 4 | es << 8] = lower_star_queue(em, fm, es);
  return t;
}

// ------------------------------------------------------------------ lower star of one vertex
// nb[dy+1][dx+1] = neighbour values, ex[..] = neighbour exists in the image. Returns
//   edge_mask (bits in L_edges order), face_mask (L_faces order), estar (0..3 or 4 = none).
// Every comparison follows the reference's candidate order (cell_compare.hxx:33-63 for
// "highest vertex", gradient.hxx:76-112 for the vertex pairing) so the result also matches when
// NaN makes the order non-transitive. Pixel ids of two neighbours compare as (dy, dx)
// lexicographically, which lets each comparison be a single float compare.
struct LowerStar {
  int emask, fmask, estar;
};

template <bool DUAL>
DMSC_HD LowerStar lower_star(const float nb[3][3], const bool ex[3][3]) {
  // Neighbour n at (dy, dx) has static rank r = 3*(dy+1) + (dx+1); ranks order like pixel ids.
  // All array indices below are compile-time constants; running "best" candidates are carried as
  // (value, rank) in registers, so nothing is indexed at run time.
  const float v = nb[1][1];
  auto hk = [](float a, int ra, float b, int rb) { return higher_known<DUAL>(a, b, ra > rb); };
  // v higher than neighbour of rank r (v has rank 4)
  const bool eL = ex[1][0], eR = ex[1][2], eU = ex[0][1], eD = ex[2][1];
  const bool gL = eL ? hk(v, 4, nb[1][0], 3) : true, gR = eR ? hk(v, 4, nb[1][2], 5) : true;
  const bool gU = eU ? hk(v, 4, nb[0][1], 1) : true, gD = eD ? hk(v, 4, nb[2][1], 7) : true;
  const int emask = (gL ? 1 : 0) | (gR ? 2 : 0) | (gU ? 4 : 0) | (gD ? 8 : 0);

  int fmask = 0;
  // UL F(y,x): candidates (-1,-1), (-1,0), (0,-1), then v; v wins iff it beats the running best.
  {
    bool have = false;
    float bv = 0.f;
    int br = 0;
    if (ex[0][0]) { have = true; bv = nb[0][0]; br = 0; }
    if (ex[0][1] && (!have || hk(nb[0][1], 1, bv, br))) { have = true; bv = nb[0][1]; br = 1; }
    if (ex[1][0] && (!have || hk(nb[1][0], 3, bv, br))) { have = true; bv = nb[1][0]; br = 3; }
    if (!have || hk(v, 4, bv, br)) fmask |= 1;
  }
  // UR F(y,x+1): (-1,0), (-1,1), v, (0,1)
  {
    bool have = false;
    float bv = 0.f;
    int br = 0;
    if (ex[0][1]) { have = true; bv = nb[0][1]; br = 1; }
    if (ex[0][2] && (!have || hk(nb[0][2], 2, bv, br))) { have = true; bv = nb[0][2]; br = 2; }
    const bool vin = !have || hk(v, 4, bv, br);
    if (vin && gR) fmask |= 2;
  }
  // DL F(y+1,x): (0,-1), v, (1,-1), (1,0)
  if (gL && (!ex[2][0] || hk(v, 4, nb[2][0], 6)) && gD) fmask |= 4;
  // DR F(y+1,x+1): v, (0,1), (1,0), (1,1)
  if (gR && gD && (!ex[2][2] || hk(v, 4, nb[2][2], 8))) fmask |= 8;

  // vertex pairing: among lower-star edges with a second vertex, in L_edges order (left, right,
  // up, down), keep the one whose other vertex is lowest (replace when the kept one is higher).
  int estar = 4;
  float sv = 0.f;
  int sr = 0;
  auto cand = [&](int e, bool exists, float ov, int orank) {
    if (!(emask >> e & 1) || !exists) return;
    if (estar == 4 || hk(sv, sr, ov, orank)) { estar = e; sv = ov; sr = orank; }
  };
  cand(0, eL, nb[1][0], 3);
  cand(1, eR, nb[1][2], 5);
  cand(2, eU, nb[0][1], 1);
  cand(3, eD, nb[2][1], 7);
  return LowerStar{emask, fmask, estar};
}

// ------------------------------------------------------------------ misc
DMSC_HD int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace dmsc
