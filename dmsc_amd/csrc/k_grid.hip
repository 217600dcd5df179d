// Grid-shaped stages: discrete gradient + critical bit planes (one LDS-tiled stencil pass),
// order-preserving critical lists, saddle records and basin/peak labels.
#include "internal.hpp"

namespace dmsc {

__constant__ constexpr PairLut c_lut = make_pair_lut();

// ------------------------------------------------------------------ K1: gradient
// One 64 x 16 tile of slots per 256-thread block (a wave per slot row). Phase 1 evaluates the lower
// star of every vertex whose star touches the tile ((16+1) x (64+1) vertices, pixels staged in
// LDS with a 2-pixel halo) and scatters the pairing of each lower-star cell into an LDS cell tile;
// every valid cell lies in exactly one lower star, so there are no write conflicts. Phase 2 writes
// each slot's four partners as one 16-byte store, the face / vertex V-path successors, and the
// four critical bit planes via wave ballots. Phase 3 follows every face and vertex V-path inside
// the tile by pointer doubling in LDS (at most 1024 steps, so 10 doublings) and writes the initial
// pointer-jumping state: (in-tile root or first node outside the tile, steps to it). The global
// pointer jumping then moves one tile crossing per step instead of one cell.
constexpr int GT_X = 64, GT_Y = 16;
constexpr int GP_W = GT_X + 3, GP_H = GT_Y + 3;

template <bool DUAL>
__global__ __launch_bounds__(256) void k_gradient(Batch bt, const int64_t* __restrict__ tile_b,
                                                  int32_t* __restrict__ pw_all, int32_t* __restrict__ fsucc_all,
                                                  int32_t* __restrict__ vsucc_all, uint32_t* __restrict__ planes,
                                                  int64_t PW, uint64_t* __restrict__ pjf_all,
                                                  uint64_t* __restrict__ pjv_all) {
  __shared__ float pix[GP_H][GP_W];
  __shared__ int32_t cell[GT_Y][GT_X][4];  // reused by phase 3 as its second buffers
  __shared__ int32_t fJ[GT_Y * GT_X], fD[GT_Y * GT_X], vJ[GT_Y * GT_X], vD[GT_Y * GT_X];
  __shared__ uint8_t lut[kLutSize];
  const int tfield = field_of(tile_b, bt.B, blockIdx.x);
  const FieldDesc fd = bt.fd[tfield];
  const int64_t tlocal = blockIdx.x - tile_b[tfield];
  const int ntx = (fd.Nx + GT_X - 1) / GT_X;
  const TileRef tr{tfield, (int)(tlocal % ntx), (int)(tlocal / ntx)};
  const float* __restrict__ f = fd.f;
  const int H = fd.H, W = fd.W, WPR = fd.WPR;
  const int Nx = W + 1, Ny = H + 1;
  const int x0 = tr.tx * GT_X, y0 = tr.ty * GT_Y;
  const int sbase = (int)fd.slot_base, pbase = (int)fd.pix_base;  // batch totals fit int32 (host check)
  int32_t* __restrict__ pw = pw_all ? pw_all + 4 * fd.slot_base : nullptr;
  int32_t* __restrict__ fsucc = fsucc_all + fd.slot_base;
  int32_t* __restrict__ vsucc = vsucc_all + fd.pix_base;
  const int64_t wbase = fd.word_base;
  const int tid = threadIdx.x;

  for (int i = tid; i < kLutSize; i += 256) lut[i] = c_lut.v[i];
  for (int i = tid; i < GP_H * GP_W; i += 256) {
    int r = i / GP_W, c = i - r * GP_W;
    int py = y0 - 2 + r, px = x0 - 2 + c;
    pix[r][c] = (py >= 0 && py < H && px >= 0 && px < W) ? f[(int64_t)py * W + px] : 0.0f;
  }
  for (int i = tid; i < GT_Y * GT_X * 4; i += 256) (&cell[0][0][0])[i] = -1;
  __syncthreads();

  for (int i = tid; i < (GT_Y + 1) * (GT_X + 1); i += 256) {
    const int ly = i / (GT_X + 1), lx = i - ly * (GT_X + 1);
    const int vy = y0 - 1 + ly, vx = x0 - 1 + lx;
    if (vy < 0 || vy >= H || vx < 0 || vx >= W) continue;
    float nb[3][3];
    bool ex[3][3];
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        int py = vy + dy, px = vx + dx;
        ex[dy + 1][dx + 1] = (py >= 0 && py < H && px >= 0 && px < W);
        nb[dy + 1][dx + 1] = pix[ly + dy + 1][lx + dx + 1];
      }
    const LowerStar ls = lower_star<DUAL>(nb, ex);
    const int code = lut[ls.emask | ls.fmask << 4 | ls.estar << 8];
    const int q = vy * Nx + vx;
    // edge e of v: left H(y,x), right H(y,x+1), up V(y,x), down V(y+1,x); face j: UL, UR, DL, DR
    auto eid = [&](int e) { return 4 * (q + (e == 1 ? 1 : 0) + (e == 3 ? Nx : 0)) + (e < 2 ? 1 : 2); };
    auto fid = [&](int j) { return 4 * (q + (j & 1) + (j >> 1) * Nx) + 3; };
    const int sy = vy - y0, sx = vx - x0;
    if (sy >= 0 && sx >= 0) {
      int p = -1;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (ls.estar == e) p = eid(e);
      cell[sy][sx][0] = p;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (!(ls.emask >> e & 1)) continue;
      const int ty = sy + (e == 3 ? 1 : 0), tx = sx + (e == 1 ? 1 : 0);
      if (ty < 0 || ty >= GT_Y || tx < 0 || tx >= GT_X) continue;
      int partner = -1;
      if (e == ls.estar) partner = 4 * q;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int cj = (code >> (2 * j)) & 3;
        if (cj && face_edge(j, cj - 1) == e) partner = fid(j);
      }
      cell[ty][tx][e < 2 ? 1 : 2] = partner;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!(ls.fmask >> j & 1)) continue;
      const int ty = sy + (j >> 1), tx = sx + (j & 1);
      if (ty < 0 || ty >= GT_Y || tx < 0 || tx >= GT_X) continue;
      const int cj = (code >> (2 * j)) & 3;
      int partner = -1;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (cj && face_edge(j, cj - 1) == e) partner = eid(e);
      cell[ty][tx][3] = partner;
    }
  }
  __syncthreads();

  const int lane = tid & 63, wv = tid >> 6;
  for (int r = wv; r < GT_Y; r += 4) {
    const int y = y0 + r, x = x0 + lane;
    const bool in = (y < Ny) && (x < Nx);
    const int c0 = cell[r][lane][0], c1 = cell[r][lane][1], c2 = cell[r][lane][2], c3 = cell[r][lane][3];
    const bool k0 = in && c0 < 0 && y < H && x < W;
    const bool k1 = in && c1 < 0 && y < H;
    const bool k2 = in && c2 < 0 && x < W;
    const bool k3 = in && c3 < 0;
    // phase-3 state: local successor index inside the tile, -(global id) - 1 outside, self = root
    const int e = r * GT_X + lane;
    int fje = e, fde = 0, vje = e, vde = 0;
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
        fde = 1;
        fje = (rr >= 0 && rr < GT_Y && ll >= 0 && ll < GT_X) ? rr * GT_X + ll : -(sbase + fs) - 1;
      }
      fsucc[q] = sbase + fs;
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
          vde = 1;
          vje = (rr >= 0 && rr < GT_Y && ll >= 0 && ll < GT_X) ? rr * GT_X + ll : -(pbase + vs) - 1;
        }
        vsucc[p] = pbase + vs;
      }
    }
    fJ[e] = fje;
    fD[e] = fde;
    vJ[e] = vje;
    vD[e] = vde;
    const uint64_t b0 = __ballot(k0), b1 = __ballot(k1), b2 = __ballot(k2), b3 = __ballot(k3);
    if (y < Ny && lane < 8) {
      const int t = lane >> 1, half = lane & 1;
      const uint64_t bb = t == 0 ? b0 : t == 1 ? b1 : t == 2 ? b2 : b3;
      const int wi = (x0 >> 5) + half;
      if (wi < WPR) planes[t * PW + wbase + (int64_t)y * WPR + wi] = (uint32_t)(bb >> (32 * half));
    }
  }
  __syncthreads();

  // phase 3: pointer doubling inside the tile (double-buffered through the phase-1 cell tile)
  constexpr int TN = GT_Y * GT_X;
  int32_t* fJb = &cell[0][0][0];
  int32_t* fDb = fJb + TN;
  int32_t* vJb = fJb + 2 * TN;
  int32_t* vDb = fJb + 3 * TN;
  int32_t *fJa = fJ, *fDa = fD, *vJa = vJ, *vDa = vD;
#pragma unroll 1
  for (int it = 0; it < 10; ++it) {
    int moved = 0;  // a pass in which no pointer moves leaves every chain resolved
    for (int e = tid; e < TN; e += 256) {
      int j = fJa[e], d = fDa[e];
      if (j >= 0 && j != e) {
        const int jj = fJa[j];
        moved |= jj != j;
        d += fDa[j];
        j = jj;
      }
      fJb[e] = j;
      fDb[e] = d;
      j = vJa[e];
      d = vDa[e];
      if (j >= 0 && j != e) {
        const int jj = vJa[j];
        moved |= jj != j;
        d += vDa[j];
        j = jj;
      }
      vJb[e] = j;
      vDb[e] = d;
    }
    const int any = __syncthreads_or(moved);
    int32_t* t;
    t = fJa; fJa = fJb; fJb = t;
    t = fDa; fDa = fDb; fDb = t;
    t = vJa; vJa = vJb; vJb = t;
    t = vDa; vDa = vDb; vDb = t;
    if (!any) break;
  }
  uint64_t* __restrict__ pjf = pjf_all + fd.slot_base;
  uint64_t* __restrict__ pjv = pjv_all + fd.pix_base;
  for (int e = tid; e < TN; e += 256) {
    const int r = e / GT_X, c = e - r * GT_X;
    const int y = y0 + r, x = x0 + c;
    if (y >= Ny || x >= Nx) continue;
    const int q = y * Nx + x;
    int j = fJa[e];
    uint32_t target = j >= 0 ? (uint32_t)(sbase + (y0 + j / GT_X) * Nx + x0 + j % GT_X) : (uint32_t)(-j - 1);
    pjf[q] = (uint64_t)target | ((uint64_t)(uint32_t)fDa[e] << 32);
    if (y < H && x < W) {
      j = vJa[e];
      target = j >= 0 ? (uint32_t)(pbase + (y0 + j / GT_X) * W + x0 + j % GT_X) : (uint32_t)(-j - 1);
      pjv[y * W + x] = (uint64_t)target | ((uint64_t)(uint32_t)vDa[e] << 32);
    }
  }
}

void launch_gradient(bool dual, const Batch& bt, const int64_t* tile_b, int64_t ntiles, int32_t* pw, int32_t* fsucc,
                     int32_t* vsucc, uint32_t* planes, int64_t PW, uint64_t* pjf, uint64_t* pjv, hipStream_t st) {
  KScope ks_("gradient", st, true);
  if (ntiles == 0) return;
  if (dual)
    hipLaunchKernelGGL(k_gradient<true>, dim3((unsigned)ntiles), dim3(256), 0, st, bt, tile_b, pw, fsucc, vsucc, planes, PW,
                       pjf, pjv);
  else
    hipLaunchKernelGGL(k_gradient<false>, dim3((unsigned)ntiles), dim3(256), 0, st, bt, tile_b, pw, fsucc, vsucc, planes, PW,
                       pjf, pjv);
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
// entries, a wave prefix sum places them, the lanes write the entries' cell ids into an LDS stage,
// and the wave then writes the chunk's part of each list with consecutive lanes on consecutive
// entries (one lane per word would scatter 4-byte stores over 64 cache lines). An entry's field
// comes from its list position (per-field list bases), its value from its cell id.
constexpr int kCompactChunk = 64;  // words per wave
constexpr int kCompactWaves = 2;   // waves per block (the stage holds 2 saddles per slot: 16 KB a wave)

template <bool DUAL>
__global__ __launch_bounds__(64 * kCompactWaves) void k_compact_lists(RankView rv, const int32_t* __restrict__ list_b,
                                                       int32_t* __restrict__ min_ids, uint64_t* __restrict__ min_key,
                                                       int32_t* __restrict__ sad_ids, float* __restrict__ sad_val,
                                                       int32_t* __restrict__ sad_field, int32_t* __restrict__ max_ids,
                                                       uint64_t* __restrict__ max_key) {
  __shared__ int32_t stage[kCompactWaves][kCompactChunk * 32 * 2];  // up to 2 saddles per slot
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int B = rv.bt.B;
  const int64_t nchunks = (rv.PW + kCompactChunk - 1) / kCompactChunk;
  // block-uniform loop: every wave runs the same iterations (the block synchronises in between)
  for (int64_t cb = (int64_t)blockIdx.x * kCompactWaves; cb < nchunks; cb += (int64_t)gridDim.x * kCompactWaves) {
    const int64_t chunk = cb + wid;
    const int64_t w = chunk * kCompactChunk + lane;
    const bool valid = chunk < nchunks && w < rv.PW;
    uint32_t m[4] = {0, 0, 0, 0};
    int Nx = 0, xb = 0, yrow = 0;
    if (valid) {
      const int fi = field_of(rv.bt.word_b, B, w);
      const FieldDesc& fd = rv.bt.fd[fi];
      const int64_t lw = w - fd.word_base;
      yrow = (int)(lw / fd.WPR);
      xb = (int)(lw - (int64_t)yrow * fd.WPR) * 32;
      Nx = fd.Nx;
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
        const int q = yrow * Nx + xb + bit;
        if (list == 0) stage[wid][o++] = 4 * q;
        else if (list == 2) stage[wid][o++] = 4 * q + 3;
        else {
          if (ma >> bit & 1) stage[wid][o++] = 4 * q + 1;
          if (mb >> bit & 1) stage[wid][o++] = 4 * q + 2;
        }
      }
      __syncthreads();
      if (chunk < nchunks && total > 0) {
        const int32_t* pre = list == 0 ? rv.pre_min : list == 1 ? rv.pre_sad : rv.pre_max;
        const int32_t* lb = list_b + list * (B + 1);
        const int64_t base = pre[w0];
        for (int t = lane; t < total; t += 64) {
          const int id = stage[wid][t];
          const int64_t oi = base + t;
          const int fi = field_of64(lb, B, oi);
          const FieldDesc& fd = rv.bt.fd[fi];
          const int q = id >> 2, ty = q / fd.Nx, tx = q - ty * fd.Nx;
          if (list == 0) {
            min_ids[oi] = id;
            min_key[oi] = pack_key(fd.f[(int64_t)ty * fd.W + tx], id);
          } else if (list == 2) {
            max_ids[oi] = id;
            max_key[oi] = pack_key(fd.f[top_pixel<DUAL>(3, ty, tx, fd.H, fd.W, fd.f)], id);
          } else {
            sad_ids[oi] = id;
            sad_val[oi] = fd.f[top_pixel<DUAL>(id & 3, ty, tx, fd.H, fd.W, fd.f)];
            if (sad_field) sad_field[oi] = fi;
          }
        }
      }
      __syncthreads();
    }
  }
}

void launch_compact_lists(bool dual, const RankView& rv, const int32_t* list_b, int32_t* min_ids, uint64_t* min_key,
                          int32_t* sad_ids, float* sad_val, int32_t* sad_field, int32_t* max_ids, uint64_t* max_key,
                          hipStream_t st) {
  KScope ks_("compact_lists", st);
  if (rv.PW == 0) return;
  const int g = grid_for((rv.PW + kCompactChunk - 1) / kCompactChunk, kCompactWaves, 16384);
  if (dual)
    hipLaunchKernelGGL(k_compact_lists<true>, dim3(g), dim3(64 * kCompactWaves), 0, st, rv, list_b, min_ids, min_key, sad_ids, sad_val,
                       sad_field, max_ids, max_key);
  else
    hipLaunchKernelGGL(k_compact_lists<false>, dim3(g), dim3(64 * kCompactWaves), 0, st, rv, list_b, min_ids, min_key, sad_ids,
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
__global__ void k_labels_faces(int64_t S, RankView rv, const uint64_t* __restrict__ pj,
                               const int32_t* __restrict__ root_out, const int32_t* __restrict__ out_base,
                               int32_t* __restrict__ out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < S; q += (int64_t)gridDim.x * blockDim.x) {
    const int fi = field_of(rv.bt.slot_b, rv.bt.B, q);
    out[q] = label_of(root_out[rank_max_slot(rv, fi, (int)(uint32_t)pj[q])], out_base[fi]);
  }
}
__global__ void k_labels_verts(int64_t N, RankView rv, const uint64_t* __restrict__ pj,
                               const int32_t* __restrict__ root_out, const int32_t* __restrict__ out_base,
                               int32_t* __restrict__ out) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    const int fi = field_of(rv.bt.pix_b, rv.bt.B, p);
    out[p] = label_of(root_out[rank_min_pixel(rv, fi, (int)(uint32_t)pj[p])], out_base[fi]);
  }
}
void launch_labels_faces(const RankView& rv, int64_t S, const uint64_t* pj_face, const int32_t* root_out,
                         const int32_t* out_base, int32_t* out, hipStream_t st) {
  KScope ks_("labels_faces", st);
  if (S == 0) return;
  hipLaunchKernelGGL(k_labels_faces, dim3(grid_for(S, 256, 16384)), dim3(256), 0, st, S, rv, pj_face, root_out,
                     out_base, out);
  DMSC_LAUNCH_CHECK();
}
void launch_labels_verts(const RankView& rv, int64_t N, const uint64_t* pj_vert, const int32_t* root_out,
                         const int32_t* out_base, int32_t* out, hipStream_t st) {
  KScope ks_("labels_verts", st);
  if (N == 0) return;
  hipLaunchKernelGGL(k_labels_verts, dim3(grid_for(N, 256, 16384)), dim3(256), 0, st, N, rv, pj_vert, root_out,
                     out_base, out);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
