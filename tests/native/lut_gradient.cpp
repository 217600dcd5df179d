// Host build of the kernel's lower-star formulation (dmsc_amd/csrc/dmsc_common.hpp: per-vertex
// comparisons + the compile-time FIFO-queue table), scattered into a full paired_with array the way
// the gradient kernel's phase 1 does. tests/test_gradient_lut.py compares it with the oracle's literal
// queue restatement on exhaustive and random small fields. Test helper only.
#include <cstdint>

#include "dmsc_common.hpp"

using namespace dmsc;

static constexpr PairLut kLut = make_pair_lut();

template <bool DUAL>
static void run(const float* f, int H, int W, int32_t* pw) {
  const int Nx = W + 1;
  const int64_t C = 4LL * (H + 1) * (W + 1);
  for (int64_t i = 0; i < C; ++i) pw[i] = -1;
  for (int vy = 0; vy < H; ++vy)
    for (int vx = 0; vx < W; ++vx) {
      float nb[3][3];
      bool ex[3][3];
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int py = vy + dy, px = vx + dx;
          bool e = py >= 0 && py < H && px >= 0 && px < W;
          ex[dy + 1][dx + 1] = e;
          nb[dy + 1][dx + 1] = e ? f[py * W + px] : 0.0f;
        }
      LowerStar ls = lower_star<DUAL>(nb, ex);
      int code = kLut.v[ls.emask | ls.fmask << 4 | ls.estar << 8];
      int q = vy * Nx + vx;
      int eid[4] = {4 * q + 1, 4 * (q + 1) + 1, 4 * q + 2, 4 * (q + Nx) + 2};
      int fid[4] = {4 * q + 3, 4 * (q + 1) + 3, 4 * (q + Nx) + 3, 4 * (q + Nx + 1) + 3};
      pw[4 * q] = ls.estar < 4 ? eid[ls.estar] : -1;
      for (int e = 0; e < 4; ++e) {
        if (!(ls.emask >> e & 1)) continue;
        int partner = e == ls.estar ? 4 * q : -1;
        for (int j = 0; j < 4; ++j) {
          int cj = (code >> (2 * j)) & 3;
          if (cj && face_edge(j, cj - 1) == e) partner = fid[j];
        }
        pw[eid[e]] = partner;
      }
      for (int j = 0; j < 4; ++j) {
        if (!(ls.fmask >> j & 1)) continue;
        int cj = (code >> (2 * j)) & 3;
        pw[fid[j]] = cj ? eid[face_edge(j, cj - 1)] : -1;
      }
    }
}

extern "C" void lut_gradient(const float* f, int H, int W, int dual, int32_t* pw) {
  if (dual) run<true>(f, H, W, pw);
  else run<false>(f, H, W, pw);
}
