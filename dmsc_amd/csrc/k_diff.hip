// Backward of persistence-based losses: d(loss)/d(field) for
//     loss = sum over pairs (saddle s, extremum e) with 0 <= pers < max_pers of pers^2,
// where pers = +-(f(px(s)) - f(px(e))) and px(c) is the pixel that carries the value of cell c
// (its highest vertex in the primal complex, its lowest in the dual; reference
// cpu/cell_compare.hxx:33-87). On either complex the min-side persistence rises with the saddle's
// value and falls with the minimum's; the max-side persistence rises with the maximum's value and
// falls with the saddle's. So a pair contributes +2*pers*g at one pixel and -2*pers*g at the other.
//
// No float atomics: pass 1 writes every critical cell's total contribution into a per-cell array
// (each cell has exactly one writer: a saddle sums its two sides, an extremum dies in one pair
// when the pairs come from the unsimplified complex); pass 2 lets every pixel add the
// contributions of the cells whose value it carries, in a fixed order. The result is therefore
// deterministic.
#include "internal.hpp"

namespace dmsc {


__global__ void k_pers_contrib(PairView pv, float max_pers, const float* __restrict__ gscale,
                               float* __restrict__ contrib, int64_t ncells) {
  const float g = *gscale;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < pv.n_sad; k += (int64_t)gridDim.x * blockDim.x) {
    float cs = 0.0f;
    const float pm = pv.p_min[k];
    const int em = pv.pp_min[k];
    if (em >= 0 && em < pv.n_min && pm >= 0.0f && pm < max_pers) {
      const float c = 2.0f * pm * g;
      cs += c;
      const int cell = pv.min_pts[em];
      if (cell >= 0 && cell < ncells) contrib[cell] = -c;
    }
    const float pM = pv.p_max[k];
    const int eM = pv.pp_max[k];
    if (eM >= 0 && eM < pv.n_max && pM >= 0.0f && pM < max_pers) {
      const float c = 2.0f * pM * g;
      cs -= c;
      const int cell = pv.max_pts[eM];
      if (cell >= 0 && cell < ncells) contrib[cell] = c;
    }
    const int sc = pv.sad_pts[k];
    if (sc >= 0 && sc < ncells) contrib[sc] = cs;
  }
}

// pixel (y, x): its vertex, the four edges and four faces that contain it, in that order
template <bool DUAL>
__global__ void k_pers_gather(const float* __restrict__ f, int H, int W, const float* __restrict__ contrib,
                              float* __restrict__ grad) {
  const int Nx = W + 1;
  const int64_t N = (int64_t)H * W;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / W), x = (int)(p - (int64_t)y * W);
    float acc = 0.0f;
    auto add = [&](int t, int cy, int cx) {
      const float c = contrib[4 * ((int64_t)cy * Nx + cx) + t];
      if (c != 0.0f && top_pixel<DUAL>(t, cy, cx, H, W, f) == (int)p) acc += c;
    };
    add(0, y, x);
    add(1, y, x);
    add(1, y, x + 1);
    add(2, y, x);
    add(2, y + 1, x);
    add(3, y, x);
    add(3, y, x + 1);
    add(3, y + 1, x);
    add(3, y + 1, x + 1);
    grad[p] = acc;
  }
}

void fill_f32_zero(float* p, int64_t n, hipStream_t st);

void launch_persistence_backward(bool dual, const float* f, int H, int W, const PairView& pv, float max_pers,
                                 const float* gscale, float* contrib, float* grad, hipStream_t st) {
  KScope ks_("persistence_backward", st, true);
  const int64_t S = (int64_t)(H + 1) * (W + 1), N = (int64_t)H * W;
  fill_f32_zero(contrib, 4 * S, st);
  if (pv.n_sad > 0) {
    hipLaunchKernelGGL(k_pers_contrib, dim3(grid_for(pv.n_sad, 256, 16384)), dim3(256), 0, st, pv, max_pers, gscale,
                       contrib, 4 * S);
    DMSC_LAUNCH_CHECK();
  }
  if (dual)
    hipLaunchKernelGGL(k_pers_gather<true>, dim3(grid_for(N, 256, 16384)), dim3(256), 0, st, f, H, W, contrib, grad);
  else
    hipLaunchKernelGGL(k_pers_gather<false>, dim3(grid_for(N, 256, 16384)), dim3(256), 0, st, f, H, W, contrib, grad);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
