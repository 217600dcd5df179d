// Persistence losses and their backward: for each field,
//     loss = sum over pairs (saddle s, extremum e) with 0 <= pers < max_pers of pers^2,
// where pers = +-(f(px(s)) - f(px(e))) and px(c) is the pixel that carries the value of cell c
// (its highest vertex in the primal complex, its lowest in the dual; reference
// cpu/cell_compare.hxx:33-87). On either complex the min-side persistence rises with the saddle's
// value and falls with the minimum's; the max-side persistence rises with the maximum's value and
// falls with the saddle's. So a pair contributes +2*pers*g at one pixel and -2*pers*g at the other.
//
// Fields go through the kernels in chunks of up to kBwdChunk, their descriptors passed by value;
// each kernel finds its field from block-uniform per-field block bases. No float atomics: the
// scatter pass writes every critical cell's total contribution into a per-cell scratch array (each
// cell has one writer: a saddle sums its two sides, an extremum dies in one pair when the pairs
// come from the unsimplified complex); the gather pass lets every pixel add the contributions of
// the cells whose value it carries, in a fixed order; the clear pass writes zeros back at exactly
// the cells the scatter wrote. The scratch is therefore all zeros between calls and no pass over
// the whole cell array is needed, and the result is deterministic.
#include "k_diff.hpp"

namespace dmsc {

namespace {

constexpr int kBwdThreads = 256;
constexpr int kPixPerBlock = 4 * kBwdThreads;  // each thread of the gather takes 4 pixels
constexpr int kSadPerBlock = kBwdThreads;

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// field of block b: the last field whose first block is <= b (block-uniform, from kernel arguments)
template <int Which>
__device__ __forceinline__ int chunk_field(const BwdChunk& c, int b) {
  int f = 0;
  for (int k = 1; k < c.n; ++k)
    if ((Which == 0 ? c.fld[k].blk_sad : c.fld[k].blk_pix) <= b) f = k;
  return f;
}

}  // namespace

// CLEAR = false: write each critical cell's contribution; CLEAR = true: write zeros at those cells
template <bool CLEAR>
__global__ __launch_bounds__(kBwdThreads) void k_pers_scatter(BwdChunk c, float max_pers, float* __restrict__ scratch) {
  const int fi = chunk_field<0>(c, blockIdx.x);
  const BwdField& F = c.fld[fi];
  const int64_t k = (int64_t)(blockIdx.x - F.blk_sad) * kSadPerBlock + threadIdx.x;
  if (k >= F.n_sad) return;
  const int64_t ncells = 4 * (int64_t)(F.H + 1) * (F.W + 1);
  float* contrib = scratch + F.cell_base;
  const float g = CLEAR ? 0.0f : *F.gscale;
  float cs = 0.0f;
  const float pm = F.p_min[k];
  const int em = F.pp_min[k];
  if (em >= 0 && em < F.n_min && pm >= 0.0f && pm < max_pers) {
    const float v = 2.0f * pm * g;
    cs += v;
    const int cell = F.min_pts[em];
    if (cell >= 0 && cell < ncells) contrib[cell] = CLEAR ? 0.0f : -v;
  }
  const float pM = F.p_max[k];
  const int eM = F.pp_max[k];
  if (eM >= 0 && eM < F.n_max && pM >= 0.0f && pM < max_pers) {
    const float v = 2.0f * pM * g;
    cs -= v;
    const int cell = F.max_pts[eM];
    if (cell >= 0 && cell < ncells) contrib[cell] = CLEAR ? 0.0f : v;
  }
  const int sc = F.sad_pts[k];
  if (sc >= 0 && sc < ncells) contrib[sc] = CLEAR ? 0.0f : cs;
}

// pixel (y, x): its vertex, the four edges and four faces that contain it, in that order
template <bool DUAL>
__global__ __launch_bounds__(kBwdThreads) void k_pers_gather(BwdChunk c, const float* __restrict__ scratch) {
  const int fi = chunk_field<1>(c, blockIdx.x);
  const BwdField& F = c.fld[fi];
  const int H = F.H, W = F.W, Nx = W + 1;
  const int64_t N = (int64_t)H * W;
  const float* __restrict__ f = F.f;
  const float* __restrict__ contrib = scratch + F.cell_base;
  const int64_t first = (int64_t)(blockIdx.x - F.blk_pix) * kPixPerBlock + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kPixPerBlock / kBwdThreads; ++r) {
    const int64_t p = first + (int64_t)r * kBwdThreads;
    if (p >= N) break;
    const int y = (int)(p / W), x = (int)(p - (int64_t)y * W);
    float acc = 0.0f;
    auto add = [&](int t, int cy, int cx) {
      const float v = contrib[4 * ((int64_t)cy * Nx + cx) + t];
      if (v != 0.0f && top_pixel<DUAL>(t, cy, cx, H, W, f) == (int)p) acc += v;
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
    F.grad[p] = acc;
  }
}

// loss of each field of the chunk: one block per field sums pers^2 over its pairs in double, in a
// fixed order (thread strides, then a tree over the block), so the result is reproducible
__global__ __launch_bounds__(kBwdThreads) void k_pers_loss(BwdChunk c, float max_pers, float* __restrict__ loss) {
  const BwdField& F = c.fld[blockIdx.x];
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < F.n_sad; k += kBwdThreads) {
    const float pm = F.p_min[k];
    if (F.pp_min[k] >= 0 && pm >= 0.0f && pm < max_pers) acc += (double)pm * (double)pm;
    const float pM = F.p_max[k];
    if (F.pp_max[k] >= 0 && pM >= 0.0f && pM < max_pers) acc += (double)pM * (double)pM;
  }
  __shared__ double s[kBwdThreads];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBwdThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[c.loss_base + blockIdx.x] = (float)s[0];
}

void bwd_assign_blocks(BwdChunk& c) {
  int bs = 0, bp = 0;
  int64_t cells = 0;
  for (int k = 0; k < c.n; ++k) {
    BwdField& F = c.fld[k];
    F.blk_sad = bs;
    F.blk_pix = bp;
    F.cell_base = cells;
    bs += (int)cdiv(F.n_sad, kSadPerBlock);
    bp += (int)cdiv((int64_t)F.H * F.W, kPixPerBlock);
    cells += 4 * (int64_t)(F.H + 1) * (F.W + 1);
  }
  c.blocks_sad = bs;
  c.blocks_pix = bp;
  c.cells = cells;
}

int64_t bwd_blocks_for(int H, int W, int64_t n_sad) {
  return std::max(cdiv((int64_t)H * W, kPixPerBlock), cdiv(n_sad, kSadPerBlock));
}

void launch_persistence_backward_chunk(bool dual, const BwdChunk& c, float max_pers, float* scratch, hipStream_t st) {
  KScope ks_("persistence_backward", st, true);
  if (c.blocks_sad > 0) {
    hipLaunchKernelGGL(k_pers_scatter<false>, dim3(c.blocks_sad), dim3(kBwdThreads), 0, st, c, max_pers, scratch);
    DMSC_LAUNCH_CHECK();
  }
  if (c.blocks_pix > 0) {
    if (dual)
      hipLaunchKernelGGL(k_pers_gather<true>, dim3(c.blocks_pix), dim3(kBwdThreads), 0, st, c,
                         (const float*)scratch);
    else
      hipLaunchKernelGGL(k_pers_gather<false>, dim3(c.blocks_pix), dim3(kBwdThreads), 0, st, c,
                         (const float*)scratch);
    DMSC_LAUNCH_CHECK();
  }
  if (c.blocks_sad > 0) {
    hipLaunchKernelGGL(k_pers_scatter<true>, dim3(c.blocks_sad), dim3(kBwdThreads), 0, st, c, max_pers, scratch);
    DMSC_LAUNCH_CHECK();
  }
}

void launch_persistence_loss_chunk(const BwdChunk& c, float max_pers, float* loss, hipStream_t st) {
  if (c.n <= 0) return;
  hipLaunchKernelGGL(k_pers_loss, dim3(c.n), dim3(kBwdThreads), 0, st, c, max_pers, loss);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
