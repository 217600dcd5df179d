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
// scatter pass writes every critical cell's total contribution to its value pixel, into one of nine
// per-pixel scratch planes chosen by where the cell lies around that pixel (each entry has one
// writer: a saddle sums its two sides, an extremum dies in one pair when the pairs come from the
// unsimplified complex); the gather pass lets every pixel add its nine entries in a fixed order and
// zero them. The scratch is therefore all zeros between calls (no pass over the whole of it), and
// the result is deterministic.
#include "k_diff.hpp"

namespace dmsc {

namespace {

constexpr int kBwdThreads = 256;
constexpr int kPixPerBlock = 4 * kBwdThreads;  // each thread of the gather takes 4 pixels
constexpr int kSadPerBlock = kBwdThreads;

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
__device__ __forceinline__ int64_t cdiv_d(int64_t a, int64_t b) { return (a + b - 1) / b; }

// field of block b: the last field whose first block is <= b (block-uniform, from kernel arguments)
template <int Which>
__device__ __forceinline__ int chunk_field(const BwdChunk& c, int b) {
  int f = 0;
  for (int k = 1; k < c.n; ++k)
    if ((Which == 0 ? c.fld[k].blk_sad : c.fld[k].blk_pix) <= b) f = k;
  return f;
}

}  // namespace

// A cell id of an H x W field that is not a padding slot (vertices need y < H and x < W, type-1
// edges y < H, type-2 edges x < W; ids below 4 (H+1)(W+1) have y <= H): every such cell has at least
// one pixel among its vertices, so its value pixel exists.
__device__ __forceinline__ bool valid_cell(int cell, int H, int W) {
  const int Nx = W + 1, t = cell & 3, q = cell >> 2, y = q / Nx, x = q - y * Nx;
  return t == 3 || (t == 1 && y < H) || (t == 2 && x < W) || (t == 0 && y < H && x < W);
}

// Position of cell (t, cy, cx) among the nine cells that contain its value pixel (py, px), in the
// order the gather adds them: the vertex, the edges (y, x) and (y, x+1) of type 1, the edges (y, x)
// and (y+1, x) of type 2, the faces (y, x), (y, x+1), (y+1, x), (y+1, x+1).
__device__ __forceinline__ int star_slot(int t, int cy, int cx, int py, int px) {
  if (t == 0) return 0;
  if (t == 1) return px == cx ? 1 : 2;
  if (t == 2) return py == cy ? 3 : 4;
  return py == cy ? (px == cx ? 5 : 6) : (px == cx ? 7 : 8);
}

// Writes a critical cell's contribution into plane star_slot of its value pixel.
template <bool DUAL>
__device__ __forceinline__ void put(const BwdField& F, float* planes, uint32_t N, int cell, float v) {
  const int Nx = F.W + 1;
  const int t = cell & 3, q = cell >> 2;
  const int cy = q / Nx, cx = q - cy * Nx;
  const int vp = top_pixel<DUAL>(t, cy, cx, F.H, F.W, F.f);
  const int py = vp / F.W, px = vp - py * F.W;
  planes[(uint32_t)star_slot(t, cy, cx, py, px) * N + (uint32_t)vp] = v;
}

// Each critical cell's total contribution (a saddle's two sides summed), placed by its value pixel.
template <bool DUAL>
__global__ __launch_bounds__(kBwdThreads) void k_pers_scatter(BwdChunk c, float max_pers, float* __restrict__ scratch) {
  const int fi = chunk_field<0>(c, blockIdx.x);
  const BwdField& F = c.fld[fi];
  const int64_t k = (int64_t)(blockIdx.x - F.blk_sad) * kSadPerBlock + threadIdx.x;
  if (k >= F.n_sad) return;
  const int ncells = 4 * (F.H + 1) * (F.W + 1);
  const uint32_t N = (uint32_t)F.H * (uint32_t)F.W;
  float* planes = scratch + F.scratch_base;
  const float g = *F.gscale;
  float cs = 0.0f;
  const float pm = F.p_min[k];
  const int em = F.pp_min[k];
  if (em >= 0 && em < F.n_min && pm >= 0.0f && pm < max_pers) {
    const float v = 2.0f * pm * g;
    cs += v;
    const int cell = F.min_pts[em];
    if (cell >= 0 && cell < ncells && valid_cell(cell, F.H, F.W)) put<DUAL>(F, planes, N, cell, -v);
  }
  const float pM = F.p_max[k];
  const int eM = F.pp_max[k];
  if (eM >= 0 && eM < F.n_max && pM >= 0.0f && pM < max_pers) {
    const float v = 2.0f * pM * g;
    cs -= v;
    const int cell = F.max_pts[eM];
    if (cell >= 0 && cell < ncells && valid_cell(cell, F.H, F.W)) put<DUAL>(F, planes, N, cell, v);
  }
  const int sc = F.sad_pts[k];
  if (sc >= 0 && sc < ncells && valid_cell(sc, F.H, F.W)) put<DUAL>(F, planes, N, sc, cs);
}

// Pixel p adds its nine planes in order and zeroes what was set, so the scratch is all zeros again
// when the kernel ends. Adding an unset (zero) entry leaves the sum unchanged: it starts at +0 and
// can never become -0, since x + (-x) rounds to +0.
__global__ __launch_bounds__(kBwdThreads) void k_pers_gather(BwdChunk c, float* __restrict__ scratch) {
  const int fi = chunk_field<1>(c, blockIdx.x);
  const BwdField& F = c.fld[fi];
  const uint32_t N = (uint32_t)F.H * (uint32_t)F.W;
  float* planes = scratch + F.scratch_base;
  const uint32_t first = (uint32_t)(blockIdx.x - F.blk_pix) * kPixPerBlock + threadIdx.x;
  constexpr int R = kPixPerBlock / kBwdThreads;
  float v[R][9];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t p = first + (uint32_t)r * kBwdThreads;
#pragma unroll
    for (int k = 0; k < 9; ++k) v[r][k] = p < N ? planes[(uint32_t)k * N + p] : 0.0f;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t p = first + (uint32_t)r * kBwdThreads;
    if (p < N) {
      float acc = 0.0f;
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        acc += v[r][k];
        if (v[r][k] != 0.0f) planes[(uint32_t)k * N + p] = 0.0f;
      }
      F.grad[p] = acc;
    }
  }
}

// Partial loss sums: kLossSplit blocks per field of the chunk, block k of a field summing pers^2 over
// its k-th slice of the saddles in double (thread strides, then a tree over the block).
constexpr int kLossSplit = 64;
__global__ __launch_bounds__(kBwdThreads) void k_pers_loss_part(BwdChunk c, float max_pers, double* __restrict__ part) {
  const BwdField& F = c.fld[blockIdx.x / kLossSplit];
  const int split = blockIdx.x % kLossSplit;
  const int64_t len = cdiv_d(F.n_sad, kLossSplit);
  const int64_t k0 = split * len, k1 = min((int64_t)F.n_sad, k0 + len);
  double acc = 0.0;
  for (int64_t k = k0 + threadIdx.x; k < k1; k += kBwdThreads) {
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
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// loss of each field: its kLossSplit partial sums added in order
__global__ void k_pers_loss_sum(BwdChunk c, const double* __restrict__ part, float* __restrict__ loss) {
  const int fi = threadIdx.x;
  if (fi >= c.n) return;
  double acc = 0.0;
  for (int k = 0; k < kLossSplit; ++k) acc += part[fi * kLossSplit + k];
  loss[c.loss_base + fi] = (float)acc;
}

void bwd_assign_blocks(BwdChunk& c) {
  int bs = 0, bp = 0;
  int64_t cells = 0;
  for (int k = 0; k < c.n; ++k) {
    BwdField& F = c.fld[k];
    F.blk_sad = bs;
    F.blk_pix = bp;
    F.scratch_base = cells;
    bs += (int)cdiv(F.n_sad, kSadPerBlock);
    bp += (int)cdiv((int64_t)F.H * F.W, kPixPerBlock);
    cells += bwd_scratch_floats(F.H, F.W);
  }
  c.blocks_sad = bs;
  c.blocks_pix = bp;
  c.scratch = cells;
}

int64_t bwd_blocks_for(int H, int W, int64_t n_sad) {
  return std::max(cdiv((int64_t)H * W, kPixPerBlock), cdiv(n_sad, kSadPerBlock));
}

void launch_persistence_backward_chunk(bool dual, const BwdChunk& c, float max_pers, float* scratch, hipStream_t st) {
  KScope ks_("persistence_backward", st, true);
  if (c.blocks_sad > 0) {
    if (dual)
      hipLaunchKernelGGL(k_pers_scatter<true>, dim3(c.blocks_sad), dim3(kBwdThreads), 0, st, c, max_pers, scratch);
    else
      hipLaunchKernelGGL(k_pers_scatter<false>, dim3(c.blocks_sad), dim3(kBwdThreads), 0, st, c, max_pers, scratch);
    DMSC_LAUNCH_CHECK();
  }
  if (c.blocks_pix > 0) {
    hipLaunchKernelGGL(k_pers_gather, dim3(c.blocks_pix), dim3(kBwdThreads), 0, st, c, scratch);
    DMSC_LAUNCH_CHECK();
  }
}

int64_t loss_partials(int n_fields) { return (int64_t)n_fields * kLossSplit; }

void launch_persistence_loss_chunk(const BwdChunk& c, float max_pers, double* part, float* loss, hipStream_t st) {
  if (c.n <= 0) return;
  hipLaunchKernelGGL(k_pers_loss_part, dim3(c.n * kLossSplit), dim3(kBwdThreads), 0, st, c, max_pers, part);
  DMSC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pers_loss_sum, dim3(1), dim3(64), 0, st, c, (const double*)part, loss);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
