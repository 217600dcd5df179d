// Device-wide primitives written for wave64: exclusive scans (reduce / scan-of-sums / down-sweep),
// a stable LSD radix sort of (uint32 key, int32 value) pairs, and asynchronous pointer jumping
// over successor forests.
#include "internal.hpp"

namespace dmsc {

constexpr int SCAN_THREADS = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// exclusive scan of one value per thread over a 256-thread block
template <class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds4, T& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  T inc = wave_inclusive_scan(v);
  if (lane == 63) lds4[wid] = inc;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    T s = lds4[w];
    if (w < wid) base += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + inc - v;
}

// Scan inputs. Each thread of a tile owns SCAN_ITEMS consecutive elements; load8 fetches them with
// 16-byte loads when the source is 16-byte aligned and the run is whole (vec is set by the host).
static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <class T>
struct LoadArr {
  const T* p;
  bool vec;
  __device__ T operator()(int64_t i) const { return p[i]; }
  __device__ void load8(int64_t base, int64_t n, T (&v)[8]) const {
    if (vec && base + 8 <= n) {
      if constexpr (sizeof(T) == 4) {
        const int4 a = *reinterpret_cast<const int4*>(p + base), b = *reinterpret_cast<const int4*>(p + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
      } else {
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
          const longlong2 a = *reinterpret_cast<const longlong2*>(p + base + j);
          v[j] = a.x;
          v[j + 1] = a.y;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = base + j < n ? p[base + j] : (T)0;
    }
  }
};
struct LoadPop {
  const uint32_t* a;
  const uint32_t* b;
  bool vec;
  __device__ int32_t operator()(int64_t i) const { return __popc(a[i]) + (b ? __popc(b[i]) : 0); }
  __device__ void load8(int64_t base, int64_t n, int32_t (&v)[8]) const {
    if (vec && base + 8 <= n) {
      const uint4 x = *reinterpret_cast<const uint4*>(a + base), y = *reinterpret_cast<const uint4*>(a + base + 4);
      v[0] = __popc(x.x); v[1] = __popc(x.y); v[2] = __popc(x.z); v[3] = __popc(x.w);
      v[4] = __popc(y.x); v[5] = __popc(y.y); v[6] = __popc(y.z); v[7] = __popc(y.w);
      if (b) {
        const uint4 z = *reinterpret_cast<const uint4*>(b + base), w = *reinterpret_cast<const uint4*>(b + base + 4);
        v[0] += __popc(z.x); v[1] += __popc(z.y); v[2] += __popc(z.z); v[3] += __popc(z.w);
        v[4] += __popc(w.x); v[5] += __popc(w.y); v[6] += __popc(w.z); v[7] += __popc(w.w);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = base + j < n ? (*this)(base + j) : 0;
    }
  }
};
// Raw arc i = 2 s + k (length blen[i]) counts when it lies in the arena: longer than min_len - 1 cells
// and used by the outputs (saddle s survives, or this side cancels it).
// A thread's 8 arcs are the arcs of 4 consecutive saddles: with aligned arrays, two 16-byte loads of
// lengths, one 4-byte load of removal flags and one 16-byte load of cancellations.
struct LoadArcLen {
  const int32_t* blen;
  const uint8_t* removed;
  const int32_t* cancel;
  int32_t min_len;
  bool vec;
  __device__ int64_t keep(int32_t len, uint32_t rm, int32_t cd) const {
    return (len >= min_len && (!rm || cd >= 0)) ? (int64_t)len : 0;
  }
  __device__ int64_t operator()(int64_t i) const { return keep(blen[i], removed[i >> 1], cancel[i >> 1]); }
  __device__ void load8(int64_t base, int64_t n, int64_t (&v)[8]) const {
    if (vec && base + 8 <= n) {
      const int4 a = *reinterpret_cast<const int4*>(blen + base), b = *reinterpret_cast<const int4*>(blen + base + 4);
      const uint32_t rm = *reinterpret_cast<const uint32_t*>(removed + (base >> 1));
      const int4 cd = *reinterpret_cast<const int4*>(cancel + (base >> 1));
      v[0] = keep(a.x, rm & 0xFF, cd.x); v[1] = keep(a.y, rm & 0xFF, cd.x);
      v[2] = keep(a.z, rm >> 8 & 0xFF, cd.y); v[3] = keep(a.w, rm >> 8 & 0xFF, cd.y);
      v[4] = keep(b.x, rm >> 16 & 0xFF, cd.z); v[5] = keep(b.y, rm >> 16 & 0xFF, cd.z);
      v[6] = keep(b.z, rm >> 24, cd.w); v[7] = keep(b.w, rm >> 24, cd.w);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = base + j < n ? (*this)(base + j) : 0;
    }
  }
};
// 0/1 flags from a byte array (nonzero = 1); 8-byte loads when aligned
struct LoadU8 {
  const uint8_t* p;
  bool vec;
  __device__ int32_t operator()(int64_t i) const { return p[i] ? 1 : 0; }
  __device__ void load8(int64_t base, int64_t n, int32_t (&v)[8]) const {
    if (vec && base + 8 <= n) {
      const uint2 w = *reinterpret_cast<const uint2*>(p + base);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = (w.x >> (8 * j)) & 0xFF ? 1 : 0;
        v[4 + j] = (w.y >> (8 * j)) & 0xFF ? 1 : 0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = base + j < n ? (*this)(base + j) : 0;
    }
  }
};
static_assert(SCAN_ITEMS == 8, "load8 / store8 move 8 elements per thread");

// out[base + j] = v[j] for the thread's run, with 16-byte stores when allowed
template <class T>
__device__ __forceinline__ void store8(T* __restrict__ out, int64_t base, int64_t n, const T (&v)[8], bool vec) {
  if (vec && base + 8 <= n) {
    if constexpr (sizeof(T) == 4) {
      *reinterpret_cast<int4*>(out + base) = make_int4(v[0], v[1], v[2], v[3]);
      *reinterpret_cast<int4*>(out + base + 4) = make_int4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
      for (int j = 0; j < 8; j += 2) {
        longlong2 a;
        a.x = v[j];
        a.y = v[j + 1];
        *reinterpret_cast<longlong2*>(out + base + j) = a;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (base + j < n) out[base + j] = v[j];
  }
}

template <class T, class Load>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(Load ld, int64_t n, T* __restrict__ bsums) {
  __shared__ T lds4[4];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  ld.load8(base, n, v);
  T s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) s += v[j];
  T tot;
  block_exclusive_scan(s, lds4, tot);
  if (threadIdx.x == 0) bsums[blockIdx.x] = tot;
}

// single block: in-place exclusive scan of the block sums; grand total to *total
template <class T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(T* __restrict__ bsums, int64_t nb, T* __restrict__ total) {
  __shared__ T lds4[4];
  T carry = 0;
  for (int64_t c = 0; c < nb; c += SCAN_TILE) {
    const int64_t base = c + threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
      v[j] = (base + j < nb) ? bsums[base + j] : (T)0;
      s += v[j];
    }
    T tot;
    T ex = block_exclusive_scan(s, lds4, tot) + carry;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
      if (base + j < nb) bsums[base + j] = ex;
      ex += v[j];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class T, class Load>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(Load ld, int64_t n, const T* __restrict__ bsums,
                                                            T* __restrict__ out, bool vec_out) {
  __shared__ T lds4[4];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  ld.load8(base, n, v);
  T s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) s += v[j];
  T tot;
  T ex = block_exclusive_scan(s, lds4, tot) + bsums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const T x = v[j];
    v[j] = ex;
    ex += x;
  }
  store8(out, base, n, v, vec_out);
}

// Down-sweep that adds up the block sums before its own tile itself (no separate scan of the block
// sums: for a few thousand tiles that costs each block a few microseconds of L2 reads, less than a
// launch); the last block writes the grand total.
template <class T, class Load>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down2(Load ld, int64_t n, const T* __restrict__ bsums,
                                                             T* __restrict__ out, T* __restrict__ total,
                                                             bool vec_out) {
  __shared__ T lds4[4];
  T pre = 0;
  for (int64_t b = threadIdx.x; b < blockIdx.x; b += SCAN_THREADS) pre += bsums[b];
  T base_sum;
  block_exclusive_scan(pre, lds4, base_sum);
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  ld.load8(base, n, v);
  T s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) s += v[j];
  T tot;
  T ex = block_exclusive_scan(s, lds4, tot) + base_sum;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const T x = v[j];
    v[j] = ex;
    ex += x;
  }
  store8(out, base, n, v, vec_out);
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = base_sum + tot;
}
constexpr int64_t kScanFewTiles = 4096;  // up to this many tiles: reduce + down2 (two launches)

template <class T, class Load>
static void scan_impl(Load ld, T* out, int64_t n, T* d_total, Arena& ar, int slot, hipStream_t st) {
  KScope ks_("scan", st);
  const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  T* bsums = ar.get<T>(slot, (size_t)(nb > 0 ? nb : 1));
  if (nb <= kScanFewTiles) {
    if (n > 0) {
      hipLaunchKernelGGL((k_scan_reduce<T, Load>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, ld, n, bsums);
      DMSC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_scan_down2<T, Load>), dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(SCAN_THREADS), 0, st, ld,
                       n, bsums, out, d_total, aligned16(out));
    DMSC_LAUNCH_CHECK();
    return;
  }
  if (n > 0) {
    hipLaunchKernelGGL((k_scan_reduce<T, Load>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, ld, n, bsums);
    DMSC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(SCAN_THREADS), 0, st, bsums, nb, d_total);
  DMSC_LAUNCH_CHECK();
  if (n > 0) {
    hipLaunchKernelGGL((k_scan_down<T, Load>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, ld, n, bsums, out,
                       aligned16(out));
    DMSC_LAUNCH_CHECK();
  }
}

template <class T>
void scan_exclusive(const T* in, T* out, int64_t n, T* d_total, Arena& ar, int slot, hipStream_t st) {
  scan_impl<T>(LoadArr<T>{in, aligned16(in)}, out, n, d_total, ar, slot, st);
}
template void scan_exclusive<int32_t>(const int32_t*, int32_t*, int64_t, int32_t*, Arena&, int, hipStream_t);
template void scan_exclusive<int64_t>(const int64_t*, int64_t*, int64_t, int64_t*, Arena&, int, hipStream_t);

void scan_plane_popcounts(const uint32_t* plane_a, const uint32_t* plane_b, int32_t* out, int64_t n, int32_t* d_total,
                          Arena& ar, int slot, hipStream_t st) {
  scan_impl<int32_t>(LoadPop{plane_a, plane_b, aligned16(plane_a) && (!plane_b || aligned16(plane_b))}, out, n,
                     d_total, ar, slot, st);
}

void scan_u8_flags(const uint8_t* in, int32_t* out, int64_t n, int32_t* d_total, Arena& ar, int slot, hipStream_t st) {
  scan_impl<int32_t>(LoadU8{in, (reinterpret_cast<uintptr_t>(in) & 7) == 0}, out, n, d_total, ar, slot, st);
}

void scan_arc_offsets(const int32_t* blen, const uint8_t* removed, const int32_t* cancel, int32_t min_len, int64_t* out,
                      int64_t n, int64_t* d_total, Arena& ar, int slot, hipStream_t st) {
  const bool vec = aligned16(blen) && aligned16(cancel) && (reinterpret_cast<uintptr_t>(removed) & 3) == 0;
  scan_impl<int64_t>(LoadArcLen{blen, removed, cancel, min_len, vec}, out, n, d_total, ar, slot, st);
}

// ------------------------------------------------------------------ radix sort
// Stable LSD sort, 8-bit digits, up to 4 passes. Per pass: per-tile digit histograms (digit-major),
// one exclusive scan, then a scatter that ranks equal digits in element order: chunks of 256 in
// order, within a chunk wave-level peer masks from 8 ballots, waves in order via LDS counts.
// Between passes the pairs travel packed as one 64-bit word (key << 32 | value): one scattered
// 8-byte store per element instead of two 4-byte stores to two arrays.
constexpr int RS_THREADS = 256, RS_ITEMS = 16, RS_TILE = RS_THREADS * RS_ITEMS;

template <bool PK>
__device__ __forceinline__ uint32_t rs_key(const void* in, int64_t i) {
  if (PK) return (uint32_t)(static_cast<const uint64_t*>(in)[i] >> 32);
  return static_cast<const uint32_t*>(in)[i];
}

template <bool PK>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const void* __restrict__ keys, int64_t n, int shift,
                                                        int32_t* __restrict__ counts, int64_t nb) {
  __shared__ int32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int j = 0; j < RS_ITEMS; ++j) {
    int64_t i = base + j * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(rs_key<PK>(keys, i) >> shift) & 255], 1);
  }
  __syncthreads();
  counts[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// PIN / POUT: input / output packed (one uint64 array) or separate (key array, value array).
// RANK: no sorted output; vout[value] = output position instead (the inverse permutation).
// A tile is RS_ITEMS chunks of 256 elements in order; chunks are ranked RS_GROUP at a time: each wave
// ranks its 64 elements of each chunk (peer masks from 8 ballots) and records per-digit counts per
// (chunk, wave); one thread per digit then turns the counts into output offsets in element order;
// four barriers per group of chunks.
constexpr int RS_GROUP = 4;
static_assert(RS_ITEMS % RS_GROUP == 0, "whole groups of chunks per tile");
template <bool PIN, bool POUT, bool RANK = false>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const void* __restrict__ kin, const int32_t* __restrict__ vin,
                                                           void* __restrict__ kout, int32_t* __restrict__ vout,
                                                           int64_t n, int shift, const int32_t* __restrict__ offs,
                                                           int64_t nb) {
  __shared__ int32_t next_off[256];                  // per digit: output offset of the next element
  __shared__ int32_t slot[RS_GROUP][4][256];         // per (chunk, wave, digit): count, then offset
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  next_off[tid] = offs[(int64_t)tid * nb + blockIdx.x];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int g = 0; g < RS_ITEMS; g += RS_GROUP) {
#pragma unroll
    for (int c = 0; c < RS_GROUP; ++c)
#pragma unroll
      for (int w = 0; w < 4; ++w) slot[c][w][tid] = 0;
    __syncthreads();
    uint32_t k[RS_GROUP];
    int32_t v[RS_GROUP];
    int rank[RS_GROUP];
#pragma unroll
    for (int c = 0; c < RS_GROUP; ++c) {
      const int64_t i = base + (int64_t)(g + c) * RS_THREADS + tid;
      const bool valid = i < n;
      k[c] = 0u;
      v[c] = 0;
      if (valid) {
        if (PIN) {
          const uint64_t p = static_cast<const uint64_t*>(kin)[i];
          k[c] = (uint32_t)(p >> 32);
          v[c] = (int32_t)(uint32_t)p;
        } else {
          k[c] = static_cast<const uint32_t*>(kin)[i];
          v[c] = vin ? vin[i] : (int32_t)i;  // no value array: the values are the indices
        }
      }
      const int d = (k[c] >> shift) & 255;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      rank[c] = valid ? __popcll(peers & lt) : -1;
      if (valid && rank[c] == 0) slot[c][wid][d] = __popcll(peers);
    }
    __syncthreads();
    {  // thread tid owns digit tid: offsets in (chunk, wave) order
      int32_t run = next_off[tid];
#pragma unroll
      for (int c = 0; c < RS_GROUP; ++c)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          const int32_t t = slot[c][w][tid];
          slot[c][w][tid] = run;
          run += t;
        }
      next_off[tid] = run;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < RS_GROUP; ++c) {
      if (rank[c] < 0) continue;
      const int d = (k[c] >> shift) & 255;
      const int pos = slot[c][wid][d] + rank[c];
      if (RANK) {
        vout[v[c]] = pos;
      } else if (POUT) {
        static_cast<uint64_t*>(kout)[pos] = ((uint64_t)k[c] << 32) | (uint32_t)v[c];
      } else {
        static_cast<uint32_t*>(kout)[pos] = k[c];
        vout[pos] = v[c];
      }
    }
    __syncthreads();
  }
}

__global__ void k_copy_pairs(const uint32_t* __restrict__ ks, const int32_t* __restrict__ vs, uint32_t* __restrict__ kd,
                             int32_t* __restrict__ vd, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

// Sorts on the low `bits` bits of the keys (ceil(bits / 8) passes); the result is in (keys, vals).
// iota_vals: the input values are 0 .. n-1 (vals is then only written). Scratch: arena slots `slot`
// (counts), slot + 1 (their scan), pslot and pslot + 1 (packed pairs).
void radix_sort_pairs_bits(uint32_t* keys, int32_t* vals, uint32_t* keys_tmp, int32_t* vals_tmp, int64_t n, int bits,
                           Arena& ar, int slot, int pslot, hipStream_t st, bool iota_vals) {
  KScope ks_("radix_sort", st);
  if (n <= 1 || bits <= 0) {
    if (iota_vals) iota_i32(vals, n, st);
    return;
  }
  const int32_t* const vals_in = iota_vals ? nullptr : vals;
  const int npass = (bits + 7) / 8 > 4 ? 4 : (bits + 7) / 8;
  const int64_t nb = (n + RS_TILE - 1) / RS_TILE;
  int32_t* counts = ar.get<int32_t>(slot, (size_t)(256 * nb));
  auto pass = [&](int p, const void* kin, const int32_t* vin, void* kout, int32_t* vout, bool pin, bool pout) {
    const int shift = 8 * p;
    if (pin) hipLaunchKernelGGL(k_rs_hist<true>, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, kin, n, shift, counts, nb);
    else hipLaunchKernelGGL(k_rs_hist<false>, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, kin, n, shift, counts, nb);
    DMSC_LAUNCH_CHECK();
    scan_exclusive<int32_t>(counts, counts, 256 * nb, nullptr, ar, slot + 1, st);
    const dim3 g((unsigned)nb), b(RS_THREADS);
    if (pin && pout) hipLaunchKernelGGL((k_rs_scatter<true, true>), g, b, 0, st, kin, vin, kout, vout, n, shift, counts, nb);
    else if (pin) hipLaunchKernelGGL((k_rs_scatter<true, false>), g, b, 0, st, kin, vin, kout, vout, n, shift, counts, nb);
    else if (pout) hipLaunchKernelGGL((k_rs_scatter<false, true>), g, b, 0, st, kin, vin, kout, vout, n, shift, counts, nb);
    else hipLaunchKernelGGL((k_rs_scatter<false, false>), g, b, 0, st, kin, vin, kout, vout, n, shift, counts, nb);
    DMSC_LAUNCH_CHECK();
  };
  if (npass == 1) {
    pass(0, keys, vals_in, keys_tmp, vals_tmp, false, false);
    hipLaunchKernelGGL(k_copy_pairs, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, keys_tmp, vals_tmp, keys, vals, n);
    DMSC_LAUNCH_CHECK();
    return;
  }
  uint64_t* pk[2] = {ar.get<uint64_t>(pslot, n), ar.get<uint64_t>(pslot + 1, n)};
  pass(0, keys, vals_in, pk[0], nullptr, false, true);
  for (int p = 1; p + 1 < npass; ++p) pass(p, pk[(p - 1) & 1], nullptr, pk[p & 1], nullptr, true, true);
  pass(npass - 1, pk[(npass - 2) & 1], nullptr, keys, vals, true, false);
}

// rank[i] = position of element i in the stable ascending order of keys (keys are only read); 32-bit
// keys, four passes, the last one scattering positions instead of pairs. Scratch as above.
void radix_rank(const uint32_t* keys, int32_t* rank, int64_t n, Arena& ar, int slot, int pslot, hipStream_t st) {
  KScope ks_("radix_sort", st);
  if (n <= 1) {
    iota_i32(rank, n, st);
    return;
  }
  const int64_t nb = (n + RS_TILE - 1) / RS_TILE;
  int32_t* counts = ar.get<int32_t>(slot, (size_t)(256 * nb));
  uint64_t* pk[2] = {ar.get<uint64_t>(pslot, n), ar.get<uint64_t>(pslot + 1, n)};
  const dim3 g((unsigned)nb), b(RS_THREADS);
  for (int p = 0; p < 4; ++p) {
    const void* kin = p == 0 ? (const void*)keys : (const void*)pk[(p - 1) & 1];
    if (p == 0) hipLaunchKernelGGL(k_rs_hist<false>, g, b, 0, st, kin, n, 0, counts, nb);
    else hipLaunchKernelGGL(k_rs_hist<true>, g, b, 0, st, kin, n, 8 * p, counts, nb);
    DMSC_LAUNCH_CHECK();
    scan_exclusive<int32_t>(counts, counts, 256 * nb, nullptr, ar, slot + 1, st);
    if (p == 0)
      hipLaunchKernelGGL((k_rs_scatter<false, true>), g, b, 0, st, kin, nullptr, pk[0], nullptr, n, 0, counts, nb);
    else if (p < 3)
      hipLaunchKernelGGL((k_rs_scatter<true, true>), g, b, 0, st, kin, nullptr, pk[p & 1], nullptr, n, 8 * p, counts, nb);
    else
      hipLaunchKernelGGL((k_rs_scatter<true, false, true>), g, b, 0, st, kin, nullptr, nullptr, rank, n, 24, counts, nb);
    DMSC_LAUNCH_CHECK();
  }
}

void radix_sort_pairs(uint32_t* keys, int32_t* vals, uint32_t* keys_tmp, int32_t* vals_tmp, int64_t n, Arena& ar,
                      int slot, int pslot, hipStream_t st, bool iota_vals) {
  radix_sort_pairs_bits(keys, vals, keys_tmp, vals_tmp, n, 32, ar, slot, pslot, st, iota_vals);
}

// ------------------------------------------------------------------ pointer jumping
// pj[i] packs (ancestor J in the low 32 bits, distance D in the high 32 bits); a root has J == i.
// Each thread repeatedly replaces (J, D) by (J', D + D') read from pj[J] until J is a root. Reads
// and writes are single 64-bit atomics, so every observed pair is a consistent (ancestor,
// distance); each step moves strictly closer to the root, so every thread terminates on its own
// without waiting for any other (no barrier, no spin).
__global__ void k_pj_init(const int32_t* __restrict__ succ, uint64_t* __restrict__ pj, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t s = (uint32_t)succ[i];
    pj[i] = (uint64_t)s | ((uint64_t)(s != (uint32_t)i ? 1u : 0u) << 32);
  }
}

// The loads and stores are single 64-bit accesses without a coherence scope: a read may return an
// older entry (from this CU's caches), which is still a valid (ancestor, distance) pair, so it only
// costs extra steps; roots never change, so every walk still ends at the root.
__global__ void k_pj(uint64_t* __restrict__ pj, int64_t n, int32_t* __restrict__ err) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t cur = __hip_atomic_load(&pj[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    uint32_t J = (uint32_t)cur, D = (uint32_t)(cur >> 32);
    if (J == (uint32_t)i) continue;
    for (int it = 0;; ++it) {
      if (it >= kLoopCap) {  // only a cyclic successor array can get here
        report(err, ERR_PJ_CAP);
        break;
      }
      const uint64_t nx = __hip_atomic_load(&pj[J], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      const uint32_t JJ = (uint32_t)nx;
      if (JJ == J) break;
      D += (uint32_t)(nx >> 32);
      J = JJ;
      __hip_atomic_store(&pj[i], (uint64_t)J | ((uint64_t)D << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  }
}

void pointer_jump_init(const int32_t* succ, uint64_t* pj, int64_t n, hipStream_t st) {
  KScope ks_("pointer_jump_init", st);
  if (n == 0) return;
  hipLaunchKernelGGL(k_pj_init, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, succ, pj, n);
  DMSC_LAUNCH_CHECK();
}
void pointer_jump(uint64_t* pj, int64_t n, int32_t* err, hipStream_t st) {
  KScope ks_("pointer_jump", st);
  if (n == 0) return;
  hipLaunchKernelGGL(k_pj, dim3(grid_for(n, 256, 1 << 20)), dim3(256), 0, st, pj, n, err);
  DMSC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ per-field bases
// out[f] = pre[starts[f]] (an exclusive prefix sampled at each field's first element), or *total
// when the field starts at the end; f = 0..B.
template <class T, class S>
__global__ void k_bases_at(const T* __restrict__ pre, const S* __restrict__ starts, int64_t n, const T* __restrict__ total,
                           int B, T* __restrict__ out) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f <= B; f += gridDim.x * blockDim.x) {
    const int64_t s = (int64_t)starts[f];
    out[f] = s < n ? pre[s] : *total;
  }
}
template <class T, class S>
void bases_at(const T* pre, const S* starts, int64_t n, const T* total, int B, T* out, hipStream_t st) {
  hipLaunchKernelGGL((k_bases_at<T, S>), dim3(grid_for(B + 1, 256, 64)), dim3(256), 0, st, pre, starts, n, total, B, out);
  DMSC_LAUNCH_CHECK();
}
template void bases_at<int32_t, int64_t>(const int32_t*, const int64_t*, int64_t, const int32_t*, int, int32_t*, hipStream_t);
template void bases_at<int32_t, int32_t>(const int32_t*, const int32_t*, int64_t, const int32_t*, int, int32_t*, hipStream_t);
template void bases_at<int64_t, int32_t>(const int64_t*, const int32_t*, int64_t, const int64_t*, int, int64_t*, hipStream_t);

// ------------------------------------------------------------------ fills
__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, uint8_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
void fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, p, v, n);
  DMSC_LAUNCH_CHECK();
}
__global__ void k_iota_i32(int32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}
__global__ void k_fill_f32_zero(float* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}
void fill_f32_zero(float* p, int64_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_f32_zero, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, p, n);
  DMSC_LAUNCH_CHECK();
}
void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, p, v, n);
  DMSC_LAUNCH_CHECK();
}
void fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, p, v, n);
  DMSC_LAUNCH_CHECK();
}
void iota_i32(int32_t* p, int64_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, p, n);
  DMSC_LAUNCH_CHECK();
}

}  // namespace dmsc
