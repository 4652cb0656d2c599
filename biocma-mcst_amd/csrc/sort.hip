// sort.hip -- stable reorder of one run's particles by compartment.
//
// The particle pass reads a per-compartment record for every particle and
// adds its contributions to a per-compartment accumulator.  With the slots of
// a compartment stored together, each block touches a small neighbourhood of
// compartments: the records stay in L1 and the contribution window of large
// networks covers the block's particles.  Moves and newborns (appended at the
// end) scatter that order over time; this pass restores it.  The order is a
// stable sort by position (slot order among equal positions), the same
// permutation the CPU oracle applies, so runs stay comparable step by step.
//
// The sort is a least-significant-digit radix sort of (position, slot) pairs,
// 8 bits per pass (two passes for up to 65 536 compartments), each pass three
// steps over tiles of 4096 pairs: a digit histogram per tile, an exclusive
// scan of the histograms in digit-major order (digit d of tile b starts after
// every smaller digit and after digit d of every earlier tile), and a stable
// scatter.  The scatter ranks a tile's pairs round by round, 256 at a time in
// slot order: a wave finds the lanes that share its digit with 8 ballots, its
// lowest such lane counts them into LDS, and one thread per digit turns the
// waves' counts into offsets.  Then every field is gathered through the
// permutation: for the only live run of an engine, in one pass straight into
// a second set of the batch arrays, which then replaces the first (the
// engine swaps the pointers); otherwise into scratch and copied back (a run
// owns a slot range of arrays the whole batch shares).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace mcst {

namespace {

constexpr uint32_t RADIX_BITS = 8, RADIX = 1u << RADIX_BITS;
constexpr uint32_t SORT_THREADS = 256, SORT_WAVES = SORT_THREADS / 64;
constexpr uint32_t SORT_ROUNDS = 16, SORT_TILE = SORT_THREADS * SORT_ROUNDS;  // pairs per tile
constexpr uint32_t SCAN_SEG = 4096;  // histogram entries per block of the scan
static_assert(SORT_THREADS == RADIX, "one thread per digit in the histogram and the scatter");
static_assert(SCAN_SEG % SORT_THREADS == 0, "whole entries per thread in the scan");

__device__ __forceinline__ uint64_t lanes_below(uint32_t lane) { return lane ? (~0ull >> (64u - lane)) : 0ull; }

// Digit histogram of tile b of keys_in: hist[d * n_tiles + b].
__global__ void __launch_bounds__(SORT_THREADS) k_digit_hist(const uint32_t* __restrict__ keys_in, uint64_t n,
                                                             uint32_t shift, uint32_t* __restrict__ hist,
                                                             uint32_t n_tiles) {
  __shared__ uint32_t cnt[RADIX];
  const uint32_t tid = threadIdx.x, b = blockIdx.x;
  cnt[tid] = 0u;
  __syncthreads();
  const uint64_t base = (uint64_t)b * SORT_TILE;
  for (uint32_t r = 0; r < SORT_ROUNDS; ++r) {
    const uint64_t i = base + r * SORT_THREADS + tid;
    if (i < n) atomicAdd(&cnt[(keys_in[i] >> shift) & (RADIX - 1u)], 1u);
  }
  __syncthreads();
  hist[(uint64_t)tid * n_tiles + b] = cnt[tid];
}

// Exclusive scan of x[0, m) in place, in three launches: the sum of every
// segment of SCAN_SEG entries, an exclusive scan of those sums (one block),
// and each segment's own scan from its offset.
__global__ void __launch_bounds__(SORT_THREADS) k_seg_sum(const uint32_t* __restrict__ x, uint64_t m,
                                                          uint32_t* __restrict__ seg) {
  __shared__ uint32_t s_w[SORT_WAVES];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_SEG;
  uint32_t v = 0;
  for (uint32_t k = tid; k < SCAN_SEG; k += SORT_THREADS)
    if (base + k < m) v += x[base + k];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (lane == 0) s_w[wave] = v;
  __syncthreads();
  if (tid == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < SORT_WAVES; ++w) t += s_w[w];
    seg[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(SORT_THREADS) k_seg_scan(uint32_t* __restrict__ seg, uint32_t n_seg) {
  __shared__ uint32_t s_w[SORT_WAVES];
  __shared__ uint32_t s_carry;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  if (tid == 0) s_carry = 0u;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < n_seg; c0 += SORT_THREADS) {
    const uint32_t i = c0 + tid;
    const uint32_t v = i < n_seg ? seg[i] : 0u;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(inc, d, 64);
      if (lane >= (uint32_t)d) inc += y;
    }
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t pre = s_carry, tot = 0;
    for (uint32_t w = 0; w < SORT_WAVES; ++w) {
      if (w < wave) pre += s_w[w];
      tot += s_w[w];
    }
    if (i < n_seg) seg[i] = pre + inc - v;
    __syncthreads();
    if (tid == 0) s_carry += tot;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SORT_THREADS) k_seg_apply(uint32_t* __restrict__ x, uint64_t m,
                                                            const uint32_t* __restrict__ seg) {
  constexpr uint32_t PER = SCAN_SEG / SORT_THREADS;  // consecutive entries per thread
  __shared__ uint32_t s_w[SORT_WAVES];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_SEG + (uint64_t)tid * PER;
  uint32_t v[PER], t = 0;
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    v[k] = base + k < m ? x[base + k] : 0u;
    t += v[k];
  }
  uint32_t inc = t;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= (uint32_t)d) inc += y;
  }
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t run = seg[blockIdx.x] + inc - t;
  for (uint32_t w = 0; w < wave; ++w) run += s_w[w];
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    if (base + k < m) x[base + k] = run;
    run += v[k];
  }
}

// Stable scatter of tile b: pair i goes to off[digit][b] + (pairs of the tile
// before i with the same digit).  vals_in null: the value of pair i is i.
__global__ void __launch_bounds__(SORT_THREADS) k_digit_scatter(const uint32_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in, uint64_t n,
                                                                uint32_t shift, const uint32_t* __restrict__ off,
                                                                uint32_t n_tiles, uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t s_base[RADIX];              // where the tile's digit d goes next
  __shared__ uint32_t s_wave[SORT_WAVES][RADIX];  // this round: counts per wave, then their offsets
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6, b = blockIdx.x;
  s_base[tid] = off[(uint64_t)tid * n_tiles + b];
  const uint64_t base = (uint64_t)b * SORT_TILE;
  for (uint32_t r = 0; r < SORT_ROUNDS; ++r) {
#pragma unroll
    for (uint32_t w = 0; w < SORT_WAVES; ++w) s_wave[w][tid] = 0u;
    __syncthreads();
    const uint64_t i = base + r * SORT_THREADS + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t val = valid ? (vals_in ? vals_in[i] : (uint32_t)i) : 0u;
    const uint32_t d = (key >> shift) & (RADIX - 1u);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t bit = 0; bit < RADIX_BITS; ++bit) {
      const uint64_t set = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? set : ~set;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lanes_below(lane));
    if (valid && rank == 0) s_wave[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {  // thread tid owns digit tid: offsets of the waves' groups, in wave order
      uint32_t p = s_base[tid];
#pragma unroll
      for (uint32_t w = 0; w < SORT_WAVES; ++w) {
        const uint32_t c = s_wave[w][tid];
        s_wave[w][tid] = p;
        p += c;
      }
      s_base[tid] = p;
    }
    __syncthreads();
    if (valid) {
      const uint32_t dst = s_wave[wave][d] + rank;
      keys_out[dst] = key;
      vals_out[dst] = val;
    }
    __syncthreads();  // every offset of the round read before the next round clears them
  }
}

__global__ void __launch_bounds__(256) k_gather_pos(uint64_t n, const uint32_t* __restrict__ sorted_keys,
                                                    uint32_t* __restrict__ pos) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    pos[i] = sorted_keys[i];
}

template <class T>
__global__ void __launch_bounds__(256) k_gather(uint64_t n, const uint32_t* __restrict__ perm, const T* __restrict__ src,
                                                T* __restrict__ dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

// Every field of the sorted order at once: one read of the permutation per
// slot, the fields' loads issued four at a time before their stores.
struct FieldSet {
  const double* src_f64[MAX_VAR + 2];
  double* dst_f64[MAX_VAR + 2];
  const uint32_t* src_u32[2];
  uint32_t* dst_u32[2];
  const uint8_t* src_status;
  uint8_t* dst_status;
  uint32_t* dst_pos;
  uint32_t n_f64, n_u32;
};

__global__ void __launch_bounds__(256) k_gather_fields(uint64_t n, const uint32_t* __restrict__ sorted_keys,
                                                       const uint32_t* __restrict__ perm, FieldSet f) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = perm[i];
    const uint8_t st = f.src_status[s];
    uint32_t u[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) u[k] = k < (int)f.n_u32 ? f.src_u32[k][s] : 0u;
    for (uint32_t k0 = 0; k0 < f.n_f64; k0 += 4) {
      double v[4];
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) v[j] = k0 + j < f.n_f64 ? f.src_f64[k0 + j][s] : 0.0;
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j)
        if (k0 + j < f.n_f64) f.dst_f64[k0 + j][i] = v[j];
    }
    f.dst_pos[i] = sorted_keys[i];
    f.dst_status[i] = st;
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (k < (int)f.n_u32) f.dst_u32[k][i] = u[k];
  }
}

int grid_for(uint64_t n) { return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384)); }

struct SortLayout {
  uint64_t n_tiles, m, n_seg;
  size_t keys_alt, hist, seg, bytes;  // byte offsets into the scratch
};
SortLayout sort_layout(uint64_t n) {
  SortLayout L{};
  L.n_tiles = (n + SORT_TILE - 1) / SORT_TILE;
  L.m = L.n_tiles * RADIX;
  L.n_seg = (L.m + SCAN_SEG - 1) / SCAN_SEG;
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  L.keys_alt = 0;
  L.hist = up(n * sizeof(uint32_t));
  L.seg = L.hist + up(L.m * sizeof(uint32_t));
  L.bytes = L.seg + up(L.n_seg * sizeof(uint32_t));
  return L;
}

}  // namespace

hipError_t sort_temp_bytes(uint64_t n, uint32_t, size_t* bytes) {
  *bytes = sort_layout(std::max<uint64_t>(n, 1)).bytes;
  return hipSuccess;
}

hipError_t sort_run_particles(const SortArgs& a, hipStream_t s) {
  const uint64_t n = a.n;
  if (n < 2) return hipSuccess;
  if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
  const SortLayout L = sort_layout(n);
  if (a.temp_bytes < L.bytes) return hipErrorInvalidValue;
  char* t = (char*)a.temp;
  uint32_t* keys_alt = (uint32_t*)(t + L.keys_alt);
  uint32_t* hist = (uint32_t*)(t + L.hist);
  uint32_t* seg = (uint32_t*)(t + L.seg);
  const uint32_t passes = std::max<uint32_t>(1, (a.end_bit + RADIX_BITS - 1) / RADIX_BITS);
  // the last pass writes (keys, perm); earlier ones alternate with
  // (keys_alt, iota) backwards from it; the first reads the positions and
  // takes the slot index as its value
  const uint32_t* kin = a.pos;
  const uint32_t* vin = nullptr;
  for (uint32_t p = 0; p < passes; ++p) {
    const bool to_final = ((passes - 1u - p) % 2u) == 0u;
    uint32_t* kout = to_final ? a.keys : keys_alt;
    uint32_t* vout = to_final ? a.perm : a.iota;
    const uint32_t shift = p * RADIX_BITS;
    hipLaunchKernelGGL(k_digit_hist, dim3((uint32_t)L.n_tiles), dim3(SORT_THREADS), 0, s, kin, n, shift, hist,
                       (uint32_t)L.n_tiles);
    hipLaunchKernelGGL(k_seg_sum, dim3((uint32_t)L.n_seg), dim3(SORT_THREADS), 0, s, (const uint32_t*)hist, L.m, seg);
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(SORT_THREADS), 0, s, seg, (uint32_t)L.n_seg);
    hipLaunchKernelGGL(k_seg_apply, dim3((uint32_t)L.n_seg), dim3(SORT_THREADS), 0, s, hist, L.m, (const uint32_t*)seg);
    hipLaunchKernelGGL(k_digit_scatter, dim3((uint32_t)L.n_tiles), dim3(SORT_THREADS), 0, s, kin, vin, n, shift,
                       (const uint32_t*)hist, (uint32_t)L.n_tiles, kout, vout);
    kin = kout;
    vin = vout;
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (a.to_dst) {  // one pass into the destination arrays
    if (a.n_f64 > MAX_VAR + 2 || a.n_u32 > 2) return hipErrorInvalidValue;
    FieldSet f{};
    for (uint32_t k = 0; k < a.n_f64; ++k) {
      f.src_f64[k] = a.f64[k];
      f.dst_f64[k] = a.dst_f64[k];
    }
    for (uint32_t k = 0; k < a.n_u32; ++k) {
      f.src_u32[k] = a.u32[k];
      f.dst_u32[k] = a.dst_u32[k];
    }
    f.src_status = a.status;
    f.dst_status = a.dst_status;
    f.dst_pos = a.dst_pos;
    f.n_f64 = a.n_f64;
    f.n_u32 = a.n_u32;
    hipLaunchKernelGGL(k_gather_fields, dim3(grid_for(n)), dim3(256), 0, s, n, (const uint32_t*)a.keys,
                       (const uint32_t*)a.perm, f);
    return hipGetLastError();
  }
  // positions: the sorted keys; every other field through the permutation
  hipLaunchKernelGGL(k_gather_pos, dim3(grid_for(n)), dim3(256), 0, s, n, (const uint32_t*)a.keys, a.pos);
  hipLaunchKernelGGL(k_gather<uint8_t>, dim3(grid_for(n)), dim3(256), 0, s, n, (const uint32_t*)a.perm,
                     (const uint8_t*)a.status, (uint8_t*)a.buf);
  if ((err = hipMemcpyAsync(a.status, a.buf, n, hipMemcpyDeviceToDevice, s)) != hipSuccess) return err;
  for (uint32_t f = 0; f < a.n_f64; ++f) {
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(256), 0, s, n, (const uint32_t*)a.perm,
                       (const double*)a.f64[f], (double*)a.buf);
    if ((err = hipMemcpyAsync(a.f64[f], a.buf, n * sizeof(double), hipMemcpyDeviceToDevice, s)) != hipSuccess) return err;
  }
  for (uint32_t f = 0; f < a.n_u32; ++f) {
    hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid_for(n)), dim3(256), 0, s, n, (const uint32_t*)a.perm,
                       (const uint32_t*)a.u32[f], (uint32_t*)a.buf);
    if ((err = hipMemcpyAsync(a.u32[f], a.buf, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s)) != hipSuccess) return err;
  }
  return hipGetLastError();
}

}  // namespace mcst
