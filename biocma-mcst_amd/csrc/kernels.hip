// kernels.hip -- the step of a batch of runs as gfx950 kernels.
//
// One step of every active run is, in stream order:
//   k_transport      liquid transport of all runs      (implScalar.cpp:250-264);
//                    its last block lays out the step's tiles: the tile -> run
//                    table (k_layout alone at commit)
//   k_particle<M>    fused per-particle pass: model update + division requests
//                    + contribution scatter + move test + leave
//                    (model_kernel.hpp:163-268, contribution_kernel.hpp:144-186,
//                     move_kernel.hpp:209-273,347-437,586-648)
//   k_resolve<M>     neighbour pick + exit test of the particles that left their
//                    compartment; division acceptance in slot order over the
//                    run, mother halving, daughters written to [n, n + births)
//                    (particles_container.hpp:559-599); its last block keeps the
//                    books: counters, removal decision (particles_container.hpp:539-557)
//   k_gap_count / k_gap_scan / k_gap_list / k_gap_fill
//                    removal of Exit/Dead slots by gap filling over [0, n + births)
//                    (particles_container.hpp:292-385,735-796), when a particle
//                    can become inactive
// Every kernel reads particle counts from device memory, so the sequence can
// be replayed without host round trips.  The model-dependent kernels
// (k_transport, k_particle, k_resolve, k_properties) are templates in
// step_kernels.hpp; this file instantiates them for the built-in models and
// holds the others.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "step_kernels.hpp"

namespace mcst {

namespace {

// Exclusive scan of in[0..n) into out (optional) by one block of SCAN_THREADS
// threads, in ascending or descending index order; returns the total.
//  - up to RB items per thread: each thread owns a contiguous run, loaded in
//    one round; one block-wide scan of the per-thread sums;
//  - more: the items are read in coalesced chunks of SCAN_THREADS (item
//    base + j*SCAN_THREADS + tid), RB chunks at a time with all their loads
//    in flight; the waves scan each chunk, publish their totals for all RB
//    chunks at once, and every thread adds up the totals before it (two
//    barriers per RB chunks).  One CU streams the 2e5 tile counts of a
//    2e8-particle run in about 0.1 ms; contiguous runs of 200 items per
//    thread took 0.45 ms (each load instruction touched 64 lines).
__device__ uint32_t block_scan(const uint32_t* in, uint32_t* out, uint32_t n, bool descending) {
  constexpr uint32_t RB = 16;
  __shared__ uint32_t s_w[RB][SCAN_WAVES];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  if (n <= RB * SCAN_THREADS) {
    const uint32_t per = (n + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint32_t k0 = tid * per;
    const uint32_t k1 = (k0 + per < n) ? k0 + per : n;
    uint32_t v[RB], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < RB; ++j) {
      const uint32_t k = k0 + j;
      v[j] = k < k1 ? in[descending ? n - 1u - k : k] : 0u;
      sum += v[j];
    }
    const uint32_t inc = wave_incl_scan(sum, lane);
    __syncthreads();  // earlier readers of s_w are done
    if (lane == 63) s_w[0][wave] = inc;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
    for (uint32_t w = 0; w < SCAN_WAVES; ++w) {
      const uint32_t x = s_w[0][w];
      if (w < wave) wpre += x;
      tot += x;
    }
    if (out) {
      uint32_t run = wpre + inc - sum;
#pragma unroll
      for (uint32_t j = 0; j < RB; ++j) {
        const uint32_t k = k0 + j;
        if (k < k1) out[descending ? n - 1u - k : k] = run;
        run += v[j];
      }
    }
    return tot;
  }
  uint32_t carry = 0;  // the items before the current chunks (the same in every thread)
  for (uint32_t base = 0; base < n; base += RB * SCAN_THREADS) {
    uint32_t v[RB], inc[RB];
#pragma unroll
    for (uint32_t j = 0; j < RB; ++j) {
      const uint32_t k = base + j * SCAN_THREADS + tid;
      v[j] = k < n ? in[descending ? n - 1u - k : k] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < RB; ++j) inc[j] = wave_incl_scan(v[j], lane);
    __syncthreads();  // earlier readers of s_w are done
    if (lane == 63)
#pragma unroll
      for (uint32_t j = 0; j < RB; ++j) s_w[j][wave] = inc[j];
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < RB; ++j) {
      uint32_t wpre = 0, tot = 0;
      for (uint32_t w = 0; w < SCAN_WAVES; ++w) {
        const uint32_t x = s_w[j][w];
        if (w < wave) wpre += x;
        tot += x;
      }
      const uint32_t k = base + j * SCAN_THREADS + tid;
      if (out && k < n) out[descending ? n - 1u - k : k] = carry + wpre + inc[j] - v[j];
      carry += tot;
    }
  }
  return carry;
}

}  // namespace

// ------------------------------------------------------------------ removal
// Over the tiles of [0, n_end) of every compacting run (the removal layout
// rm_tile_off, written by the bookkeeping): gaps are the non-Idle slots below
// n_keep (ascending); fillers the Idle slots in [n_keep, n_end) (descending),
// the step's newborns among them.  The i-th gap receives the i-th filler.
__global__ void __launch_bounds__(BLOCK) k_gap_count(BatchDev b) {
  __shared__ uint32_t s_g[WAVES], s_f[WAVES];
  const uint32_t* off = b.rm_tile_off;
  const uint32_t T = off[b.R];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    const uint32_t r = find_run(off, b.R, t);
    const RunState* rs = b.rs + r;
    const RunDev& rd = b.runs[r];
    const uint32_t slot0 = (t - off[r]) * TILE;
    const uint32_t n_end = rs->n_end, n_keep = rs->n_keep;
    uint32_t ng = 0, nf = 0;
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t s = slot0 + k * BLOCK + tid;
      if (s < n_end) {
        const bool live = b.status[rd.p_off + s] == IDLE;
        ng += (s < n_keep && !live) ? 1u : 0u;
        nf += (s >= n_keep && live) ? 1u : 0u;
      }
    }
    ng = wave_sum(ng);
    nf = wave_sum(nf);
    if (lane == 0) {
      s_g[wave] = ng;
      s_f[wave] = nf;
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t a = 0, c = 0;
      for (uint32_t wv = 0; wv < WAVES; ++wv) {
        a += s_g[wv];
        c += s_f[wv];
      }
      b.tile_gap[t] = a;
      b.tile_fill[t] = c;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_gap_scan(BatchDev b) {
  const uint32_t r = blockIdx.x;
  RunState* rs = b.rs + r;
  if (rs->compact != 1u) return;
  const uint32_t t0 = b.rm_tile_off[r];
  const uint32_t nt = b.rm_tile_off[r + 1] - t0;
  const uint32_t g = block_scan(b.tile_gap + t0, b.tile_gap_off + t0, nt, false);
  const uint32_t f = block_scan(b.tile_fill + t0, b.tile_fill_off + t0, nt, true);
  // gaps and fillers are equinumerous when the inactive count is right; if
  // not, fill only as many gaps as there are listed fillers
  if (threadIdx.x == 0) rs->gap_total = g < f ? g : f;
}

// rank of this thread's flagged slots among the flagged slots of the tile,
// in ascending slot order (slot = k*BLOCK + tid); returns the tile total.
__device__ __forceinline__ uint32_t tile_ranks(const bool (&f)[PPT], uint32_t (&rank)[PPT],
                                               uint32_t (*s_cnt)[WAVES], uint32_t lane, uint32_t wave) {
  uint64_t mk[PPT];
#pragma unroll
  for (uint32_t k = 0; k < PPT; ++k) {
    mk[k] = __ballot(f[k]);
    if (lane == 0) s_cnt[k][wave] = (uint32_t)__popcll(mk[k]);
  }
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (uint32_t k = 0; k < PPT; ++k) {
    uint32_t rk = before + (uint32_t)__popcll(mk[k] & lanemask_lt(lane));
#pragma unroll
    for (uint32_t wv = 0; wv < WAVES; ++wv) {
      rk += (wv < wave) ? s_cnt[k][wv] : 0u;
      before += s_cnt[k][wv];
    }
    rank[k] = rk;
  }
  __syncthreads();
  return before;
}

__global__ void __launch_bounds__(BLOCK) k_gap_list(BatchDev b) {
  __shared__ uint32_t s_cnt[PPT][WAVES];
  const uint32_t* off = b.rm_tile_off;
  const uint32_t T = off[b.R];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    if (b.tile_fill[t] == 0) continue;
    const uint32_t r = find_run(off, b.R, t);
    const RunState* rs = b.rs + r;
    const RunDev& rd = b.runs[r];
    const uint32_t slot0 = (t - off[r]) * TILE;
    const uint32_t n_end = rs->n_end, n_keep = rs->n_keep;
    bool f[PPT];
    uint32_t rank[PPT];
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t s = slot0 + k * BLOCK + tid;
      f[k] = s < n_end && s >= n_keep && b.status[rd.p_off + s] == IDLE;
    }
    const uint32_t tot = tile_ranks(f, rank, s_cnt, lane, wave);
    const uint32_t fo = b.tile_fill_off[t];
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k)
      if (f[k]) b.fill_src[rd.p_off + fo + (tot - 1u - rank[k])] = slot0 + k * BLOCK + tid;
  }
}

__global__ void __launch_bounds__(BLOCK) k_gap_fill(BatchDev b, uint32_t n_var) {
  __shared__ uint32_t s_cnt[PPT][WAVES];
  const uint32_t* off = b.rm_tile_off;
  const uint32_t T = off[b.R];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    if (b.tile_gap[t] == 0) continue;
    const uint32_t r = find_run(off, b.R, t);
    const RunState* rs = b.rs + r;
    const RunDev& rd = b.runs[r];
    const uint32_t slot0 = (t - off[r]) * TILE;
    const uint32_t n_keep = rs->n_keep;
    bool f[PPT];
    uint32_t rank[PPT];
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t s = slot0 + k * BLOCK + tid;
      f[k] = s < n_keep && b.status[rd.p_off + s] != IDLE;
    }
    tile_ranks(f, rank, s_cnt, lane, wave);
    const uint32_t go = b.tile_gap_off[t];
    const uint32_t n_fill = rs->gap_total;
    const uint32_t n_end = rs->n_end;
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      if (!f[k] || go + rank[k] >= n_fill) continue;
      const uint32_t src_slot = b.fill_src[rd.p_off + go + rank[k]];
      if (src_slot >= n_end) continue;
      const uint64_t dst = rd.p_off + slot0 + k * BLOCK + tid;
      const uint64_t src = rd.p_off + src_slot;
      b.pos[dst] = b.pos[src];
      b.status[dst] = IDLE;
      b.age_hyd[dst] = b.age_hyd[src];
      b.age_div[dst] = b.age_div[src];
      b.clk_hyd[dst] = b.clk_hyd[src];
      b.clk_div[dst] = b.clk_div[src];
      for (uint32_t v = 0; v < n_var; ++v) b.state[(uint64_t)v * b.P + dst] = b.state[(uint64_t)v * b.P + src];
    }
  }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_layout(BatchDev b) { layout_pass<SCAN_THREADS>(b); }

// ------------------------------------------------------------------ histogram
// Idle particles per compartment of run r (getRepartition, unit.cpp:190-230).
__global__ void __launch_bounds__(256) k_repartition(BatchDev b, uint32_t r, unsigned long long* hist) {
  const RunDev& rd = b.runs[r];
  const uint32_t n = b.rs[r].n;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const uint64_t g = rd.p_off + s;
    if (b.status[g] == IDLE) atomicAdd(hist + b.pos[g], 1ull);
  }
}

// Clears n counters (a kernel rather than a memset node, so that a caller
// may capture it in a graph).
__global__ void __launch_bounds__(256) k_clear_u64(unsigned long long* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// ------------------------------------------------------------------ launchers
const ModelOps* builtin_model_ops(int model_id) {
  static const ModelOps ops[] = {model_ops<MonodMass>(), model_ops<DualMonod>(), model_ops<Monod>(),
                                 model_ops<FixedLength>(), model_ops<SimpleAcetate>(), model_ops<QuadMonod>()};
  static_assert(MonodMass::ID == 0 && DualMonod::ID == 1 && Monod::ID == 2 && FixedLength::ID == 3 &&
                    SimpleAcetate::ID == 4 && QuadMonod::ID == 5,
                "built-in model ids index the table");
  if (model_id < 0 || model_id >= (int)(sizeof(ops) / sizeof(ops[0]))) return nullptr;
  return &ops[model_id];
}

void launch_removal(const BatchDev& b, uint32_t n_var, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_gap_count, dim3(grid), dim3(BLOCK), 0, s, b);
  hipLaunchKernelGGL(k_gap_scan, dim3(b.R), dim3(SCAN_THREADS), 0, s, b);
  hipLaunchKernelGGL(k_gap_list, dim3(grid), dim3(BLOCK), 0, s, b);
  hipLaunchKernelGGL(k_gap_fill, dim3(grid), dim3(BLOCK), 0, s, b, n_var);
}

void launch_layout(const BatchDev& b, hipStream_t s) {
  hipLaunchKernelGGL(k_layout, dim3(1), dim3(SCAN_THREADS), 0, s, b);
}

void launch_clear_u64(unsigned long long* p, uint64_t n, hipStream_t s) {
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_clear_u64, dim3((unsigned)std::max<uint64_t>(blocks, 1)), dim3(256), 0, s, p, n);
}

void launch_repartition(const BatchDev& b, uint32_t run, unsigned long long* hist, hipStream_t s) {
  hipLaunchKernelGGL(k_repartition, dim3(1024), dim3(256), 0, s, b, run, hist);
}

}  // namespace mcst
