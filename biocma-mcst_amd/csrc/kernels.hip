// kernels.hip -- the step of a batch of runs as gfx950 kernels.
//
// One step of every active run is, in stream order:
//   k_transport      liquid transport of all runs      (implScalar.cpp:250-264)
//   k_particle<M>    fused per-particle pass: model update + division staging
//                    + contribution scatter + move + leave
//                    (model_kernel.hpp:163-268, contribution_kernel.hpp:144-186,
//                     move_kernel.hpp:209-273,347-437,586-648)
//   k_bookkeep       per-run counters, removal decision  (particles_container.hpp:539-557)
//   k_gap_count / k_gap_scan / k_gap_list / k_gap_fill
//                    removal of Exit/Dead slots by gap filling (particles_container.hpp:292-385,735-796)
//   k_merge          append staged newborns                (particles_container.hpp:403-443,575-599)
//   k_layout         tile -> run table and division budgets for the next step
// Every kernel reads particle counts from device memory, so the sequence can
// be replayed without host round trips.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "models.hpp"

namespace mcst {

namespace {

// Pointers read from the run tables are generic; tell the compiler they point
// to global memory so it emits global_* rather than flat_* accesses.
#define GLOBAL_AS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GLOBAL_AS T* gp(T* p) {
  return (GLOBAL_AS T*)p;
}

__device__ __forceinline__ uint64_t lanemask_lt(uint32_t lane) { return lane ? (~0ull >> (64u - lane)) : 0ull; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

// largest r in [0, R) with off[r] <= t, given off[0] = 0 <= t < off[R]
__device__ __forceinline__ uint32_t find_run(const uint32_t* off, uint32_t R, uint32_t t) {
  uint32_t lo = 0, hi = R;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Exclusive scan of in[0..n) into out (optional) by one block of SCAN_THREADS
// threads, in ascending or descending index order; returns the total.
constexpr uint32_t SCAN_THREADS = 1024;
constexpr uint32_t SCAN_WAVES = SCAN_THREADS / 64;

__device__ uint32_t block_scan(const uint32_t* in, uint32_t* out, uint32_t n, bool descending) {
  __shared__ uint32_t s_w[SCAN_WAVES];
  __shared__ uint32_t s_carry;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  __syncthreads();  // earlier readers of s_carry are done
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < n; b0 += SCAN_THREADS) {
    const uint32_t k = b0 + tid;
    const uint32_t idx = descending ? (n - 1u - k) : k;
    const uint32_t v = (k < n) ? in[idx] : 0u;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
    for (uint32_t w = 0; w < SCAN_WAVES; ++w) {
      const uint32_t x = s_w[w];
      if (w < wave) wpre += x;
      tot += x;
    }
    const uint32_t carry = s_carry;
    if (k < n && out) out[idx] = carry + wpre + inc - v;
    __syncthreads();
    if (tid == 0) s_carry = carry + tot;
    __syncthreads();
  }
  return s_carry;
}

}  // namespace

// ------------------------------------------------------------------ transport
// m += dt*(C*M - C*diag(sink) + S + F);  C = m*diag(1/V)  for every
// compartment of every active run; also refreshes the per-compartment record
// {C_0..C_{n_s-1}, p_move, q_exit} the particle pass gathers, and clears S.
__global__ void __launch_bounds__(256) k_transport(BatchDev b) {
  const uint32_t total = b.n_comp_total;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    const uint32_t r = find_run(b.comp_off, b.R, g);
    const RunState* rs = b.rs + r;
    if (!rs->active) continue;
    const RunDev& rd = b.runs[r];
    const uint32_t j = g - b.comp_off[r];
    const uint32_t n_c = rd.n_c, n_s = rd.n_s;
    const uint32_t step = rs->step;
    const MapDev& mp = b.maps[rd.map_base + (step / rd.steps_per_map) % rd.n_maps];
    const double* Co = rd.C[step & 1u];
    double* Cn = rd.C[(step + 1u) & 1u];
    double* rec = rd.rec + (size_t)j * rd.rec_w;
    const bool on = (b.stages & STAGE_TRANSPORT) != 0;
    const uint32_t e0 = mp.col_ptr[j], e1 = mp.col_ptr[j + 1];
    const double sink = rd.sink[j];
    const double inv_v = mp.inv_v[j];
    const double dt = rd.dt;
    for (uint32_t s = 0; s < n_s; ++s) {
      const double* c = Co + (size_t)s * n_c;
      const size_t o = (size_t)s * n_c + j;
      double cn = c[j];
      if (on) {
        double acc = 0.0;
        for (uint32_t e = e0; e < e1; ++e) acc += c[mp.src[e]] * mp.val[e];
        const double dmdt = (acc - c[j] * sink) + (rd.S[o] + rd.F[o]);
        const double mm = rd.m[o] + dt * dmdt;
        rd.m[o] = mm;
        cn = mm * inv_v;
      }
      Cn[o] = cn;
      rec[s] = cn;
      rd.S[o] = 0.0;
    }
    rec[n_s] = mp.p_move[j];
    rec[n_s + 1] = mp.q_exit[j];
  }
}

// ------------------------------------------------------------------ particles
template <class M>
__global__ void __launch_bounds__(BLOCK) k_particle(BatchDev b) {
  constexpr int NV = M::NV, NS = M::NS, REC = rec_stride(M::NS);
  __shared__ uint32_t s_div[PPT][WAVES];
  __shared__ uint32_t s_red[4][WAVES];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t T = b.tile_off[b.R];
  const uint32_t stages = b.stages;
  const uint64_t P = b.P;

  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    const uint32_t r = find_run(b.tile_off, b.R, t);
    const RunDev& rd = b.runs[r];
    RunState* rs = b.rs + r;
    const uint32_t n = rs->n;
    const uint32_t step = rs->step;
    const uint32_t budget = rs->budget;
    const uint32_t t_local = t - b.tile_off[r];
    const uint32_t slot0 = t_local * TILE;
    const uint32_t cnt = (n - slot0 < TILE) ? n - slot0 : TILE;
    const uint64_t base = rd.p_off + slot0;
    const MapDev& mp = b.maps[rd.map_base + (step / rd.steps_per_map) % rd.n_maps];
    const double dt = rd.dt;
    const double w = rd.weight;
    const uint32_t n_c = rd.n_c;
    double p[M::NP > 0 ? M::NP : 1];
#pragma unroll
    for (int q = 0; q < M::NP; ++q) p[q] = rd.params[q];
    const bool do_model = (stages & STAGE_MODEL) != 0;
    const bool do_contrib = (stages & STAGE_CONTRIB) != 0;
    const bool do_move = rd.enable_move && (stages & STAGE_MOVE);
    const bool do_leave = rd.enable_leave && (stages & STAGE_LEAVE);

    uint8_t st[PPT];
    uint32_t pos0[PPT];
    double x[PPT][NV];
    double ad[PPT], ah[PPT];
    bool idle[PPT], want[PPT];
    uint32_t n_exit = 0, n_dead = 0, n_move = 0, n_proc = 0;

#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t i = k * BLOCK + tid;
      const bool valid = i < cnt;
      st[k] = valid ? b.status[base + i] : (uint8_t)EXIT;
      pos0[k] = valid ? b.pos[base + i] : 0u;
      idle[k] = st[k] == IDLE;
      n_proc += idle[k] ? 1u : 0u;
      want[k] = false;
    }
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint64_t g = base + k * BLOCK + tid;
      if (idle[k]) {
#pragma unroll
        for (int v = 0; v < NV; ++v) x[k][v] = b.state[(uint64_t)v * P + g];
        ad[k] = b.age_div[g];
        ah[k] = do_leave ? b.age_hyd[g] : 0.0;
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      if (!idle[k]) continue;
      const uint32_t i = k * BLOCK + tid;
      const uint64_t g = base + i;
      const uint32_t slot = slot0 + i;
      const GLOBAL_AS double* rp = gp(rd.rec) + (size_t)pos0[k] * REC;
      double rr[NS + 2];
#pragma unroll
      for (int q = 0; q < NS + 2; ++q) rr[q] = rp[q];
      double contrib[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) contrib[s] = 0.0;
      int code = IDLE;
      if (do_model) {
        ad[k] += dt;
        code = M::update(p, dt, x[k], rr, contrib);
      }
      if (code == DEAD) {  // a dead cell stops contributing, moving and leaving
        b.status[g] = DEAD;
        idle[k] = false;
        ++n_dead;
#pragma unroll
        for (int v = 0; v < NV; ++v) b.state[(uint64_t)v * P + g] = x[k][v];
        b.age_div[g] = ad[k];
        continue;
      }
      want[k] = code == DIVISION;
      if (do_contrib) {
#pragma unroll
        for (int s = 0; s < NS; ++s) unsafeAtomicAdd((double*)(gp(rd.S) + (size_t)s * n_c + pos0[k]), w * contrib[s]);
      }
      const U4 u = draw4(rd.seed, slot, step, TAG_TRANSPORT, 0u);
      uint32_t pos = pos0[k];
      bool moved = false;
      if (do_move && rr[NS] > u01(u.x)) {  // dt*Q/V > u  (probability_leaving.hpp:313-325)
        const double u2 = u01(u.y);
        const uint32_t m = mp.m;
        const GLOBAL_AS double* cp = gp(mp.cdf) + (size_t)pos * m;
        int left = 0, right = (int)m - 1;
        while (left < right) {  // first column with cp >= u  (move_kernel.hpp:89-98)
          const int mid = (left + right) >> 1;
          if (u2 > cp[mid])
            left = mid + 1;
          else
            right = mid;
        }
        pos = gp(mp.nbr)[(size_t)pos * m + left];
        moved = true;
        ++n_move;
        b.pos[g] = pos;
      }
      if (do_leave) {
        ah[k] += dt;
        const double q = moved ? gp(rd.rec)[(size_t)pos * REC + NS + 1] : rr[NS + 1];
        if (q < 1.0 && u01(u.z) > q) {  // u > exp(-dt*Q/V)  <=>  dt*Q > -ln(u)*V
          if (rd.probe_exit) {
            const unsigned long long e = atomicAdd(&rs->probe_exit_n, 1ull);
            if (e < rd.probe_cap) rd.probe_exit[e] = ah[k];
          }
          ah[k] = 0.0;
          b.status[g] = EXIT;
          ++n_exit;
        }
        b.age_hyd[g] = ah[k];
      }
    }

    // rank the division requests of the tile in slot order
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint64_t mk = __ballot(want[k]);
      if (lane == 0) s_div[k][wave] = (uint32_t)__popcll(mk);
    }
    n_exit = wave_sum(n_exit);
    n_dead = wave_sum(n_dead);
    n_move = wave_sum(n_move);
    n_proc = wave_sum(n_proc);
    if (lane == 0) {
      s_red[0][wave] = n_exit;
      s_red[1][wave] = n_dead;
      s_red[2][wave] = n_move;
      s_red[3][wave] = n_proc;
    }
    __syncthreads();
    uint32_t total = 0;
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k)
#pragma unroll
      for (uint32_t wv = 0; wv < WAVES; ++wv) total += s_div[k][wv];
    const uint32_t births = total < budget ? total : budget;

    uint32_t before = 0;  // requests in earlier k-rows
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t i = k * BLOCK + tid;
      const uint64_t g = base + i;
      const uint64_t mk = __ballot(want[k]);
      uint32_t rank = before + (uint32_t)__popcll(mk & lanemask_lt(lane));
#pragma unroll
      for (uint32_t wv = 0; wv < WAVES; ++wv) rank += (wv < wave) ? s_div[k][wv] : 0u;
#pragma unroll
      for (uint32_t wv = 0; wv < WAVES; ++wv) before += s_div[k][wv];
      if (want[k] && rank < budget) {
        if (rd.probe_div) {
          const unsigned long long e = atomicAdd(&rs->probe_div_n, 1ull);
          if (e < rd.probe_cap) rd.probe_div[e] = ad[k];
        }
        double xd[NV];
        M::division(p, x[k], xd, RngKey{rd.seed, slot0 + i, step});
        const uint64_t sidx = (uint64_t)t * b.stage_stride + rank;
#pragma unroll
        for (int v = 0; v < NV; ++v) b.st_state[(uint64_t)v * b.ST + sidx] = xd[v];
        b.st_pos[sidx] = pos0[k];
        ad[k] = 0.0;
      }
      if (idle[k]) {
#pragma unroll
        for (int v = 0; v < NV; ++v) b.state[(uint64_t)v * P + g] = x[k][v];
        if (do_model) b.age_div[g] = ad[k];
      }
    }
    if (tid == 0) {
      b.tile_births[t] = births;
      uint32_t e = 0, d = 0, mv = 0, pr = 0;
      for (uint32_t wv = 0; wv < WAVES; ++wv) {
        e += s_red[0][wv];
        d += s_red[1][wv];
        mv += s_red[2][wv];
        pr += s_red[3][wv];
      }
      if (pr) atomicAdd(&rs->step_processed, (unsigned long long)pr);
      if (e) atomicAdd(&rs->step_exits, (unsigned long long)e);
      if (d) atomicAdd(&rs->step_deaths, (unsigned long long)d);
      if (mv) atomicAdd(&rs->step_moves, (unsigned long long)mv);
      if (total > births) atomicAdd(&rs->step_overflow, (unsigned long long)(total - births));
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ bookkeeping
// One block per run: births of the pass (scan of tile counts), counters, and
// the removal decision of update_and_remove_inactive.
__global__ void __launch_bounds__(SCAN_THREADS) k_bookkeep(BatchDev b) {
  const uint32_t r = blockIdx.x;
  RunState* rs = b.rs + r;
  if (!rs->active) return;
  const RunDev& rd = b.runs[r];
  const uint32_t t0 = b.tile_off[r];
  const uint32_t nt = b.tile_off[r + 1] - t0;
  const uint32_t births = block_scan(b.tile_births + t0, b.tile_birth_off + t0, nt, false);
  if (threadIdx.x != 0) return;
  const uint32_t n = rs->n;
  const unsigned long long exits = rs->step_exits, deaths = rs->step_deaths;
  rs->counters[C_BIRTHS] += births;
  rs->counters[C_EXITS] += exits;
  rs->counters[C_DEATHS] += deaths;
  rs->counters[C_MOVES] += rs->step_moves;
  rs->counters[C_OVERFLOW] += rs->step_overflow;
  rs->counters[C_PROCESSED] += rs->step_processed;
  rs->step_exits = rs->step_deaths = rs->step_moves = rs->step_overflow = rs->step_processed = 0;
  uint64_t inactive = (uint64_t)rs->inactive + exits + deaths;
  const uint64_t thr_ratio = (uint64_t)((double)n * rd.ratio);
  const uint64_t thr = rd.min_removal > thr_ratio ? rd.min_removal : thr_ratio;
  uint32_t n_mid = n, compact = 0;
  if ((b.stages & STAGE_REMOVE) && inactive > thr) {
    rs->counters[C_COMPACTIONS] += 1;
    if (inactive >= n) {
      n_mid = 0;
      compact = 2;  // everything removed, nothing to move
    } else {
      n_mid = n - (uint32_t)inactive;
      compact = 1;
    }
    inactive = 0;
  }
  rs->n_prev = n;
  rs->n_mid = n_mid;
  rs->compact = compact;
  rs->births = births;
  rs->inactive = (uint32_t)inactive;
  rs->n = n_mid + births;
  rs->step = rs->step + 1u;
  rs->counters[C_STEPS] = rs->step;
  rs->counters[C_N] = rs->n;
  rs->counters[C_INACTIVE] = rs->inactive;
}

// ------------------------------------------------------------------ removal
// Gaps are the non-Idle slots below n_mid (ascending); fillers the Idle slots
// in [n_mid, n_prev) (descending).  The i-th gap receives the i-th filler.
__global__ void __launch_bounds__(BLOCK) k_gap_count(BatchDev b) {
  __shared__ uint32_t s_g[WAVES], s_f[WAVES];
  const uint32_t T = b.tile_off[b.R];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    const uint32_t r = find_run(b.tile_off, b.R, t);
    const RunState* rs = b.rs + r;
    if (rs->compact != 1u) continue;
    const RunDev& rd = b.runs[r];
    const uint32_t slot0 = (t - b.tile_off[r]) * TILE;
    const uint32_t n_prev = rs->n_prev, n_mid = rs->n_mid;
    uint32_t ng = 0, nf = 0;
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t s = slot0 + k * BLOCK + tid;
      if (s < n_prev) {
        const bool live = b.status[rd.p_off + s] == IDLE;
        ng += (s < n_mid && !live) ? 1u : 0u;
        nf += (s >= n_mid && live) ? 1u : 0u;
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
  const uint32_t t0 = b.tile_off[r];
  const uint32_t nt = b.tile_off[r + 1] - t0;
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
  const uint32_t T = b.tile_off[b.R];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    const uint32_t r = find_run(b.tile_off, b.R, t);
    const RunState* rs = b.rs + r;
    if (rs->compact != 1u || b.tile_fill[t] == 0) continue;
    const RunDev& rd = b.runs[r];
    const uint32_t slot0 = (t - b.tile_off[r]) * TILE;
    const uint32_t n_prev = rs->n_prev, n_mid = rs->n_mid;
    bool f[PPT];
    uint32_t rank[PPT];
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t s = slot0 + k * BLOCK + tid;
      f[k] = s < n_prev && s >= n_mid && b.status[rd.p_off + s] == IDLE;
    }
    const uint32_t tot = tile_ranks(f, rank, s_cnt, lane, wave);
    const uint32_t off = b.tile_fill_off[t];
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k)
      if (f[k]) b.fill_src[rd.p_off + off + (tot - 1u - rank[k])] = slot0 + k * BLOCK + tid;
  }
}

__global__ void __launch_bounds__(BLOCK) k_gap_fill(BatchDev b, uint32_t n_var) {
  __shared__ uint32_t s_cnt[PPT][WAVES];
  const uint32_t T = b.tile_off[b.R];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    const uint32_t r = find_run(b.tile_off, b.R, t);
    const RunState* rs = b.rs + r;
    if (rs->compact != 1u || b.tile_gap[t] == 0) continue;
    const RunDev& rd = b.runs[r];
    const uint32_t slot0 = (t - b.tile_off[r]) * TILE;
    const uint32_t n_mid = rs->n_mid;
    bool f[PPT];
    uint32_t rank[PPT];
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      const uint32_t s = slot0 + k * BLOCK + tid;
      f[k] = s < n_mid && b.status[rd.p_off + s] != IDLE;
    }
    tile_ranks(f, rank, s_cnt, lane, wave);
    const uint32_t off = b.tile_gap_off[t];
    const uint32_t n_fill = rs->gap_total;
    const uint32_t n_prev = rs->n_prev;
#pragma unroll
    for (uint32_t k = 0; k < PPT; ++k) {
      if (!f[k] || off + rank[k] >= n_fill) continue;
      const uint32_t src_slot = b.fill_src[rd.p_off + off + rank[k]];
      if (src_slot >= n_prev) continue;
      const uint64_t dst = rd.p_off + slot0 + k * BLOCK + tid;
      const uint64_t src = rd.p_off + src_slot;
      b.pos[dst] = b.pos[src];
      b.status[dst] = IDLE;
      b.age_hyd[dst] = b.age_hyd[src];
      b.age_div[dst] = b.age_div[src];
      for (uint32_t v = 0; v < n_var; ++v) b.state[(uint64_t)v * b.P + dst] = b.state[(uint64_t)v * b.P + src];
    }
  }
}

// ------------------------------------------------------------------ merge
__global__ void __launch_bounds__(BLOCK) k_merge(BatchDev b, uint32_t n_var) {
  const uint32_t T = b.tile_off[b.R];
  for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
    const uint32_t nb = b.tile_births[t];
    if (nb == 0) continue;
    const uint32_t r = find_run(b.tile_off, b.R, t);
    const RunDev& rd = b.runs[r];
    const RunState* rs = b.rs + r;
    const uint64_t dst0 = rd.p_off + rs->n_mid + b.tile_birth_off[t];
    const uint64_t src0 = (uint64_t)t * b.stage_stride;
    for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x) {
      const uint64_t d = dst0 + j, s = src0 + j;
      b.pos[d] = b.st_pos[s];
      b.status[d] = IDLE;
      b.age_hyd[d] = 0.0;
      b.age_div[d] = 0.0;
      for (uint32_t v = 0; v < n_var; ++v) b.state[(uint64_t)v * b.P + d] = b.st_state[(uint64_t)v * b.ST + s];
    }
  }
}

// ------------------------------------------------------------------ layout
// Tile -> run table and per-tile division budget for the next pass:
// budget = min(stage_cap, (cap - n) / tiles), so staged newborns always fit.
__global__ void __launch_bounds__(SCAN_THREADS) k_layout(BatchDev b) {
  __shared__ uint32_t s_w[SCAN_WAVES];
  __shared__ uint32_t s_carry;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t r0 = 0; r0 < b.R; r0 += SCAN_THREADS) {
    const uint32_t r = r0 + tid;
    uint32_t tiles = 0;
    if (r < b.R) {
      RunState* rs = b.rs + r;
      const RunDev& rd = b.runs[r];
      const uint32_t active = rs->step < rd.step_target ? 1u : 0u;
      rs->active = active;
      tiles = active ? (rs->n + TILE - 1u) / TILE : 0u;
      uint32_t budget = 0;
      if (tiles) {
        const uint32_t q = (rd.cap - rs->n) / tiles;
        budget = q < rd.stage_cap ? q : rd.stage_cap;
      }
      rs->budget = budget;
      rs->n_tiles = tiles;
    }
    const uint32_t inc = wave_incl_scan(tiles, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
    for (uint32_t w = 0; w < SCAN_WAVES; ++w) {
      if (w < wave) wpre += s_w[w];
      tot += s_w[w];
    }
    const uint32_t carry = s_carry;
    if (r < b.R) {
      b.tile_off[r] = carry + wpre + inc - tiles;
      b.rs[r].tile_begin = carry + wpre + inc - tiles;
    }
    __syncthreads();
    if (tid == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (tid == 0) b.tile_off[b.R] = s_carry;
}

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

// ------------------------------------------------------------------ launchers
template <class M>
static void launch_particle_t(const BatchDev& b, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_particle<M>, dim3(grid), dim3(BLOCK), 0, s, b);
}

void launch_particle(int model_id, const BatchDev& b, int grid, hipStream_t s) {
  switch (model_id) {
    case MonodMass::ID: launch_particle_t<MonodMass>(b, grid, s); break;
    case DualMonod::ID: launch_particle_t<DualMonod>(b, grid, s); break;
    case Monod::ID: launch_particle_t<Monod>(b, grid, s); break;
    case FixedLength::ID: launch_particle_t<FixedLength>(b, grid, s); break;
    case SimpleAcetate::ID: launch_particle_t<SimpleAcetate>(b, grid, s); break;
    default: break;
  }
}

const void* particle_kernel_ptr(int model_id) {
  switch (model_id) {
    case MonodMass::ID: return (const void*)&k_particle<MonodMass>;
    case DualMonod::ID: return (const void*)&k_particle<DualMonod>;
    case Monod::ID: return (const void*)&k_particle<Monod>;
    case FixedLength::ID: return (const void*)&k_particle<FixedLength>;
    case SimpleAcetate::ID: return (const void*)&k_particle<SimpleAcetate>;
    default: return nullptr;
  }
}

void launch_transport(const BatchDev& b, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_transport, dim3(grid), dim3(256), 0, s, b);
}
void launch_bookkeep(const BatchDev& b, hipStream_t s) {
  hipLaunchKernelGGL(k_bookkeep, dim3(b.R), dim3(SCAN_THREADS), 0, s, b);
}
void launch_removal(const BatchDev& b, uint32_t n_var, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_gap_count, dim3(grid), dim3(BLOCK), 0, s, b);
  hipLaunchKernelGGL(k_gap_scan, dim3(b.R), dim3(SCAN_THREADS), 0, s, b);
  hipLaunchKernelGGL(k_gap_list, dim3(grid), dim3(BLOCK), 0, s, b);
  hipLaunchKernelGGL(k_gap_fill, dim3(grid), dim3(BLOCK), 0, s, b, n_var);
}
void launch_merge(const BatchDev& b, uint32_t n_var, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_merge, dim3(grid), dim3(BLOCK), 0, s, b, n_var);
}
void launch_layout(const BatchDev& b, hipStream_t s) {
  hipLaunchKernelGGL(k_layout, dim3(1), dim3(SCAN_THREADS), 0, s, b);
}
void launch_repartition(const BatchDev& b, uint32_t run, unsigned long long* hist, hipStream_t s) {
  hipLaunchKernelGGL(k_repartition, dim3(1024), dim3(256), 0, s, b, run, hist);
}

}  // namespace mcst
