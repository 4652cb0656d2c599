// kernels.hip -- the step of a batch of runs as gfx950 kernels.
//
// One step of every active run is, in stream order:
//   k_transport      liquid transport of all runs      (implScalar.cpp:250-264);
//                    its last block lays out the step's tiles: tile -> run
//                    table and division budgets (k_layout alone at commit)
//   k_particle<M>    fused per-particle pass: model update + division staging
//                    + contribution scatter + move + leave
//                    (model_kernel.hpp:163-268, contribution_kernel.hpp:144-186,
//                     move_kernel.hpp:209-273,347-437,586-648)
//   k_move           neighbour pick + exit test of the particles that left their compartment
//   k_bookkeep       per-run counters, removal decision  (particles_container.hpp:539-557)
//   k_gap_count / k_gap_scan / k_gap_list / k_gap_fill
//                    removal of Exit/Dead slots by gap filling (particles_container.hpp:292-385,735-796)
//   k_merge          append staged newborns                (particles_container.hpp:403-443,575-599)
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

// A value every lane of the wave holds identically, moved to scalar registers.
template <class T>
__device__ __forceinline__ T uniform(T v) {
  static_assert(sizeof(T) % 4 == 0, "uniform() works on whole dwords");
  union {
    T t;
    uint32_t w[sizeof(T) / 4];
  } u;
  u.t = v;
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) u.w[i] = __builtin_amdgcn_readfirstlane(u.w[i]);
  return u.t;
}

// Streaming accesses to the particle arrays: non-temporal, so the particle
// stream does not evict the flowmap and compartment tables from L2.
template <class T>
__device__ __forceinline__ T ld_stream(const GLOBAL_AS T* p) {
  return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void st_stream(GLOBAL_AS T* p, T v) {
  __builtin_nontemporal_store(v, p);
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
  // each thread owns a contiguous run of `per` items: one read pass, one
  // block-wide scan of the per-thread sums, one write pass
  __shared__ uint32_t s_w[SCAN_WAVES];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t per = (n + SCAN_THREADS - 1) / SCAN_THREADS;
  const uint32_t k0 = tid * per;
  const uint32_t k1 = (k0 + per < n) ? k0 + per : n;
  uint32_t sum = 0;
  for (uint32_t k = k0; k < k1; ++k) sum += in[descending ? n - 1u - k : k];
  const uint32_t inc = wave_incl_scan(sum, lane);
  __syncthreads();  // earlier readers of s_w are done
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t wpre = 0, tot = 0;
  for (uint32_t w = 0; w < SCAN_WAVES; ++w) {
    const uint32_t x = s_w[w];
    if (w < wave) wpre += x;
    tot += x;
  }
  if (out) {
    uint32_t run = wpre + inc - sum;
    for (uint32_t k = k0; k < k1; ++k) {
      const uint32_t idx = descending ? n - 1u - k : k;
      const uint32_t v = in[idx];
      out[idx] = run;
      run += v;
    }
  }
  return tot;
}

}  // namespace

// ------------------------------------------------------------------ transport
// m += dt*(C*M - C*diag(sink) + S + F);  C = m*diag(1/V)  for every
// (compartment, species) of every active run, one thread each; also refreshes
// the per-compartment record {C_0..C_{n_s-1}, p_move, q_exit} the particle
// pass gathers, and clears S.  The extra last block lays out the step's
// tiles (layout_pass), which needs only the counts of the previous step.
template <uint32_t NT>
__device__ void layout_pass(const BatchDev& b);

__global__ void __launch_bounds__(256) k_transport(BatchDev b) {
  if (blockIdx.x == gridDim.x - 1) {
    layout_pass<256>(b);
    return;
  }
  const uint32_t total = b.n_comp_total;
  const uint32_t stride = (gridDim.x - 1) * blockDim.x;
  const uint32_t pairs = total * b.ns_max;
  for (uint32_t gs = blockIdx.x * blockDim.x + threadIdx.x; gs < pairs; gs += stride) {
    // gs enumerates (compartment g, species s) pairs, species-major
    const uint32_t g = gs % total;
    const uint32_t s = gs / total;
    const uint32_t r = find_run(b.comp_off, b.R, g);
    const RunDev& rd = b.runs[r];
    if (s >= rd.n_s && s != 0) continue;
    const RunState* rs = b.rs + r;
    // the run's activity is recomputed here rather than read from the
    // layout block, which writes it concurrently
    if (rs->step >= rd.step_target) continue;
    const uint32_t j = g - b.comp_off[r];
    const uint32_t n_c = rd.n_c, n_s = rd.n_s;
    const uint32_t step = rs->step;
    const MapDev& mp = b.maps[rd.map_base + (step / rd.steps_per_map) % rd.n_maps];
    double* rec = rd.rec + (size_t)j * rd.rec_w;
    if (s == 0) {
      rec[n_s] = mp.p_move[j];
      rec[n_s + 1] = mp.q_exit[j];
    }
    if (s >= n_s) continue;
    const double* c = rd.C[step & 1u] + (size_t)s * n_c;
    double* Cn = rd.C[(step + 1u) & 1u];
    const size_t o = (size_t)s * n_c + j;
    double cn = c[j];
    if (b.stages & STAGE_TRANSPORT) {
      const uint32_t e0 = mp.col_ptr[j], e1 = mp.col_ptr[j + 1];
      double acc = 0.0;
      for (uint32_t e = e0; e < e1; ++e) acc += c[mp.src[e]] * mp.val[e];
      const double dmdt = (acc - c[j] * rd.sink[j]) + (rd.S[o] + rd.F[o]);
      const double mm = rd.m[o] + rd.dt * dmdt;
      rd.m[o] = mm;
      cn = mm * mp.inv_v[j];
    }
    Cn[o] = cn;
    rec[s] = cn;
    rd.S[o] = 0.0;
  }
}

// ------------------------------------------------------------------ particles
// Counters of the tiles a block processed for one run, flushed with one
// atomic each when the block moves on to another run (or finishes).
struct TileCounts {
  uint32_t exits = 0, deaths = 0, moves = 0, overflow = 0, processed = 0;
};

__device__ __forceinline__ void flush_counts(RunState* rs, TileCounts& c) {
  if (c.exits) atomicAdd(&rs->step_exits, (unsigned long long)c.exits);
  if (c.deaths) atomicAdd(&rs->step_deaths, (unsigned long long)c.deaths);
  if (c.moves) atomicAdd(&rs->step_moves, (unsigned long long)c.moves);
  if (c.overflow) atomicAdd(&rs->step_overflow, (unsigned long long)c.overflow);
  if (c.processed) atomicAdd(&rs->step_processed, (unsigned long long)c.processed);
  c = TileCounts{};
}

// First column of a cumulative row with cp[c] >= u, clamped to m-1: for a
// non-decreasing row this is #{c < m-1 : cp[c] < u}, the result of the
// reference's bisection (move_kernel.hpp:89-102).  The count runs over chunks
// of 8 independent loads, one memory round trip each, and stops at the first
// chunk holding a column with cp[c] >= u: in a non-decreasing row no later
// column can count.  Most rows end within the first chunk.
__device__ __forceinline__ uint32_t cdf_column(const GLOBAL_AS double* cp, uint32_t m, double u) {
  constexpr int CH = 8;
  uint32_t left = 0;
  const int last = (int)m - 1;
  for (int c0 = 0; c0 < last; c0 += CH) {
    double v[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) v[j] = (c0 + j < last) ? cp[c0 + j] : 2.0;
    uint32_t k = 0;
#pragma unroll
    for (int j = 0; j < CH; ++j) k += (u > v[j]) ? 1u : 0u;
    left += k;
    if (k < (uint32_t)CH) break;
  }
  return left;
}

// The particle pass.  BLK threads process a TILE = 1024-slot tile, NPT = TILE/BLK
// slots each.  Each block takes a contiguous range of tiles, so it meets few
// runs, and runs it as a two-stage pipeline: the slot loads of tile t+1 are
// in flight while tile t is computed.  With PRIV the block accumulates its
// contributions in LDS (ds_add_f64) and flushes them to the run's S with
// coalesced atomics when it moves to another run; without PRIV every
// particle adds to S directly.
template <class M, int BLK, bool PRIV>
__global__ void __launch_bounds__(BLK)
    k_particle(BatchDev b, const RunDev* __restrict__ runs, const MapDev* __restrict__ maps,
               const uint32_t* __restrict__ tile_off) {
  constexpr int NV = M::NV, NS = M::NS, REC = rec_stride(M::NS);
  constexpr int NP = M::NP > 0 ? M::NP : 1;
  constexpr uint32_t NPT = TILE / BLK;   // slots per thread
  constexpr uint32_t NW = BLK / 64;      // waves per block
  extern __shared__ double s_acc[];      // [n_s][n_c] of the current run (PRIV)
  __shared__ uint32_t s_div[2][NPT][NW];  // double-buffered by tile parity:
  __shared__ uint32_t s_mov[2][NPT][NW];  // one barrier per tile suffices
  __shared__ uint32_t s_red[2][4][NW];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t stages = b.stages;
  const uint64_t P = b.P;
  const uint64_t ST = b.ST;
  const uint32_t stride = b.stage_stride;
  const uint32_t R = b.R;
  const uint32_t T = tile_off[R];
  const uint32_t chunk = (T + gridDim.x - 1) / gridDim.x;
  const uint32_t t_begin = blockIdx.x * chunk;
  const uint32_t t_end = (t_begin + chunk < T) ? t_begin + chunk : T;
  if (t_begin >= t_end) return;
  const bool do_model = (stages & STAGE_MODEL) != 0;
  const bool do_contrib = (stages & STAGE_CONTRIB) != 0;

  // run whose step/count/budget are cached in scalar registers
  uint32_t c_run = 0xFFFFFFFFu, c_n = 0, c_step = 0, c_budget = 0;
  auto run_state = [&](uint32_t r) {
    if (r != c_run) {
      const RunState* rs = b.rs + r;
      c_n = uniform(rs->n);
      c_step = uniform(rs->step);
      c_budget = uniform(rs->budget);
      c_run = r;
    }
  };

  // slots of one tile, as loaded
  struct Slots {
    uint8_t st[NPT];
    uint32_t pos[NPT];
    double x[NPT][NV];
    double ad[NPT], ah[NPT];
  };
  struct TileRef {
    uint32_t r, slot0, cnt;
    uint64_t base;
  };
  auto tile_ref = [&](uint32_t t) {
    TileRef q;
    q.r = find_run(tile_off, R, t);
    run_state(q.r);
    q.slot0 = (t - tile_off[q.r]) * TILE;
    q.cnt = (c_n - q.slot0 < TILE) ? c_n - q.slot0 : TILE;
    q.base = runs[q.r].p_off + q.slot0;
    return q;
  };
  auto load = [&](const TileRef& q, Slots& a) {
    const bool leave = runs[q.r].enable_leave && (stages & STAGE_LEAVE);
    const GLOBAL_AS uint8_t* st_b = gp(b.status) + q.base;
    const GLOBAL_AS uint32_t* pos_b = gp(b.pos) + q.base;
    const GLOBAL_AS double* ad_b = gp(b.age_div) + q.base;
    const GLOBAL_AS double* ah_b = gp(b.age_hyd) + q.base;
#pragma unroll
    for (uint32_t k = 0; k < NPT; ++k) {
      const uint32_t i = k * BLK + tid;
      const uint32_t li = i < q.cnt ? i : 0u;  // clamped: no load depends on another
      const uint8_t stv = ld_stream(st_b + li);
      a.st[k] = i < q.cnt ? stv : (uint8_t)EXIT;
      a.pos[k] = ld_stream(pos_b + li);
#pragma unroll
      for (int v = 0; v < NV; ++v) a.x[k][v] = ld_stream(gp(b.state) + (uint64_t)v * P + q.base + li);
      a.ad[k] = ld_stream(ad_b + li);
      a.ah[k] = leave ? ld_stream(ah_b + li) : 0.0;
    }
  };

  TileCounts cnt_acc;
  uint32_t acc_run = 0xFFFFFFFFu;
  GLOBAL_AS double* acc_S = nullptr;
  uint32_t acc_len = 0;

  TileRef cur = tile_ref(t_begin);
  Slots A;
  load(cur, A);
  for (uint32_t t = t_begin; t < t_end; ++t) {
    const uint32_t par = t & 1u;
    const uint32_t r = cur.r;
    const RunDev& rd = runs[r];
    if (r != acc_run) {  // block moves to another run: flush the last one's sums
      if (PRIV && acc_run != 0xFFFFFFFFu) {
        __syncthreads();
        for (uint32_t i = tid; i < acc_len; i += BLK) {
          const double v = s_acc[i];
          if (v != 0.0) unsafeAtomicAdd((double*)(acc_S + i), v);
        }
      }
      if (tid == 0 && acc_run != 0xFFFFFFFFu) flush_counts(b.rs + acc_run, cnt_acc);
      acc_run = r;
      if (PRIV) {
        acc_S = gp(rd.S);
        acc_len = rd.n_s * rd.n_c;
        __syncthreads();
        for (uint32_t i = tid; i < acc_len; i += BLK) s_acc[i] = 0.0;
        __syncthreads();
      }
    }
    run_state(r);
    const uint32_t step = c_step;
    const uint32_t budget = c_budget;
    const GLOBAL_AS double* __restrict__ rec = gp(rd.rec);

    // ---- compartment records of the current tile (needs its positions)
    double rr[NPT][NS + 2];
#pragma unroll
    for (uint32_t k = 0; k < NPT; ++k) {
      const GLOBAL_AS double* rp = rec + (size_t)A.pos[k] * REC;
#pragma unroll
      for (int q = 0; q < NS + 2; ++q) rr[k][q] = rp[q];
    }
    // ---- prefetch the next tile's slots
    Slots B;
    TileRef nxt = cur;
    if (t + 1 < t_end) {
      nxt = tile_ref(t + 1);
      load(nxt, B);
      run_state(r);
    }

    // ---- per-particle work on the current tile
    const double dt = rd.dt;
    const double w = rd.weight;
    const uint32_t n_c = rd.n_c;
    const uint64_t seed = rd.seed;
    GLOBAL_AS double* __restrict__ S = gp(rd.S);
    double p[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) p[q] = rd.params[q];
    const bool do_move = rd.enable_move && (stages & STAGE_MOVE);
    const bool do_leave = rd.enable_leave && (stages & STAGE_LEAVE);
    const uint64_t base = cur.base;
    const uint32_t slot0 = cur.slot0;
    GLOBAL_AS uint8_t* __restrict__ st_b = gp(b.status) + base;
    GLOBAL_AS double* __restrict__ ah_b = gp(b.age_hyd) + base;
    GLOBAL_AS double* __restrict__ ad_b = gp(b.age_div) + base;
    bool idle[NPT], want[NPT], mover[NPT];
    uint32_t n_exit = 0, n_dead = 0, n_move = 0, n_proc = 0;
#pragma unroll
    for (uint32_t k = 0; k < NPT; ++k) {
      idle[k] = A.st[k] == IDLE;
      want[k] = false;
      mover[k] = false;
      if (!idle[k]) continue;
      ++n_proc;
      const uint32_t i = k * BLK + tid;
      double contrib[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) contrib[s] = 0.0;
      int code = IDLE;
      if (do_model) {
        A.ad[k] += dt;
        code = M::update(p, dt, A.x[k], rr[k], contrib);
      }
      if (code == DEAD) {  // a dead cell stops contributing, moving and leaving
        st_b[i] = DEAD;
        idle[k] = false;
        ++n_dead;
#pragma unroll
        for (int v = 0; v < NV; ++v) (gp(b.state) + (uint64_t)v * P + base)[i] = A.x[k][v];
        ad_b[i] = A.ad[k];
        continue;
      }
      want[k] = code == DIVISION;
      if (do_contrib) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const double c = w * contrib[s];
          if (PRIV)
            atomicAdd(s_acc + s * n_c + A.pos[k], c);
          else
            unsafeAtomicAdd((double*)(S + (size_t)s * n_c + A.pos[k]), c);
        }
      }
      const U4 u = draw4(seed, slot0 + i, step, TAG_TRANSPORT, 0u);
      // a particle that leaves its compartment (dt*Q/V > u, probability_leaving.hpp:313-325)
      // is queued; k_move finds its neighbour and runs its exit test
      mover[k] = do_move && rr[k][NS] > u01(u.x);
      n_move += mover[k] ? 1u : 0u;
      if (do_leave) {
        A.ah[k] += dt;
        const double q = rr[k][NS + 1];
        if (!mover[k] && q < 1.0 && u01(u.z) > q) {  // u > exp(-dt*Q/V)  <=>  dt*Q > -ln(u)*V
          if (rd.probe_exit) {
            const unsigned long long e = atomicAdd(&b.rs[r].probe_exit_n, 1ull);
            if (e < rd.probe_cap) rd.probe_exit[e] = A.ah[k];
          }
          A.ah[k] = 0.0;
          st_b[i] = EXIT;
          ++n_exit;
        }
        st_stream(ah_b + i, A.ah[k]);
      }
      if (!want[k]) {  // final state now; dividers are written after ranking
#pragma unroll
        for (int v = 0; v < NV; ++v) st_stream(gp(b.state) + (uint64_t)v * P + base + i, A.x[k][v]);
        if (do_model) st_stream(ad_b + i, A.ad[k]);
      }
    }

    // ---- rank the division requests of the tile in slot order
    uint64_t mk[NPT], mv[NPT];
#pragma unroll
    for (uint32_t k = 0; k < NPT; ++k) {
      mk[k] = __ballot(want[k]);
      mv[k] = __ballot(mover[k]);
      if (lane == 0) {
        s_div[par][k][wave] = (uint32_t)__popcll(mk[k]);
        s_mov[par][k][wave] = (uint32_t)__popcll(mv[k]);
      }
    }
    n_exit = wave_sum(n_exit);
    n_dead = wave_sum(n_dead);
    n_move = wave_sum(n_move);
    n_proc = wave_sum(n_proc);
    if (lane == 0) {
      s_red[par][0][wave] = n_exit;
      s_red[par][1][wave] = n_dead;
      s_red[par][2][wave] = n_move;
      s_red[par][3][wave] = n_proc;
    }
    __syncthreads();
    uint32_t total = 0;
#pragma unroll
    for (uint32_t k = 0; k < NPT; ++k)
#pragma unroll
      for (uint32_t wv = 0; wv < NW; ++wv) total += s_div[par][k][wv];
    const uint32_t births = total < budget ? total : budget;
    if (total) {
      uint32_t before = 0;  // requests in earlier k-rows
#pragma unroll
      for (uint32_t k = 0; k < NPT; ++k) {
        uint32_t rank = before + (uint32_t)__popcll(mk[k] & lanemask_lt(lane));
#pragma unroll
        for (uint32_t wv = 0; wv < NW; ++wv) {
          rank += (wv < wave) ? s_div[par][k][wv] : 0u;
          before += s_div[par][k][wv];
        }
        if (!want[k]) continue;
        const uint32_t i = k * BLK + tid;
        if (rank < budget) {
          if (rd.probe_div) {
            const unsigned long long e = atomicAdd(&b.rs[r].probe_div_n, 1ull);
            if (e < rd.probe_cap) rd.probe_div[e] = A.ad[k];
          }
          double xd[NV];
          M::division(p, A.x[k], xd, RngKey{seed, slot0 + i, step});
          const uint64_t sidx = (uint64_t)t * stride + rank;
#pragma unroll
          for (int v = 0; v < NV; ++v) gp(b.st_state)[(uint64_t)v * ST + sidx] = xd[v];
          gp(b.st_pos)[sidx] = A.pos[k];
          A.ad[k] = 0.0;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) (gp(b.state) + (uint64_t)v * P + base)[i] = A.x[k][v];
        if (do_model) ad_b[i] = A.ad[k];
      }
    }
    // ---- the tile's movers, in slot order, for k_move
    uint32_t n_moves = 0;
#pragma unroll
    for (uint32_t k = 0; k < NPT; ++k) {
      uint32_t rank = n_moves + (uint32_t)__popcll(mv[k] & lanemask_lt(lane));
#pragma unroll
      for (uint32_t wv = 0; wv < NW; ++wv) {
        rank += (wv < wave) ? s_mov[par][k][wv] : 0u;
        n_moves += s_mov[par][k][wv];
      }
      if (mover[k]) {
        b.move_slot[(uint64_t)t * TILE + rank] = (uint16_t)(k * BLK + tid);
        b.move_pos[(uint64_t)t * TILE + rank] = A.pos[k];
      }
    }
    if (tid == 0) {
      b.tile_moves[t] = n_moves;
      b.tile_births[t] = births;
      for (uint32_t wv = 0; wv < NW; ++wv) {
        cnt_acc.exits += s_red[par][0][wv];
        cnt_acc.deaths += s_red[par][1][wv];
        cnt_acc.moves += s_red[par][2][wv];
        cnt_acc.processed += s_red[par][3][wv];
      }
      cnt_acc.overflow += total - births;
    }
    cur = nxt;
#pragma unroll
    for (uint32_t k = 0; k < NPT; ++k) {
      A.st[k] = B.st[k];
      A.pos[k] = B.pos[k];
#pragma unroll
      for (int v = 0; v < NV; ++v) A.x[k][v] = B.x[k][v];
      A.ad[k] = B.ad[k];
      A.ah[k] = B.ah[k];
    }
  }
  if (PRIV) {
    __syncthreads();
    for (uint32_t i = tid; i < acc_len; i += BLK) {
      const double v = s_acc[i];
      if (v != 0.0) unsafeAtomicAdd((double*)(acc_S + i), v);
    }
  }
  if (tid == 0) flush_counts(b.rs + acc_run, cnt_acc);
}

// ------------------------------------------------------------------ moves
// Neighbour pick for the particles the pass listed: the first CDF column at or
// above u, then the exit test in the new compartment (move_kernel.hpp:392-437,
// 586-648).  A block takes MOVE_GROUP tiles and spreads their movers over its
// threads; every mover is independent, so the processing order is free.

__global__ void __launch_bounds__(256) k_move(BatchDev b, const RunDev* __restrict__ runs,
                                              const MapDev* __restrict__ maps,
                                              const uint32_t* __restrict__ tile_off) {
  __shared__ uint32_t s_pre[MOVE_GROUP + 1];
  __shared__ uint32_t s_run[MOVE_GROUP];
  const uint32_t T = tile_off[b.R];
  const uint32_t n_groups = (T + MOVE_GROUP - 1) / MOVE_GROUP;
  const bool leave_stage = (b.stages & STAGE_LEAVE) != 0;
  for (uint32_t grp = blockIdx.x; grp < n_groups; grp += gridDim.x) {
    const uint32_t t0 = grp * MOVE_GROUP;
    if (threadIdx.x < 64) {  // the group's tiles, one lane each
      const uint32_t lane = threadIdx.x, t = t0 + lane;
      uint32_t cnt = 0, r = 0;
      if (lane < MOVE_GROUP && t < T) {
        cnt = b.tile_moves[t];
        r = find_run(tile_off, b.R, t);
      }
      const uint32_t inc = wave_incl_scan(cnt, lane);
      if (lane < MOVE_GROUP) {
        s_pre[lane] = inc - cnt;
        s_run[lane] = r;
      }
      if (lane == MOVE_GROUP - 1) s_pre[MOVE_GROUP] = inc;
    }
    __syncthreads();
    const uint32_t total = s_pre[MOVE_GROUP];
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      uint32_t j = 0;
      while (s_pre[j + 1] <= e) ++j;
      const uint32_t t = t0 + j;
      const uint32_t r = s_run[j];
      const RunDev& rd = runs[r];
      RunState* rs = b.rs + r;
      const uint64_t q = (uint64_t)t * TILE + (e - s_pre[j]);
      const uint32_t slot = (t - tile_off[r]) * TILE + b.move_slot[q];
      uint32_t pos = b.move_pos[q];
      const uint64_t g = rd.p_off + slot;
      const uint32_t step = rs->step;
      const MapDev& mp = maps[rd.map_base + (step / rd.steps_per_map) % rd.n_maps];
      const U4 u = draw4(rd.seed, slot, step, TAG_TRANSPORT, 0u);
      const uint32_t m = mp.m;
      const uint32_t col = cdf_column(gp(mp.cdf) + (size_t)pos * m, m, u01(u.y));
      pos = gp(mp.nbr)[(size_t)pos * m + col];
      b.pos[g] = pos;
      if (rd.enable_leave && leave_stage) {
        const double qx = gp(rd.rec)[(size_t)pos * rd.rec_w + rd.n_s + 1];
        if (qx < 1.0 && u01(u.z) > qx) {
          const double age = b.age_hyd[g];
          if (rd.probe_exit) {
            const unsigned long long k = atomicAdd(&rs->probe_exit_n, 1ull);
            if (k < rd.probe_cap) rd.probe_exit[k] = age;
          }
          b.age_hyd[g] = 0.0;
          b.status[g] = EXIT;
          atomicAdd(&rs->step_exits, 1ull);
        }
      }
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

// ------------------------------------------------------------------ layout
// Tile -> run table and per-tile division budget for the next pass:
// budget = min(stage_cap, (cap - n) / tiles), so staged newborns always fit.
template <uint32_t NT>
__device__ void layout_pass(const BatchDev& b) {
  constexpr uint32_t NWV = NT / 64;
  __shared__ uint32_t s_w[NWV];
  __shared__ uint32_t s_carry;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t r0 = 0; r0 < b.R; r0 += NT) {
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
    for (uint32_t w = 0; w < NWV; ++w) {
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

__global__ void __launch_bounds__(SCAN_THREADS) k_layout(BatchDev b) { layout_pass<SCAN_THREADS>(b); }

// ------------------------------------------------------------------ merge
// Appends staged newborns in tile order.
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
template <class M, int BLK, bool PRIV>
static void launch_particle_t(const BatchDev& b, int grid, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_particle<M, BLK, PRIV>), dim3(grid), dim3(BLK), lds, s, b, b.runs, b.maps,
                     (const uint32_t*)b.tile_off);
}

template <class M>
static const void* kernel_ptr_t(int blk, bool priv) {
  if (priv) {
    if (blk == 256) return (const void*)&k_particle<M, 256, true>;
    if (blk == 512) return (const void*)&k_particle<M, 512, true>;
    return (const void*)&k_particle<M, 1024, true>;
  }
  return (const void*)&k_particle<M, 256, false>;
}

template <class M>
static void launch_t(const BatchDev& b, const ParticleLaunch& L, hipStream_t s) {
  if (!L.priv)
    launch_particle_t<M, 256, false>(b, L.grid, 0, s);
  else if (L.block == 256)
    launch_particle_t<M, 256, true>(b, L.grid, L.lds, s);
  else if (L.block == 512)
    launch_particle_t<M, 512, true>(b, L.grid, L.lds, s);
  else
    launch_particle_t<M, 1024, true>(b, L.grid, L.lds, s);
}

void launch_particle(int model_id, const BatchDev& b, const ParticleLaunch& L, hipStream_t s) {
  switch (model_id) {
    case MonodMass::ID: launch_t<MonodMass>(b, L, s); break;
    case DualMonod::ID: launch_t<DualMonod>(b, L, s); break;
    case Monod::ID: launch_t<Monod>(b, L, s); break;
    case FixedLength::ID: launch_t<FixedLength>(b, L, s); break;
    case SimpleAcetate::ID: launch_t<SimpleAcetate>(b, L, s); break;
    default: break;
  }
}

const void* particle_kernel_ptr(int model_id, int block, bool priv) {
  switch (model_id) {
    case MonodMass::ID: return kernel_ptr_t<MonodMass>(block, priv);
    case DualMonod::ID: return kernel_ptr_t<DualMonod>(block, priv);
    case Monod::ID: return kernel_ptr_t<Monod>(block, priv);
    case FixedLength::ID: return kernel_ptr_t<FixedLength>(block, priv);
    case SimpleAcetate::ID: return kernel_ptr_t<SimpleAcetate>(block, priv);
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
void launch_move(const BatchDev& b, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_move, dim3(grid), dim3(256), 0, s, b, b.runs, b.maps, (const uint32_t*)b.tile_off);
}
void launch_layout(const BatchDev& b, hipStream_t s) {
  hipLaunchKernelGGL(k_layout, dim3(1), dim3(SCAN_THREADS), 0, s, b);
}
void launch_repartition(const BatchDev& b, uint32_t run, unsigned long long* hist, hipStream_t s) {
  hipLaunchKernelGGL(k_repartition, dim3(1024), dim3(256), 0, s, b, run, hist);
}

}  // namespace mcst
