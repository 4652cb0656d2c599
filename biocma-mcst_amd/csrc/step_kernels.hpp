// step_kernels.hpp -- the model-dependent kernels of the step (k_transport,
// k_particle, k_resolve, k_properties) as templates over the model, and the
// ModelOps table that launches them.  kernels.hip instantiates them for the
// built-in models; udf_plugin.hip for a model compiled into a plugin and
// loaded at run time (mcst_model_load).
#pragma once

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

constexpr uint32_t SCAN_THREADS = 1024;
constexpr uint32_t SCAN_WAVES = SCAN_THREADS / 64;

}  // namespace

// ------------------------------------------------------------------ transport
// m += dt*(C*M - C*diag(sink) + S + F);  C = m*diag(1/V)  for every
// compartment of every active run, one thread each, all species at once (the
// flow matrix entries are read once for all of them); also refreshes the
// per-compartment record {M::record(C), p_move, q_exit} the particle pass
// gathers, and clears S.  The extra last block lays out the step's tiles
// (layout_pass), which needs only the counts of the previous step.
template <uint32_t NT>
__device__ void layout_pass(const BatchDev& b);

// sum_i C[s][i] * M(i, j) over column j of a transport matrix stored by
// destination column, for every species: entries in groups of 8 whose loads
// are all in flight together (a column holds the degree plus the diagonal,
// rarely more than 8); the sums still run in entry order.
template <int NS>
__device__ __forceinline__ void column_inflow(const uint32_t* __restrict__ col_ptr, const uint32_t* __restrict__ src,
                                              const double* __restrict__ val, const double* c, uint32_t n_c,
                                              uint32_t j, double (&acc)[NS]) {
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const uint32_t e0 = col_ptr[j], e1 = col_ptr[j + 1];
  constexpr uint32_t EG = NS >= 4 ? 4u : 8u;  // entries in flight: fewer for many species (registers)
  for (uint32_t e = e0; e < e1; e += EG) {
    uint32_t i[EG];
    double v[EG], cv[EG][NS];
#pragma unroll
    for (uint32_t k = 0; k < EG; ++k) {
      const uint32_t ek = e + k < e1 ? e + k : e1 - 1u;
      i[k] = src[ek];
      v[k] = val[ek];
    }
#pragma unroll
    for (uint32_t k = 0; k < EG; ++k)
#pragma unroll
      for (int s = 0; s < NS; ++s) cv[k][s] = c[(size_t)s * n_c + i[k]];
#pragma unroll
    for (uint32_t k = 0; k < EG; ++k)
      if (e + k < e1)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[s] += cv[k][s] * v[k];
  }
}

// The transport of batch compartment g for the step `ahead` steps after its
// run's current one (0: the step about to run; 1: the next one, computed
// during this step's resolve kernel).  It reads C and the contributions of
// the last pass, from S (s_in null; S is then cleared) or from s_in (the
// compartment's NS summed contribution rows), and writes the other C buffer,
// m and the record table of its step's parity; the resolve kernel reads none
// of these but the other record table.
template <class M>
__device__ void transport_one(const BatchDev& b, uint32_t g, uint32_t ahead, const double* s_in) {
  constexpr int NS = M::NS;
  {
    const uint32_t r = find_run(b.comp_off, b.R, g);
    const RunDev& rd = b.runs[r];
    const RunState* rs = b.rs + r;
    // the step being prepared; the run's activity is recomputed here rather
    // than read from the layout, which other blocks write concurrently
    const uint32_t step = rs->step + ahead;
    if (step >= rd.step_target) return;
    const uint32_t j = g - b.comp_off[r];
    const uint32_t n_c = rd.n_c;
    const MapDev& mp = b.maps[rd.map_base + (step / rd.steps_per_map) % rd.n_maps];
    constexpr int NR = record_terms<M>();
    double* rec = rd.rec + ((size_t)(step & 1u) * n_c + j) * rd.rec_w;
    rec[NR] = mp.p_move[j];
    double qx = mp.q_exit[j];
    if (rd.n_vf && rd.vx_slot[j] >= 0) qx = rd.vf_q[rd.n_vf + (uint32_t)rd.vx_slot[j]];
    rec[NR + 1] = qx;
    const double* c = rd.C[step & 1u];
    double* Cn = rd.C[(step + 1u) & 1u];
    double cn[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) cn[s] = c[(size_t)s * n_c + j];
    if (b.stages & STAGE_TRANSPORT) {
      double acc[NS];
      column_inflow<NS>(mp.col_ptr, mp.src, mp.val, c, n_c, j, acc);
      double sink = rd.sink[j];
      const double inv_v = mp.inv_v[j];
      const bool gas = rd.Cg[0] != nullptr;
      double fv[NS], fvg[NS];  // time-dependent feeds of this compartment, in feed order
      double sink_g = gas ? rd.sink_g[j] : 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) fv[s] = fvg[s] = 0.0;
      if (rd.n_vf) {
        for (uint32_t q = rd.vf_sink_ptr[j]; q < rd.vf_sink_ptr[j + 1]; ++q) {
          const uint32_t k = rd.vf_sink_k[q];
          if (k & VF_GAS)
            sink_g = sink_g + rd.vf_q[k & ~VF_GAS];
          else
            sink = sink + rd.vf_q[k];
        }
        for (uint32_t q = rd.vf_src_ptr[j]; q < rd.vf_src_ptr[j + 1]; ++q) {
          const double add = rd.vf_q[rd.vf_src_k[q]] * rd.vf_src_c[q];
          const uint32_t sp = rd.vf_src_s[q];
#pragma unroll
          for (int s = 0; s < NS; ++s) {
            if ((uint32_t)s == sp) fv[s] += add;
            if (((uint32_t)s | VF_GAS) == sp) fvg[s] += add;
          }
        }
      }
      if (gas) {
        // two-phase step (SimulationUnit::ode_step, simulation.model.cpp:131-150):
        // mtr = kla*(H*C_gas - C)*V from the old concentrations, the gas
        // step with -mtr and the liquid step with +mtr (performStepGL,
        // implScalar.cpp:228-248), then clearNegs on the liquid
        // (implScalar.cpp:266-297): a value in (-1e-4*5e-3, 0) becomes 0
        const double* cg = rd.Cg[step & 1u];
        double* Cgn = rd.Cg[(step + 1u) & 1u];
        double accg[NS];
        column_inflow<NS>(mp.g_col_ptr, mp.g_src, mp.g_val, cg, n_c, j, accg);
        const double inv_vg = mp.g_inv_v[j];
        const double tol = 1e-4 * 5e-3;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const size_t o = (size_t)s * n_c + j;
          const double cgs = cg[o];
          const double mtr = (mp.g_kla[o] * ((cgs * rd.henry[s]) - cn[s])) * mp.vol[j];
          const double dg = ((accg[s] - cgs * sink_g) + (rd.Fg[o] + fvg[s])) - mtr;
          const double mgo = rd.mg[o] + rd.dt * dg;
          rd.mg[o] = mgo;
          Cgn[o] = mgo * inv_vg;
          const double sv = s_in ? s_in[s] : rd.S[o];  // the contributions of the last pass
          const double dl = ((acc[s] - cn[s] * sink) + (sv + (rd.F[o] + fv[s]))) + mtr;
          const double mm = rd.m[o] + rd.dt * dl;
          rd.m[o] = mm;
          double cl = mm * inv_v;
          if (cl < 0.0 && fabs(cl) < tol) cl = 0.0;
          cn[s] = cl;
        }
      } else {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const size_t o = (size_t)s * n_c + j;
          const double sv = s_in ? s_in[s] : rd.S[o];  // the contributions of the last pass
          double dmdt = (acc[s] - cn[s] * sink) + (sv + (rd.F[o] + fv[s]));
          // gas-liquid transfer from a fixed gas phase: kla*(C* - C)*V
          if (rd.kla) dmdt = dmdt + (rd.kla[o] * (rd.cstar[o] - cn[s])) * mp.vol[j];
          const double mm = rd.m[o] + rd.dt * dmdt;
          rd.m[o] = mm;
          cn[s] = mm * inv_v;
        }
      }
    }
    double rt[NR];
    M::record(rd.params, cn, rt);
#pragma unroll
    for (int k = 0; k < NR; ++k) rec[k] = rt[k];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const size_t o = (size_t)s * n_c + j;
      Cn[o] = cn[s];
      if (!s_in) rd.S[o] = 0.0;
    }
  }
}

// The transport of every active run's compartments (transport_one from S),
// by blocks `blk` of `n_blk`.
template <class M>
__device__ void transport_compartments(const BatchDev& b, uint32_t blk, uint32_t n_blk, uint32_t ahead) {
  const uint32_t total = b.n_comp_total;
  const uint32_t stride = n_blk * blockDim.x;
  for (uint32_t g = blk * blockDim.x + threadIdx.x; g < total; g += stride) transport_one<M>(b, g, ahead, nullptr);
}

// With contribution rows: block `blk` of the resolve kernel's aux blocks sums
// the rows of compartments [blk*CPB, (blk+1)*CPB) of the run, every species,
// in a fixed order (each lane its rows in turn, then the wave's four lane
// groups by a butterfly, then the waves in turn), then transports those compartments for the
// next step with the sums (ahead) or stores the sums in S (the last step of a
// call or of the run, so that S holds the step's contributions whenever the
// host can read it).
template <class M>
__device__ void sum_rows(const BatchDev& b, uint32_t blk, bool ahead) {
  constexpr int NS = M::NS;
  constexpr uint32_t CPB = part_cpb<NS>();
  constexpr uint32_t NE = (uint32_t)NS * CPB;  // 16: entries of the block, one per lane of a group of 16
  constexpr uint32_t NWV = 256 / 64;
  static_assert(NE <= 16, "at most 16 entries per block");
  __shared__ double s_wsum[NWV][16];
  __shared__ double s_sum[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t r = b.part_run;
  const RunDev& rd = b.runs[r];
  const uint32_t n_c = rd.n_c;
  const uint32_t c0 = blk * CPB;
  // lane group g (16 lanes) takes rows g, g + 16, ...; lane e of the group
  // the entry e = species s, compartment c0 + k: a wave's load covers four
  // rows, 64 contiguous bytes of each per species
  const uint32_t e = tid & 15u, grp16 = tid >> 4;
  const uint32_t sp = e / CPB, k = e % CPB;
  const bool on = e < NE && c0 + k < n_c;
  const size_t off = on ? (size_t)sp * n_c + c0 + k : 0;
  double acc = 0.0;
#pragma unroll 8
  for (uint32_t row = grp16; row < b.part_rows; row += 16u)
    if (on) acc += b.part[(size_t)row * b.part_w + off];
  acc += __shfl_xor(acc, 16, 64);
  acc += __shfl_xor(acc, 32, 64);
  if (lane < 16) s_wsum[wave][lane] = acc;
  __syncthreads();
  if (tid < 16) {
    double t = 0.0;
#pragma unroll
    for (uint32_t w = 0; w < NWV; ++w) t += s_wsum[w][tid];
    s_sum[tid] = t;
  }
  __syncthreads();
  if (tid < CPB && c0 + tid < n_c) {
    double sv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) sv[s] = s_sum[s * CPB + tid];
    const uint32_t step = b.rs[r].step;  // the books advance it after every aux block is done
    if (ahead && step + 1u < rd.step_target) {
      transport_one<M>(b, b.comp_off[r] + c0 + tid, 1u, sv);
    } else if (step < rd.step_target) {  // the run took this step, its last of the call or its last
#pragma unroll
      for (int s = 0; s < NS; ++s) rd.S[(size_t)s * n_c + c0 + tid] = sv[s];
    }
  }
}

template <class M>
__global__ void __launch_bounds__(256) k_transport(BatchDev b) {
  if (blockIdx.x == gridDim.x - 1) {
    layout_pass<256>(b);
    return;
  }
  transport_compartments<M>(b, blockIdx.x, gridDim.x - 1, 0u);
}

// ------------------------------------------------------------------ particles
// Counters of the tiles a block processed for one run, flushed with one
// atomic each when the block moves on to another run (or finishes).
struct TileCounts {
  uint32_t exits = 0, deaths = 0, moves = 0, requests = 0, processed = 0;
};

__device__ __forceinline__ void flush_counts(RunState* rs, TileCounts& c) {
  if (c.exits) atomicAdd(&rs->step_exits, (unsigned long long)c.exits);
  if (c.deaths) atomicAdd(&rs->step_deaths, (unsigned long long)c.deaths);
  if (c.moves) atomicAdd(&rs->step_moves, (unsigned long long)c.moves);
  if (c.requests) atomicAdd(&rs->step_requests, (unsigned long long)c.requests);
  if (c.processed) atomicAdd(&rs->step_processed, (unsigned long long)c.processed);
  c = TileCounts{};
}

// Adds a block's LDS contribution accumulator, [n_s][win] for compartments
// [w0, w0 + win) of a run with n_c compartments, to S: one atomic per
// nonzero entry.
template <int BLK>
__device__ __forceinline__ void flush_window(const double* s_acc, GLOBAL_AS double* S, uint32_t ns, uint32_t win,
                                             uint32_t n_c, uint32_t w0, uint32_t tid) {
  const uint32_t len = ns * win;
  if (win == n_c) {  // the whole run: the layouts coincide
    for (uint32_t i = tid; i < len; i += BLK) {
      const double v = s_acc[i];
      if (v != 0.0) unsafeAtomicAdd((double*)(S + i), v);
    }
    return;
  }
  for (uint32_t i = tid; i < len; i += BLK) {
    const double v = s_acc[i];
    if (v != 0.0) {
      const uint32_t q = i / win;
      unsafeAtomicAdd((double*)(S + (size_t)q * n_c + w0 + (i - q * win)), v);
    }
  }
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

// Loads and stores of the NL consecutive slots of one lane: 8-byte fields
// move in 16-byte accesses per lane when NL > 1 (aligned: every run's slots
// start on a tile boundary), 4-byte fields in one access of 4*NL bytes.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int NL>
__device__ __forceinline__ void ld_lane(const GLOBAL_AS double* p, double (&o)[NL]) {
  if constexpr (NL == 1) {
    o[0] = ld_stream(p);
  } else {
#pragma unroll
    for (int h = 0; h < NL / 2; ++h) {
      const f64x2 v = ld_stream((const GLOBAL_AS f64x2*)p + h);
      o[2 * h] = v[0];
      o[2 * h + 1] = v[1];
    }
  }
}
template <int NL>
__device__ __forceinline__ void ld_lane(const GLOBAL_AS uint32_t* p, uint32_t (&o)[NL]) {
  if constexpr (NL == 1) {
    o[0] = ld_stream(p);
  } else if constexpr (NL == 2) {
    const u32x2 v = ld_stream((const GLOBAL_AS u32x2*)p);
    o[0] = v[0];
    o[1] = v[1];
  } else {
    const u32x4 v = ld_stream((const GLOBAL_AS u32x4*)p);
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = v[j];
  }
}
template <int NL>
__device__ __forceinline__ uint32_t ld_lane(const GLOBAL_AS uint8_t* p) {  // NL status bytes
  if constexpr (NL == 1)
    return ld_stream(p);
  else if constexpr (NL == 2)
    return ld_stream((const GLOBAL_AS uint16_t*)p);
  else
    return ld_stream((const GLOBAL_AS uint32_t*)p);
}
template <int NL>
__device__ __forceinline__ void st_lane(GLOBAL_AS double* p, const double (&v)[NL]) {
  if constexpr (NL == 1) {
    st_stream(p, v[0]);
  } else {
#pragma unroll
    for (int h = 0; h < NL / 2; ++h) {
      f64x2 w;
      w[0] = v[2 * h];
      w[1] = v[2 * h + 1];
      st_stream((GLOBAL_AS f64x2*)p + h, w);
    }
  }
}
template <int NL>
__device__ __forceinline__ void st_status(GLOBAL_AS uint8_t* p, uint32_t v) {
  if constexpr (NL == 1)
    *p = (uint8_t)v;
  else if constexpr (NL == 2)
    *(GLOBAL_AS uint16_t*)p = (uint16_t)v;
  else
    *(GLOBAL_AS uint32_t*)p = v;
}

// The particle pass.  The unit of work is a sub-chunk of SUB = 64*NL slots;
// lane l holds the slots j*64 + l (ILV: every field access of the wave is a
// contiguous 64-slot run) or NL*l + j (16-byte accesses per lane).  A
// sub-chunk lists its movers and its division requests in its own segment of
// the tile's lists, so sub-chunks are independent: the block's waves take
// the sub-chunks of one group (one per wave) and groups are dealt to blocks G
// at a time (b.tile_group; 0: one contiguous range per block).  Which
// requests the run accepts (the first min(cap - n, div_cap) in slot order)
// is decided after the pass (k_resolve).  With PRIV the block accumulates the
// contributions of one run in LDS (ds_add_f64), switching runs only between
// groups; sub-chunks of another run add to S directly.
//
// Memory pipeline of a wave, per sub-chunk: issue the next sub-chunk's loads;
// compute on the current one with its compartment records (gathered one step
// earlier); gather the records of the next sub-chunk; store the current one.
// Every load is issued for every lane (slots past the count read valid
// memory of the tile and are masked where used), so the waits the compiler
// places for a load's operands do not depend on the data.
//
// LV: some run of the batch has a leaving flow.  Without one (LV false) the
// pass issues no load or store of hydraulic ages and gathers no exit
// threshold: the loads a run without an outlet would issue anyway, to a
// single address, took 3 % of the pass on the 1e7-particle network, and a
// run-time branch around them cost more than it saved.
template <class M, int BLK, int NL, bool ILV, bool PRIV, bool PROBE, bool WIN, bool LV>
__global__ void __launch_bounds__(BLK)
    k_particle(BatchDev b, const RunDev* __restrict__ runs, const RunState* __restrict__ rsr,
               const uint32_t* __restrict__ tile_off) {
  constexpr int NV = M::NV, NS = M::NS, NR = record_terms<M>(), REC = rec_stride(NR);
  constexpr int NQ = NR + 2;             // record fields: M::record terms r_0..r_{NR-1}, p_move, q_exit
  constexpr int NG = LV ? NQ : NQ - 1;   // fields gathered (q_exit only when some run leaves)
  constexpr int NP = M::NP > 0 ? M::NP : 1;
  constexpr uint32_t WPB = BLK / 64;     // waves per block = sub-chunks per group
  constexpr uint32_t SUB = 64 * NL;       // slots per sub-chunk
  static_assert(NL == 1 || NL == 2 || NL == 4, "1, 2 or 4 slots per lane");
  static_assert(NR >= 1 && NR + 2 <= 16, "1..14 record terms");
  constexpr uint32_t NSUB = TILE / SUB;  // sub-chunks per tile (b.nsub is set to the same)
  extern __shared__ double s_acc[];      // [n_s][acc_win] of run acc_run (PRIV)
  __shared__ uint32_t s_w0;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = uniform(tid >> 6);  // wave-uniform: scalar bookkeeping
  const uint32_t stages = b.stages;
  const uint64_t P = b.P;
  const uint32_t R = b.R;
  const uint32_t T = tile_off[R];
  const uint32_t n_sc = T * NSUB;        // sub-chunks of the pass
  const uint32_t n_groups = (n_sc + WPB - 1) / WPB;
  const uint32_t chunk = (n_groups + gridDim.x - 1) / gridDim.x;
  const uint32_t G = b.tile_group ? b.tile_group : (chunk ? chunk : 1u);
  uint32_t grp = blockIdx.x * G;
  // contribution rows: this block's row of the pass (rows of all blocks are
  // summed by the resolve kernel)
  GLOBAL_AS double* const part_row =
      (PRIV && !WIN && b.part && blockIdx.x < b.part_rows) ? gp(b.part) + (size_t)blockIdx.x * b.part_w : nullptr;
  if (grp >= n_groups) {  // the whole block: its row holds nothing
    if (part_row)
      for (uint32_t i = threadIdx.x; i < b.part_w; i += BLK) part_row[i] = 0.0;
    return;
  }
  auto next_group = [&](uint32_t g) -> uint32_t {
    const uint32_t g0 = g - g % G;
    return (g + 1 < g0 + G) ? g + 1 : g0 + gridDim.x * G;
  };
  const bool do_model = (stages & STAGE_MODEL) != 0;
  const bool do_contrib = (stages & STAGE_CONTRIB) != 0;

  // one sub-chunk as the wave sees it; sc >= n_sc is a phantom: no slots,
  // no stores.  cnt counts the sub-chunk's slots below the run's count.
  struct SubRef {
    uint32_t sc, t, r, slot0, cnt;  // slot0: run-relative slot of the sub-chunk's first slot
    uint64_t base;                   // its batch-wide slot
    const double* rec;               // the run's record table of its current step
  };
  auto sub_ref = [&](uint32_t sc) -> SubRef {
    SubRef q{sc, 0u, 0u, 0u, 0u, 0ull, nullptr};
    // a phantom stands at the first slot of the pass's last tile: the run of
    // that tile is live, so the loads of the phantom (particle arrays, the
    // run's compartment record) stay inside memory the run owns; the run of
    // index 0 may have been retired and its arrays freed
    const uint32_t t = sc < n_sc ? sc / NSUB : T - 1u;
    q.t = t;
    q.r = find_run(tile_off, R, t);
    q.slot0 = (t - tile_off[q.r]) * TILE + (sc < n_sc ? (sc % NSUB) * SUB : 0u);
    if (sc < n_sc) {
      const uint32_t n = rsr[q.r].n;
      q.cnt = n > q.slot0 ? (n - q.slot0 < SUB ? n - q.slot0 : SUB) : 0u;
    }
    q.base = runs[q.r].p_off + q.slot0;
    q.rec = runs[q.r].rec + ((rsr[q.r].step & 1u) ? (size_t)runs[q.r].n_c * runs[q.r].rec_w : (size_t)0);
    return q;
  };
  // sub-chunk-relative slot of the lane's j-th slot: interleaved (64 apart,
  // every field access a contiguous 64-lane run) or NL consecutive slots
  auto off = [&](int j) -> uint32_t { return ILV ? (uint32_t)j * 64u + lane : NL * lane + (uint32_t)j; };
  // the NL slots of a lane in one sub-chunk, as loaded
  struct Sub {
    uint32_t st;       // NL status bytes
    uint32_t pos[NL];
    double x[NV][NL];
  };
  auto load_sub = [&](const SubRef& q, Sub& a) {
    if constexpr (ILV) {
      uint32_t st = 0;
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        const uint64_t ij = q.base + (q.cnt ? off(j) : 0u);  // an empty sub-chunk reads its first slot
        st |= (uint32_t)ld_stream(gp(b.status) + ij) << (8 * j);
        a.pos[j] = ld_stream(gp(b.pos) + ij);
#pragma unroll
        for (int v = 0; v < NV; ++v) a.x[v][j] = ld_stream(gp(b.state) + (uint64_t)v * P + ij);
      }
      a.st = st;
    } else {
      const uint64_t i0 = q.base + (q.cnt ? NL * lane : 0u);
      a.st = ld_lane<NL>(gp(b.status) + i0);
      ld_lane<NL>(gp(b.pos) + i0, a.pos);
#pragma unroll
      for (int v = 0; v < NV; ++v) ld_lane<NL>(gp(b.state) + (uint64_t)v * P + i0, a.x[v]);
    }
  };
  auto gather = [&](const SubRef& q, const Sub& a, double (&rr)[NL][NQ]) {
    const GLOBAL_AS double* __restrict__ rec = gp(q.rec);
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      // slots past the count hold no particle: their position may be anything
      const uint32_t pj = off(j) < q.cnt ? a.pos[j] : 0u;
      const GLOBAL_AS double* rp = rec + (size_t)pj * REC;
#pragma unroll
      for (int f = 0; f < NG; ++f) rr[j][f] = rp[f];
      if constexpr (!LV) rr[j][NQ - 1] = 1.0;  // no exit test without a leaving flow
    }
  };

  // per-wave counters of the run cnt_run, flushed when the wave changes runs
  TileCounts cnt_acc;
  uint32_t cnt_run = 0xFFFFFFFFu;
  // the wave's division requests summed per block of each level of the
  // request counts (div_l1/2/3: blocks of 2^6, 2^12, 2^18 sub-chunks),
  // added with one atomic when the wave leaves the block: its sub-chunks
  // come in increasing order.  An atomic per sub-chunk with requests put
  // most of a step's atomics on the few top-level entries when a tenth of
  // the cells divide together (the lattice tank: 1e6 requests per step)
  uint32_t dv_i[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, dv_n[3] = {0u, 0u, 0u};
  auto flush_requests = [&](int lvl) {
    if (dv_n[lvl] && lane == 0) {
      uint32_t* const base = lvl == 0 ? b.div_l1 : lvl == 1 ? b.div_l2 : b.div_l3;
      atomicAdd(base + dv_i[lvl], dv_n[lvl]);
    }
    dv_n[lvl] = 0u;
  };
  // the LDS accumulator holds compartments [acc_w0, acc_w0 + acc_win) of run
  // acc_run: the whole run when it fits (acc_win = n_c), else a window
  // centred on the block's first slot of the run (particles are stored in
  // compartment order, so a block's slots see few compartments); particles
  // outside the window add to S directly
  uint32_t acc_run = 0xFFFFFFFFu;
  GLOBAL_AS double* acc_S = nullptr;
  uint32_t acc_ns = 0, acc_nc = 0, acc_win = 0, acc_w0 = 0;

  SubRef cur = sub_ref(grp * WPB + wave);
  Sub A;
  double rr[NL][NQ];
  load_sub(cur, A);
  gather(cur, A, rr);
  // Later sub-chunks enter with their slots copied from the prefetch and
  // their records in flight; taking the first one's registers through empty
  // asm statements makes the entry path match.
  asm volatile("" : "+v"(A.st));
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    asm volatile("" : "+v"(A.pos[j]));
#pragma unroll
    for (int v = 0; v < NV; ++v) asm volatile("" : "+v"(A.x[v][j]));
#pragma unroll
    for (int f = 0; f < NQ; ++f) asm volatile("" : "+v"(rr[j][f]));
  }

  for (;;) {
    // ---- block-uniform: the run whose contributions the LDS accumulates
    if (PRIV) {
      const uint32_t r0 = find_run(tile_off, R, (grp * WPB) / NSUB);
      if (r0 != acc_run) {
        __syncthreads();
        flush_window<BLK>(s_acc, acc_S, acc_ns, acc_win, acc_nc, acc_w0, tid);
        acc_run = r0;
        acc_S = gp(runs[r0].S);
        acc_ns = runs[r0].n_s;
        acc_nc = runs[r0].n_c;
        acc_win = acc_nc;
        if constexpr (WIN) {
          acc_win = b.acc_w < acc_nc ? b.acc_w : acc_nc;
          if (tid == 0) {
            // wave 0 holds the group's first sub-chunk; its lane 0 the first slot
            const uint32_t c0 = cur.cnt ? A.pos[0] : 0u;
            const uint32_t half = acc_win / 2;
            const uint32_t lo = c0 > half ? c0 - half : 0u;
            s_w0 = lo + acc_win <= acc_nc ? lo : acc_nc - acc_win;
          }
          __syncthreads();
          acc_w0 = s_w0;
        }
        for (uint32_t i = tid; i < acc_ns * acc_win; i += BLK) s_acc[i] = 0.0;
        __syncthreads();
      } else if constexpr (WIN) {
        // the window follows the block's slots: when the group's first
        // compartment leaves the window's lower half (slots are in
        // compartment order, newborns at the end of the run in parent
        // order), flush it and re-centre a quarter ahead.  Checked on every
        // other group (grp is the same in every wave): the check costs two
        // block barriers, and the window only decides which contributions
        // take the LDS and which take global atomics.  On the 50 000-
        // compartment lattice at 1e8 particles, checking every group cost 6 %
        // of the step; every 8th or 16th let the window lag behind the
        // newborns and cost more
        if (acc_win < acc_nc && (grp & 1u) == 0) {
          if (tid == 0) s_w0 = cur.cnt ? A.pos[0] : acc_w0;
          __syncthreads();
          const uint32_t c0 = s_w0;
          __syncthreads();  // every wave has read s_w0
          if (c0 - acc_w0 >= acc_win / 2) {  // wraps when below the window
            flush_window<BLK>(s_acc, acc_S, acc_ns, acc_win, acc_nc, acc_w0, tid);
            const uint32_t q = acc_win / 4;
            const uint32_t lo = c0 > q ? c0 - q : 0u;
            acc_w0 = lo + acc_win <= acc_nc ? lo : acc_nc - acc_win;
            __syncthreads();  // the flush has read every entry
            for (uint32_t i = tid; i < acc_ns * acc_win; i += BLK) s_acc[i] = 0.0;
            __syncthreads();
          }
        }
      }
    }
    // ---- the wave's sub-chunk of this group
    const uint32_t sc = cur.sc;
    const uint32_t r = cur.r;
    const RunDev& rd = runs[r];
    if (cur.cnt && r != cnt_run) {  // wave moves to another run
      if (cnt_run != 0xFFFFFFFFu && lane == 0) flush_counts(b.rs + cnt_run, cnt_acc);
      cnt_acc = TileCounts{};
      cnt_run = r;
    }
    const uint32_t step = rsr[r].step;
    // the run's age clocks: the division clock of this step (after its tick
    // when the model runs), the hydraulic clock before this step's tick and
    // after it
    const uint32_t clk_d = rsr[r].clock_div + (do_model ? 1u : 0u);
    const uint32_t clk_h0 = rsr[r].clock_hyd;
    const double dt = rd.dt;
    const double w = rd.weight;
    const uint32_t n_c = rd.n_c;
    const uint64_t seed = rd.seed;
    GLOBAL_AS double* __restrict__ S = gp(rd.S);
    const bool to_lds = PRIV && r == acc_run;
    double p[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) p[q] = rd.params[q];
    const bool do_move = rd.enable_move && (stages & STAGE_MOVE);
    const bool do_leave = LV && rd.enable_leave && (stages & STAGE_LEAVE);
    const uint32_t clk_h = clk_h0 + (do_leave ? 1u : 0u);
    const uint32_t g_next = next_group(grp);

    // ---- prefetch: the wave's sub-chunk of the next group (a phantom past
    // the last group)
    const SubRef nxt = sub_ref(g_next < n_groups ? g_next * WPB + wave : n_sc);
    __builtin_amdgcn_sched_barrier(0);
    Sub B;
    load_sub(nxt, B);
    __builtin_amdgcn_sched_barrier(0);

    // ---- per-particle work: arithmetic only
    const uint32_t i0 = NL * lane;  // contiguous layout: sub-chunk-relative slot of j = 0
    const bool pair_valid = i0 < cur.cnt;
    bool valid[NL], idle[NL], want[NL], mover[NL], exited[NL], dead[NL];
    double x[NL][NV];
    double contrib[NL][NS];
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      const uint32_t stj = (A.st >> (8 * j)) & 0xFFu;
      valid[j] = off(j) < cur.cnt;
      idle[j] = valid[j] && stj == IDLE;
      want[j] = mover[j] = exited[j] = dead[j] = false;
#pragma unroll
      for (int v = 0; v < NV; ++v) x[j][v] = A.x[v][j];
#pragma unroll
      for (int q = 0; q < NS; ++q) contrib[j][q] = 0.0;
      if (!valid[j]) continue;
      if (idle[j]) {  // its division age ticks with the run's clock
        int code = IDLE;
        if (do_model) code = M::update(p, dt, x[j], rr[j], contrib[j]);
        if constexpr (M::CAN_DIE) dead[j] = code == DEAD;  // stops contributing and leaving
        want[j] = code == DIVISION;
      }
      // every slot below the count takes the move test, whatever its status
      // (move_kernel.hpp:209-273: Exit and Dead slots waiting for removal
      // keep moving); a particle that leaves its compartment (dt*Q/V > u,
      // probability_leaving.hpp:33-46) is listed, and k_resolve finds its
      // neighbour and, for an Idle one, runs its exit test there
      const U4 u = draw4(seed, cur.slot0 + off(j), step, TAG_TRANSPORT, 0u);
      mover[j] = do_move && rr[j][NR] > u01(u.x);
      if (do_leave && idle[j] && !dead[j]) {  // its hydraulic age ticks with the run's clock
        const double qx = rr[j][NR + 1];
        if (!mover[j] && qx < 1.0 && u01(u.z) > qx) {  // u > exp(-dt*Q/V)  <=>  dt*Q > -ln(u)*V
          if constexpr (PROBE) {
            if (rd.probe_exit) {
              const uint64_t g = cur.base + off(j);
              const double age = clock_age(gp(b.age_hyd)[g], clk_h, gp(b.clk_hyd)[g], dt);
              const unsigned long long e = atomicAdd(&b.rs[r].probe_exit_n, 1ull);
              if (e < rd.probe_cap) rd.probe_exit[e] = age;
            }
          }
          exited[j] = true;
        }
      }
    }
    // contributions of the Idle particles (a dead cell contributes nothing)
    if (do_contrib) {
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        if (!idle[j] || dead[j]) continue;
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          const double c = w * contrib[j][q];
          if constexpr (WIN) {
            const uint32_t k = A.pos[j] - acc_w0;  // wraps above the window when below it
            if (to_lds && k < acc_win)
              atomicAdd(s_acc + q * acc_win + k, c);
            else
              unsafeAtomicAdd((double*)(S + (size_t)q * n_c + A.pos[j]), c);
          } else {
            if (to_lds)
              atomicAdd(s_acc + q * n_c + A.pos[j], c);
            else
              unsafeAtomicAdd((double*)(S + (size_t)q * n_c + A.pos[j]), c);
          }
        }
      }
    }

    // ---- the records of the next sub-chunk (its positions have arrived)
    double rn[NL][NQ];
    gather(nxt, B, rn);

    // ---- ranks of the sub-chunk's movers and division requests in slot order
    const uint64_t lt = lanemask_lt(lane);
    const uint64_t seg = (uint64_t)sc * SUB;  // the sub-chunk's list segment
    const uint32_t slot_t = cur.slot0 % TILE;  // tile-relative slot of the sub-chunk's first slot
    uint32_t n_mov = 0, n_div = 0;
    uint32_t rm[NL], rd_[NL];  // rank of slot off(j) among the flagged slots of the sub-chunk
    {
      uint64_t mv[NL], wd[NL];
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        mv[j] = __ballot(mover[j]);
        wd[j] = __ballot(want[j]);
      }
      if constexpr (ILV) {
        uint32_t am = 0, adv = 0;
#pragma unroll
        for (int j = 0; j < NL; ++j) {
          rm[j] = am + (uint32_t)__popcll(mv[j] & lt);
          rd_[j] = adv + (uint32_t)__popcll(wd[j] & lt);
          am += (uint32_t)__popcll(mv[j]);
          adv += (uint32_t)__popcll(wd[j]);
        }
        n_mov = am;
        n_div = adv;
      } else {
        uint32_t bm = 0, bd = 0;
#pragma unroll
        for (int j = 0; j < NL; ++j) {
          bm += (uint32_t)__popcll(mv[j] & lt);
          bd += (uint32_t)__popcll(wd[j] & lt);
          n_mov += (uint32_t)__popcll(mv[j]);
          n_div += (uint32_t)__popcll(wd[j]);
        }
#pragma unroll
        for (int j = 0; j < NL; ++j) {
          rm[j] = bm;
          rd_[j] = bd;
          bm += mover[j] ? 1u : 0u;
          bd += want[j] ? 1u : 0u;
        }
      }
    }
    // ---- divisions in registers (rare): the first DIV_STAGE requests of the
    // sub-chunk are divided here, where the mother's state is at hand.  Her
    // state before division is staged, one contiguous record per request; the
    // halved mother is stored below with everything else.  k_resolve decides
    // acceptance: an accepted request writes the daughter (M::division of the
    // staged state again: the same draws, so the same daughter) and restarts
    // the mother's division age, a refused one restores the mother from the
    // stage.  Later requests of a crowded sub-chunk are divided by k_resolve
    // from the stored state.
    if (n_div) {
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        if (!(want[j] && rd_[j] < DIV_STAGE)) continue;
        GLOBAL_AS double* stg = gp(b.div_stage) + ((uint64_t)sc * DIV_STAGE + rd_[j]) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) stg[v] = x[j][v];
        double xd[NV];
        M::division(p, x[j], xd, RngKey{seed, cur.slot0 + off(j), step});
      }
    }

    // ---- stores: the lane's final state, status changes, and the
    // sub-chunk's movers and division requests
    const uint64_t gi = cur.base + i0;
    if constexpr (ILV) {
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        if (!valid[j]) continue;
        const uint64_t gj = cur.base + off(j);
#pragma unroll
        for (int v = 0; v < NV; ++v) st_stream(gp(b.state) + (uint64_t)v * P + gj, x[j][v]);
        if (exited[j] || dead[j]) gp(b.status)[gj] = dead[j] ? (uint8_t)DEAD : (uint8_t)EXIT;
      }
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        double xv[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) xv[j] = x[j][v];
        if (pair_valid) st_lane<NL>(gp(b.state) + (uint64_t)v * P + gi, xv);
      }
      uint32_t stn = A.st;
      bool changed = false;
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        const uint32_t nv = dead[j] ? (uint32_t)DEAD : exited[j] ? (uint32_t)EXIT : ((A.st >> (8 * j)) & 0xFFu);
        stn = (stn & ~(0xFFu << (8 * j))) | (nv << (8 * j));
        changed = changed || exited[j] || dead[j];
      }
      if (changed && pair_valid) st_status<NL>(gp(b.status) + gi, stn);
    }
    // a slot that stops being Idle keeps its ages from now on: the hydraulic
    // age (0 after an exit; a dead cell's without this step's tick) and, in
    // clk_hyd, the division clock it stopped at (rare)
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      if (!(exited[j] || dead[j])) continue;
      const uint64_t g = cur.base + off(j);
      GLOBAL_AS double* a_h = gp(b.age_hyd) + g;
      GLOBAL_AS uint32_t* c_h = gp(b.clk_hyd) + g;
      *a_h = dead[j] ? clock_age(*a_h, clk_h0, *c_h, dt) : 0.0;
      *c_h = clk_d;
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      if (mover[j]) {
        const bool leaves = idle[j] && !dead[j];
        const uint32_t pv = A.pos[j] < LIST_POS_FAR ? A.pos[j] : LIST_POS_FAR;
        gp(b.move_rec)[seg + rm[j]] = (slot_t + off(j)) | (leaves ? 0u : MOVER_INACTIVE) | (pv << LIST_POS_SHIFT);
      }
      if (n_div && want[j]) {  // rare
        const uint32_t pv = A.pos[j] < LIST_POS_FAR ? A.pos[j] : LIST_POS_FAR;
        gp(b.div_rec)[seg + rd_[j]] = (slot_t + off(j)) | (pv << LIST_POS_SHIFT);
        gp(b.div_pos)[seg + rd_[j]] = A.pos[j];
      }
    }
    if (lane == 0 && sc < n_sc) {
      gp(b.sub_moves)[sc] = n_mov;
      gp(b.sub_divs)[sc] = n_div;
    }
    if (n_div) {  // the block counts the resolve kernel's prefixes start from
#pragma unroll
      for (int lvl = 0; lvl < 3; ++lvl) {
        const uint32_t i = sc >> (6 * (lvl + 1));
        if (i != dv_i[lvl]) {
          flush_requests(lvl);
          dv_i[lvl] = i;
        }
        dv_n[lvl] += n_div;
      }
    }
    // ---- counters of the sub-chunk (wave-uniform)
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      cnt_acc.processed += (uint32_t)__popcll(__ballot(idle[j]));
      cnt_acc.exits += (uint32_t)__popcll(__ballot(exited[j]));
      cnt_acc.deaths += (uint32_t)__popcll(__ballot(dead[j]));
    }
    cnt_acc.moves += n_mov;
    cnt_acc.requests += n_div;

    __builtin_amdgcn_sched_barrier(0);
    cur = nxt;
    A = B;
#pragma unroll
    for (int j = 0; j < NL; ++j)
#pragma unroll
      for (int f = 0; f < NQ; ++f) rr[j][f] = rn[j][f];
    grp = g_next;
    if (grp >= n_groups) break;
  }
  if (PRIV) {
    __syncthreads();
    if (part_row) {  // the accumulator of the one run, every entry
      for (uint32_t i = tid; i < b.part_w; i += BLK) part_row[i] = s_acc[i];
    } else {
      flush_window<BLK>(s_acc, acc_S, acc_ns, acc_win, acc_nc, acc_w0, tid);
    }
  }
  if (cnt_run != 0xFFFFFFFFu && lane == 0) flush_counts(b.rs + cnt_run, cnt_acc);
#pragma unroll
  for (int lvl = 0; lvl < 3; ++lvl) flush_requests(lvl);
}

// The books of every run after its step, kept by the last block of
// k_resolve: counters, births (the accepted requests), the removal decision
// of update_and_remove_inactive (particles_container.hpp:539-557, threshold
// from the pass's count) over [0, n_end) where the newborns already sit, and
// the removal's tile layout; then the request block counts are cleared for
// the next pass (n_sc: the sub-chunks of this step's layout).  With `layout`
// it also lays out the next step's tiles (layout_pass), from the counts it
// has just computed.
__device__ void bookkeep_pass(const BatchDev& b, const RunDev* __restrict__ runs, uint32_t n_sc, bool layout) {
  constexpr uint32_t NT = 256, NWV = NT / 64;
  __shared__ uint32_t s_w[NWV], s_wl[NWV];
  __shared__ uint32_t s_carry, s_carry_l;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  if (tid == 0) s_carry = s_carry_l = 0;
  __syncthreads();
  for (uint32_t r0 = 0; r0 < b.R; r0 += NT) {
    const uint32_t r = r0 + tid;
    uint32_t tiles = 0;   // tiles of [0, n_end) when the run compacts
    uint32_t tiles_l = 0; // tiles of the next step's layout
    if (r < b.R) {
      RunState* rs = b.rs + r;
      const RunDev& rd = runs[r];
      uint32_t step_now, n_now;  // after this step's books
      if (rs->active) {
        const uint32_t n = rs->n;
        // read and reset with atomics: the resolve kernel's blocks added to
        // the exit counter with atomics up to a moment ago
        const unsigned long long req = atomicExch(&rs->step_requests, 0ull);
        const unsigned long long exits = atomicExch(&rs->step_exits, 0ull);
        const unsigned long long deaths = atomicExch(&rs->step_deaths, 0ull);
        const unsigned long long moves = atomicExch(&rs->step_moves, 0ull);
        const unsigned long long processed = atomicExch(&rs->step_processed, 0ull);
        const uint32_t room = rd.cap - n;
        const uint32_t budget = room < rd.div_cap ? room : rd.div_cap;
        const uint32_t births = req < budget ? (uint32_t)req : budget;
        unsigned long long* cnt = rs->counters;
        cnt[C_BIRTHS] += births;
        cnt[C_EXITS] += exits;
        cnt[C_DEATHS] += deaths;
        cnt[C_MOVES] += moves;
        cnt[C_OVERFLOW] += req - births;
        cnt[C_PROCESSED] += processed;
        uint64_t inactive = (uint64_t)rs->inactive + exits + deaths;
        const uint64_t thr_ratio = (uint64_t)((double)n * rd.ratio);
        const uint64_t thr = rd.min_removal > thr_ratio ? rd.min_removal : thr_ratio;
        const uint32_t n_end = n + births;
        uint32_t n_keep = n_end, compact = 0;
        if ((b.stages & STAGE_REMOVE) && inactive > thr) {
          cnt[C_COMPACTIONS] += 1;
          if (inactive >= n_end) {
            n_keep = 0;
            compact = 2;  // everything removed, nothing to move
          } else {
            n_keep = n_end - (uint32_t)inactive;
            compact = 1;
            tiles = (n_end + TILE - 1u) / TILE;
          }
          inactive = 0;
        }
        const uint32_t step = rs->step + 1u;
        // the age clocks tick with the stages the step ran
        if (b.stages & STAGE_MODEL) rs->clock_div += 1u;
        if (rd.enable_leave && (b.stages & STAGE_LEAVE)) rs->clock_hyd += 1u;
        rs->n_end = n_end;
        rs->n_keep = n_keep;
        rs->compact = compact;
        rs->births = births;
        rs->inactive = (uint32_t)inactive;
        rs->n = n_keep;
        rs->step = step;
        cnt[C_STEPS] = step;
        cnt[C_N] = n_keep;
        cnt[C_INACTIVE] = (uint32_t)inactive;
        step_now = step;
        n_now = n_keep;
      } else {
        rs->compact = 0;
        step_now = rs->step;
        n_now = rs->n;
      }
      if (layout) {  // layout_pass's table, from the values at hand
        const uint32_t act = step_now < rd.step_target ? 1u : 0u;
        tiles_l = act ? (n_now + TILE - 1u) / TILE : 0u;
        rs->active = act;
        rs->n_tiles = tiles_l;
      }
    }
    const uint32_t inc = wave_incl_scan(tiles, lane);
    const uint32_t inc_l = layout ? wave_incl_scan(tiles_l, lane) : 0u;
    if (lane == 63) {
      s_w[wave] = inc;
      s_wl[wave] = inc_l;
    }
    __syncthreads();
    uint32_t wpre = 0, tot = 0, wpre_l = 0, tot_l = 0;
    for (uint32_t w = 0; w < NWV; ++w) {
      if (w < wave) {
        wpre += s_w[w];
        wpre_l += s_wl[w];
      }
      tot += s_w[w];
      tot_l += s_wl[w];
    }
    const uint32_t carry = s_carry, carry_l = s_carry_l;
    if (r < b.R) {
      b.rm_tile_off[r] = carry + wpre + inc - tiles;
      if (layout) {
        b.tile_off[r] = carry_l + wpre_l + inc_l - tiles_l;
        b.rs[r].tile_begin = carry_l + wpre_l + inc_l - tiles_l;
      }
    }
    __syncthreads();
    if (tid == 0) {
      s_carry = carry + tot;
      s_carry_l = carry_l + tot_l;
    }
    __syncthreads();
  }
  if (tid == 0) {
    b.rm_tile_off[b.R] = s_carry;
    if (layout) b.tile_off[b.R] = s_carry_l;
  }
  for (uint32_t i = tid; i < (n_sc + 63u) >> 6; i += NT) b.div_l1[i] = 0u;
  for (uint32_t i = tid; i < (n_sc + 4095u) >> 12; i += NT) b.div_l2[i] = 0u;
  for (uint32_t i = tid; i <= (n_sc >> 18); i += NT) b.div_l3[i] = 0u;
  for (uint32_t i = tid; i <= RESOLVE_TICKETS; i += NT) b.ticket[i] = 0u;
}

// After the particle pass, per group of 32 or 64 sub-chunks (one per thread
// of the block for the prefix sums): (1) the movers the pass listed take
// their neighbour, the first CDF column at or above u, and an Idle one runs
// the exit test in its new compartment (move_kernel.hpp:392-437, 586-648);
// (2) the division requests are accepted in slot order over the whole run:
// the k-th request of the step (k counted from the run's first slot)
// succeeds iff k < min(cap - n, div_cap), the reference's division buffer
// (particles_container.hpp:559-573) made deterministic; the daughter is
// written straight to slot n + k, where the reference's merge_buffer would
// append her (:575-599), from the mother's state the pass staged (the pass
// stored the halved mother; a refused request restores her) or, past the
// first DIV_STAGE requests of a sub-chunk, from the mother's slot (halved
// here, M::division).  k is the requests before the group (the per-block
// counts of the pass, summed over at most 64 entries per level) plus the
// requests before it in the group.
// Every mover and every request is independent, so the processing order is
// free.  (3) The last block to finish keeps the books of every run: counters,
// births, the removal decision (update_and_remove_inactive,
// particles_container.hpp:539-557) and the removal's tile layout.
constexpr uint32_t RES_SC = RESOLVE_GROUP_MAX;  // shared arrays; a group is GS <= RES_SC sub-chunks
static_assert(RES_SC <= 256 && 32 % NSUB_MAX == 0, "a group is at most one sub-chunk per thread, of whole tiles");

// Requests of the pass in sub-chunks [0, sc), summed by one wave from the
// block counts (div_l3, div_l2, div_l1: blocks of 2^18, 2^12 and 2^6
// sub-chunks) and the sub-chunk counts of sc's own block of 64.
__device__ __forceinline__ uint32_t request_prefix(const BatchDev& b, uint32_t sc, uint32_t lane) {
  const uint32_t a3 = sc >> 18, a2 = sc >> 12, a1 = sc >> 6;
  uint32_t v = 0;
  for (uint32_t i = lane; i < a3; i += 64) v += b.div_l3[i];
  const uint32_t i2 = (a3 << 6) + lane, i1 = (a2 << 6) + lane, i0 = (a1 << 6) + lane;
  const uint32_t x2 = i2 < a2 ? b.div_l2[i2] : 0u;
  const uint32_t x1 = i1 < a1 ? b.div_l1[i1] : 0u;
  const uint32_t x0 = i0 < sc ? b.sub_divs[i0] : 0u;
  return wave_sum(v + x2 + x1 + x0);
}

// The resolve kernel.  The grid's first tr_blocks blocks (dispatched first:
// their chain is the shorter) are aux blocks.  With `ahead` they transport
// the liquid (and gas) for the next step meanwhile: the two touch disjoint
// data (the records of the two steps are separate tables, the concentrations
// are double-buffered), and the block that keeps the books also lays out the
// next step's tiles, so the next step starts with its particle pass.  With
// contribution rows (ROWS: the variant launched when b.part is set; the
// other one keeps no code of it, which would cost it registers) they first
// sum the pass's rows, for that transport or, without `ahead`, into S.
template <class M, bool ROWS>
__global__ void __launch_bounds__(256) k_resolve(BatchDev b, const RunDev* __restrict__ runs,
                                                 const MapDev* __restrict__ maps,
                                                 const uint32_t* __restrict__ tile_off, uint32_t tr_blocks,
                                                 uint32_t ahead) {
  constexpr int NV = M::NV;
  constexpr int NP = M::NP > 0 ? M::NP : 1;
  __shared__ uint32_t s_pre[RES_SC + 1];   // movers: prefix over the group's sub-chunks
  __shared__ uint32_t s_run[RES_SC];
  __shared__ uint32_t s_dpre[RES_SC + 1];  // division requests: prefix over the group's sub-chunks
  __shared__ uint32_t s_p0, s_pr0;         // requests before the group, before its first run
  __shared__ uint32_t s_wm[4], s_wd[4];    // movers and requests of each wave's sub-chunks
  __shared__ uint32_t s_last;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t R = b.R;
  const uint32_t nsub = b.nsub;          // sub-chunks per tile (divides the group)
  const uint32_t sub = TILE / nsub;      // slots per sub-chunk
  const uint32_t T = tile_off[R];
  const uint32_t n_sc = T * nsub;
  // groups of 32 sub-chunks while they fit the grid (one group per block:
  // shorter chains); with more sub-chunks, groups twice as large as long as
  // they still fill the grid, up to RES_SC: a mover takes a chain of
  // dependent loads, and a larger group keeps more of them in flight per
  // block, but one block works through all requests of its group, and where
  // newborns divide together (the end of the run) a large group's requests
  // become a long serial chain.  On the 50 000-compartment lattice at 1e8
  // particles groups of 64, 128 and 256 took 225, 227 and 266 us per step on
  // average.  Threads GS..255 hold no sub-chunk.
  const uint32_t n_res = gridDim.x - tr_blocks;  // blocks of the resolve work
  const bool tr_block = blockIdx.x < tr_blocks;
  const uint32_t rb = tr_block ? 0u : blockIdx.x - tr_blocks;  // index among the resolve blocks
  uint32_t GS = 32;
  while (GS < RES_SC && (n_sc + 2 * GS - 1) / (2 * GS) >= n_res) GS *= 2;
  const uint32_t n_groups = (n_sc + GS - 1) / GS;
  const bool leave_stage = (b.stages & STAGE_LEAVE) != 0;
  if (tr_block) {
    if constexpr (ROWS)
      sum_rows<M>(b, blockIdx.x, ahead != 0);
    else
      transport_compartments<M>(b, blockIdx.x, tr_blocks, 1u);
  }
  for (uint32_t grp = rb; !tr_block && grp < n_groups; grp += n_res) {
    const uint32_t sc0 = grp * GS;
    {  // the group's sub-chunks, one thread each: block-wide prefix sums
      const uint32_t sc = sc0 + tid;
      uint32_t cm = 0, cd = 0, r = 0;
      if (tid < GS && sc < n_sc) {
        cm = b.sub_moves[sc];
        cd = b.sub_divs[sc];
        r = find_run(tile_off, R, sc / nsub);
      }
      const uint32_t inc = wave_incl_scan(cm, lane);
      const uint32_t incd = wave_incl_scan(cd, lane);
      if (lane == 63) {
        s_wm[wave] = inc;
        s_wd[wave] = incd;
      }
      __syncthreads();
      uint32_t om = 0, od = 0;  // the earlier waves' totals
      for (uint32_t w = 0; w < wave; ++w) {
        om += s_wm[w];
        od += s_wd[w];
      }
      if (tid < GS) {
        s_pre[tid] = om + inc - cm;
        s_dpre[tid] = od + incd - cd;
        s_run[tid] = r;
      }
      if (tid == GS - 1) {
        s_pre[GS] = om + inc;
        s_dpre[GS] = od + incd;
      }
    }
    __syncthreads();
    // ---- movers
    const uint32_t total = s_pre[GS];
    for (uint32_t e = tid; e < total; e += blockDim.x) {
      uint32_t lo = 0, hi = GS;  // the sub-chunk j with s_pre[j] <= e < s_pre[j+1]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_pre[mid] <= e) lo = mid; else hi = mid;
      }
      const uint32_t j = lo;
      const uint32_t sc = sc0 + j;
      const uint32_t t = sc / nsub;
      const uint32_t r = s_run[j];
      const RunDev& rd = runs[r];
      RunState* rs = b.rs + r;
      const uint64_t q = (uint64_t)sc * sub + (e - s_pre[j]);
      const uint32_t ms = b.move_rec[q];
      const uint32_t slot = (t - tile_off[r]) * TILE + (ms & (TILE - 1u));
      const uint64_t g = rd.p_off + slot;
      uint32_t pos = ms >> LIST_POS_SHIFT;
      if (pos == LIST_POS_FAR) pos = b.pos[g];  // a compartment id of 21 bits or more
      const uint32_t step = rs->step;
      const MapDev& mp = maps[rd.map_base + (step / rd.steps_per_map) % rd.n_maps];
      const U4 u = draw4(rd.seed, slot, step, TAG_TRANSPORT, 0u);
      const uint32_t m = mp.m;
      if (mp.move_row) {  // the compartment's 64-byte record: 5 cdf columns, 6 neighbours
        const GLOBAL_AS u32x4* rw = (const GLOBAL_AS u32x4*)(gp(mp.move_row) + (size_t)pos * 16);
        const u32x4 q0 = rw[0], q1 = rw[1], q2 = rw[2], q3 = rw[3];
        const double uy = u01(u.y);
        const uint32_t col = (uy > __hiloint2double((int)q0.y, (int)q0.x) ? 1u : 0u) +
                             (uy > __hiloint2double((int)q0.w, (int)q0.z) ? 1u : 0u) +
                             (uy > __hiloint2double((int)q1.y, (int)q1.x) ? 1u : 0u) +
                             (uy > __hiloint2double((int)q1.w, (int)q1.z) ? 1u : 0u) +
                             (uy > __hiloint2double((int)q2.y, (int)q2.x) ? 1u : 0u);
        pos = col == 0 ? q2.z : col == 1 ? q2.w : col == 2 ? q3.x : col == 3 ? q3.y : col == 4 ? q3.z : q3.w;
      } else {
        const uint32_t col = cdf_column(gp(mp.cdf) + (size_t)pos * m, m, u01(u.y));
        pos = gp(mp.nbr)[(size_t)pos * m + col];
      }
      b.pos[g] = pos;
      bool exits = false;
      if (rd.enable_leave && leave_stage && !(ms & MOVER_INACTIVE)) {
        const double qx = gp(rd.rec)[((size_t)(step & 1u) * rd.n_c + pos) * rd.rec_w + record_terms<M>() + 1];
        if (qx < 1.0 && u01(u.z) > qx) {
          if (rd.probe_exit) {  // its hydraulic age after this step's tick
            const double age = clock_age(b.age_hyd[g], rs->clock_hyd + 1u, b.clk_hyd[g], rd.dt);
            const unsigned long long k = atomicAdd(&rs->probe_exit_n, 1ull);
            if (k < rd.probe_cap) rd.probe_exit[k] = age;
          }
          // no longer Idle: hydraulic age 0, the division clock it stopped at
          b.age_hyd[g] = 0.0;
          b.clk_hyd[g] = rs->clock_div + ((b.stages & STAGE_MODEL) ? 1u : 0u);
          b.status[g] = EXIT;
          exits = true;
        }
      }
      // the exit counter takes one atomic per wave for the run of the wave's
      // first active lane (the wave's movers are in slot order: nearly always
      // one run), and one per exit of any other run; an atomic per exit
      // queued tens of thousands of them on one address per step
      const uint32_t r_lead = __builtin_amdgcn_readfirstlane(r);
      const uint64_t ex_lead = __ballot(exits && r == r_lead);
      if (exits && r != r_lead) atomicAdd(&rs->step_exits, 1ull);
      if (ex_lead && lane == (uint32_t)__ffsll((long long)ex_lead) - 1u)
        atomicAdd(&b.rs[r_lead].step_exits, (unsigned long long)__popcll(ex_lead));
    }
    // ---- divisions: every request of the group at once, one thread each
    const uint32_t n_req = s_dpre[GS];
    if (n_req) {
      // requests before the group and before the start of its first run
      if (wave == 0) {
        const uint32_t p0 = request_prefix(b, sc0, lane);
        if (lane == 0) s_p0 = p0;
      } else if (wave == 1) {
        const uint32_t s_r = tile_off[s_run[0]] * nsub;
        const uint32_t pr = request_prefix(b, s_r, lane);
        if (lane == 0) s_pr0 = pr;
      }
      __syncthreads();
      for (uint32_t f = tid; f < n_req; f += blockDim.x) {
        uint32_t lo = 0, hi = GS;  // the sub-chunk j with s_dpre[j] <= f < s_dpre[j+1]
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_dpre[mid] <= f) lo = mid; else hi = mid;
        }
        const uint32_t j = lo;
        const uint32_t sc = sc0 + j;
        const uint32_t t = sc / nsub;
        const uint32_t r = s_run[j];
        const RunDev& rd = runs[r];
        RunState* rs = b.rs + r;
        // rank of the request among the run's requests of the step, in slot order
        const uint32_t first = tile_off[r] * nsub;  // the run's first sub-chunk
        const uint32_t k = first <= sc0 ? s_p0 - s_pr0 + f : f - s_dpre[first - sc0];
        const uint32_t n = rs->n;
        const uint32_t room = rd.cap - n;
        const bool accepted = k < (room < rd.div_cap ? room : rd.div_cap);  // else refused: Overflow
        const uint32_t rk = f - s_dpre[j];  // rank of the request in its sub-chunk
        if (rk >= DIV_STAGE && !accepted) continue;
        const uint64_t q = (uint64_t)sc * sub + rk;
        const uint32_t dw = b.div_rec[q];
        const uint32_t slot = (t - tile_off[r]) * TILE + (dw & (TILE - 1u));
        const uint64_t g = rd.p_off + slot;
        double p[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) p[i] = rd.params[i];
        double x[NV], xd[NV];
        if (rk < DIV_STAGE) {
          // divided by the pass: the mother's slot holds her halved state, the
          // stage her state before division
          const double* stg = b.div_stage + ((uint64_t)sc * DIV_STAGE + rk) * NV;
#pragma unroll
          for (int v = 0; v < NV; ++v) x[v] = stg[v];
          if (!accepted) {  // the mother as she was before the pass divided her
#pragma unroll
            for (int v = 0; v < NV; ++v) b.state[(uint64_t)v * b.P + g] = x[v];
            continue;
          }
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) x[v] = b.state[(uint64_t)v * b.P + g];
        }
        // requests come from the model stage: the division clock has ticked
        const uint32_t clk_d = rs->clock_div + 1u;
        if (rd.probe_div) {  // the mother's division age
          const uint32_t w = b.clk_div[g];
          const double age = div_age((w & CLK_FRESH) ? 0.0 : b.age_div[g], clk_d, w, rd.dt);
          const unsigned long long i = atomicAdd(&rs->probe_div_n, 1ull);
          if (i < rd.probe_cap) rd.probe_div[i] = age;
        }
        M::division(p, x, xd, RngKey{rd.seed, slot, rs->step});
        const uint64_t d = rd.p_off + n + k;
#pragma unroll
        for (int v = 0; v < NV; ++v) b.state[(uint64_t)v * b.P + d] = xd[v];
        if (rk >= DIV_STAGE) {  // the mother, divided here
#pragma unroll
          for (int v = 0; v < NV; ++v) b.state[(uint64_t)v * b.P + g] = x[v];
        }
        // the mother's division age restarts (one store: an origin word with
        // CLK_FRESH); the daughter's ages start at 0
        b.clk_div[g] = clk_d | CLK_FRESH;
        uint32_t dpos = dw >> LIST_POS_SHIFT;
        if (dpos == LIST_POS_FAR) dpos = b.div_pos[q];  // a compartment id of 21 bits or more
        b.pos[d] = dpos;
        b.status[d] = IDLE;
        b.age_hyd[d] = 0.0;
        b.clk_hyd[d] = rs->clock_hyd + ((rd.enable_leave && leave_stage) ? 1u : 0u);
        b.clk_div[d] = clk_d | CLK_FRESH;
      }
    }
    __syncthreads();
  }
  // ---- the last block to finish keeps the books.  The books read, of this
  // kernel's work, only the exit counters, which are updated by atomics:
  // every thread waits for its memory operations to complete before its
  // block counts itself done, and the last block reads the counters with
  // atomics.  No cache writeback is needed (an agent-scope fence per block
  // writes back this XCD's L2 each time: 60 us over a 2048-block grid).
  // Blocks count themselves done on RESOLVE_TICKETS sub-counters (block b on
  // b % RESOLVE_TICKETS), and the last of each on the main counter, so no counter
  // takes more than a few dozen atomics; resolve blocks without a group take no
  // part, transport blocks do (the books advance the step count they read).
  const uint32_t nb = n_groups < n_res ? n_groups : n_res;  // resolve blocks with work
  const uint32_t np = nb + tr_blocks;                       // blocks that count themselves done
  if (np == 0) {
    if (blockIdx.x == 0) bookkeep_pass(b, runs, n_sc, ahead != 0);
    return;
  }
  if (!tr_block && rb >= nb) return;
  const uint32_t id = tr_block ? blockIdx.x : tr_blocks + rb;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) {
    const uint32_t j = id % RESOLVE_TICKETS;
    const uint32_t in_j = (np - j + RESOLVE_TICKETS - 1u) / RESOLVE_TICKETS;  // blocks counting on sub-counter j
    const uint32_t n_sub = np < RESOLVE_TICKETS ? np : RESOLVE_TICKETS;        // sub-counters in use
    uint32_t last = 0;
    if (atomicAdd(b.ticket + 1u + j, 1u) == in_j - 1u) last = atomicAdd(b.ticket, 1u) == n_sub - 1u ? 1u : 0u;
    s_last = last;
  }
  __syncthreads();
  if (!s_last) return;
  bookkeep_pass(b, runs, n_sc, ahead != 0);
}

// ------------------------------------------------------------------ layout
// Tile -> run table of the next pass (the tiles of [0, n) of every active run).
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


// ------------------------------------------------------------------ properties
// Per-particle export values of run r (get_properties, post_process.hpp:54-98):
// for every Idle slot the model state and M::mass, row-major [n_var+1][n]
// (zero for inactive slots), and their sums per compartment [n_var+1][n_c].
// values may be null: the sums only.
template <class M>
__global__ void __launch_bounds__(256) k_properties(BatchDev b, uint32_t r, uint64_t n, double* values,
                                                    double* spatial) {
  constexpr int NV = M::NV;
  const RunDev& rd = b.runs[r];
  const uint32_t n_c = rd.n_c;
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n; s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = rd.p_off + s;
    if (b.status[g] != IDLE) {
      if (values)
        for (int v = 0; v <= NV; ++v) values[(uint64_t)v * n + s] = 0.0;
      continue;
    }
    const uint32_t c = b.pos[g];
    double x[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      x[v] = b.state[(uint64_t)v * b.P + g];
      if (values) values[(uint64_t)v * n + s] = x[v];
      unsafeAtomicAdd(spatial + (uint64_t)v * n_c + c, x[v]);
    }
    const double m = M::mass(rd.params, x);
    if (values) values[(uint64_t)NV * n + s] = m;
    unsafeAtomicAdd(spatial + (uint64_t)NV * n_c + c, m);
  }
}

// ------------------------------------------------------------------ launchers
template <class M, int BLK, bool PRIV, bool PROBE, bool WIN, bool LV>
static void launch_particle_t(const BatchDev& b, int grid, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_particle<M, BLK, M::NL, lane_interleave(M::NL), PRIV, PROBE, WIN, LV>), dim3(grid), dim3(BLK), lds, s,
                     b, b.runs, (const RunState*)b.rs, (const uint32_t*)b.tile_off);
}

// The variants: direct global atomics (256 threads); an LDS accumulator of
// whole runs (256/512/1024) or of a compartment window (256/512); each with
// and without the leaving-flow work; and the 256-thread probe variants of
// each accumulator kind (with it).
template <class M, bool LV>
static const void* kernel_ptr_lv(int blk, bool priv, bool win) {
  if (!priv) return blk == 256 ? (const void*)&k_particle<M, 256, M::NL, lane_interleave(M::NL), false, false, false, LV> : nullptr;
  if (win) {
    if (blk == 256) return (const void*)&k_particle<M, 256, M::NL, lane_interleave(M::NL), true, false, true, LV>;
    if (blk == 512) return (const void*)&k_particle<M, 512, M::NL, lane_interleave(M::NL), true, false, true, LV>;
    return nullptr;
  }
  if (blk == 256) return (const void*)&k_particle<M, 256, M::NL, lane_interleave(M::NL), true, false, false, LV>;
  if (blk == 512) return (const void*)&k_particle<M, 512, M::NL, lane_interleave(M::NL), true, false, false, LV>;
  if (blk == 1024) return (const void*)&k_particle<M, 1024, M::NL, lane_interleave(M::NL), true, false, false, LV>;
  return nullptr;
}
template <class M>
static const void* kernel_ptr_t(int blk, bool priv, bool probe, bool win, bool leave) {
  if (probe) {
    if (!priv) return (const void*)&k_particle<M, 256, M::NL, lane_interleave(M::NL), false, true, false, true>;
    return win ? (const void*)&k_particle<M, 256, M::NL, lane_interleave(M::NL), true, true, true, true>
               : (const void*)&k_particle<M, 256, M::NL, lane_interleave(M::NL), true, true, false, true>;
  }
  return leave ? kernel_ptr_lv<M, true>(blk, priv, win) : kernel_ptr_lv<M, false>(blk, priv, win);
}

// Runs with probes take a 256-thread variant that records exit and division
// ages; the atomics it needs (with returned values) are kept out of the
// production variants, where they would make every wait for a load in the
// tile loop a wait for all outstanding memory operations.
template <class M, bool LV>
static void launch_lv(const BatchDev& b, const ParticleLaunch& L, hipStream_t s) {
  if (!L.priv) {
    launch_particle_t<M, 256, false, false, false, LV>(b, L.grid, 0, s);
  } else if (L.win) {
    if (L.block == 512)
      launch_particle_t<M, 512, true, false, true, LV>(b, L.grid, L.lds, s);
    else
      launch_particle_t<M, 256, true, false, true, LV>(b, L.grid, L.lds, s);
  } else if (L.block == 256) {
    launch_particle_t<M, 256, true, false, false, LV>(b, L.grid, L.lds, s);
  } else if (L.block == 512) {
    launch_particle_t<M, 512, true, false, false, LV>(b, L.grid, L.lds, s);
  } else {
    launch_particle_t<M, 1024, true, false, false, LV>(b, L.grid, L.lds, s);
  }
}
template <class M>
static void launch_t(const BatchDev& b, const ParticleLaunch& L, hipStream_t s) {
  if (L.probe) {
    if (!L.priv)
      launch_particle_t<M, 256, false, true, false, true>(b, L.grid, 0, s);
    else if (L.win)
      launch_particle_t<M, 256, true, true, true, true>(b, L.grid, L.lds, s);
    else
      launch_particle_t<M, 256, true, true, false, true>(b, L.grid, L.lds, s);
  } else if (L.leave) {
    launch_lv<M, true>(b, L, s);
  } else {
    launch_lv<M, false>(b, L, s);
  }
}

template <class M>
void transport_t(const BatchDev& b, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_transport<M>, dim3(grid), dim3(256), 0, s, b);
}
template <class M>
void resolve_t(const BatchDev& b, int grid, uint32_t aux, bool ahead, hipStream_t s) {
  if (b.part)
    hipLaunchKernelGGL((k_resolve<M, true>), dim3(grid + (int)aux), dim3(256), 0, s, b, b.runs, b.maps,
                       (const uint32_t*)b.tile_off, aux, ahead ? 1u : 0u);
  else
    hipLaunchKernelGGL((k_resolve<M, false>), dim3(grid + (int)aux), dim3(256), 0, s, b, b.runs, b.maps,
                       (const uint32_t*)b.tile_off, aux, ahead ? 1u : 0u);
}
template <class M>
const void* resolve_kernel_t() {
  return (const void*)&k_resolve<M, false>;
}
template <class M>
void properties_t(const BatchDev& b, uint32_t run, uint64_t n, double* values, double* spatial, hipStream_t s) {
  hipLaunchKernelGGL(k_properties<M>, dim3(1024), dim3(256), 0, s, b, run, n, values, spatial);
}

// The launchers of model M's kernels.
template <class M>
ModelOps model_ops() {
  return ModelOps{&transport_t<M>, &launch_t<M>, &kernel_ptr_t<M>, &resolve_t<M>, &resolve_kernel_t<M>,
                  &properties_t<M>};
}

}  // namespace mcst
