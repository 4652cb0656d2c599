// mcst_common.hpp -- layout of the device-resident batch of runs.
//
// A batch ("engine") holds R independent runs that share one particle model.
// Every run owns a region [p_off, p_off + cap) of the batch-wide particle SoA
// arrays, its own compartment network (flowmaps), liquid state and counters.
// Particles are processed in tiles of TILE consecutive slots of one run; the
// tile -> run map is a prefix table rebuilt on the device after every step,
// so a step never needs the host to know particle counts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcst {

constexpr uint32_t BLOCK = 256;            // threads per particle block (4 wave64)
constexpr uint32_t PPT = 4;                // particles per thread per tile
constexpr uint32_t TILE = BLOCK * PPT;     // 1024 slots per tile (the reference's 1024 particles/team)
// Particle pass layout: with 4 slots per lane a lane's slots lie 64 apart
// (every access of a wave one contiguous 64-slot run); with 2 they are
// consecutive (16-byte accesses per lane).  Measured on MI355X: 64 apart is
// 3 % faster at 4 slots per lane (pair-wise 16-byte accesses, half the load
// instructions, were slower), consecutive 4 % faster at 2.
constexpr bool lane_interleave(int nl) { return nl >= 4; }
constexpr uint32_t NSUB_MAX = TILE / 128;  // sub-chunks (the pass's unit of work) per tile: TILE / (64 * M::NL)
constexpr uint32_t WAVES = BLOCK / 64;
constexpr uint32_t RESOLVE_GROUP_MAX = 64;  // sub-chunks per group of the resolve kernel: 32 or 64 (at most 256)
constexpr uint32_t RESOLVE_TICKETS = 64;    // sub-counters of the resolve kernel's done-count
constexpr uint32_t DIV_STAGE = 16;          // division requests per sub-chunk divided in the particle pass
constexpr int MAX_PARAMS = 16;
constexpr int MAX_VAR = 16;
constexpr int MAX_SPECIES = 8;

// Particle status byte (apps/libs/mc/public/mc/alias.hpp:124-130).
enum Status : uint8_t { IDLE = 0, DIVISION = 1, EXIT = 2, DEAD = 3 };
// An entry of the pass's lists (u32): the slot within its tile (bits 0-9),
// MOVER_INACTIVE in a mover entry for a slot that moves but does not leave
// (not Idle: no exit test), and a compartment (bits 11-31): a mover's before
// the move, a division request's (the newborn's); LIST_POS_FAR for an id that
// does not fit.  A mover's compartment is then read from the slot's position
// (only the mover's own thread writes it, after the read), a request's from
// div_pos.
constexpr uint32_t MOVER_INACTIVE = 1u << 10;
constexpr uint32_t LIST_POS_SHIFT = 11;
constexpr uint32_t LIST_POS_FAR = 0xFFFFFFFFu >> LIST_POS_SHIFT;
static_assert(TILE <= MOVER_INACTIVE, "a tile-local slot fits below the flag");

// Stage mask of one step (values shared with the C ABI).
enum Stage : uint32_t {
  STAGE_TRANSPORT = 1u,
  STAGE_MODEL = 2u,
  STAGE_CONTRIB = 4u,
  STAGE_MOVE = 8u,
  STAGE_LEAVE = 16u,
  STAGE_REMOVE = 32u,
  STAGE_ALL = 63u
};

// Philox counter purpose tags (third counter word = tag << 16 | block).
enum RngTag : uint32_t { TAG_TRANSPORT = 0, TAG_UPDATE = 1, TAG_DIVISION = 2, TAG_INIT = 3 };

// Cumulative per-run counters (order shared with the C ABI).
enum Counter : int {
  C_STEPS = 0,
  C_BIRTHS,
  C_EXITS,
  C_DEATHS,
  C_MOVES,
  C_OVERFLOW,
  C_COMPACTIONS,
  C_N,
  C_INACTIVE,
  C_PROCESSED,  // particle-steps: Idle particles that entered a pass
  C_COUNT
};

// Tables of one flowmap of one run (device pointers).
struct MapDev {
  const double* p_move;     // [n_c]  dt*Q_out/V
  const double* q_exit;     // [n_c]  exp(-dt*Q_leave/V) at outlets, 2.0 elsewhere
  const double* cdf;        // [n_c][m] cumulative neighbour probabilities
  const uint32_t* nbr;      // [n_c][m] neighbour ids (padding repeats the last)
  const uint32_t* col_ptr;  // [n_c+1] transport matrix by destination column
  const uint32_t* src;      // [nnz]   source compartment of each column entry
  const double* val;        // [nnz]   F(i,j) off-diagonal, -Q_out(j) on the diagonal
  const double* inv_v;      // [n_c]
  const double* vol;        // [n_c]   liquid volumes V
  // gas side of the flowmap in a two-phase run (null otherwise)
  const uint32_t* g_col_ptr;  // [n_c+1] gas transport matrix by destination column
  const uint32_t* g_src;
  const double* g_val;
  const double* g_inv_v;    // [n_c]   1 / gas volume
  const double* g_kla;      // [n_s][n_c] kLa of the flowmap
  // rows of at most MOVE_ROW_M neighbours (m <= MOVE_ROW_M), else null: one
  // 64-byte record per compartment, cdf columns 0..4 (2.0 from column m-1
  // on) then neighbours 0..5 as u32 (the last repeated), so that a mover
  // reads one line where cdf and nbr take two
  const uint32_t* move_row;
  uint32_t m;
  uint32_t pad;
};
constexpr uint32_t MOVE_ROW_M = 6;

// vf_src_s / vf_sink_k entries that belong to the gas phase
constexpr uint32_t VF_GAS = 0x80000000u;

// Immutable description of one run.
struct RunDev {
  uint64_t p_off;           // first slot of the run in the batch arrays
  uint64_t seed;
  double dt;
  double weight;            // numerical weight of every particle (unit.cpp:232-257)
  double ratio;             // removal threshold ratio (particles_container.hpp:548-551)
  uint64_t min_removal;
  double params[MAX_PARAMS];
  double* C[2];             // [n_s][n_c] concentrations, ping-pong by step parity
  double* m;                // [n_s][n_c] liquid mass
  double* S;                // [n_s][n_c] particle sources of the current step
  const double* F;          // [n_s][n_c] constant feed sources
  const double* sink;       // [n_c]      outlet flow per compartment
  const double* kla;        // [n_s][n_c] gas-liquid transfer coefficient (null: no transfer)
  const double* cstar;      // [n_s][n_c] saturation concentration set by the gas phase
  // gas phase of a two-phase run (Cg[0] null otherwise): concentrations
  // (ping-pong like C), mass, constant gas feed sources, gas outlet flows,
  // Henry constants (C* = H C_gas)
  double* Cg[2];
  double* mg;
  const double* Fg;         // [n_s][n_c]
  const double* sink_g;     // [n_c]
  double henry[MAX_SPECIES];
  // time-dependent (linear, exponential) feeds: their flows this step, then
  // the exit thresholds of the outlets they lead (written by the host before
  // each step); sources and sinks by compartment (CSR); vx_slot[j] >= 0
  // selects the threshold of outlet j from vf_q[n_vf + slot]
  const double* vf_q;
  const uint32_t* vf_src_ptr;  // [n_c+1]
  const uint32_t* vf_src_k;    // feed (index into vf_q) of each source entry
  const uint32_t* vf_src_s;    // its species (| VF_GAS for a gas feed)
  const double* vf_src_c;      // its concentration
  const uint32_t* vf_sink_ptr; // [n_c+1]
  const uint32_t* vf_sink_k;   // (| VF_GAS for a gas outlet)
  const int32_t* vx_slot;      // [n_c]
  uint32_t n_vf;
  double* rec;              // [2][n_c][rec_w] per-compartment record read by the particle pass, the
                            // table of step k at k & 1 (the next step's transport may write its
                            // table while this step's resolve kernel reads the other)
  double* probe_exit;       // exit ages (residence times), may be null
  double* probe_div;        // division ages, may be null
  uint32_t cap;             // slots reserved for the run
  uint32_t n_c;
  uint32_t n_s;
  uint32_t rec_w;           // record stride in doubles: {M::record(C)_0..NR-1, p_move, q_exit, pad}
  uint32_t enable_move;     // n_c > 1                        (kernels.hpp:54-55)
  uint32_t enable_leave;    // at least one leaving flow      (kernels.hpp:56)
  uint32_t n_maps;
  uint32_t steps_per_map;
  uint32_t map_base;        // index of the run's first MapDev
  uint32_t div_cap;         // newborns per step at most (the reference's division buffer)
  uint32_t comp_off;        // first compartment of the run in the batch transport space
  uint32_t probe_cap;
  uint32_t step_target;     // the run stops stepping at this step count
};

// Mutable per-run state, updated by the device every step.
struct RunState {
  uint32_t n;               // used slots (Idle + inactive)
  uint32_t n_end;           // slots after the last step's newborns joined: the removal's range
  uint32_t n_keep;          // slots the removal keeps (n_end - inactive)
  uint32_t inactive;        // Exit/Dead slots below n
  uint32_t step;
  uint32_t compact;         // removal requested by the last bookkeeping
  uint32_t births;          // newborns of the last step
  uint32_t tile_begin;      // first global tile of the run in the current layout
  uint32_t n_tiles;         // tiles of the run in the current layout
  uint32_t gap_total;
  uint32_t active;          // 1 while step < step_target
  // age clocks: the steps so far in which the run's Idle particles aged
  // (division age: steps with the model stage; hydraulic age: steps with the
  // leave stage while the run has a leaving flow)
  uint32_t clock_div;
  uint32_t clock_hyd;
  uint32_t pad;
  unsigned long long step_exits, step_deaths, step_moves, step_requests, step_processed;
  unsigned long long counters[C_COUNT];
  unsigned long long probe_exit_n, probe_div_n;
};

// Everything a step's kernels receive.
struct BatchDev {
  // particles, batch-wide SoA
  uint32_t* pos;
  uint8_t* status;
  double* state;            // [n_var][P]
  // ages: an Idle slot's age is age_x + (clock_x - clk_x) * dt, with clock_x
  // its run's age clock, so a step writes them only when they restart;
  // a slot that is not Idle has hydraulic age age_hyd and division age
  // age_div + (clk_hyd - clk_div) * dt (clk_hyd holds the division clock at
  // which it stopped being Idle).  clk_div is an origin word (div_age): a
  // restart at 0 sets CLK_FRESH there and leaves age_div unread.
  double* age_hyd;
  double* age_div;
  uint32_t* clk_hyd;
  uint32_t* clk_div;
  uint64_t P;               // total slots
  // division requests of the pass per block of 2^6, 2^12 and 2^18
  // sub-chunks (global sub-chunk index of the layout); with sub_divs they
  // give the requests before any sub-chunk in 64 terms per level (128 at the
  // top level for pools of more than 2^31 slots)
  uint32_t* div_l1;
  uint32_t* div_l2;
  uint32_t* div_l3;
  uint32_t* ticket;         // [1 + RESOLVE_TICKETS] resolve blocks that are done (the last keeps the books)
  uint32_t* rm_tile_off;    // [R+1] removal layout: first global tile of each run's [0, n_end)
  // per-tile bookkeeping of the removal
  uint32_t* tile_gap;
  uint32_t* tile_fill;
  uint32_t* tile_gap_off;
  uint32_t* tile_fill_off;
  uint32_t* fill_src;       // [P] scratch list of filler slots during removal
  uint32_t* tile_off;       // [R+1] first global tile of each run in the current layout
  // lists of the particle pass, one segment of TILE/nsub entries per sub-chunk
  uint32_t* move_rec;       // [T_max*TILE] particles leaving their compartment, one entry each
                            // (slot, flag and compartment in one word, so that the few movers of a
                            // sub-chunk share one line)
  uint32_t* div_rec;        // [T_max*TILE] division requests, one entry each
  uint32_t* div_pos;        // [T_max*TILE] their compartment (read for ids past the entry's field only)
  double* div_stage;        // [T_max*nsub*DIV_STAGE][n_var] the first DIV_STAGE requests of each
                            // sub-chunk, divided in the pass: the mother's state before division
                            // (to restore her if the request is refused)
  uint32_t* sub_moves;      // [T_max*nsub] entries of each sub-chunk's mover segment
  uint32_t* sub_divs;       // [T_max*nsub] entries of each sub-chunk's request segment
  // runs
  const RunDev* runs;
  RunState* rs;
  const MapDev* maps;
  const uint32_t* comp_off; // [R+1]
  uint32_t R;
  uint32_t n_comp_total;
  uint32_t ns_max;          // largest species count of the batch
  uint32_t T_max;
  uint32_t nsub;            // sub-chunks per tile of the model's particle pass, TILE / (64 * M::NL)
  uint32_t tile_group;      // particle pass: consecutive tiles per block turn (0: one range per block)
  uint32_t acc_w;           // particle pass: compartments per LDS contribution window (PRIV)
  uint32_t stages;
  // contribution rows (one live run whose whole [n_s][n_c] accumulator fits
  // LDS): block k of the particle pass stores its accumulator at
  // part[k * part_w] (part_rows blocks, part_w = n_s * n_c) and the resolve
  // kernel sums the rows; null: the blocks add their accumulators to S with
  // atomics
  double* part;
  uint32_t part_rows;
  uint32_t part_w;
  uint32_t part_run;        // the run the rows belong to
};

// Compartments per block of the row sum: (NS species x PART_CPB values) per
// thread and row.
template <int NS>
constexpr uint32_t part_cpb() {
  return NS >= 16 ? 1u : 16u / (uint32_t)NS;
}

// Age of a slot from its base age, its origin on the age clock and the
// clock (the expression the host's conversion uses too).
__host__ __device__ inline double clock_age(double base, uint32_t clock, uint32_t origin, double dt) {
  return base + (double)(clock - origin) * dt;
}

// A division-age origin word: the origin on the division clock in its low 31
// bits (the age counts (clock - origin) mod 2^31 steps), and CLK_FRESH when
// the age restarted there from 0, so that a restart is one store (age_div is
// then not read).  Equal to clock_age for ages of fewer than 2^31 steps.
constexpr uint32_t CLK_FRESH = 1u << 31;
__host__ __device__ inline double div_age(double base, uint32_t clock, uint32_t word, double dt) {
  return ((word & CLK_FRESH) ? 0.0 : base) + (double)((clock - word) & ~CLK_FRESH) * dt;
}

}  // namespace mcst
