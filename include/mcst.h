/*
 * mcst.h -- C ABI of the MI355X particle Monte Carlo step (libmcst.so).
 *
 * Replaces, for the particle step of BioCMA-MCST, the internal C++ boundary
 * formed by (paths relative to the reference root):
 *   CycleFunctors<Space,Model>::update / launch_model / launch_move /
 *     get_host_reduction ......... apps/libs/simulation/public/simulation/kernels/kernels.hpp:25-224
 *   SimulationUnit::cycleProcess / post_cycle / update_feed / ode_step / advance
 *                                  apps/libs/simulation/public/simulation/simulation.hpp:183-239,
 *                                  apps/libs/simulation/src/simulation.model.cpp:42-154
 *   ScalarSimulation::performStep / set_transition / set_mass
 *                                  apps/libs/simulation/src/implScalar.cpp:250-264,316-330
 *   ReactorDomain::update / set_leaving_flow
 *                                  apps/libs/mc/src/domain.cpp:43-108
 *   ParticlesContainer<Model> (division buffer, removal, merge)
 *                                  apps/libs/mc/public/mc/particles_container.hpp:56-239
 *   MC::init / post_init_weight .. apps/libs/mc/public/mc/mcinit.hpp:67-105, apps/libs/mc/src/unit.cpp:232-306
 *   MonteCarloUnit::getRepartition apps/libs/mc/src/unit.cpp:190-230
 *
 * Conventions follow the reference's own C ABI (apps/api/public/api/api_raw.h):
 * every function returns an int, 0 on success and a negative code on error
 * (mcst_last_error() describes the last failure of the calling thread);
 * handles are opaque; host buffers are caller-owned and copied; device
 * buffers are owned by the library.  One engine is driven by one host thread;
 * its kernels run in order on the engine's HIP stream.
 *
 * An engine is a batch of independent runs that share one particle model.
 * A single simulation is a batch of one run.
 */
#ifndef MCST_H
#define MCST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCST_OK 0
#define MCST_ERR_ARG (-1)      /* invalid argument or handle */
#define MCST_ERR_STATE (-2)    /* call not valid in the engine's state */
#define MCST_ERR_DEVICE (-3)   /* HIP runtime or device failure */
#define MCST_ERR_MEMORY (-4)   /* allocation failure */

/* particle status byte (apps/libs/mc/public/mc/alias.hpp:124-130) */
#define MCST_IDLE 0
#define MCST_DIVISION 1
#define MCST_EXIT 2
#define MCST_DEAD 3

/* feed kinds (apps/libs/simulation/src/feed_descriptor.cpp:18-28) */
#define MCST_FEED_CONSTANT 0
#define MCST_FEED_LINEAR 1
#define MCST_FEED_EXPONENTIAL 2

/* kLa rules of two-phase runs (MassTransfer::Type, mass_transfer.hpp:18-38) */
#define MCST_KLA_FIXED 0       /* FixedKla: one value per species */
#define MCST_KLA_TURBULENCE 1  /* FlowmapTurbulence: from each flowmap's energy dissipation and gas fraction */

/* stage mask of mcst_engine_step_stages */
#define MCST_STAGE_TRANSPORT 1u
#define MCST_STAGE_MODEL 2u
#define MCST_STAGE_CONTRIB 4u
#define MCST_STAGE_MOVE 8u
#define MCST_STAGE_LEAVE 16u
#define MCST_STAGE_REMOVE 32u
#define MCST_STAGE_ALL 63u

/* counters of mcst_run_get_counters (the reference's MC::EventContainer
 * tallies, apps/libs/mc/public/mc/events.hpp:149-264, plus bookkeeping) */
#define MCST_C_STEPS 0
#define MCST_C_BIRTHS 1      /* successful divisions (NewParticle) */
#define MCST_C_EXITS 2       /* particles that left through an outlet (Exit) */
#define MCST_C_DEATHS 3      /* particles whose model returned Dead */
#define MCST_C_MOVES 4       /* compartment changes (Move) */
#define MCST_C_OVERFLOW 5    /* divisions refused for lack of room (Overflow) */
#define MCST_C_COMPACTIONS 6 /* removal passes */
#define MCST_C_N 7           /* used slots */
#define MCST_C_INACTIVE 8    /* Exit/Dead slots not yet removed */
#define MCST_C_PROCESSED 9   /* particle-steps: Idle particles that entered a pass */
#define MCST_C_COUNT 10

typedef struct mcst_engine mcst_engine;

int mcst_version(int* major, int* minor, int* patch);
const char* mcst_last_error(void);

/* Model registry: the reference's ModelType implementations
 * (apps/libs/models/public/models/{monod,fixed_length,simple_acetate}.hpp)
 * compiled for gfx950, plus the mass-based models of the benchmark cases.
 * Names: "monod_mass", "dual_monod", "monod", "fixed_length", "simple_acetate",
 * "quad_monod", and those of loaded plugins (mcst_model_load). */
int mcst_model_lookup(const char* name, int* model_id, int* n_var, int* n_species, int* n_params);
int mcst_model_default_params(int model_id, double* params, int n);
/* Loads a model plugin (a shared library built by `build.py --udf` from a
 * model struct; the dlopen'ed udf_model module of the reference,
 * apps/api/src/udf_handle.cpp:21-38 and udf_includes.cpp:57-78) and registers
 * its model under the struct's NAME; *model_id receives its id (100, 101, ...
 * in load order).  Loading the same path again returns the same id.  The
 * name "udf_model" (mcst_model_lookup, mcst_engine_create) loads the plugin
 * named by the environment variable BIOMC_LIB_UDF on first use, as the
 * reference does.  MCST_ERR_ARG for a file that is not a plugin, a plugin
 * built against other device structures, or a name already taken.  Plugins
 * stay loaded until the process ends. */
int mcst_model_load(const char* path, int* model_id);

/* Engine lifetime: binds HIP device `device` and creates the engine's stream. */
int mcst_engine_create(const char* model_name, int device, mcst_engine** out);
int mcst_engine_destroy(mcst_engine** engine);

/* ---- run definition (before mcst_engine_commit) ---- */
/* A run: compartments, species, time step, RNG seed.  (SimulationUnit ctor,
 * apps/libs/simulation/src/simulation.cpp:47-81) */
int mcst_run_create(mcst_engine* e, uint32_t n_compartments, uint32_t n_species, double dt, uint64_t seed,
                    int32_t* run_id);
int mcst_run_set_model_params(mcst_engine* e, int32_t run, const double* params, int n);
/* One flowmap: liquid volumes [n_c] (m3) and the compartment flow matrix as
 * COO (rows[k] -> cols[k], flows[k] m3/s; diagonal and non-positive entries
 * ignored, duplicates summed).  The library builds the transition matrix,
 * out-flows and cumulative neighbour probabilities from it (what rcmtool and
 * ReactorDomain::update provide upstream, domain.cpp:43-74). */
int mcst_run_add_flowmap(mcst_engine* e, int32_t run, const double* volumes, uint64_t nnz, const uint32_t* rows,
                         const uint32_t* cols, const double* flows);
/* One flowmap from the tables the reference's hydrodynamic state hands to
 * the step (rcmtool's output): ReactorDomain::update(volumes, neighbors_flat,
 * out_flows, proba_flat) (apps/libs/mc/src/domain.cpp:44-74) plus the
 * liquid transition matrix of ScalarSimulation::set_transition
 * (apps/libs/simulation/src/implScalar.cpp:316-320).
 *   volumes[n_c]                    liquid volumes V (m3)
 *   neighbors[n_c][n_cols]          neighbour ids, row-major, in the caller's
 *                                   order and padding (size_t in the reference)
 *   cumulative[n_c][n_cols]         cumulative leaving probabilities of those
 *                                   columns, non-decreasing along a row; a
 *                                   leaving particle takes the first column with
 *                                   cumulative >= u (move_kernel.hpp:61-103)
 *   out_flows[n_c]                  diag_transition Q_out: leaves if dt*Q/V > u
 *   rows/cols/values[nnz]           transition matrix M as COO (off-diagonal
 *                                   F(i,j), diagonal -Q_out(i)); transport is
 *                                   m += dt*(C*M - C*diag(sink) + S); duplicates
 *                                   are summed
 * Errors (MCST_ERR_ARG): an id out of range, a non-finite value, a
 * decreasing cumulative row, a non-positive volume or negative out-flow. */
int mcst_run_add_flowmap_tables(mcst_engine* e, int32_t run, const double* volumes, uint32_t n_cols,
                                const uint64_t* neighbors, const double* cumulative, const double* out_flows,
                                uint64_t nnz, const uint64_t* rows, const uint64_t* cols, const double* values);
/* The flowmap of step k is (k / steps) mod n_flowmaps. */
int mcst_run_set_steps_per_flowmap(mcst_engine* e, int32_t run, uint32_t steps);
/* A feed (FeedDescriptor, apps/libs/simulation/public/simulation/feed_descriptor.hpp:11-161):
 * flow(t) = f0 (constant), f0 + t*param (linear), f0 + exp(param*t) (exponential);
 * adds flow*conc[v] of species[v] at `input`; output >= 0 is a leaving flow.
 * A linear or exponential feed takes t = step*dt at the start of each step
 * (update_feed, simulation.model.cpp:72-122): its flow, and the exit
 * threshold of an outlet it is the first feed of, are set by the host before
 * each step, so an engine holding such a run cannot capture a graph. */
int mcst_run_add_feed(mcst_engine* e, int32_t run, int kind, double f0, double param, uint32_t input,
                      int64_t output, uint32_t n_values, const uint32_t* species, const double* conc);
/* Initial concentrations [n_species][n_c] (kg/m3); liquid mass = C * V. */
int mcst_run_set_concentrations(mcst_engine* e, int32_t run, const double* C);
/* Removal threshold: remove when inactive > max(min_removal, ratio * n)
 * (particles_container.hpp:539-557; BIOMC_MC_REMOVE_RATIO_THRESHOLD /
 * BIOMC_MC_MINIMUM_REMOVAL). */
int mcst_run_set_removal(mcst_engine* e, int32_t run, double ratio, uint64_t min_removal);
/* Gas-liquid mass transfer with a fixed gas phase (the liquid side of
 * ScalarSimulation::performStepGL, apps/libs/simulation/src/implScalar.cpp:228-248,
 * with mtr = kla*(Henry*C_gas - C)*V as in mass_transfer.cpp:140-157): the
 * liquid mass of species s in compartment j gains dt*kla[s][j]*(c_star[s][j] - C)*V_j
 * each step; c_star = Henry*C_gas is held constant (no gas transport). */
int mcst_run_set_gas_transfer(mcst_engine* e, int32_t run, const double* kla, const double* c_star);

/* ---- two-phase runs: a transported gas phase (SimulationUnit::ode_step with
 * a gas phase, apps/libs/simulation/src/simulation.model.cpp:131-150) ----
 * Each step: mtr = kla*(H*C_gas - C_liq)*V_liq from the old concentrations
 * (mass_transfer.cpp:140-157); the gas phase m_gas += dt*(C_gas*M_gas -
 * C_gas*diag(sink_gas) + F_gas - mtr) and the liquid m += dt*(C*M -
 * C*diag(sink) + S + F + mtr) (performStepGL, implScalar.cpp:228-248); then
 * liquid concentrations in (-5e-7, 0) are set to 0 (clearNegs,
 * implScalar.cpp:266-297; the mass keeps its value).  A two-phase run cannot
 * also use mcst_run_set_gas_transfer.
 * Henry constants henry[n_species] (C* = H*C_gas) and the initial gas
 * concentrations c_gas[n_species][n_c]. */
int mcst_run_set_gas_phase(mcst_engine* e, int32_t run, const double* henry, const double* c_gas);
/* kLa: MCST_KLA_FIXED takes kla[n_species] for every compartment;
 * MCST_KLA_TURBULENCE gives species `species` kL*a in each compartment of
 * each flowmap, kL = 0.3*(eps*nu)^(1/4)*Sc^(-1/2) (nu = 1.0023e-6 m2/s,
 * Sc = nu/1e-9), a = 6*alpha/(d_b*(1 - alpha)), alpha = V_gas/(V_liq + V_gas),
 * bubble diameter d_b (impl_mtr.cpp:21-50,90-134), and 0 to the others.
 * Default: fixed, all 0. */
int mcst_run_set_kla(mcst_engine* e, int32_t run, int mode, const double* kla, uint32_t species,
                     double bubble_diameter);
/* The gas side of the flowmap of the same rank: gas volumes [n_c], the gas
 * flow matrix as COO (as mcst_run_add_flowmap), and the energy dissipation
 * eps [n_c] (W/kg; may be NULL unless the kLa rule is MCST_KLA_TURBULENCE).
 * A two-phase run needs one per liquid flowmap. */
int mcst_run_add_gas_flowmap(mcst_engine* e, int32_t run, const double* gas_volumes, uint64_t nnz,
                             const uint32_t* rows, const uint32_t* cols, const double* flows,
                             const double* energy_dissipation);
/* A gas feed (SimulationFeed::gas_feeds): as mcst_run_add_feed, into the gas
 * phase; its output is a gas outlet (no particle leaves with it). */
int mcst_run_add_gas_feed(mcst_engine* e, int32_t run, int kind, double f0, double param, uint32_t input,
                          int64_t output, uint32_t n_values, const uint32_t* species, const double* conc);
/* Particle order (no counterpart in the reference, whose container keeps
 * particles in creation order): mcst_run_sort_particles reorders the run's
 * particles by compartment now, stable in slot order; with an interval k > 0
 * mcst_engine_step does so after every step whose count is a multiple of k
 * (replays of a captured graph do not).  Slots of one compartment then sit
 * together, which the particle pass's record reads and contribution window
 * profit from; moves and newborns scatter the order over time.  The reorder
 * changes which slot holds which particle, hence the slot-keyed random
 * draws: results are reproducible for a given interval, and the population
 * statistics do not depend on it. */
int mcst_run_set_sort_interval(mcst_engine* e, int32_t run, uint32_t steps);
int mcst_run_sort_particles(mcst_engine* e, int32_t run);
/* Container growth (merge_buffer / _resize, particles_container.hpp:575-643,
 * whose allocation factor is 2.5): before every `every`-th step of the run
 * (counted from its step 0), a run with fewer free slots than particles gets
 * capacity ceil(n * factor), factor in (1, 16]; its particles keep their
 * slots, moving to a new range of the engine's pool (grown if needed) when
 * the slots after them are taken.  The checks fall on fixed steps, so
 * results do not depend on timing; each check waits for the stream.  Between
 * checks a full run refuses divisions (Overflow), as with a fixed capacity.
 * every = 0 (the default) keeps the capacity fixed.  An engine with a growing
 * run cannot capture a graph. */
int mcst_run_set_growth(mcst_engine* e, int32_t run, uint32_t every, double factor);
/* The division buffer (particles_container.hpp:559-573, whose size is
 * buffer_ratio x the allocated slots): at most `cap` (>= 1) newborns per
 * step.  Division requests are taken in slot order over the whole run and
 * the first min(capacity - n, cap) succeed; the others are refused
 * (Overflow) and ask again next step.  Default: no cap beyond the free
 * slots. */
int mcst_run_set_division_cap(mcst_engine* e, int32_t run, uint32_t cap);
/* Keep up to `cap` exit ages and division ages (the reference's probes). */
int mcst_run_set_probes(mcst_engine* e, int32_t run, uint32_t cap);
/* The run stops stepping once it has done `steps` steps (default: never). */
int mcst_run_set_step_target(mcst_engine* e, int32_t run, uint32_t steps);
/* Initial particles: n particles uniformly over compartments with
 * Model::init, capacity slots reserved, weight = x0 * V_total / sum(mass)
 * (MC::init + post_init_weight).  Particle i is drawn from Philox stream i;
 * the drawn particles are then stored in compartment order (stable in i).
 * Needs a flowmap. */
int mcst_run_init_particles(mcst_engine* e, int32_t run, uint64_t n, uint64_t capacity, double x0);
/* Or explicit particles: pos[n], status[n], state[n_var][n], ages[n]. */
int mcst_run_set_particles(mcst_engine* e, int32_t run, uint64_t n, uint64_t capacity, const uint32_t* pos,
                           const uint8_t* status, const double* state, const double* age_hyd,
                           const double* age_div, double weight);

/* Allocate device memory, upload every run; the engine can then step. */
int mcst_engine_commit(mcst_engine* e);

/* ---- serving: runs that join and leave a stepping engine ----
 * No counterpart in the reference (its SimulationUnit holds one run for the
 * whole program); it serves the request loop that admits simulations by
 * device-memory footprint and keeps each run resident until it completes.
 * Before commit: reserve a particle-slot pool and a number of run entries
 * for runs defined later.  The engine may then be committed with no run. */
int mcst_engine_reserve(mcst_engine* e, uint64_t slots, uint32_t runs);
/* After commit: put the runs defined since (in id order) on the device; they
 * take part from the next step on, starting at their own step 0.  Stops at
 * the first run that does not fit (MCST_ERR_MEMORY; it and later runs stay
 * defined and can be admitted once room is free).  *n_admitted counts the
 * runs admitted by this call. */
int mcst_engine_admit(mcst_engine* e, uint32_t* n_admitted);
/* Take a run off the device (read its results first): its slots and arrays
 * return to the pool and its id becomes invalid for result calls. */
int mcst_engine_retire(mcst_engine* e, int32_t run);

/* ---- stepping ---- */
/* n steps of every active run: feed, transport, particle cycle, newborns
 * appended, removal (the body of main_loop,
 * apps/core/src/host_specific.cpp:280-291). */
int mcst_engine_step(mcst_engine* e, uint32_t n_steps);
/* One step restricted to the MCST_STAGE_* bits in `stages`. */
int mcst_engine_step_stages(mcst_engine* e, uint32_t stages);
/* Capture `n_steps` steps into a HIP graph (replaces any earlier one).
 * MCST_ERR_STATE while a live run has a time-dependent feed or grows its
 * capacity (mcst_run_set_growth). */
int mcst_engine_capture(mcst_engine* e, uint32_t n_steps);
/* Replay the captured graph `times` times (times * n_steps steps). */
int mcst_engine_replay(mcst_engine* e, uint32_t times);
int mcst_engine_synchronize(mcst_engine* e);
/* Take n steps, as mcst_engine_step, timed with HIP events on the engine
 * stream: total ms and the mean duration of each step kernel, in the order
 * transport (with the tile layout), particle pass, mover and division
 * resolution (with the bookkeeping), removal (4 kernels together, skipped
 * when no particle can become inactive) -> per_kernel_ms[4]. */
int mcst_engine_profile(mcst_engine* e, uint32_t n_steps, double* total_ms, double* per_kernel_ms);
/* Autotuner of the particle pass (no counterpart in the reference; the
 * launch shapes of its Kokkos kernels are fixed).  Runs `steps` real steps
 * with each candidate shape (threads per block 256/512/1024 x consecutive
 * tiles per block turn 0/1/2/4/8), times the particle pass with HIP events,
 * and keeps the fastest for the engine's later steps (also across admit and
 * retire); the untuned shape stays unless another is at least 2 % faster.  The runs advance by the trial steps; their results do not depend
 * on the shape.  Drops a captured graph. */
int mcst_engine_autotune(mcst_engine* e, uint32_t steps, int32_t* block, int32_t* group, double* particle_ms);

/* ---- results ---- */
/* Step counts of runs 0..n_runs-1 with one read of the device (0 for runs
 * not on the device): what a request loop polls between chunks of steps.
 * No counterpart in the reference (one simulation per process there). */
int mcst_engine_get_steps(mcst_engine* e, uint32_t* steps, uint32_t n_runs);
int mcst_run_get_counters(mcst_engine* e, int32_t run, uint64_t* counters /* MCST_C_COUNT */);
int mcst_run_get_n(mcst_engine* e, int32_t run, uint64_t* n);
int mcst_run_get_weight(mcst_engine* e, int32_t run, double* weight);
/* The first n slots (n from mcst_run_get_n); any pointer may be NULL.
 * Ages are computed on the host from the device's base ages and age clocks:
 * an age of k steps reads k*dt (one rounding) where the reference's running
 * sum of dt differs in the last bits. */
int mcst_run_get_particles(mcst_engine* e, int32_t run, uint32_t* pos, uint8_t* status, double* state,
                           double* age_hyd, double* age_div);
int mcst_run_get_concentrations(mcst_engine* e, int32_t run, double* C);
int mcst_run_get_mass(mcst_engine* e, int32_t run, double* m);
/* Particle sources of the last step, [n_species][n_c] (kg/s). */
int mcst_run_get_contributions(mcst_engine* e, int32_t run, double* S);
/* Gas concentrations and gas mass of a two-phase run, [n_species][n_c]. */
int mcst_run_get_gas_concentrations(mcst_engine* e, int32_t run, double* C);
int mcst_run_get_gas_mass(mcst_engine* e, int32_t run, double* m);
int mcst_run_get_repartition(mcst_engine* e, int32_t run, uint64_t* hist /* [n_c] */);
/* mcst_run_get_repartition without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its counters live in one library-owned device array per GPU (8 MiB, held
 * until the process ends), so a run may have at most 2^20 compartments
 * (MCST_ERR_ARG beyond).  Calls on one GPU share that array: calls that
 * overlap on one GPU (from several threads or engines at once) race. */
int mcst_run_get_repartition_noalloc(mcst_engine* e, int32_t run, uint64_t* hist /* [n_c] */);
/* Export values (PostProcessing::get_properties, apps/core/public/core/post_process.hpp:54-98,173-190):
 * values[n_var+1][n]: each Idle slot's model state then Model::mass (zero
 * for inactive slots); spatial[n_var+1][n_c]: their sums per compartment
 * over Idle particles.  Either output may be NULL. */
int mcst_run_get_properties(mcst_engine* e, int32_t run, double* values, double* spatial);
/* Exit / division ages recorded so far (up to the probe capacity). */
int mcst_run_get_exit_ages(mcst_engine* e, int32_t run, double* out, uint64_t max, uint64_t* n);
int mcst_run_get_division_ages(mcst_engine* e, int32_t run, double* out, uint64_t max, uint64_t* n);

/* ---- checkpoint / resume (replaces the reference's serde of a run:
 * apps/core/src/serde.cpp:65-104 and the container's save/load,
 * apps/libs/mc/public/mc/particles_container.hpp:491-537).
 * mcst_run_checkpoint writes everything a step carries to the next into a
 * caller buffer: particles (positions, status, state, both ages), liquid
 * concentrations, liquid mass and the pending contributions, the step count,
 * the counters and the recorded probes.  With buf == NULL (or size too small:
 * then MCST_ERR_ARG) it only reports the size in *needed.
 * mcst_run_restore loads such a buffer into a defined, not yet committed or
 * admitted run created with the same compartment count, species count, dt
 * and seed; the caller defines its flowmaps, feeds and parameters as for the
 * original run.  Stepping it then continues the original run exactly.
 * A later set_concentrations / init_particles / set_particles on that run
 * discards the restored state. */
int mcst_run_checkpoint(mcst_engine* e, int32_t run, void* buf, uint64_t size, uint64_t* needed);
int mcst_run_restore(mcst_engine* e, int32_t run, const void* buf, uint64_t size);

#ifdef __cplusplus
}
#endif
#endif
