/*
 * mcst_oracle.h -- CPU oracle for the particle Monte Carlo step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a deterministic FP64 restatement, in
 * plain C, of BioCMA-MCST's per-step pipeline.  It is the checker that the
 * HIP path is compared against and the timed CPU baseline; it is never linked
 * into, loaded by, or called from the product library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * The reference functions each part follows are cited in mcst_oracle.c.
 */
#ifndef MCST_ORACLE_H
#define MCST_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum orc_model {
  ORC_MODEL_MONOD_MASS = 0,
  ORC_MODEL_DUAL_MONOD = 1,
  ORC_MODEL_MONOD = 2,
  ORC_MODEL_FIXED_LENGTH = 3,
  ORC_MODEL_SIMPLE_ACETATE = 4,
  ORC_MODEL_QUAD_MONOD = 5,
  ORC_MODEL_UDF_MINIMAL = 6,    /* example model plugins: the reference's minimal udf_model, */
  ORC_MODEL_UDF_STARVATION = 7, /* and Monod growth with starvation death */
  ORC_MODEL_COUNT
};

enum orc_status { ORC_IDLE = 0, ORC_DIVISION = 1, ORC_EXIT = 2, ORC_DEAD = 3 };

enum orc_feed_type { ORC_FEED_CONSTANT = 0, ORC_FEED_LINEAR = 1, ORC_FEED_EXPONENTIAL = 2 };

/* kLa rules of two-phase runs */
enum orc_kla_mode { ORC_KLA_FIXED = 0, ORC_KLA_TURBULENCE = 1 };

/* counters returned by orc_get_counters, in this order */
enum orc_counter {
  ORC_C_STEPS = 0,
  ORC_C_BIRTHS,
  ORC_C_EXITS,
  ORC_C_DEATHS,
  ORC_C_MOVES,
  ORC_C_OVERFLOW,
  ORC_C_COMPACTIONS,
  ORC_C_N,
  ORC_C_INACTIVE,
  ORC_C_PROCESSED,
  ORC_C_COUNT
};

/* stage mask bits for orc_step_stages (same values as the product's) */
enum orc_stage {
  ORC_STAGE_TRANSPORT = 1u,
  ORC_STAGE_MODEL = 2u,
  ORC_STAGE_CONTRIB = 4u,
  ORC_STAGE_MOVE = 8u,
  ORC_STAGE_LEAVE = 16u,
  ORC_STAGE_REMOVE = 32u,
  ORC_STAGE_ALL = 63u
};

typedef struct orc_run orc_run;

/* ---- building blocks (pinned individually by the tests) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u01(uint32_t x);
double orc_u01_52(uint32_t a, uint32_t b);
double orc_norm_inv(double p);
double orc_norm_cdf(double x);
int orc_model_info(int model, int* n_var, int* n_species, int* n_params);
double orc_feed_flow(int type, double f0, double param, double t);
int orc_model_default_params(int model, double* params, int n);

/* Transition matrix of a flow matrix given as COO, written dense (n x n,
 * row-major): off-diagonal F(i,j), diagonal -sum_{j!=i} F(i,j). */
int orc_transition_dense(uint32_t n, uint64_t nnz, const uint32_t* rows, const uint32_t* cols,
                         const double* flows, double* out);

/* Per-compartment move tables for one flowmap: out_flow[n], p_move[n],
 * cdf[n*m], nbr[n*m]; *m_out receives the padded row width.  cdf/nbr may be
 * NULL to query m. */
int orc_build_move_tables(uint32_t n, const double* V, uint64_t nnz, const uint32_t* rows,
                          const uint32_t* cols, const double* flows, double dt, uint32_t* m_out,
                          double* out_flow, double* p_move, double* cdf, uint32_t* nbr);

/* One explicit transport step on dense inputs (used by the KAT against
 * scipy.sparse): C,m,S are [n_s][n] row-major; sink[n]; writes C_out, m. */
int orc_transport_once(uint32_t n, uint32_t n_s, const double* V, uint64_t nnz, const uint32_t* rows,
                       const uint32_t* cols, const double* flows, const double* sink, const double* S,
                       double dt, const double* C_in, double* m_inout, double* C_out);

/* ---- a simulation run ---- */
orc_run* orc_create(int model, uint32_t n_c, uint32_t n_s, double dt, uint64_t seed);
void orc_destroy(orc_run* r);
int orc_set_model_params(orc_run* r, const double* p, int n);
int orc_add_flowmap(orc_run* r, const double* V, uint64_t nnz, const uint32_t* rows,
                    const uint32_t* cols, const double* flows);
/* A flowmap from ready-made tables (the reference's ReactorDomain::update and
 * set_transition): neighbour and cumulative-probability rows [n][n_cols] as
 * given, out-flows [n], and the transition matrix as COO triplets. */
int orc_add_flowmap_tables(orc_run* r, const double* V, uint32_t n_cols, const uint64_t* nbr, const double* cdf,
                           const double* out_flow, uint64_t nnz, const uint64_t* rows, const uint64_t* cols,
                           const double* vals);
int orc_set_steps_per_map(orc_run* r, uint32_t k);
int orc_add_feed(orc_run* r, int type, double f0, double param, uint32_t input, int64_t output,
                 uint32_t n_values, const uint32_t* species, const double* conc);
int orc_set_concentrations(orc_run* r, const double* C);
int orc_set_compaction(orc_run* r, double ratio, uint64_t min_removal);
int orc_set_division_cap(orc_run* r, uint32_t cap);
int orc_set_gas_transfer(orc_run* r, const double* kla, const double* cstar);
/* Two-phase runs: Henry constants [n_s] and initial gas concentrations
 * [n_s][n_c]; the kLa rule (fixed per species, or from the turbulence of the
 * flowmap for one species with bubble diameter d_b); the gas side of each
 * flowmap (gas volumes, gas flows as COO, energy dissipation [n_c] or NULL);
 * gas feeds.  A two-phase run needs one gas flowmap per liquid flowmap. */
int orc_set_gas_phase(orc_run* r, const double* henry, const double* c_gas);
int orc_set_kla(orc_run* r, int mode, const double* kla, uint32_t species, double bubble_d);
int orc_add_gas_flowmap(orc_run* r, const double* Vg, uint64_t nnz, const uint32_t* rows, const uint32_t* cols,
                        const double* flows, const double* eps);
int orc_add_gas_feed(orc_run* r, int type, double f0, double param, uint32_t input, int64_t output,
                     uint32_t n_values, const uint32_t* species, const double* conc);
int orc_get_gas_concentrations(const orc_run* r, double* C);
int orc_get_gas_mass(const orc_run* r, double* m);
int orc_set_sort_interval(orc_run* r, uint32_t steps);
int orc_sort_particles(orc_run* r);
int orc_set_growth(orc_run* r, uint32_t every, double factor);
int orc_init_particles(orc_run* r, uint64_t n, uint64_t capacity, double x0);
int orc_set_particles(orc_run* r, uint64_t n, uint64_t capacity, const uint32_t* pos,
                      const uint8_t* status, const double* state, const double* age_hyd,
                      const double* age_div, double weight);
int orc_step(orc_run* r, uint32_t n_steps, int n_threads);
int orc_step_stages(orc_run* r, uint32_t n_steps, int n_threads, uint32_t stages);

uint64_t orc_n(const orc_run* r);
double orc_weight(const orc_run* r);
int orc_get_particles(const orc_run* r, uint32_t* pos, uint8_t* status, double* state,
                      double* age_hyd, double* age_div);
int orc_get_concentrations(const orc_run* r, double* C);
int orc_get_mass(const orc_run* r, double* m);
int orc_get_contributions(const orc_run* r, double* S);
int orc_get_counters(const orc_run* r, uint64_t* c);
int orc_get_repartition(const orc_run* r, uint64_t* hist);
int orc_get_properties(const orc_run* r, double* values, double* spatial);
uint64_t orc_get_exit_ages(const orc_run* r, double* out, uint64_t max);
uint64_t orc_get_division_ages(const orc_run* r, double* out, uint64_t max);

#ifdef __cplusplus
}
#endif
#endif
