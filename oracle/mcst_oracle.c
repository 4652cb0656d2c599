/*
 * mcst_oracle.c -- CPU oracle for the particle Monte Carlo step.
 *
 * TEST INFRASTRUCTURE ONLY (see mcst_oracle.h).  Deterministic FP64
 * restatement of BioCMA-MCST's per-step pipeline; compiled with gcc
 * -ffp-contract=off and no fast-math so its arithmetic is plain IEEE double.
 *
 * Reference functions restated here (paths relative to the reference root):
 *   step order ............ apps/core/src/host_specific.cpp:280-291,
 *                           apps/libs/simulation/public/simulation/simulation.hpp:183-239
 *   feeds ................. apps/libs/simulation/src/simulation.model.cpp:54-122,
 *                           apps/libs/simulation/src/feed_descriptor.cpp:18-28
 *   transport ............. apps/libs/simulation/src/implScalar.cpp:250-264,316-320
 *   transition matrix ..... documentation/pages/03_cma.md:33-53 (diag = -sum out-flows)
 *   move tables ........... documentation/pages/03_cma.md:55-64 (P(j|leave)=F(i,j)/sum_k F(i,k)),
 *                           apps/libs/mc/src/domain.cpp:43-74
 *   model cycle ........... apps/libs/simulation/public/simulation/kernels/model_kernel.hpp:163-268
 *   contribution .......... apps/libs/simulation/public/simulation/kernels/contribution_kernel.hpp:144-186
 *   move .................. apps/libs/simulation/public/simulation/kernels/move_kernel.hpp:61-103,392-437
 *   leave ................. apps/libs/simulation/public/simulation/kernels/move_kernel.hpp:347-359,586-648,
 *                           apps/libs/simulation/public/simulation/probability_leaving.hpp:16-46
 *   division slots ........ apps/libs/mc/public/mc/particles_container.hpp:559-573
 *   removal / compaction .. apps/libs/mc/public/mc/particles_container.hpp:292-385,539-557,735-796
 *   merge of newborns ..... apps/libs/mc/public/mc/particles_container.hpp:403-443,575-599
 *   init + weight ......... apps/libs/mc/src/unit.cpp:102-163,232-257
 *   repartition ........... apps/libs/mc/src/unit.cpp:67-100,190-230
 *   models ................ apps/libs/models/public/models/monod.hpp:80-193,
 *                           fixed_length.hpp:108-161, simple_acetate.hpp:132-248, utils.hpp:63-67
 *                           (monod_mass, dual_monod, quad_monod: the benchmark cases' models;
 *                           udf_minimal: apps/udf_model/minimal.cpp:23-107, and udf_starvation,
 *                           the example plugins; the latter's Dead status: mc/alias.hpp:124-130)
 *   gas-liquid transfer ... apps/libs/simulation/src/simulation.model.cpp:131-150,
 *                           implScalar.cpp:228-248,266-297, hydro/mass_transfer.cpp:102-157
 *                           (two-phase step: gas transport, kLa, clearNegs), and the liquid
 *                           side alone with a fixed gas saturation
 *   flowmap tables ........ apps/libs/mc/src/domain.cpp:44-74 with implScalar.cpp:316-330
 *   export properties ..... apps/core/public/core/post_process.hpp:54-98
 *
 * Checks that pin each part (tests/test_oracle_kat.py, test_oracle_invariants.py,
 * test_oracle_models.py, test_two_phase_cpu.py, test_flow_tables_cpu.py,
 * test_export_cpu.py, test_case_io_cpu.py): Random123 and numpy Philox vectors;
 * scipy for the normal quantile and for sparse transport; the reference's
 * transition-matrix, 16-compartment flow and neighbour, feed and container
 * fixtures; closed forms of the model updates (Euler at constant substrate,
 * the Monod mu_p lag and caps, FixedLength's linear growth, SimpleAcetate's
 * uptake branches, dual_monod's and quad_monod's lags and splits), of the
 * two-phase transfer in a 0-D box and of the move/exit probabilities; exact
 * distributions (chi-square, KS) of moves, residence times, initial states and
 * division draws; mass balance; a numpy stable sort for the reorder;
 * math.fsum for the property sums.
 * Parity unpinned: which random number each particle receives in the
 * reference (Kokkos XorShift1024 pool, replaced by Philox by design), and
 * Eigen's summation order inside C*M (here: the transport matrix's column
 * entries in the caller's order).
 *
 * Deliberate, documented differences from the reference (the reference is
 * itself scheduling-dependent, so parity is defined against these rules):
 *   - uniforms come from Philox4x32-10 keyed by (run seed) and counted by
 *     (slot, step, purpose), not from a thread-indexed XorShift pool;
 *   - the leave test u > exp(-dt*Q/V) uses a per-compartment threshold
 *     computed once per flowmap/step on the host;
 *   - division requests are accepted in slot order over the whole run while
 *     it has free slots (and up to an optional newborn cap per step), the
 *     k-th accepted daughter taking slot n + k (parent slot order), written
 *     before the removal compacts [0, n + births);
 *   - removal fills the i-th gap (ascending) with the i-th live particle from
 *     the tail (descending) instead of an atomic tail offset;
 *   - per-particle state, contributions and ages are FP64;
 *   - initial particles are stored in compartment order, and a run may be
 *     reordered by compartment every k steps (the engine does the same).
 */
#include "mcst_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TILE 1024u
#define MAX_PARAMS 16
#define MAX_VAR 16
#define MAX_SPECIES 8

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"), restated from the published round function. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* (x + 1/2) / 2^32: exact in double, strictly inside (0, 1). */
double orc_u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386962890625e-10; }

/* 52-bit uniform from two words: ((a>>6)<<26 | b>>6) + 1/2, over 2^52. */
double orc_u01_52(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 6) << 26) | (uint64_t)(b >> 6);
  return ((double)k + 0.5) * 2.220446049250313080847263336181640625e-16;
}

/* purpose tags of the counter's third word */
enum { TAG_TRANSPORT = 0, TAG_UPDATE = 1, TAG_DIVISION = 2, TAG_INIT = 3 };

static inline void draw4(uint64_t seed, uint32_t slot, uint32_t step, uint32_t tag, uint32_t blk,
                         uint32_t out[4]) {
  const uint32_t ctr[4] = {slot, step, (tag << 16) | blk, 0u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------- */
/* Normal distribution helpers.  Inverse CDF: Wichura, Algorithm AS 241
 * (PPND16), Applied Statistics 37 (1988) 477-484. */
double orc_norm_inv(double p) {
  const double q = p - 0.5;
  double r, val;
  if (fabs(q) <= 0.425) {
    r = 0.180625 - q * q;
    val = q *
          (((((((r * 2509.0809287301226727 + 33430.575583588128105) * r + 67265.770927008700853) * r +
               45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r +
            133.14166789178437745) * r + 3.387132872796366608) /
          (((((((r * 5226.495278852545925 + 28729.085735721942674) * r + 39307.89580009271061) * r +
               21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r +
            42.313330701600911252) * r + 1.0);
    return val;
  }
  r = (q < 0.0) ? p : 1.0 - p;
  r = sqrt(-log(r));
  if (r <= 5.0) {
    r -= 1.6;
    val = (((((((r * 7.7454501427834140764e-4 + 0.0227238449892691845833) * r + 0.24178072517745061177) * r +
               1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r +
            4.6303378461565452959) * r + 1.42343711074968357734) /
          (((((((r * 1.05075007164441684324e-9 + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r +
               0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r +
            2.05319162663775882187) * r + 1.0);
  } else {
    r -= 5.0;
    val = (((((((r * 2.01033439929228813265e-7 + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r +
               0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r +
            5.4637849111641143699) * r + 6.6579046435011037772) /
          (((((((r * 2.04426310338993978564e-15 + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r +
               7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r +
            0.59983220655588793769) * r + 1.0);
  }
  return (q < 0.0) ? -val : val;
}

double orc_norm_cdf(double x) { return 0.5 * erfc(-x * 0.70710678118654752440); }

/* Truncated normal by inverse transform (the distribution the reference
 * draws with TruncatedNormal, prng_extension.hpp:353-440). */
static double trunc_normal(double mu, double sd, double lo, double hi, double u) {
  const double pa = isinf(lo) ? 0.0 : orc_norm_cdf((lo - mu) / sd);
  const double pb = isinf(hi) ? 1.0 : orc_norm_cdf((hi - mu) / sd);
  double p = pa + u * (pb - pa);
  if (p <= 0.0) p = 1e-300;
  if (p >= 1.0) p = 1.0 - 1e-16;
  double x = mu + sd * orc_norm_inv(p);
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

/* ------------------------------------------------------------------------- */
/* Models.  State is an FP64 vector x[n_var]; c[] the local concentrations;
 * contrib[] the per-particle source term (kg/s) for each species. */

static const int MODEL_NVAR[ORC_MODEL_COUNT] = {2, 3, 6, 2, 9, 8, 2, 3};
static const int MODEL_NS[ORC_MODEL_COUNT] = {1, 2, 1, 1, 2, 4, 1, 1};
static const int MODEL_NP[ORC_MODEL_COUNT] = {6, 10, 6, 5, 0, 15, 5, 7};

int orc_model_info(int model, int* n_var, int* n_species, int* n_params) {
  if (model < 0 || model >= ORC_MODEL_COUNT) return -1;
  if (n_var) *n_var = MODEL_NVAR[model];
  if (n_species) *n_species = MODEL_NS[model];
  if (n_params) *n_params = MODEL_NP[model];
  return 0;
}

int orc_model_default_params(int model, double* p, int n) {
  if (model < 0 || model >= ORC_MODEL_COUNT || n < MODEL_NP[model]) return -1;
  switch (model) {
    case ORC_MODEL_MONOD_MASS: /* mu_max, Ks, Y, division ratio, m0 mean, m0 cv */
      p[0] = 0.5 / 3600.0; p[1] = 0.01; p[2] = 0.5; p[3] = 2.0; p[4] = 1e-15; p[5] = 0.1;
      break;
    case ORC_MODEL_DUAL_MONOD: /* qs_max, Ks, Ko, Yxs, Yos, tau, k_st, m_div, m0 mean, m0 sd */
      p[0] = 2.8e-4; p[1] = 0.01; p[2] = 1e-4; p[3] = 0.5; p[4] = 1.0; p[5] = 100.0; p[6] = 1e-4;
      p[7] = 2e-15; p[8] = 1.4e-15; p[9] = 0.2e-15;
      break;
    case ORC_MODEL_MONOD: /* mu_max, k_s, y_s_x, l_max, d_m, tau_meta */
      p[0] = 0.77 / 3600.0; p[1] = 1e-3; p[2] = 2.0; p[3] = 2e-6; p[4] = 0.6e-6; p[5] = 3600.0 / 0.77;
      break;
    case ORC_MODEL_FIXED_LENGTH: /* l_dot_max, l_max, k, d_m, lambda (config_loader.cpp:16) */
      p[0] = 2e-6 / 3600.0; p[1] = 2e-6; p[2] = 1e-3; p[3] = 0.6e-6; p[4] = 0.69314718055994530942 / 1e-6;
      break;
    case ORC_MODEL_QUAD_MONOD: /* qs_max, Ks, Ko, Ka, qa_max, Yxs, Yxa, Yos, Yoa, Yas, tau, Yco2, m_div, m0 mean, m0 sd */
      p[0] = 2.0; p[1] = 0.01; p[2] = 1e-4; p[3] = 0.01; p[4] = 0.5; p[5] = 0.5; p[6] = 0.3; p[7] = 0.3;
      p[8] = 0.4; p[9] = 0.3; p[10] = 0.2; p[11] = 0.4; p[12] = 2e-15; p[13] = 1.4e-15; p[14] = 0.2e-15;
      break;
    case ORC_MODEL_UDF_MINIMAL: /* Y, l_dot_max, l_max, k, d_m (udf_model/minimal.cpp:26-35) */
      p[0] = 0.5; p[1] = 2e-6 / 3600.0; p[2] = 2e-6; p[3] = 1e-3; p[4] = 0.6e-6;
      break;
    case ORC_MODEL_UDF_STARVATION: /* mu_max, Ks, Y, division ratio, kd, m0 mean, m0 cv */
      p[0] = 0.5 / 3600.0; p[1] = 0.01; p[2] = 0.5; p[3] = 2.0; p[4] = 1.0 / 3600.0; p[5] = 1e-15; p[6] = 0.1;
      break;
    default:
      break;
  }
  return 0;
}

static double lin_density(double d_m) { return 1000.0 * 3.14159265358979323846 * d_m * d_m / 4.0; }

/* simple_acetate constants (simple_acetate.hpp:59-84) */
#define SA_A_MAX (2e-6 / 3600.0)
#define SA_L_MAX 2e-6
#define SA_L_MIN (SA_L_MAX / 2.0)
#define SA_D_M 0.6e-6
#define SA_K_S 1e-3
#define SA_K_A 1e-4
#define SA_Y_S 2.0
#define SA_Y_A 3.0

static void model_init(int model, const double* p, double* x, uint64_t seed, uint32_t slot) {
  uint32_t w[4];
  draw4(seed, slot, 0u, TAG_INIT, 0u, w);
  switch (model) {
    case ORC_MODEL_MONOD_MASS: {
      const double m = trunc_normal(p[4], p[5] * p[4], 0.0, INFINITY, orc_u01_52(w[0], w[1]));
      x[0] = m;
      x[1] = m / (1.0 + orc_u01(w[2]));
      break;
    }
    case ORC_MODEL_DUAL_MONOD: {
      x[0] = trunc_normal(p[8], p[9], 0.5 * p[7], p[7], orc_u01_52(w[0], w[1]));
      x[1] = 0.5 * p[0];
      x[2] = 0.0;
      break;
    }
    case ORC_MODEL_MONOD: { /* monod.hpp:80-113 */
      const double lm = p[3];
      x[0] = trunc_normal(lm * 0.75, lm * 0.75 / 4.0, lm / 2.0, lm, orc_u01_52(w[0], w[1]));
      x[1] = lm;
      x[2] = p[0];
      x[3] = 0.0;
      x[4] = (lm / 2.0) / 0.69314718055994530942;
      x[5] = 0.0;
      break;
    }
    case ORC_MODEL_FIXED_LENGTH: { /* fixed_length.hpp:108-120 with config_loader.cpp:13-36: the
                                    * Metropolis target lambda*exp(-lambda*l) on [l_max/2, l_max],
                                    * sampled exactly by its inverse CDF (lambda = 0: uniform) */
      const double a = p[1] / 2.0, b = p[1], lam = p[4];
      const double u = orc_u01_52(w[0], w[1]);
      const double l = lam != 0.0 ? a - log1p(-u * -expm1(-lam * (b - a))) / lam : a + u * (b - a);
      x[0] = l < a ? a : (l > b ? b : l);
      x[1] = p[1];
      break;
    }
    case ORC_MODEL_UDF_MINIMAL: /* minimal.cpp:55-63: length uniform on (l_max/2, l_max) (:121-140) */
      x[0] = p[2] * (0.5 + 0.5 * orc_u01(w[0]));
      x[1] = p[2];
      break;
    case ORC_MODEL_UDF_STARVATION: {
      const double m = trunc_normal(p[5], p[6] * p[5], 0.0, INFINITY, orc_u01_52(w[0], w[1]));
      x[0] = m;
      x[1] = m / (1.0 + orc_u01(w[2]));
      x[2] = 0.2 + 0.8 * orc_u01(w[3]);
      break;
    }
    case ORC_MODEL_SIMPLE_ACETATE: { /* simple_acetate.hpp:132-152 */
      x[0] = trunc_normal(SA_L_MAX * 0.75, SA_L_MAX / 10.0, 0.7 * SA_L_MIN, SA_L_MAX * 1.3, orc_u01_52(w[0], w[1]));
      x[1] = trunc_normal(SA_L_MAX * 0.75, SA_L_MAX / 10.0, 0.7 * SA_L_MIN, SA_L_MAX * 1.3, orc_u01_52(w[2], w[3]));
      x[2] = SA_A_MAX / 2.0;
      x[3] = SA_A_MAX;
      for (int i = 4; i < 9; ++i) x[i] = 0.0;
      break;
    }
    case ORC_MODEL_QUAD_MONOD:
      x[0] = trunc_normal(p[13], p[14], 0.5 * p[12], p[12], orc_u01_52(w[0], w[1]));
      x[1] = 0.5 * p[0];
      x[2] = 0.0;
      x[3] = 0.0;
      x[4] = 0.0;
      x[5] = 0.0;
      x[6] = x[0];
      x[7] = 0.0;
      break;
    default:
      break;
  }
}

static double model_mass(int model, const double* p, const double* x) {
  switch (model) {
    case ORC_MODEL_MONOD_MASS:
    case ORC_MODEL_DUAL_MONOD:
    case ORC_MODEL_QUAD_MONOD:
    case ORC_MODEL_UDF_STARVATION:
      return x[0];
    case ORC_MODEL_MONOD:
      return x[0] * lin_density(p[4]);
    case ORC_MODEL_FIXED_LENGTH:
      return x[0] * lin_density(p[3]);
    case ORC_MODEL_UDF_MINIMAL: /* minimal.cpp:101-105 */
      return x[0] * lin_density(p[4]);
    case ORC_MODEL_SIMPLE_ACETATE:
      return x[0] * lin_density(SA_D_M);
    default:
      return 0.0;
  }
}

/* returns ORC_IDLE, ORC_DIVISION or ORC_DEAD */
static int model_update(int model, const double* p, double dt, double* x, const double* c, double* contrib) {
  switch (model) {
    case ORC_MODEL_MONOD_MASS: {
      const double s = c[0] > 0.0 ? c[0] : 0.0;
      const double mu = p[0] * s / (p[1] + s);
      const double g = mu * x[0];
      contrib[0] = -(g / p[2]);
      x[0] = x[0] + dt * g;
      return (x[0] >= p[3] * x[1]) ? ORC_DIVISION : ORC_IDLE;
    }
    case ORC_MODEL_DUAL_MONOD: {
      const double S = c[0] > 0.0 ? c[0] : 0.0;
      const double O = c[1] > 0.0 ? c[1] : 0.0;
      const double qt = p[0] * S / (p[1] + S);
      const double U = x[1] * x[0];
      const double fo = O / (p[2] + O);
      const double G = p[3] * U * fo;
      const double Fst = U * (1.0 - fo);
      const double R = p[6] * x[2] * fo;
      const double Gt = G + p[3] * R;
      contrib[0] = -U;
      contrib[1] = -(p[4] * Gt);
      x[0] = x[0] + dt * Gt;
      x[2] = x[2] + dt * (Fst - R);
      x[1] = x[1] + (dt / p[5]) * (qt - x[1]);
      return (x[0] >= p[7]) ? ORC_DIVISION : ORC_IDLE;
    }
    case ORC_MODEL_MONOD: { /* monod.hpp:122-160 */
      const double s = c[0] > 0.0 ? c[0] : 0.0;
      const double mu = p[0] * s / (p[1] + s);
      const double mu_eff = x[2] < mu ? x[2] : mu;
      x[0] = x[0] + dt * (mu_eff * x[4]);
      x[2] = x[2] + dt * (1.0 / p[5]) * (mu - x[2]);
      x[3] = mu_eff;
      x[5] = -mu_eff * p[2] * (x[0] * lin_density(p[4]));
      contrib[0] = x[5];
      return (x[0] >= x[1]) ? ORC_DIVISION : ORC_IDLE;
    }
    case ORC_MODEL_FIXED_LENGTH: { /* fixed_length.hpp:123-142 */
      const double ld = lin_density(p[3]);
      const double phi_s_max = (ld * p[0]) / 0.5;
      const double s = c[0];
      const double g = s / (p[2] + s);
      const double phi = phi_s_max * g;
      const double ldot = p[0] * g;
      x[0] = x[0] + dt * ldot;
      contrib[0] = -phi;
      return (x[0] >= x[1]) ? ORC_DIVISION : ORC_IDLE;
    }
    case ORC_MODEL_UDF_STARVATION: { /* viability falls at kd (1 - g); Dead at zero */
      const double s = c[0] > 0.0 ? c[0] : 0.0;
      const double g = s / (p[1] + s);
      x[2] = x[2] - dt * (p[4] * (1.0 - g));
      if (x[2] <= 0.0) return ORC_DEAD;
      const double gm = (p[0] * g) * x[0];
      contrib[0] = -(gm / p[2]);
      x[0] = x[0] + dt * gm;
      return (x[0] >= p[3] * x[1]) ? ORC_DIVISION : ORC_IDLE;
    }
    case ORC_MODEL_UDF_MINIMAL: { /* minimal.cpp:65-83 */
      const double s = c[0];
      const double g = s / (p[3] + s);
      const double phi_s_max = (p[1] * lin_density(p[4])) / p[0];
      const double ldot = p[1] * g;
      const double d_length = dt * ldot;
      x[0] = x[0] + d_length / (1.0 + dt * ldot);
      contrib[0] = -(phi_s_max * g);
      return (x[0] >= x[1]) ? ORC_DIVISION : ORC_IDLE;
    }
    case ORC_MODEL_SIMPLE_ACETATE: { /* simple_acetate.hpp:154-204 */
      const double ld = lin_density(SA_D_M);
      const double adm0 = x[3], adm1 = x[3] / 3.0;
      const double c0 = c[0], c1 = c[1];
      const double D0 = adm0 * c0 * (1.0 / (c0 + SA_K_S));
      const double D1 = adm1 * c1 * (1.0 / (c1 + SA_K_A));
      const double U0 = D0 < x[2] ? D0 : x[2];
      const double pa = D0 - x[2];
      const double mask = (pa < 0.0) ? 1.0 : 0.0;
      const double mpa = -pa;
      const double U1 = mask * (D1 < mpa ? D1 : mpa) + (1.0 - mask) * 0.0;
      x[4] = 0.0 + U0;
      x[4] = x[4] + U1;
      x[0] = x[0] + dt * x[4];
      x[5] = U0;
      x[6] = U1;
      const double phi_s = -1.0 * D0 * ld * SA_Y_S;
      const double phi_a = mask * (-U1 * ld * SA_Y_A) + (1.0 - mask) * (pa * ld * SA_Y_S / SA_Y_A);
      x[7] = phi_s;
      x[8] = phi_a;
      contrib[0] = phi_s;
      contrib[1] = phi_a;
      return (x[0] >= x[1]) ? ORC_DIVISION : ORC_IDLE;
    }
    case ORC_MODEL_QUAD_MONOD: {
      /* glucose S, oxygen O, acetate A, CO2: saturations of the liquid */
      const double S = c[0] > 0.0 ? c[0] : 0.0;
      const double O = c[1] > 0.0 ? c[1] : 0.0;
      const double A = c[2] > 0.0 ? c[2] : 0.0;
      const double fs = S / (p[1] + S);
      const double fo = O / (p[2] + O);
      const double fa = A / (p[3] + A);
      const double rep = p[1] / (p[1] + S);
      const double qs_t = p[0] * fs;
      const double qa_t = p[4] * fa * rep;
      const double Us = x[1] * x[0];
      const double Ua = x[2] * x[0] * fo;
      const double Resp = Us * fo;
      const double Ovf = Us - Resp;
      const double G = p[5] * Resp + p[6] * Ua;
      const double O2 = p[7] * Resp + p[8] * Ua;
      contrib[0] = -Us;
      contrib[1] = -O2;
      contrib[2] = p[9] * Ovf - Ua;
      contrib[3] = p[11] * (Resp + Ua);
      x[0] = x[0] + dt * G;
      x[3] = x[3] + dt * (Ovf * (1.0 - p[9]));
      x[1] = x[1] + (dt / p[10]) * (qs_t - x[1]);
      x[2] = x[2] + (dt / p[10]) * (qa_t - x[2]);
      x[4] = O2;
      x[5] = Ovf;
      return (x[0] >= p[12]) ? ORC_DIVISION : ORC_IDLE;
    }
    default:
      return ORC_IDLE;
  }
}

static void model_division(int model, const double* p, double* x, double* xd, uint64_t seed, uint32_t slot,
                           uint32_t step) {
  switch (model) {
    case ORC_MODEL_MONOD_MASS: {
      const double h = x[0] * 0.5;
      x[0] = h;
      x[1] = h;
      xd[0] = h;
      xd[1] = h;
      break;
    }
    case ORC_MODEL_DUAL_MONOD:
      x[0] = x[0] * 0.5;
      x[2] = x[2] * 0.5;
      xd[0] = x[0];
      xd[1] = x[1];
      xd[2] = x[2];
      break;
    case ORC_MODEL_MONOD: { /* monod.hpp:162-193 */
      const double h = x[0] / 2.0;
      x[0] = h;
      xd[0] = h;
      xd[1] = p[3];
      xd[2] = x[2];
      xd[3] = 0.0;
      xd[4] = x[4];
      xd[5] = 0.0;
      break;
    }
    case ORC_MODEL_FIXED_LENGTH: { /* fixed_length.hpp:144-161 */
      const double h = x[0] / 2.0;
      x[0] = h;
      x[1] = p[1];
      xd[0] = h;
      xd[1] = p[1];
      break;
    }
    case ORC_MODEL_UDF_STARVATION: {
      const double h = x[0] * 0.5;
      x[0] = h;
      x[1] = h;
      x[2] = 1.0;
      xd[0] = h;
      xd[1] = h;
      xd[2] = 1.0;
      break;
    }
    case ORC_MODEL_UDF_MINIMAL: { /* minimal.cpp:85-99 */
      const double h = x[0] / 2.0;
      x[0] = h;
      x[1] = p[2];
      xd[0] = h;
      xd[1] = p[2];
      break;
    }
    case ORC_MODEL_SIMPLE_ACETATE: { /* simple_acetate.hpp:206-248 */
      x[0] = x[0] / 2.0;
      for (int i = 0; i < 4; ++i) xd[i] = x[i];
      for (int i = 4; i < 9; ++i) xd[i] = 0.0;
      uint32_t w[4], v[4];
      draw4(seed, slot, step, TAG_DIVISION, 0u, w);
      draw4(seed, slot, step, TAG_DIVISION, 1u, v);
      const double sigma = 0.2;
      const double average = log(x[4]) - sigma * sigma / 2.0;
      const double g1 = exp(average + sigma * orc_norm_inv(orc_u01_52(w[0], w[1])));
      const double g2 = exp(average + sigma * orc_norm_inv(orc_u01_52(w[2], w[3])));
      const double l1 = trunc_normal(SA_L_MAX, SA_L_MAX / 10.0, SA_L_MAX * 0.7, 1.3 * SA_L_MAX, orc_u01_52(v[0], v[1]));
      const double l2 = trunc_normal(SA_L_MAX, SA_L_MAX / 10.0, SA_L_MAX * 0.7, 1.3 * SA_L_MAX, orc_u01_52(v[2], v[3]));
      x[2] = g1;
      x[1] = l1;
      xd[2] = g2;
      xd[1] = l2;
      break;
    }
    case ORC_MODEL_QUAD_MONOD: {
      x[0] = x[0] * 0.5;
      x[3] = x[3] * 0.5;
      x[6] = x[0];
      x[7] = x[7] + 1.0;
      for (int i = 0; i < 8; ++i) xd[i] = x[i];
      break;
    }
    default:
      break;
  }
}

/* ------------------------------------------------------------------------- */
/* Flowmap tables (what rcmtool + ReactorDomain::update provide upstream). */

typedef struct {
  uint32_t n;
  double* V;
  double* invV;
  double* out_flow;
  double* p_move;
  uint32_t m;
  double* cdf;
  uint32_t* nbr;
  uint64_t* col_ptr; /* transport matrix by destination column, n+1 */
  uint32_t* src;
  double* val;
} flowmap;

typedef struct {
  uint32_t i, j;
  uint64_t ord; /* input position: duplicates are summed in input order */
  double q;
} coo_t;

static int cmp_row_col(const void* a, const void* b) {
  const coo_t* x = (const coo_t*)a;
  const coo_t* y = (const coo_t*)b;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  if (x->ord != y->ord) return x->ord < y->ord ? -1 : 1;
  return 0;
}
static int cmp_col_row(const void* a, const void* b) {
  const coo_t* x = (const coo_t*)a;
  const coo_t* y = (const coo_t*)b;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  return 0;
}

static int cmp_col_row_ord(const void* a, const void* b) {
  const coo_t* x = (const coo_t*)a;
  const coo_t* y = (const coo_t*)b;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->ord != y->ord) return x->ord < y->ord ? -1 : 1;
  return 0;
}

/* Canonical flow list: off-diagonal, positive, duplicates summed in input
 * order, sorted by (row, col). */
static coo_t* canonical_flows(uint32_t n, uint64_t nnz, const uint32_t* rows, const uint32_t* cols,
                              const double* flows, uint64_t* out_nnz) {
  coo_t* e = (coo_t*)malloc(sizeof(coo_t) * (nnz ? nnz : 1));
  uint64_t k = 0;
  for (uint64_t t = 0; t < nnz; ++t) {
    if (rows[t] >= n || cols[t] >= n) {
      free(e);
      return NULL;
    }
    if (rows[t] == cols[t]) continue;
    e[k].i = rows[t];
    e[k].j = cols[t];
    e[k].ord = t;
    e[k].q = flows[t];
    ++k;
  }
  /* duplicates of (i,j) end up adjacent, in input order */
  qsort(e, k, sizeof(coo_t), cmp_row_col);
  uint64_t w = 0;
  for (uint64_t t = 0; t < k;) {
    coo_t acc = e[t];
    uint64_t u = t + 1;
    while (u < k && e[u].i == acc.i && e[u].j == acc.j) {
      acc.q += e[u].q;
      ++u;
    }
    if (acc.q > 0.0) e[w++] = acc;
    t = u;
  }
  *out_nnz = w;
  return e;
}

static void flowmap_free(flowmap* f) {
  free(f->V);
  free(f->invV);
  free(f->out_flow);
  free(f->p_move);
  free(f->cdf);
  free(f->nbr);
  free(f->col_ptr);
  free(f->src);
  free(f->val);
  memset(f, 0, sizeof(*f));
}

static int flowmap_build(flowmap* f, uint32_t n, const double* V, uint64_t nnz, const uint32_t* rows,
                         const uint32_t* cols, const double* flows, double dt) {
  memset(f, 0, sizeof(*f));
  uint64_t k = 0;
  coo_t* e = canonical_flows(n, nnz, rows, cols, flows, &k);
  if (!e) return -2;
  f->n = n;
  f->V = (double*)malloc(sizeof(double) * n);
  f->invV = (double*)malloc(sizeof(double) * n);
  f->out_flow = (double*)calloc(n, sizeof(double));
  f->p_move = (double*)malloc(sizeof(double) * n);
  uint32_t* deg = (uint32_t*)calloc(n, sizeof(uint32_t));
  uint64_t* rs = (uint64_t*)calloc((size_t)n + 1, sizeof(uint64_t));
  for (uint64_t t = 0; t < k; ++t) deg[e[t].i]++;
  uint32_t m = 1;
  for (uint32_t i = 0; i < n; ++i) {
    if (deg[i] > m) m = deg[i];
    rs[i + 1] = rs[i] + deg[i];
  }
  f->m = m;
  f->cdf = (double*)malloc(sizeof(double) * (size_t)n * m);
  f->nbr = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n * m);
  for (uint32_t i = 0; i < n; ++i) {
    if (!(V[i] > 0.0)) {
      free(e); free(deg); free(rs);
      flowmap_free(f);
      return -3;
    }
    f->V[i] = V[i];
    f->invV[i] = 1.0 / V[i];
    double acc = 0.0;
    for (uint64_t t = rs[i]; t < rs[i + 1]; ++t) acc += e[t].q;
    f->out_flow[i] = acc;
    f->p_move[i] = (dt * acc) / V[i];
    double run = 0.0;
    uint32_t c = 0;
    for (uint64_t t = rs[i]; t < rs[i + 1]; ++t, ++c) {
      run += e[t].q;
      f->cdf[(size_t)i * m + c] = run / acc;
      f->nbr[(size_t)i * m + c] = e[t].j;
    }
    const uint32_t last = deg[i] ? e[rs[i + 1] - 1].j : i;
    for (; c < m; ++c) {
      f->cdf[(size_t)i * m + c] = deg[i] ? 1.0 : 0.0;
      f->nbr[(size_t)i * m + c] = last;
    }
  }
  /* transport matrix by destination column: off-diagonal F(i,j) + diagonal */
  const uint64_t tn = k + n;
  coo_t* T = (coo_t*)malloc(sizeof(coo_t) * tn);
  for (uint64_t t = 0; t < k; ++t) T[t] = e[t];
  for (uint32_t i = 0; i < n; ++i) {
    T[k + i].i = i;
    T[k + i].j = i;
    T[k + i].ord = 0;
    T[k + i].q = -f->out_flow[i];
  }
  qsort(T, tn, sizeof(coo_t), cmp_col_row);
  f->col_ptr = (uint64_t*)calloc((size_t)n + 1, sizeof(uint64_t));
  f->src = (uint32_t*)malloc(sizeof(uint32_t) * tn);
  f->val = (double*)malloc(sizeof(double) * tn);
  for (uint64_t t = 0; t < tn; ++t) {
    f->col_ptr[T[t].j + 1]++;
    f->src[t] = T[t].i;
    f->val[t] = T[t].q;
  }
  for (uint32_t j = 0; j < n; ++j) f->col_ptr[j + 1] += f->col_ptr[j];
  free(T);
  free(e);
  free(deg);
  free(rs);
  return 0;
}

/* A flowmap from the tables the reference's hydrodynamic state hands over
 * (ReactorDomain::update, domain.cpp:44-74, and set_transition,
 * implScalar.cpp:316-320): the neighbour rows and cumulative probability rows
 * as given (n x n_cols, any neighbour order, any padding), the out-flows
 * (diag_transition: p_move = dt*Q/V), and the transition matrix as COO
 * triplets (off-diagonal F(i,j), diagonal -Q_out(i)).  The triplets are taken
 * by destination column, ascending source, duplicates summed in input order
 * (as sparse_from_coo sums them). */
static int flowmap_from_tables(flowmap* f, uint32_t n, const double* V, uint32_t n_cols, const uint64_t* nbr,
                               const double* cdf, const double* out_flow, uint64_t nnz, const uint64_t* rows,
                               const uint64_t* cols, const double* vals, double dt) {
  memset(f, 0, sizeof(*f));
  if (n_cols == 0) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    if (!(V[i] > 0.0) || !isfinite(V[i]) || !(out_flow[i] >= 0.0) || !isfinite(out_flow[i])) return -3;
    for (uint32_t k = 0; k < n_cols; ++k) {
      const size_t o = (size_t)i * n_cols + k;
      if (nbr[o] >= n || !isfinite(cdf[o])) return -3;
      if (k && cdf[o] < cdf[o - 1]) return -3; /* rows must be cumulative */
    }
  }
  coo_t* T = (coo_t*)malloc(sizeof(coo_t) * (nnz ? nnz : 1));
  for (uint64_t t = 0; t < nnz; ++t) {
    if (rows[t] >= n || cols[t] >= n || !isfinite(vals[t])) {
      free(T);
      return -3;
    }
    T[t].i = (uint32_t)rows[t];
    T[t].j = (uint32_t)cols[t];
    T[t].ord = t;
    T[t].q = vals[t];
  }
  qsort(T, nnz, sizeof(coo_t), cmp_col_row_ord);
  f->n = n;
  f->m = n_cols;
  f->V = (double*)malloc(sizeof(double) * n);
  f->invV = (double*)malloc(sizeof(double) * n);
  f->out_flow = (double*)malloc(sizeof(double) * n);
  f->p_move = (double*)malloc(sizeof(double) * n);
  f->cdf = (double*)malloc(sizeof(double) * (size_t)n * n_cols);
  f->nbr = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n * n_cols);
  for (uint32_t i = 0; i < n; ++i) {
    f->V[i] = V[i];
    f->invV[i] = 1.0 / V[i];
    f->out_flow[i] = out_flow[i];
    f->p_move[i] = (dt * out_flow[i]) / V[i];
  }
  for (size_t o = 0; o < (size_t)n * n_cols; ++o) {
    f->cdf[o] = cdf[o];
    f->nbr[o] = (uint32_t)nbr[o];
  }
  f->col_ptr = (uint64_t*)calloc((size_t)n + 1, sizeof(uint64_t));
  f->src = (uint32_t*)malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
  f->val = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
  uint64_t w = 0;
  for (uint64_t t = 0; t < nnz;) {
    coo_t acc = T[t];
    uint64_t u = t + 1;
    for (; u < nnz && T[u].i == acc.i && T[u].j == acc.j; ++u) acc.q += T[u].q;
    f->col_ptr[acc.j + 1]++;
    f->src[w] = acc.i;
    f->val[w] = acc.q;
    ++w;
    t = u;
  }
  for (uint32_t j = 0; j < n; ++j) f->col_ptr[j + 1] += f->col_ptr[j];
  free(T);
  return 0;
}

int orc_transition_dense(uint32_t n, uint64_t nnz, const uint32_t* rows, const uint32_t* cols,
                         const double* flows, double* out) {
  uint64_t k = 0;
  coo_t* e = canonical_flows(n, nnz, rows, cols, flows, &k);
  if (!e) return -2;
  memset(out, 0, sizeof(double) * (size_t)n * n);
  for (uint64_t t = 0; t < k; ++t) {
    out[(size_t)e[t].i * n + e[t].j] = e[t].q;
    out[(size_t)e[t].i * n + e[t].i] -= e[t].q;
  }
  free(e);
  return 0;
}

int orc_build_move_tables(uint32_t n, const double* V, uint64_t nnz, const uint32_t* rows,
                          const uint32_t* cols, const double* flows, double dt, uint32_t* m_out,
                          double* out_flow, double* p_move, double* cdf, uint32_t* nbr) {
  flowmap f;
  const int rc = flowmap_build(&f, n, V, nnz, rows, cols, flows, dt);
  if (rc) return rc;
  *m_out = f.m;
  if (out_flow) memcpy(out_flow, f.out_flow, sizeof(double) * n);
  if (p_move) memcpy(p_move, f.p_move, sizeof(double) * n);
  if (cdf) memcpy(cdf, f.cdf, sizeof(double) * (size_t)n * f.m);
  if (nbr) memcpy(nbr, f.nbr, sizeof(uint32_t) * (size_t)n * f.m);
  flowmap_free(&f);
  return 0;
}

/* m += dt*(C*M - C*diag(sink) + S); C = m*diag(1/V)   (implScalar.cpp:250-264).
   With gas-liquid transfer (kla != NULL) the liquid also receives
   mtr = kla*(C* - C)*V, C* the saturation set by a fixed gas phase
   (performStepGL, implScalar.cpp:228-248; mass_transfer.cpp:140-157). */
static void transport(const flowmap* f, uint32_t n_s, const double* sink, const double* S_total, double dt,
                      const double* C_in, double* m, double* C_out, const double* kla, const double* cstar) {
  const uint32_t n = f->n;
  for (uint32_t s = 0; s < n_s; ++s) {
    const double* c = C_in + (size_t)s * n;
    for (uint32_t j = 0; j < n; ++j) {
      double acc = 0.0;
      for (uint64_t t = f->col_ptr[j]; t < f->col_ptr[j + 1]; ++t) acc += c[f->src[t]] * f->val[t];
      const size_t o = (size_t)s * n + j;
      double dmdt = (acc - c[j] * sink[j]) + S_total[o];
      if (kla) dmdt = dmdt + (kla[o] * (cstar[o] - c[j])) * f->V[j];
      m[o] = m[o] + dt * dmdt;
      C_out[o] = m[o] * f->invV[j];
    }
  }
}

int orc_transport_once(uint32_t n, uint32_t n_s, const double* V, uint64_t nnz, const uint32_t* rows,
                       const uint32_t* cols, const double* flows, const double* sink, const double* S,
                       double dt, const double* C_in, double* m_inout, double* C_out) {
  flowmap f;
  const int rc = flowmap_build(&f, n, V, nnz, rows, cols, flows, dt);
  if (rc) return rc;
  transport(&f, n_s, sink, S, dt, C_in, m_inout, C_out, NULL, NULL);
  flowmap_free(&f);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Run state */

typedef struct {
  int type;
  int gas; /* a feed of the gas phase (SimulationFeed::gas_feeds): its output is a gas sink only */
  double f0, param;
  uint32_t in;
  int64_t out;
  uint32_t nv;
  uint32_t sp[MAX_SPECIES];
  double conc[MAX_SPECIES];
} feed_t;

typedef struct {
  double* v;
  uint64_t n, cap;
} dvec;

static void dvec_push(dvec* d, double x) {
  if (d->n == d->cap) {
    d->cap = d->cap ? 2 * d->cap : 1024;
    d->v = (double*)realloc(d->v, sizeof(double) * d->cap);
  }
  d->v[d->n++] = x;
}

struct orc_run {
  int model, n_var, n_s, n_params;
  double params[MAX_PARAMS];
  uint32_t n_c;
  double dt;
  uint64_t seed;
  flowmap* maps;
  uint32_t n_maps, steps_per_map;
  feed_t* feeds;
  uint32_t n_feeds;
  double *C, *C2, *m, *S, *S_total, *sink, *q_exit;
  double *kla, *cstar; /* gas-liquid transfer [n_s][n_c], NULL without */
  /* two-phase runs: the gas phase (gas side of each flowmap, its state and
     feeds), Henry constants and the kLa rule */
  int two_phase;
  double henry[MAX_SPECIES];
  double *Cg, *Cg2, *mg, *Sg_total, *sink_g, *Cg0;
  flowmap* gmaps;
  double** g_eps; /* energy dissipation per gas map [n_c], NULL if not given */
  double** g_kla; /* kLa per gas map [n_s][n_c], computed at the first step */
  uint32_t n_gmaps;
  int kla_mode;
  double kla_fixed[MAX_SPECIES];
  uint32_t kla_species;
  double bubble_d;
  int gas_ready;
  uint64_t n, cap;
  uint32_t* pos;
  uint8_t* status;
  double* state; /* [n_var][cap] */
  double* age_hyd;
  double* age_div;
  double weight;
  uint64_t inactive, min_removal;
  double ratio;
  uint64_t div_cap;       /* newborns per step at most (the reference's buffer size) */
  uint32_t sort_interval; /* reorder by compartment every k steps (0: never) */
  uint32_t grow_every;    /* capacity check every k steps (0: fixed capacity) */
  double grow_factor;     /* new capacity = ceil(n * grow_factor) */
  uint64_t counters[ORC_C_COUNT];
  dvec exit_ages, div_ages;
};

orc_run* orc_create(int model, uint32_t n_c, uint32_t n_s, double dt, uint64_t seed) {
  if (model < 0 || model >= ORC_MODEL_COUNT || n_c == 0 || !(dt > 0.0)) return NULL;
  if ((int)n_s != MODEL_NS[model]) return NULL;
  orc_run* r = (orc_run*)calloc(1, sizeof(orc_run));
  r->model = model;
  r->n_var = MODEL_NVAR[model];
  r->n_s = (int)n_s;
  r->n_params = MODEL_NP[model];
  orc_model_default_params(model, r->params, MAX_PARAMS);
  r->n_c = n_c;
  r->dt = dt;
  r->seed = seed;
  r->steps_per_map = 1;
  const size_t ns = (size_t)n_s * n_c;
  r->C = (double*)calloc(ns, sizeof(double));
  r->C2 = (double*)calloc(ns, sizeof(double));
  r->m = (double*)calloc(ns, sizeof(double));
  r->S = (double*)calloc(ns, sizeof(double));
  r->S_total = (double*)calloc(ns, sizeof(double));
  r->sink = (double*)calloc(n_c, sizeof(double));
  r->q_exit = (double*)calloc(n_c, sizeof(double));
  r->ratio = 0.01;
  r->min_removal = 0;
  r->div_cap = UINT64_MAX;
  return r;
}

void orc_destroy(orc_run* r) {
  if (!r) return;
  for (uint32_t i = 0; i < r->n_maps; ++i) flowmap_free(&r->maps[i]);
  free(r->maps);
  free(r->feeds);
  free(r->C); free(r->C2); free(r->m); free(r->S); free(r->S_total); free(r->sink); free(r->q_exit);
  free(r->kla); free(r->cstar);
  for (uint32_t i = 0; i < r->n_gmaps; ++i) {
    flowmap_free(&r->gmaps[i]);
    free(r->g_eps[i]);
    free(r->g_kla[i]);
  }
  free(r->gmaps); free(r->g_eps); free(r->g_kla);
  free(r->Cg); free(r->Cg2); free(r->mg); free(r->Sg_total); free(r->sink_g); free(r->Cg0);
  free(r->pos); free(r->status); free(r->state); free(r->age_hyd); free(r->age_div);
  free(r->exit_ages.v);
  free(r->div_ages.v);
  free(r);
}

int orc_set_model_params(orc_run* r, const double* p, int n) {
  if (n > MAX_PARAMS || n < 0) return -1;
  for (int i = 0; i < n; ++i) r->params[i] = p[i];
  return 0;
}

int orc_add_flowmap(orc_run* r, const double* V, uint64_t nnz, const uint32_t* rows, const uint32_t* cols,
                    const double* flows) {
  r->maps = (flowmap*)realloc(r->maps, sizeof(flowmap) * (r->n_maps + 1));
  const int rc = flowmap_build(&r->maps[r->n_maps], r->n_c, V, nnz, rows, cols, flows, r->dt);
  if (rc) return rc;
  r->n_maps++;
  return 0;
}

int orc_add_flowmap_tables(orc_run* r, const double* V, uint32_t n_cols, const uint64_t* nbr, const double* cdf,
                           const double* out_flow, uint64_t nnz, const uint64_t* rows, const uint64_t* cols,
                           const double* vals) {
  flowmap f;
  const int rc = flowmap_from_tables(&f, r->n_c, V, n_cols, nbr, cdf, out_flow, nnz, rows, cols, vals, r->dt);
  if (rc) {
    flowmap_free(&f);
    return rc;
  }
  r->maps = (flowmap*)realloc(r->maps, sizeof(flowmap) * (r->n_maps + 1));
  r->maps[r->n_maps++] = f;
  return 0;
}

int orc_set_steps_per_map(orc_run* r, uint32_t k) {
  if (k == 0) return -1;
  r->steps_per_map = k;
  return 0;
}

static int add_feed_phase(orc_run* r, int gas, int type, double f0, double param, uint32_t input, int64_t output,
                          uint32_t n_values, const uint32_t* species, const double* conc) {
  if (input >= r->n_c || output >= (int64_t)r->n_c || n_values > MAX_SPECIES || f0 < 0.0) return -1;
  r->feeds = (feed_t*)realloc(r->feeds, sizeof(feed_t) * (r->n_feeds + 1));
  feed_t* f = &r->feeds[r->n_feeds];
  memset(f, 0, sizeof(*f));
  f->gas = gas;
  f->type = type;
  f->f0 = f0;
  f->param = param;
  f->in = input;
  f->out = output;
  f->nv = n_values;
  for (uint32_t i = 0; i < n_values; ++i) {
    if (species[i] >= (uint32_t)r->n_s) return -1;
    f->sp[i] = species[i];
    f->conc[i] = conc[i];
  }
  r->n_feeds++;
  return 0;
}

int orc_add_feed(orc_run* r, int type, double f0, double param, uint32_t input, int64_t output,
                 uint32_t n_values, const uint32_t* species, const double* conc) {
  return add_feed_phase(r, 0, type, f0, param, input, output, n_values, species, conc);
}

/* ------------------------------------------------------------------------- */
/* Gas phase of two-phase runs: SimulationUnit::ode_step with a gas phase
 * (simulation.model.cpp:131-150): mtr = kla*(H*C_gas - C_liq)*V_liq
 * (mass_transfer.cpp:140-157), the gas step with -mtr and the liquid step
 * with +mtr (performStepGL, implScalar.cpp:228-248), then clearNegs on the
 * liquid concentrations (implScalar.cpp:266-297).  kLa is fixed per species
 * (MassTransfer::Type::FixedKla, mass_transfer.cpp:26-37) or, for one
 * species, kL(eps)*a(alpha_g) from the flowmap's energy dissipation and gas
 * fraction (FlowmapTurbulence, impl_mtr.cpp:21-50,90-134). */
int orc_add_gas_feed(orc_run* r, int type, double f0, double param, uint32_t input, int64_t output,
                     uint32_t n_values, const uint32_t* species, const double* conc) {
  return add_feed_phase(r, 1, type, f0, param, input, output, n_values, species, conc);
}

int orc_set_gas_phase(orc_run* r, const double* henry, const double* c_gas) {
  const size_t nsc = (size_t)r->n_s * r->n_c;
  if (!henry || !c_gas || r->kla) return -1;
  for (int s = 0; s < r->n_s; ++s)
    if (!(henry[s] >= 0.0) || !isfinite(henry[s])) return -1;
  for (size_t o = 0; o < nsc; ++o)
    if (!(c_gas[o] >= 0.0) || !isfinite(c_gas[o])) return -1;
  r->two_phase = 1;
  for (int s = 0; s < r->n_s; ++s) r->henry[s] = henry[s];
  free(r->Cg0);
  r->Cg0 = (double*)malloc(sizeof(double) * nsc);
  memcpy(r->Cg0, c_gas, sizeof(double) * nsc);
  r->gas_ready = 0;
  return 0;
}

int orc_set_kla(orc_run* r, int mode, const double* kla, uint32_t species, double bubble_d) {
  if (mode == ORC_KLA_FIXED) {
    if (!kla) return -1;
    for (int s = 0; s < r->n_s; ++s) {
      if (!(kla[s] >= 0.0) || !isfinite(kla[s])) return -1;
      r->kla_fixed[s] = kla[s];
    }
  } else if (mode == ORC_KLA_TURBULENCE) {
    if (species >= (uint32_t)r->n_s || !(bubble_d > 0.0) || !isfinite(bubble_d)) return -1;
    r->kla_species = species;
    r->bubble_d = bubble_d;
  } else {
    return -1;
  }
  r->kla_mode = mode;
  r->gas_ready = 0;
  return 0;
}

int orc_add_gas_flowmap(orc_run* r, const double* Vg, uint64_t nnz, const uint32_t* rows, const uint32_t* cols,
                        const double* flows, const double* eps) {
  flowmap g;
  const int rc = flowmap_build(&g, r->n_c, Vg, nnz, rows, cols, flows, r->dt);
  if (rc) return rc;
  if (eps)
    for (uint32_t j = 0; j < r->n_c; ++j)
      if (!(eps[j] >= 0.0) || !isfinite(eps[j])) {
        flowmap_free(&g);
        return -3;
      }
  r->gmaps = (flowmap*)realloc(r->gmaps, sizeof(flowmap) * (r->n_gmaps + 1));
  r->g_eps = (double**)realloc(r->g_eps, sizeof(double*) * (r->n_gmaps + 1));
  r->g_kla = (double**)realloc(r->g_kla, sizeof(double*) * (r->n_gmaps + 1));
  r->gmaps[r->n_gmaps] = g;
  r->g_eps[r->n_gmaps] = NULL;
  if (eps) {
    r->g_eps[r->n_gmaps] = (double*)malloc(sizeof(double) * r->n_c);
    memcpy(r->g_eps[r->n_gmaps], eps, sizeof(double) * r->n_c);
  }
  r->g_kla[r->n_gmaps] = NULL;
  r->n_gmaps++;
  r->gas_ready = 0;
  return 0;
}

/* Kinematic viscosity of water at 20 C as the reference's correlation gives
   it (impl_mtr.cpp:58-84 at temperature 20, impl_mtr.cpp:98-99), and the
   oxygen diffusion constant (impl_mtr.cpp:97). */
#define GL_WATER_NU_20C 1.0023e-6
#define GL_O2_DIFFUSION 1e-9

/* kLa [n_s][n] of gas map g with liquid map f (FixedKla rows, or the
   turbulence correlation kL = 0.3 (eps nu)^(1/4) Sc^(-1/2) times the
   interfacial area a = 6 alpha/(d_b (1 - alpha)), alpha = Vg/(Vl + Vg)). */
static void gas_kla(const orc_run* r, const flowmap* f, const flowmap* g, const double* eps, double* kla) {
  const uint32_t n = r->n_c;
  for (int s = 0; s < r->n_s; ++s)
    for (uint32_t j = 0; j < n; ++j)
      kla[(size_t)s * n + j] = r->kla_mode == ORC_KLA_FIXED ? r->kla_fixed[s] : 0.0;
  if (r->kla_mode != ORC_KLA_TURBULENCE) return;
  const double nu = GL_WATER_NU_20C;
  const double sc = nu / GL_O2_DIFFUSION;
  for (uint32_t j = 0; j < n; ++j) {
    const double e = eps ? eps[j] : 0.0;
    const double kl = (0.3 * pow(e * nu, 0.25)) * pow(sc, -0.5);
    const double alpha = g->V[j] / (f->V[j] + g->V[j]);
    const double a = (6.0 * alpha) / (r->bubble_d * (1.0 - alpha));
    kla[(size_t)r->kla_species * n + j] = kl * a;
  }
}

/* Checks and first-step set-up of a two-phase run: one gas map per liquid
   map, the gas mass from the initial gas concentrations and the first gas
   volumes, the kLa of every map. */
static int gas_prepare(orc_run* r) {
  if (r->gas_ready) return 0;
  if (r->n_gmaps != r->n_maps || r->kla) return -4;
  const uint32_t n = r->n_c;
  const size_t nsc = (size_t)r->n_s * n;
  if (!r->Cg) {
    r->Cg = (double*)calloc(nsc, sizeof(double));
    r->Cg2 = (double*)calloc(nsc, sizeof(double));
    r->mg = (double*)calloc(nsc, sizeof(double));
    r->Sg_total = (double*)calloc(nsc, sizeof(double));
    r->sink_g = (double*)calloc(n, sizeof(double));
    for (int s = 0; s < r->n_s; ++s)
      for (uint32_t j = 0; j < n; ++j) {
        const size_t o = (size_t)s * n + j;
        r->Cg[o] = r->Cg0[o];
        r->mg[o] = r->Cg0[o] * r->gmaps[0].V[j];
      }
  }
  for (uint32_t k = 0; k < r->n_gmaps; ++k) {
    free(r->g_kla[k]);
    r->g_kla[k] = (double*)malloc(sizeof(double) * nsc);
    gas_kla(r, &r->maps[k], &r->gmaps[k], r->g_eps[k], r->g_kla[k]);
  }
  r->gas_ready = 1;
  return 0;
}

/* One two-phase transport step: mtr from the old concentrations, the gas
   step with -mtr, the liquid step with +mtr, clearNegs on the liquid. */
static void transport_gl(orc_run* r, const flowmap* f, const flowmap* g, const double* kla) {
  const uint32_t n = r->n_c;
  const double dt = r->dt;
  const double tol = 1e-4 * 5e-3; /* clearNegs: O(dt) relative error x the species' magnitude */
  for (int s = 0; s < r->n_s; ++s) {
    const double* cl = r->C + (size_t)s * n;
    const double* cg = r->Cg + (size_t)s * n;
    for (uint32_t j = 0; j < n; ++j) {
      const size_t o = (size_t)s * n + j;
      const double mtr = (kla[o] * ((cg[j] * r->henry[s]) - cl[j])) * f->V[j];
      double ag = 0.0;
      for (uint64_t t = g->col_ptr[j]; t < g->col_ptr[j + 1]; ++t) ag += cg[g->src[t]] * g->val[t];
      const double dg = ((ag - cg[j] * r->sink_g[j]) + r->Sg_total[o]) - mtr;
      r->mg[o] = r->mg[o] + dt * dg;
      r->Cg2[o] = r->mg[o] * g->invV[j];
      double al = 0.0;
      for (uint64_t t = f->col_ptr[j]; t < f->col_ptr[j + 1]; ++t) al += cl[f->src[t]] * f->val[t];
      const double dl = ((al - cl[j] * r->sink[j]) + r->S_total[o]) + mtr;
      r->m[o] = r->m[o] + dt * dl;
      double c = r->m[o] * f->invV[j];
      if (c < 0.0 && fabs(c) < tol) c = 0.0;
      r->C2[o] = c;
    }
  }
}

int orc_get_gas_concentrations(const orc_run* r, double* C) {
  if (!r->two_phase) return -1;
  const size_t nsc = (size_t)r->n_s * r->n_c;
  memcpy(C, r->Cg ? r->Cg : r->Cg0, sizeof(double) * nsc);
  return 0;
}

int orc_get_gas_mass(const orc_run* r, double* m) {
  if (!r->two_phase) return -1;
  const size_t nsc = (size_t)r->n_s * r->n_c;
  if (r->mg) {
    memcpy(m, r->mg, sizeof(double) * nsc);
  } else {
    if (r->n_gmaps == 0) return -1;
    for (int s = 0; s < r->n_s; ++s)
      for (uint32_t j = 0; j < r->n_c; ++j) m[(size_t)s * r->n_c + j] = r->Cg0[(size_t)s * r->n_c + j] * r->gmaps[0].V[j];
  }
  return 0;
}

int orc_set_concentrations(orc_run* r, const double* C) {
  if (r->n_maps == 0) return -1;
  const uint32_t n = r->n_c;
  for (int s = 0; s < r->n_s; ++s)
    for (uint32_t j = 0; j < n; ++j) {
      const size_t o = (size_t)s * n + j;
      if (C[o] < 0.0) return -2;
      r->C[o] = C[o];
      r->m[o] = C[o] * r->maps[0].V[j]; /* set_mass: total_mass = C * diag(V) */
    }
  return 0;
}

int orc_set_compaction(orc_run* r, double ratio, uint64_t min_removal) {
  r->ratio = ratio;
  r->min_removal = min_removal;
  return 0;
}

int orc_set_gas_transfer(orc_run* r, const double* kla, const double* cstar) {
  const size_t nsc = (size_t)r->n_s * r->n_c;
  if (!kla || !cstar) return -1;
  if (!r->kla) r->kla = (double*)malloc(sizeof(double) * nsc);
  if (!r->cstar) r->cstar = (double*)malloc(sizeof(double) * nsc);
  if (!r->kla || !r->cstar) return -2;
  memcpy(r->kla, kla, sizeof(double) * nsc);
  memcpy(r->cstar, cstar, sizeof(double) * nsc);
  return 0;
}

int orc_set_division_cap(orc_run* r, uint32_t cap) {
  if (cap == 0) return -1;
  r->div_cap = cap;
  return 0;
}

static int alloc_particles(orc_run* r, uint64_t n, uint64_t cap) {
  if (cap < n || cap > 0xFFFFFFFFull) return -1;
  free(r->pos); free(r->status); free(r->state); free(r->age_hyd); free(r->age_div);
  r->cap = cap;
  r->n = n;
  r->pos = (uint32_t*)calloc(cap ? cap : 1, sizeof(uint32_t));
  r->status = (uint8_t*)calloc(cap ? cap : 1, 1);
  r->state = (double*)calloc((size_t)(cap ? cap : 1) * r->n_var, sizeof(double));
  r->age_hyd = (double*)calloc(cap ? cap : 1, sizeof(double));
  r->age_div = (double*)calloc(cap ? cap : 1, sizeof(double));
  r->inactive = 0;
  return 0;
}

/* unit.cpp:102-163 (position uniform over compartments, Model::init, sum of
 * mass) and unit.cpp:232-257 (w = X0 * V_total / sum mass). */
/* Stable reorder of the first n slots by compartment (slot order among equal
   positions), every per-particle field; the engine's mcst_run_sort_particles
   applies the same permutation. */
static int sort_by_compartment(orc_run* r) {
  const uint64_t n = r->n;
  if (n < 2 || r->n_c < 2) return 0;
  uint64_t* next = (uint64_t*)calloc((size_t)r->n_c + 1, sizeof(uint64_t));
  uint64_t* dst = (uint64_t*)malloc(sizeof(uint64_t) * n);
  double* tmp = (double*)malloc(sizeof(double) * n);
  if (!next || !dst || !tmp) {
    free(next), free(dst), free(tmp);
    return -2;
  }
  for (uint64_t p = 0; p < n; ++p) next[r->pos[p] + 1]++;
  for (uint32_t c = 0; c < r->n_c; ++c) next[c + 1] += next[c];
  for (uint64_t p = 0; p < n; ++p) dst[p] = next[r->pos[p]]++;
  uint32_t* u = (uint32_t*)tmp;
  for (uint64_t p = 0; p < n; ++p) u[dst[p]] = r->pos[p];
  memcpy(r->pos, u, sizeof(uint32_t) * n);
  uint8_t* b = (uint8_t*)tmp;
  for (uint64_t p = 0; p < n; ++p) b[dst[p]] = r->status[p];
  memcpy(r->status, b, n);
  for (int v = 0; v < r->n_var + 2; ++v) {
    double* x = v < r->n_var ? r->state + (size_t)v * r->cap : (v == r->n_var ? r->age_hyd : r->age_div);
    for (uint64_t p = 0; p < n; ++p) tmp[dst[p]] = x[p];
    memcpy(x, tmp, sizeof(double) * n);
  }
  free(next), free(dst), free(tmp);
  return 0;
}

int orc_sort_particles(orc_run* r) { return sort_by_compartment(r); }

int orc_set_sort_interval(orc_run* r, uint32_t steps) {
  r->sort_interval = steps;
  return 0;
}

/* Container growth (merge_buffer / _resize, particles_container.hpp:575-643,
 * allocation factor 2.5), checked on fixed steps: before every
 * grow_every-th step, a run with fewer free slots than particles gets
 * capacity ceil(n * grow_factor).  Slots keep their particles. */
int orc_set_growth(orc_run* r, uint32_t every, double factor) {
  if (every && !(factor > 1.0 && factor <= 16.0)) return -1;
  r->grow_every = every;
  r->grow_factor = factor;
  return 0;
}

static int grow_capacity(orc_run* r, uint64_t cap) {
  if (cap > 0xFFFFFFFFull) return -1;
  const uint64_t n = r->n, old = r->cap;
  uint32_t* pos = (uint32_t*)calloc(cap, sizeof(uint32_t));
  uint8_t* st = (uint8_t*)calloc(cap, 1);
  double* x = (double*)calloc((size_t)cap * r->n_var, sizeof(double));
  double* ah = (double*)calloc(cap, sizeof(double));
  double* ad = (double*)calloc(cap, sizeof(double));
  if (!pos || !st || !x || !ah || !ad) {
    free(pos), free(st), free(x), free(ah), free(ad);
    return -2;
  }
  memcpy(pos, r->pos, n * sizeof(uint32_t));
  memcpy(st, r->status, n);
  memcpy(ah, r->age_hyd, n * sizeof(double));
  memcpy(ad, r->age_div, n * sizeof(double));
  for (int v = 0; v < r->n_var; ++v) memcpy(x + (size_t)v * cap, r->state + (size_t)v * old, n * sizeof(double));
  free(r->pos), free(r->status), free(r->state), free(r->age_hyd), free(r->age_div);
  r->pos = pos, r->status = st, r->state = x, r->age_hyd = ah, r->age_div = ad;
  r->cap = cap;
  return 0;
}

int orc_init_particles(orc_run* r, uint64_t n, uint64_t capacity, double x0) {
  if (r->n_maps == 0) return -1;
  if (alloc_particles(r, n, capacity)) return -2;
  double total_mass = 0.0;
  double x[MAX_VAR];
  for (uint64_t p = 0; p < n; ++p) {
    uint32_t w[4];
    draw4(r->seed, (uint32_t)p, 0u, TAG_INIT, 1u, w);
    r->pos[p] = (uint32_t)(((uint64_t)w[0] * r->n_c) >> 32);
    model_init(r->model, r->params, x, r->seed, (uint32_t)p);
    for (int v = 0; v < r->n_var; ++v) r->state[(size_t)v * r->cap + p] = x[v];
    total_mass += model_mass(r->model, r->params, x);
  }
  double vt = 0.0;
  for (uint32_t j = 0; j < r->n_c; ++j) vt += r->maps[0].V[j];
  r->weight = n ? (x0 * vt) / total_mass : 0.0;
  /* the drawn particles are stored in compartment order, draw order among
     equal positions (the weight above is summed in draw order) */
  return sort_by_compartment(r);
}

int orc_set_particles(orc_run* r, uint64_t n, uint64_t capacity, const uint32_t* pos, const uint8_t* status,
                      const double* state, const double* age_hyd, const double* age_div, double weight) {
  if (alloc_particles(r, n, capacity)) return -2;
  uint64_t inactive = 0;
  for (uint64_t p = 0; p < n; ++p) {
    if (pos[p] >= r->n_c) return -3;
    r->pos[p] = pos[p];
    r->status[p] = status[p];
    inactive += status[p] != ORC_IDLE;
    r->age_hyd[p] = age_hyd[p];
    r->age_div[p] = age_div[p];
    for (int v = 0; v < r->n_var; ++v) r->state[(size_t)v * capacity + p] = state[(size_t)v * n + p];
  }
  r->inactive = inactive;
  r->weight = weight;
  return 0;
}

/* feed_descriptor.cpp:18-28 */
double orc_feed_flow(int type, double f0, double param, double t) {
  switch (type) {
    case ORC_FEED_LINEAR:
      return f0 + t * param;
    case ORC_FEED_EXPONENTIAL:
      return f0 + exp(param * t);
    default:
      return f0;
  }
}

static double feed_flow(const feed_t* f, double t) { return orc_feed_flow(f->type, f->f0, f->param, t); }

typedef struct {
  uint32_t exits, deaths, moves, processed;
  uint32_t n_req;         /* division requests of the tile, in slot order */
  uint32_t* req_slot;
  uint32_t* req_pos;      /* the mother's compartment when she asked (before her move) */
  dvec exit_ages;         /* probes of this tile (threaded path), appended in tile order */
} tile_out;

/* One 1024-slot tile of the particle cycle, in the reference's stage order:
 * the model pass over the Idle slots (model_kernel.hpp:163-268), their
 * contributions (contribution_kernel.hpp:144-186), the move pass over EVERY
 * slot below n whatever its status (move_kernel.hpp:209-273,392-437: Exit
 * and Dead slots waiting for removal keep moving and count Move events),
 * then the leave pass over the Idle slots (move_kernel.hpp:347-359,586-648).
 * Division requests are only listed here: which are accepted depends on the
 * whole run's requests (see accept_divisions). */
static void particle_tile(orc_run* r, const flowmap* f, uint64_t step, uint32_t tile, uint32_t stages,
                          int enable_move, int enable_leave, double* S, tile_out* o, dvec* exit_ages) {
  const uint64_t p0 = (uint64_t)tile * TILE;
  const uint64_t p1 = (p0 + TILE < r->n) ? p0 + TILE : r->n;
  const uint32_t n_c = r->n_c;
  const int nv = r->n_var, ns = r->n_s;
  const double dt = r->dt;
  double x[MAX_VAR], c[MAX_SPECIES], contrib[MAX_SPECIES];
  for (uint64_t p = p0; p < p1; ++p) {
    uint32_t pos = r->pos[p];
    int idle = r->status[p] == ORC_IDLE;
    if (idle) {
      o->processed++;
      for (int v = 0; v < nv; ++v) x[v] = r->state[(size_t)v * r->cap + p];
      for (int s = 0; s < ns; ++s) contrib[s] = 0.0;
      int st = ORC_IDLE;
      if (stages & ORC_STAGE_MODEL) {
        r->age_div[p] += dt;
        for (int s = 0; s < ns; ++s) c[s] = r->C[(size_t)s * n_c + pos];
        st = model_update(r->model, r->params, dt, x, c, contrib);
      }
      for (int v = 0; v < nv; ++v) r->state[(size_t)v * r->cap + p] = x[v];
      if (st == ORC_DEAD) { /* a dead cell stops contributing and leaving */
        r->status[p] = ORC_DEAD;
        o->deaths++;
        idle = 0;
      } else {
        if (st == ORC_DIVISION) {
          o->req_slot[o->n_req] = (uint32_t)p;
          o->req_pos[o->n_req] = pos;
          o->n_req++;
        }
        /* contribution: S(j,pos) += w * contrib(j) */
        if (stages & ORC_STAGE_CONTRIB)
          for (int s = 0; s < ns; ++s) S[(size_t)s * n_c + pos] += r->weight * contrib[s];
      }
    }
    uint32_t w[4];
    draw4(r->seed, (uint32_t)p, (uint32_t)step, TAG_TRANSPORT, 0u, w);
    /* move, every slot: the linear test of probability_leaving.hpp:33-46,
     * then the first CDF column >= u (move_kernel.hpp:61-103) */
    if (enable_move && (stages & ORC_STAGE_MOVE)) {
      const double u1 = orc_u01(w[0]);
      if (f->p_move[pos] > u1) {
        const double u2 = orc_u01(w[1]);
        const double* cp = f->cdf + (size_t)pos * f->m;
        int left = 0, right = (int)f->m - 1;
        while (left < right) {
          const int mid = (left + right) >> 1;
          if (u2 > cp[mid])
            left = mid + 1;
          else
            right = mid;
        }
        pos = f->nbr[(size_t)pos * f->m + left];
        r->pos[p] = pos;
        o->moves++;
      }
    }
    /* leave, Idle slots: move_kernel.hpp:347-359, 586-648 */
    if (idle && enable_leave && (stages & ORC_STAGE_LEAVE)) {
      r->age_hyd[p] += dt;
      const double q = r->q_exit[pos];
      if (q < 1.0 && orc_u01(w[2]) > q) {
        if (exit_ages) dvec_push(exit_ages, r->age_hyd[p]);
        r->age_hyd[p] = 0.0;
        r->status[p] = ORC_EXIT;
        o->exits++;
      }
    }
  }
}

/* The division buffer (particles_container.hpp:559-573, merge_buffer
 * :575-599), made deterministic: the step's requests are taken in slot
 * order and the first min(cap - n, div_cap) are accepted (the reference
 * refuses a division only when its buffer is full); the k-th accepted
 * mother divides (Model::division) and her daughter takes slot n + k, at
 * the mother's compartment when she asked, with both ages 0.  The others
 * are refused (Overflow) and stay as they are, to ask again next step.
 * Returns the births. */
static uint64_t accept_divisions(orc_run* r, uint64_t step, const tile_out* outs, uint32_t n_tiles) {
  const uint64_t n = r->n;
  uint64_t budget = r->cap - n;
  if (budget > r->div_cap) budget = r->div_cap;
  const int nv = r->n_var;
  double x[MAX_VAR], xd[MAX_VAR];
  uint64_t k = 0, requests = 0;
  for (uint32_t tl = 0; tl < n_tiles; ++tl) {
    requests += outs[tl].n_req;
    for (uint32_t q = 0; q < outs[tl].n_req && k < budget; ++q, ++k) {
      const uint64_t p = outs[tl].req_slot[q], d = n + k;
      dvec_push(&r->div_ages, r->age_div[p]);
      for (int v = 0; v < nv; ++v) x[v] = r->state[(size_t)v * r->cap + p];
      model_division(r->model, r->params, x, xd, r->seed, (uint32_t)p, (uint32_t)step);
      for (int v = 0; v < nv; ++v) {
        r->state[(size_t)v * r->cap + p] = x[v];
        r->state[(size_t)v * r->cap + d] = xd[v];
      }
      r->age_div[p] = 0.0;
      r->pos[d] = outs[tl].req_pos[q];
      r->status[d] = ORC_IDLE;
      r->age_hyd[d] = 0.0;
      r->age_div[d] = 0.0;
    }
  }
  r->counters[ORC_C_OVERFLOW] += requests - k;
  return k;
}

static void copy_particle(orc_run* r, uint64_t dst, uint64_t src) {
  r->pos[dst] = r->pos[src];
  r->status[dst] = r->status[src];
  r->age_hyd[dst] = r->age_hyd[src];
  r->age_div[dst] = r->age_div[src];
  for (int v = 0; v < r->n_var; ++v) r->state[(size_t)v * r->cap + dst] = r->state[(size_t)v * r->cap + src];
}

/* particles_container.hpp:539-557 (threshold) and 735-796 (removal) with a
 * deterministic pairing of gaps and tail fillers. */
static void remove_inactive(orc_run* r, uint64_t k) {
  if (k == 0) return;
  if (k >= r->n) {
    r->n = 0;
    r->inactive = 0;
    return;
  }
  const uint64_t nn = r->n - k;
  uint64_t g = 0, t = r->n;
  for (;;) {
    while (g < nn && r->status[g] == ORC_IDLE) ++g;
    if (g >= nn) break;
    do {
      --t;
    } while (r->status[t] != ORC_IDLE);
    copy_particle(r, g, t);
    ++g;
  }
  r->n = nn;
  r->inactive = 0;
}

int orc_step(orc_run* r, uint32_t n_steps, int n_threads) {
  return orc_step_stages(r, n_steps, n_threads, ORC_STAGE_ALL);
}

int orc_step_stages(orc_run* r, uint32_t n_steps, int n_threads, uint32_t stages) {
  if (r->n_maps == 0) return -1;
  const uint32_t n_c = r->n_c;
  const int ns = r->n_s;
  const size_t nsc = (size_t)ns * n_c;
  int n_leaving = 0;
  for (uint32_t i = 0; i < r->n_feeds; ++i) n_leaving += r->feeds[i].out >= 0;
  const int enable_move = n_c > 1;
  const int enable_leave = n_leaving != 0;
  if (n_threads < 1) n_threads = 1;
#ifdef _OPENMP
  omp_set_num_threads(n_threads);
#else
  n_threads = 1;
#endif
  for (uint32_t it = 0; it < n_steps; ++it) {
    const uint64_t step = r->counters[ORC_C_STEPS];
    if (r->grow_every && step % r->grow_every == 0 && r->cap - r->n < r->n) {
      const uint64_t cap = (uint64_t)ceil((double)r->n * r->grow_factor);
      if (cap > r->cap && grow_capacity(r, cap)) return -2;
    }
    const flowmap* f = &r->maps[(step / r->steps_per_map) % r->n_maps];
    const double t = (double)step * r->dt;
    /* update_feed (simulation.model.cpp:72-122).  Feed terms of one
     * (species, compartment) are summed in feed order, then added to the
     * particle sources: S_total = S + (f1 + f2 + ...). */
    if (r->two_phase && gas_prepare(r)) return -4;
    memset(r->S_total, 0, sizeof(double) * nsc);
    memset(r->sink, 0, sizeof(double) * n_c);
    if (r->two_phase) {
      memset(r->Sg_total, 0, sizeof(double) * nsc);
      memset(r->sink_g, 0, sizeof(double) * n_c);
    }
    for (uint32_t j = 0; j < n_c; ++j) r->q_exit[j] = 2.0;
    unsigned char* has_exit = (unsigned char*)calloc(n_c, 1);
    for (uint32_t i = 0; i < r->n_feeds; ++i) {
      const feed_t* fd = &r->feeds[i];
      const double q = feed_flow(fd, t);
      if (fd->gas) { /* the gas scalar's feed and sink; no particle leaves with it */
        if (!r->two_phase) continue;
        for (uint32_t v = 0; v < fd->nv; ++v) r->Sg_total[(size_t)fd->sp[v] * n_c + fd->in] += q * fd->conc[v];
        if (fd->out >= 0) r->sink_g[(uint32_t)fd->out] += q;
        continue;
      }
      for (uint32_t v = 0; v < fd->nv; ++v) r->S_total[(size_t)fd->sp[v] * n_c + fd->in] += q * fd->conc[v];
      if (fd->out >= 0) {
        const uint32_t o = (uint32_t)fd->out;
        r->sink[o] += q;
        if (!has_exit[o]) {
          has_exit[o] = 1;
          r->q_exit[o] = exp(-(r->dt * q) / f->V[o]);
        }
      }
    }
    free(has_exit);
    for (size_t o = 0; o < nsc; ++o) r->S_total[o] = r->S[o] + r->S_total[o];
    /* ode_step then clear contributions */
    if (stages & ORC_STAGE_TRANSPORT) {
      if (r->two_phase) {
        const uint32_t k = (uint32_t)((step / r->steps_per_map) % r->n_maps);
        transport_gl(r, f, &r->gmaps[k], r->g_kla[k]);
        double* tg = r->Cg;
        r->Cg = r->Cg2;
        r->Cg2 = tg;
      } else {
        transport(f, (uint32_t)ns, r->sink, r->S_total, r->dt, r->C, r->m, r->C2, r->kla, r->cstar);
      }
      double* tmp = r->C;
      r->C = r->C2;
      r->C2 = tmp;
    }
    memset(r->S, 0, sizeof(double) * nsc);

    /* particle cycle */
    const uint64_t N = r->n;
    const uint32_t n_tiles = (uint32_t)((N + TILE - 1) / TILE);
    tile_out* outs = (tile_out*)calloc(n_tiles ? n_tiles : 1, sizeof(tile_out));
    for (uint32_t tl = 0; tl < n_tiles; ++tl) {
      outs[tl].req_slot = (uint32_t*)malloc(sizeof(uint32_t) * TILE);
      outs[tl].req_pos = (uint32_t*)malloc(sizeof(uint32_t) * TILE);
    }
    if (n_threads == 1) {
      for (uint32_t tl = 0; tl < n_tiles; ++tl)
        particle_tile(r, f, step, tl, stages, enable_move, enable_leave, r->S, &outs[tl], &r->exit_ages);
    } else {
      double* Sp = (double*)calloc(nsc * (size_t)n_threads, sizeof(double));
#pragma omp parallel num_threads(n_threads)
      {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* S_loc = Sp + nsc * (size_t)tid;
#pragma omp for schedule(static)
        for (int64_t tl = 0; tl < (int64_t)n_tiles; ++tl)
          particle_tile(r, f, step, (uint32_t)tl, stages, enable_move, enable_leave, S_loc, &outs[tl],
                        &outs[tl].exit_ages);
      }
      for (int th = 0; th < n_threads; ++th)
        for (size_t o = 0; o < nsc; ++o) r->S[o] += Sp[nsc * (size_t)th + o];
      free(Sp);
      for (uint32_t tl = 0; tl < n_tiles; ++tl) {
        for (uint64_t k = 0; k < outs[tl].exit_ages.n; ++k) dvec_push(&r->exit_ages, outs[tl].exit_ages.v[k]);
        free(outs[tl].exit_ages.v);
      }
    }
    uint64_t exits = 0, deaths = 0;
    for (uint32_t tl = 0; tl < n_tiles; ++tl) {
      exits += outs[tl].exits;
      deaths += outs[tl].deaths;
      r->counters[ORC_C_MOVES] += outs[tl].moves;
      r->counters[ORC_C_PROCESSED] += outs[tl].processed;
    }
    const uint64_t births = accept_divisions(r, step, outs, n_tiles);
    for (uint32_t tl = 0; tl < n_tiles; ++tl) {
      free(outs[tl].req_slot);
      free(outs[tl].req_pos);
    }
    free(outs);
    r->counters[ORC_C_BIRTHS] += births;
    r->counters[ORC_C_EXITS] += exits;
    r->counters[ORC_C_DEATHS] += deaths;
    /* post_cycle: the newborns join the container, then
     * update_and_remove_inactive with the threshold of the pass's count */
    r->n = N + births;
    r->inactive += exits + deaths;
    const uint64_t thr_ratio = (uint64_t)((double)N * r->ratio);
    const uint64_t thr = r->min_removal > thr_ratio ? r->min_removal : thr_ratio;
    if ((stages & ORC_STAGE_REMOVE) && r->inactive > thr) {
      remove_inactive(r, r->inactive);
      r->counters[ORC_C_COMPACTIONS]++;
    }
    r->counters[ORC_C_STEPS]++;
    if (r->sort_interval && r->counters[ORC_C_STEPS] % r->sort_interval == 0 && sort_by_compartment(r)) return -2;
  }
  return 0;
}

uint64_t orc_n(const orc_run* r) { return r->n; }
double orc_weight(const orc_run* r) { return r->weight; }

int orc_get_particles(const orc_run* r, uint32_t* pos, uint8_t* status, double* state, double* age_hyd,
                      double* age_div) {
  const uint64_t n = r->n;
  if (pos) memcpy(pos, r->pos, sizeof(uint32_t) * n);
  if (status) memcpy(status, r->status, n);
  if (age_hyd) memcpy(age_hyd, r->age_hyd, sizeof(double) * n);
  if (age_div) memcpy(age_div, r->age_div, sizeof(double) * n);
  if (state)
    for (int v = 0; v < r->n_var; ++v) memcpy(state + (size_t)v * n, r->state + (size_t)v * r->cap, sizeof(double) * n);
  return 0;
}

int orc_get_concentrations(const orc_run* r, double* C) {
  memcpy(C, r->C, sizeof(double) * (size_t)r->n_s * r->n_c);
  return 0;
}
int orc_get_mass(const orc_run* r, double* m) {
  memcpy(m, r->m, sizeof(double) * (size_t)r->n_s * r->n_c);
  return 0;
}
int orc_get_contributions(const orc_run* r, double* S) {
  memcpy(S, r->S, sizeof(double) * (size_t)r->n_s * r->n_c);
  return 0;
}
int orc_get_counters(const orc_run* r, uint64_t* c) {
  memcpy(c, r->counters, sizeof(uint64_t) * ORC_C_COUNT);
  c[ORC_C_N] = r->n;
  c[ORC_C_INACTIVE] = r->inactive;
  return 0;
}
int orc_get_repartition(const orc_run* r, uint64_t* hist) {
  memset(hist, 0, sizeof(uint64_t) * r->n_c);
  for (uint64_t p = 0; p < r->n; ++p)
    if (r->status[p] == ORC_IDLE) hist[r->pos[p]]++;
  return 0;
}
/* post_process.hpp:54-98: per Idle particle its state and Model::mass
   (values[n_var+1][n], zero for inactive slots), summed per compartment in
   slot order (spatial[n_var+1][n_c]). */
int orc_get_properties(const orc_run* r, double* values, double* spatial) {
  const uint64_t n = r->n;
  const int nv = r->n_var;
  memset(spatial, 0, sizeof(double) * (size_t)(nv + 1) * r->n_c);
  double x[MAX_VAR];
  for (uint64_t p = 0; p < n; ++p) {
    const int idle = r->status[p] == ORC_IDLE;
    for (int v = 0; v < nv; ++v) x[v] = r->state[(size_t)v * r->cap + p];
    const double m = model_mass(r->model, r->params, x);
    for (int v = 0; v <= nv; ++v) {
      const double val = idle ? (v < nv ? x[v] : m) : 0.0;
      values[(size_t)v * n + p] = val;
      if (idle) spatial[(size_t)v * r->n_c + r->pos[p]] += val;
    }
  }
  return 0;
}
uint64_t orc_get_exit_ages(const orc_run* r, double* out, uint64_t max) {
  const uint64_t n = r->exit_ages.n < max ? r->exit_ages.n : max;
  if (out && n) memcpy(out, r->exit_ages.v, sizeof(double) * n);
  return r->exit_ages.n;
}
uint64_t orc_get_division_ages(const orc_run* r, double* out, uint64_t max) {
  const uint64_t n = r->div_ages.n < max ? r->div_ages.n : max;
  if (out && n) memcpy(out, r->div_ages.v, sizeof(double) * n);
  return r->div_ages.n;
}
