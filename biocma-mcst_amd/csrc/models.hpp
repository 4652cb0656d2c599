// models.hpp -- particle models and random draws, compiled for host and gfx950.
//
// A model is a struct of static functions (the contract of the reference's
// ModelType concept, apps/libs/mc/public/mc/traits.hpp:66-153):
//   NV, NS, NP           state variables, liquid species, parameters
//   init(p, x, seed, slot)                 host, initial state      (Model::init)
//   record(p, c, r)                        device+host, per compartment: writes r[0..NR)
//   NR (optional)                          record terms, NS when absent
//   update(p, dt, x, r, contrib) -> Status device+host               (Model::update)
//   division(p, x, xd, key)                device+host              (Model::division)
//   mass(p, x)                             device+host              (Model::mass)
//   CAN_DIE                                whether update() may return DEAD
//   NL                                     particle pass: slots per lane of a sub-chunk (4, or 2
//                                          for large states, which halves the registers)
// State is FP64; contributions (kg/s per particle) are returned in
// registers instead of a per-particle contribs row, because the step kernel
// scatters them immediately.  record() maps a compartment's concentrations
// c[0..NS) to the NR terms r[] that update() reads: the parts of the
// reference's update that depend on the liquid alone (e.g. the Monod factor
// mu_max*S/(Ks+S)), evaluated once per compartment and step by the transport
// kernel instead of once per particle, with the same operations in the same
// order, so the results are identical.  All code here is built with
// -ffp-contract=off, so each expression rounds exactly as written.
#pragma once

#include <math.h>
#include <stdint.h>

#include "mcst_common.hpp"

#define MCST_HD __host__ __device__ __forceinline__

namespace mcst {

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3",
// SC'11.  Counter-based, so a particle's draws depend only on (seed, slot,
// step, purpose) and not on which lane or thread produced them.
struct U4 {
  uint32_t x, y, z, w;
};

MCST_HD U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return U4{c0, c1, c2, c3};
}

MCST_HD U4 draw4(uint64_t seed, uint32_t slot, uint32_t step, uint32_t tag, uint32_t blk) {
  return philox4x32_10(slot, step, (tag << 16) | blk, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// (x + 1/2) * 2^-32, exact, in (0, 1)
MCST_HD double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386962890625e-10; }

// 52-bit uniform in (0, 1): ((a>>6)<<26 | b>>6) + 1/2, over 2^52
MCST_HD double u01_52(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 6) << 26) | (uint64_t)(b >> 6);
  return ((double)k + 0.5) * 2.220446049250313080847263336181640625e-16;
}

// ------------------------------------------------------ normal distribution
// Inverse CDF: Wichura, Algorithm AS 241 (PPND16), Appl. Statist. 37 (1988).
MCST_HD double norm_inv(double p) {
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

MCST_HD double norm_cdf(double x) { return 0.5 * erfc(-x * 0.70710678118654752440); }

// Truncated normal by inverse transform (the distribution of the reference's
// TruncatedNormal, apps/libs/mc/public/mc/prng/prng_extension.hpp:353-440).
MCST_HD double trunc_normal(double mu, double sd, double lo, double hi, double u) {
  const double pa = isinf(lo) ? 0.0 : norm_cdf((lo - mu) / sd);
  const double pb = isinf(hi) ? 1.0 : norm_cdf((hi - mu) / sd);
  double p = pa + u * (pb - pa);
  if (p <= 0.0) p = 1e-300;
  if (p >= 1.0) p = 1.0 - 1e-16;
  double x = mu + sd * norm_inv(p);
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

MCST_HD double lin_density(double d_m) { return 1000.0 * 3.14159265358979323846 * d_m * d_m / 4.0; }

struct RngKey {
  uint64_t seed;
  uint32_t slot;
  uint32_t step;
};

// Record row: the model's NR terms, the move probability and the exit
// threshold, padded to 4, 8 or 16 doubles.
constexpr int rec_stride(int nr) { return (nr + 2) <= 4 ? 4 : ((nr + 2) <= 8 ? 8 : 16); }

// NR of a model: its own NR member, else NS.
template <class M, class = void>
struct RecordTerms {
  static constexpr int value = M::NS;
};
template <class M>
struct RecordTerms<M, decltype((void)M::NR)> {
  static constexpr int value = M::NR;
};
template <class M>
constexpr int record_terms() {
  return RecordTerms<M>::value;
}

// ---------------------------------------------------------------- monod_mass
// Mass-based Monod uptake-growth with division at `ratio` x birth mass.
// x = {m, m_birth}; p = {mu_max, Ks, Y, ratio, m0_mean, m0_cv}.
struct MonodMass {
  static constexpr int ID = 0, NV = 2, NS = 1, NP = 6;
  static constexpr bool CAN_DIE = false;  // update() never returns DEAD
  static constexpr int NL = 4;
  static constexpr const char* NAME = "monod_mass";
  static void defaults(double* p) {
    p[0] = 0.5 / 3600.0; p[1] = 0.01; p[2] = 0.5; p[3] = 2.0; p[4] = 1e-15; p[5] = 0.1;
  }
  static void init(const double* p, double* x, uint64_t seed, uint32_t slot) {
    const U4 w = draw4(seed, slot, 0u, TAG_INIT, 0u);
    const double m = trunc_normal(p[4], p[5] * p[4], 0.0, INFINITY, u01_52(w.x, w.y));
    x[0] = m;
    x[1] = m / (1.0 + u01(w.z));
  }
  MCST_HD static double mass(const double*, const double* x) { return x[0]; }
  MCST_HD static void record(const double* p, const double* c, double* r) {
    const double s = c[0] > 0.0 ? c[0] : 0.0;
    r[0] = p[0] * s / (p[1] + s);  // mu
  }
  MCST_HD static int update(const double* p, double dt, double* x, const double* r, double* contrib) {
    const double mu = r[0];
    const double g = mu * x[0];
    contrib[0] = -(g / p[2]);
    x[0] = x[0] + dt * g;
    return (x[0] >= p[3] * x[1]) ? DIVISION : IDLE;
  }
  MCST_HD static void division(const double*, double* x, double* xd, const RngKey&) {
    const double h = x[0] * 0.5;
    x[0] = h;
    x[1] = h;
    xd[0] = h;
    xd[1] = h;
  }
};

// ---------------------------------------------------------------- dual_monod
// Dual-substrate (glucose S, oxygen O) uptake with a lagged uptake rate and an
// overflow storage pool.  x = {m, q, X}; the division age is the container's.
// p = {qs_max, Ks, Ko, Yxs, Yos, tau, k_st, m_div, m0_mean, m0_sd}.
struct DualMonod {
  static constexpr int ID = 1, NV = 3, NS = 2, NP = 10;
  static constexpr bool CAN_DIE = false;  // update() never returns DEAD
  static constexpr int NL = 4;
  static constexpr const char* NAME = "dual_monod";
  static void defaults(double* p) {
    p[0] = 2.8e-4; p[1] = 0.01; p[2] = 1e-4; p[3] = 0.5; p[4] = 1.0; p[5] = 100.0; p[6] = 1e-4;
    p[7] = 2e-15; p[8] = 1.4e-15; p[9] = 0.2e-15;
  }
  static void init(const double* p, double* x, uint64_t seed, uint32_t slot) {
    const U4 w = draw4(seed, slot, 0u, TAG_INIT, 0u);
    x[0] = trunc_normal(p[8], p[9], 0.5 * p[7], p[7], u01_52(w.x, w.y));
    x[1] = 0.5 * p[0];
    x[2] = 0.0;
  }
  MCST_HD static double mass(const double*, const double* x) { return x[0]; }
  MCST_HD static void record(const double* p, const double* c, double* r) {
    const double S = c[0] > 0.0 ? c[0] : 0.0;
    const double O = c[1] > 0.0 ? c[1] : 0.0;
    r[0] = p[0] * S / (p[1] + S);  // qt: target uptake rate
    r[1] = O / (p[2] + O);         // fo: oxygen saturation
  }
  MCST_HD static int update(const double* p, double dt, double* x, const double* r, double* contrib) {
    const double qt = r[0];
    const double U = x[1] * x[0];
    const double fo = r[1];
    const double G = p[3] * U * fo;
    const double Fst = U * (1.0 - fo);
    const double R = p[6] * x[2] * fo;
    const double Gt = G + p[3] * R;
    contrib[0] = -U;
    contrib[1] = -(p[4] * Gt);
    x[0] = x[0] + dt * Gt;
    x[2] = x[2] + dt * (Fst - R);
    x[1] = x[1] + (dt / p[5]) * (qt - x[1]);
    return (x[0] >= p[7]) ? DIVISION : IDLE;
  }
  MCST_HD static void division(const double*, double* x, double* xd, const RngKey&) {
    x[0] = x[0] * 0.5;
    x[2] = x[2] * 0.5;
    xd[0] = x[0];
    xd[1] = x[1];
    xd[2] = x[2];
  }
};

// ---------------------------------------------------------------- monod
// The reference's Monod model (apps/libs/models/public/models/monod.hpp:26-215)
// in FP64.  x = {l, l_max, mu_p, mu_eff, dl/ln2, phi_s}; the uptake phi_s is
// also returned as the particle's contribution.
// p = {mu_max, k_s, y_s_x, l_max, d_m, tau_meta}.
struct Monod {
  static constexpr int ID = 2, NV = 6, NS = 1, NP = 6;
  static constexpr bool CAN_DIE = false;  // update() never returns DEAD
  static constexpr int NL = 4;
  static constexpr const char* NAME = "monod";
  static void defaults(double* p) {
    p[0] = 0.77 / 3600.0; p[1] = 1e-3; p[2] = 2.0; p[3] = 2e-6; p[4] = 0.6e-6; p[5] = 3600.0 / 0.77;
  }
  static void init(const double* p, double* x, uint64_t seed, uint32_t slot) {  // monod.hpp:80-113
    const U4 w = draw4(seed, slot, 0u, TAG_INIT, 0u);
    const double lm = p[3];
    x[0] = trunc_normal(lm * 0.75, lm * 0.75 / 4.0, lm / 2.0, lm, u01_52(w.x, w.y));
    x[1] = lm;
    x[2] = p[0];
    x[3] = 0.0;
    x[4] = (lm / 2.0) / 0.69314718055994530942;
    x[5] = 0.0;
  }
  MCST_HD static double mass(const double* p, const double* x) { return x[0] * lin_density(p[4]); }
  MCST_HD static void record(const double* p, const double* c, double* r) {
    const double s = c[0] > 0.0 ? c[0] : 0.0;  // monod.hpp:130-160
    r[0] = p[0] * s / (p[1] + s);  // mu
  }
  MCST_HD static int update(const double* p, double dt, double* x, const double* r, double* contrib) {
    const double mu = r[0];
    const double mu_eff = x[2] < mu ? x[2] : mu;
    x[0] = x[0] + dt * (mu_eff * x[4]);
    x[2] = x[2] + dt * (1.0 / p[5]) * (mu - x[2]);
    x[3] = mu_eff;
    x[5] = -mu_eff * p[2] * (x[0] * lin_density(p[4]));
    contrib[0] = x[5];
    return (x[0] >= x[1]) ? DIVISION : IDLE;
  }
  MCST_HD static void division(const double* p, double* x, double* xd, const RngKey&) {  // monod.hpp:162-193
    const double h = x[0] / 2.0;
    x[0] = h;
    xd[0] = h;
    xd[1] = p[3];
    xd[2] = x[2];
    xd[3] = 0.0;
    xd[4] = x[4];
    xd[5] = 0.0;
  }
};

// ---------------------------------------------------------------- fixed_length
// apps/libs/models/public/models/fixed_length.hpp:108-161 in FP64.
// x = {l, l_max}; p = {l_dot_max, l_max, k, d_m, lambda}.
struct FixedLength {
  static constexpr int ID = 3, NV = 2, NS = 1, NP = 5;
  static constexpr bool CAN_DIE = false;  // update() never returns DEAD
  static constexpr int NL = 4;
  static constexpr const char* NAME = "fixed_length";
  static void defaults(double* p) {
    p[0] = 2e-6 / 3600.0; p[1] = 2e-6; p[2] = 1e-3; p[3] = 0.6e-6; p[4] = 0.69314718055994530942 / 1e-6;
  }
  // Initial length: the density lambda*exp(-lambda*l) truncated to
  // [l_max/2, l_max], the target the reference samples by Metropolis
  // (models/src/config_loader.cpp:13-36; lambda = ln2/1e-6 unless VLAMBDA
  // sets it), drawn here exactly by inverting its CDF:
  //   l = a - log1p(-u*(1 - exp(-lambda*(b - a))))/lambda.
  // lambda = 0 gives the uniform density, the limit lambda -> 0.
  static void init(const double* p, double* x, uint64_t seed, uint32_t slot) {
    const U4 w = draw4(seed, slot, 0u, TAG_INIT, 0u);
    const double a = p[1] / 2.0, b = p[1], lam = p[4];
    const double u = u01_52(w.x, w.y);
    double l = lam != 0.0 ? a - log1p(-u * -expm1(-lam * (b - a))) / lam : a + u * (b - a);
    x[0] = l < a ? a : (l > b ? b : l);
    x[1] = p[1];
  }
  MCST_HD static double mass(const double* p, const double* x) { return x[0] * lin_density(p[3]); }
  MCST_HD static void record(const double* p, const double* c, double* r) {
    const double s = c[0];
    r[0] = s / (p[2] + s);  // g: uptake saturation
  }
  MCST_HD static int update(const double* p, double dt, double* x, const double* r, double* contrib) {
    const double ld = lin_density(p[3]);
    const double phi_s_max = (ld * p[0]) / 0.5;
    const double g = r[0];
    const double phi = phi_s_max * g;
    const double ldot = p[0] * g;
    x[0] = x[0] + dt * ldot;
    contrib[0] = -phi;
    return (x[0] >= x[1]) ? DIVISION : IDLE;
  }
  MCST_HD static void division(const double* p, double* x, double* xd, const RngKey&) {
    const double h = x[0] / 2.0;
    x[0] = h;
    x[1] = p[1];
    xd[0] = h;
    xd[1] = p[1];
  }
};

// ---------------------------------------------------------------- simple_acetate
// apps/libs/models/public/models/simple_acetate.hpp:26-248 in FP64: glucose
// and acetate uptake with an overflow switch; log-normal redistribution of
// a_p and truncated-normal l_max at division.
// x = {length, l_max, a_p, a_max, a_e, a_e_s, a_e_a, phi_s, phi_a}.
struct SimpleAcetate {
  static constexpr int ID = 4, NV = 9, NS = 2, NP = 0;
  static constexpr bool CAN_DIE = false;  // update() never returns DEAD
  static constexpr int NL = 4;
  static constexpr const char* NAME = "simple_acetate";
  static constexpr double A_MAX = 2e-6 / 3600.0;
  static constexpr double L_MAX = 2e-6;
  static constexpr double L_MIN = L_MAX / 2.0;
  static constexpr double D_M = 0.6e-6;
  static constexpr double K_S = 1e-3, K_A = 1e-4, Y_S = 2.0, Y_A = 3.0;
  static void defaults(double*) {}
  static void init(const double*, double* x, uint64_t seed, uint32_t slot) {
    const U4 w = draw4(seed, slot, 0u, TAG_INIT, 0u);
    x[0] = trunc_normal(L_MAX * 0.75, L_MAX / 10.0, 0.7 * L_MIN, L_MAX * 1.3, u01_52(w.x, w.y));
    x[1] = trunc_normal(L_MAX * 0.75, L_MAX / 10.0, 0.7 * L_MIN, L_MAX * 1.3, u01_52(w.z, w.w));
    x[2] = A_MAX / 2.0;
    x[3] = A_MAX;
    for (int i = 4; i < 9; ++i) x[i] = 0.0;
  }
  MCST_HD static double mass(const double*, const double* x) { return x[0] * lin_density(D_M); }
  // D0 and D1 multiply the concentration before the saturation factor, so
  // no liquid-only term can be split off without changing their rounding:
  // the record is the concentrations themselves.
  MCST_HD static void record(const double*, const double* c, double* r) {
    r[0] = c[0];
    r[1] = c[1];
  }
  MCST_HD static int update(const double*, double dt, double* x, const double* c, double* contrib) {
    const double ld = lin_density(D_M);
    const double adm0 = x[3], adm1 = x[3] / 3.0;
    const double c0 = c[0], c1 = c[1];
    const double D0 = adm0 * c0 * (1.0 / (c0 + K_S));
    const double D1 = adm1 * c1 * (1.0 / (c1 + K_A));
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
    const double phi_s = -1.0 * D0 * ld * Y_S;
    const double phi_a = mask * (-U1 * ld * Y_A) + (1.0 - mask) * (pa * ld * Y_S / Y_A);
    x[7] = phi_s;
    x[8] = phi_a;
    contrib[0] = phi_s;
    contrib[1] = phi_a;
    return (x[0] >= x[1]) ? DIVISION : IDLE;
  }
  MCST_HD static void division(const double*, double* x, double* xd, const RngKey& key) {
    x[0] = x[0] / 2.0;
    for (int i = 0; i < 4; ++i) xd[i] = x[i];
    for (int i = 4; i < 9; ++i) xd[i] = 0.0;
    const U4 w = draw4(key.seed, key.slot, key.step, TAG_DIVISION, 0u);
    const U4 v = draw4(key.seed, key.slot, key.step, TAG_DIVISION, 1u);
    const double sigma = 0.2;
    const double average = log(x[4]) - sigma * sigma / 2.0;
    const double g1 = exp(average + sigma * norm_inv(u01_52(w.x, w.y)));
    const double g2 = exp(average + sigma * norm_inv(u01_52(w.z, w.w)));
    const double l1 = trunc_normal(L_MAX, L_MAX / 10.0, L_MAX * 0.7, 1.3 * L_MAX, u01_52(v.x, v.y));
    const double l2 = trunc_normal(L_MAX, L_MAX / 10.0, L_MAX * 0.7, 1.3 * L_MAX, u01_52(v.z, v.w));
    x[2] = g1;
    x[1] = l1;
    xd[2] = g2;
    xd[1] = l2;
  }
};

// ---------------------------------------------------------------- quad_monod
// Four-species uptake model of the 50 000-compartment lattice case: glucose
// S, oxygen O, acetate A and CO2.  Lagged glucose and acetate uptake rates;
// the glucose taken up is respired in proportion to the oxygen saturation and
// the rest overflows to acetate (a fraction Yas is excreted, the rest stored);
// acetate is taken up aerobically, repressed by glucose.
// x = {m, qs, qa, X, o2_rate, overflow, m_birth, generation};
// p = {qs_max, Ks, Ko, Ka, qa_max, Yxs, Yxa, Yos, Yoa, Yas, tau, Yco2, m_div, m0_mean, m0_sd}.
struct QuadMonod {
  static constexpr int ID = 5, NV = 8, NS = 4, NP = 15;
  static constexpr bool CAN_DIE = false;  // update() never returns DEAD
  static constexpr int NL = 2;
  static constexpr const char* NAME = "quad_monod";
  static void defaults(double* p) {
    p[0] = 2.0; p[1] = 0.01; p[2] = 1e-4; p[3] = 0.01; p[4] = 0.5; p[5] = 0.5; p[6] = 0.3; p[7] = 0.3;
    p[8] = 0.4; p[9] = 0.3; p[10] = 0.2; p[11] = 0.4; p[12] = 2e-15; p[13] = 1.4e-15; p[14] = 0.2e-15;
  }
  static void init(const double* p, double* x, uint64_t seed, uint32_t slot) {
    const U4 w = draw4(seed, slot, 0u, TAG_INIT, 0u);
    x[0] = trunc_normal(p[13], p[14], 0.5 * p[12], p[12], u01_52(w.x, w.y));
    x[1] = 0.5 * p[0];
    x[2] = 0.0;
    x[3] = 0.0;
    x[4] = 0.0;
    x[5] = 0.0;
    x[6] = x[0];
    x[7] = 0.0;
  }
  MCST_HD static double mass(const double*, const double* x) { return x[0]; }
  MCST_HD static void record(const double* p, const double* c, double* r) {
    const double S = c[0] > 0.0 ? c[0] : 0.0;
    const double O = c[1] > 0.0 ? c[1] : 0.0;
    const double A = c[2] > 0.0 ? c[2] : 0.0;
    r[0] = S / (p[1] + S);     // glucose saturation
    r[1] = O / (p[2] + O);     // oxygen saturation
    r[2] = A / (p[3] + A);     // acetate saturation
    r[3] = p[1] / (p[1] + S);  // glucose repression of acetate uptake
  }
  MCST_HD static int update(const double* p, double dt, double* x, const double* r, double* contrib) {
    const double qs_t = p[0] * r[0];
    const double qa_t = p[4] * r[2] * r[3];
    const double Us = x[1] * x[0];
    const double Ua = x[2] * x[0] * r[1];
    const double Resp = Us * r[1];
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
    return (x[0] >= p[12]) ? DIVISION : IDLE;
  }
  MCST_HD static void division(const double*, double* x, double* xd, const RngKey&) {
    x[0] = x[0] * 0.5;
    x[3] = x[3] * 0.5;
    x[6] = x[0];
    x[7] = x[7] + 1.0;
    for (int i = 0; i < NV; ++i) xd[i] = x[i];
  }
};

}  // namespace mcst
