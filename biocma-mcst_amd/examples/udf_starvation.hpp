// udf_starvation.hpp -- an example model plugin whose cells can die: Monod
// growth on one substrate, with a viability that falls while the substrate
// is short; a cell whose viability reaches zero returns Dead (the status of
// the reference's model contract, apps/libs/mc/public/mc/alias.hpp:124-130;
// the container removes dead cells with the exited ones,
// particles_container.hpp:539-557).  None of the reference's shipped models
// dies, so this one exists to exercise that path.
//
//   python biocma-mcst_amd/build.py --udf biocma-mcst_amd/examples/udf_starvation.hpp \
//       --model UdfStarvation -o biocma-mcst_amd/examples/libudf_starvation.so
#pragma once

#include <math.h>

#include "models.hpp"

// x = {m, m_birth, viability}; p = {mu_max, Ks, Y, ratio, kd, m0_mean, m0_cv}:
// dm/dt = mu m, uptake mu m / Y, viability falls at kd (1 - S/(Ks+S)),
// division at ratio x birth mass into halves, both fully viable.
struct UdfStarvation {
  static constexpr int ID = -1, NV = 3, NS = 1, NP = 7;
  static constexpr int NR = 2;  // record terms: mu and g (more than the one species)
  static constexpr bool CAN_DIE = true;
  static constexpr int NL = 4;
  static constexpr const char* NAME = "udf_starvation";
  static void defaults(double* p) {
    p[0] = 0.5 / 3600.0; p[1] = 0.01; p[2] = 0.5; p[3] = 2.0; p[4] = 1.0 / 3600.0; p[5] = 1e-15; p[6] = 0.1;
  }
  static void init(const double* p, double* x, uint64_t seed, uint32_t slot) {
    const mcst::U4 w = mcst::draw4(seed, slot, 0u, mcst::TAG_INIT, 0u);
    const double m = mcst::trunc_normal(p[5], p[6] * p[5], 0.0, INFINITY, mcst::u01_52(w.x, w.y));
    x[0] = m;
    x[1] = m / (1.0 + mcst::u01(w.z));
    x[2] = 0.2 + 0.8 * mcst::u01(w.w);
  }
  MCST_HD static double mass(const double*, const double* x) { return x[0]; }
  // NR = 2 terms per compartment: the growth rate mu = mu_max g and the
  // saturation g = S/(Ks+S)
  MCST_HD static void record(const double* p, const double* c, double* r) {
    const double s = c[0] > 0.0 ? c[0] : 0.0;
    const double g = s / (p[1] + s);
    r[0] = p[0] * g;
    r[1] = g;
  }
  MCST_HD static int update(const double* p, double dt, double* x, const double* r, double* contrib) {
    x[2] = x[2] - dt * (p[4] * (1.0 - r[1]));
    if (x[2] <= 0.0) return mcst::DEAD;
    const double gm = r[0] * x[0];
    contrib[0] = -(gm / p[2]);
    x[0] = x[0] + dt * gm;
    return (x[0] >= p[3] * x[1]) ? mcst::DIVISION : mcst::IDLE;
  }
  MCST_HD static void division(const double*, double* x, double* xd, const mcst::RngKey&) {
    const double h = x[0] * 0.5;
    x[0] = h;
    x[1] = h;
    x[2] = 1.0;
    xd[0] = h;
    xd[1] = h;
    xd[2] = 1.0;
  }
};
