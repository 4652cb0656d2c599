// udf_minimal.hpp -- an example model plugin: the length-growth model of the
// reference's minimal udf_model (apps/udf_model/minimal.cpp:23-107), in FP64.
//
//   python biocma-mcst_amd/build.py --udf biocma-mcst_amd/examples/udf_minimal.hpp \
//       --model UdfMinimal -o biocma-mcst_amd/examples/libudf_minimal.so
//
// then mcst_model_load("…/libudf_minimal.so", &id) and
// mcst_engine_create("udf_minimal", …), or BIOMC_LIB_UDF=…/libudf_minimal.so
// and the model name "udf_model".  The struct follows the model contract of
// csrc/models.hpp; the headers of csrc/ are on the include path.
#pragma once

#include "models.hpp"

// x = {length, l_max}; p = {Y (glucose to biomass yield), l_dot_max, l_max, k, d_m}.
struct UdfMinimal {
  static constexpr int ID = -1, NV = 2, NS = 1, NP = 5;  // ID: numbered by the library at load
  static constexpr bool CAN_DIE = false;
  static constexpr int NL = 4;
  static constexpr const char* NAME = "udf_minimal";
  static void defaults(double* p) {
    p[0] = 0.5; p[1] = 2e-6 / 3600.0; p[2] = 2e-6; p[3] = 1e-3; p[4] = 0.6e-6;
  }
  // minimal.cpp:55-63 with its configuration (:121-140): lengths uniform on
  // (l_max/2, l_max), l_max fixed
  static void init(const double* p, double* x, uint64_t seed, uint32_t slot) {
    const mcst::U4 w = mcst::draw4(seed, slot, 0u, mcst::TAG_INIT, 0u);
    x[0] = p[2] * (0.5 + 0.5 * mcst::u01(w.x));
    x[1] = p[2];
  }
  // minimal.cpp:101-105
  MCST_HD static double mass(const double* p, const double* x) { return x[0] * mcst::lin_density(p[4]); }
  MCST_HD static void record(const double* p, const double* c, double* r) {
    const double s = c[0];
    r[0] = s / (p[3] + s);  // g
  }
  // minimal.cpp:65-83: length grows by dt*ldot/(1 + dt*ldot); divides at l_max
  MCST_HD static int update(const double* p, double dt, double* x, const double* r, double* contrib) {
    const double g = r[0];
    const double phi_s_max = (p[1] * mcst::lin_density(p[4])) / p[0];
    const double ldot = p[1] * g;
    const double d_length = dt * ldot;
    x[0] = x[0] + d_length / (1.0 + dt * ldot);
    contrib[0] = -(phi_s_max * g);
    return (x[0] >= x[1]) ? mcst::DIVISION : mcst::IDLE;
  }
  // minimal.cpp:85-99: both halves keep l_max
  MCST_HD static void division(const double* p, double* x, double* xd, const mcst::RngKey&) {
    const double h = x[0] / 2.0;
    x[0] = h;
    x[1] = p[2];
    xd[0] = h;
    xd[1] = p[2];
  }
};
