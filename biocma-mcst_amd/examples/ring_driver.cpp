// ring_driver.cpp -- a host driver in C++ over the C ABI of libmcst.so, the
// way the reference's main loop (apps/core/src/host_specific.cpp:257-313)
// would call it: define one run, commit, step in blocks, read results.
//
// The case: a ring of N compartments (1 L each) exchanging Q in both
// directions, a feed of glucose at compartment 0 leaving at N/2, monod_mass
// particles with fast growth.  Prints one JSON line with the counters, the
// particle count, a checksum of positions and the liquid glucose mass.
//
//   ring_driver [n_compartments] [n_particles] [steps] [seed]
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "mcst.h"

namespace {

void check(int rc, const char* what) {
  if (rc != MCST_OK) throw std::runtime_error(std::string(what) + ": " + mcst_last_error());
}

}  // namespace

int main(int argc, char** argv) {
  const uint32_t n_c = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 32;
  const uint64_t n_p = argc > 2 ? (uint64_t)std::atoll(argv[2]) : 20000;
  const uint32_t steps = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 50;
  const uint64_t seed = argc > 4 ? (uint64_t)std::atoll(argv[4]) : 7;
  try {
    mcst_engine* e = nullptr;
    check(mcst_engine_create("monod_mass", 0, &e), "engine");
    int32_t run = -1;
    const double dt = 0.01;
    check(mcst_run_create(e, n_c, 1, dt, seed, &run), "run");
    const double params[6] = {0.5, 0.01, 0.5, 2.0, 1e-15, 0.1};  // mu_max, Ks, Y, ratio, m0, cv
    check(mcst_run_set_model_params(e, run, params, 6), "params");
    std::vector<double> V(n_c, 1e-3);
    std::vector<uint32_t> rows, cols;
    std::vector<double> flows;
    for (uint32_t i = 0; i < n_c; ++i) {
      const uint32_t j = (i + 1) % n_c;
      rows.push_back(i), cols.push_back(j), flows.push_back(2e-3);
      rows.push_back(j), cols.push_back(i), flows.push_back(2e-3);
    }
    check(mcst_run_add_flowmap(e, run, V.data(), flows.size(), rows.data(), cols.data(), flows.data()), "flowmap");
    const uint32_t sp = 0;
    const double conc = 10.0;
    check(mcst_run_add_feed(e, run, MCST_FEED_CONSTANT, 2e-3, 0.0, 0, n_c / 2, 1, &sp, &conc), "feed");
    std::vector<double> C0(n_c, 5.0);
    check(mcst_run_set_concentrations(e, run, C0.data()), "liquid");
    check(mcst_run_init_particles(e, run, n_p, 10 * n_p, 1.0), "particles");
    check(mcst_engine_commit(e), "commit");
    for (uint32_t done = 0; done < steps;) {  // main_loop: blocks of steps between exports
      const uint32_t k = std::min<uint32_t>(10, steps - done);
      check(mcst_engine_step(e, k), "step");
      done += k;
    }
    uint64_t c[MCST_C_COUNT];
    check(mcst_run_get_counters(e, run, c), "counters");
    uint64_t n = 0;
    check(mcst_run_get_n(e, run, &n), "n");
    std::vector<uint32_t> pos(n);
    std::vector<uint8_t> status(n);
    if (n) check(mcst_run_get_particles(e, run, pos.data(), status.data(), nullptr, nullptr, nullptr), "particles");
    uint64_t pos_sum = 0, idle = 0;
    for (uint64_t i = 0; i < n; ++i) {
      pos_sum += (uint64_t)pos[i] * (i % 7 + 1);
      idle += status[i] == 0;
    }
    std::vector<double> m(n_c);
    check(mcst_run_get_mass(e, run, m.data()), "mass");
    double m_tot = 0.0;
    for (double x : m) m_tot += x;
    std::printf("{\"steps\": %" PRIu64 ", \"births\": %" PRIu64 ", \"exits\": %" PRIu64 ", \"moves\": %" PRIu64
                ", \"n\": %" PRIu64 ", \"idle\": %" PRIu64 ", \"pos_checksum\": %" PRIu64 ", \"glucose_mass\": %.17g}\n",
                c[MCST_C_STEPS], c[MCST_C_BIRTHS], c[MCST_C_EXITS], c[MCST_C_MOVES], n, idle, pos_sum, m_tot);
    check(mcst_engine_destroy(&e), "destroy");
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "ring_driver: %s\n", ex.what());
    return 1;
  }
  return 0;
}
