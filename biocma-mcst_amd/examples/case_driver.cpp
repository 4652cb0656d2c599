// case_driver.cpp -- a host driver in C++ that loads a CMA case directory
// with the C++ case loader (host/case_loader.hpp) and runs it over the C ABI
// of libmcst.so: the reference's init_transitionner + host_init_state
// (apps/core/src/global_initaliser.cpp:200-244,521-568) then its main loop
// (apps/core/src/host_specific.cpp:257-313), the flowmap swaps done on the
// device every steps_per_flowmap steps.
//
//   case_driver --dump CASE_DIR [dt] [t_per_flowmap]
//       reads the case only (no GPU) and prints what was read as JSON
//   case_driver CASE_DIR MODEL DT T_PER_FLOWMAP N_PARTICLES STEPS SEED C0
//               [FEED_IN FEED_OUT FEED_FLOW FEED_CONC]
//       one run of MODEL (built-in model name, default parameters; every
//       species starts at C0; the feed brings species 0, FEED_OUT < 0 for no
//       outlet), prints the counters, a position checksum and the liquid mass
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../host/case_loader.hpp"
#include "mcst.h"

namespace {

void check(int rc, const char* what) {
  if (rc != MCST_OK) throw std::runtime_error(std::string(what) + ": " + mcst_last_error());
}

template <class T>
void print_list(const std::vector<T>& v, const char* fmt) {
  std::printf("[");
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) std::printf(", ");
    std::printf(fmt, v[i]);
  }
  std::printf("]");
}

int dump(int argc, char** argv) {
  const mcst_host::CmaCase c = mcst_host::load_case(argv[2]);
  const double user_dt = argc > 3 ? std::atof(argv[3]) : 0.0;
  const double t_map = argc > 4 ? std::atof(argv[4]) : 0.0;
  const double dt = mcst_host::time_step(c, user_dt);
  std::printf("{\"n_c\": %u, \"n_maps\": %zu, \"dt\": %.17g, \"steps_per_map\": %u, \"maps\": [", c.n_c(), c.n_maps(),
              dt, mcst_host::steps_per_flowmap(t_map, c.n_maps(), dt));
  for (size_t k = 0; k < c.n_maps(); ++k) {
    std::printf(k ? ", {\"V\": " : "{\"V\": ");
    print_list(c.volumes[k], "%.17g");
    std::printf(", \"rows\": ");
    print_list(c.flows[k].rows, "%u");
    std::printf(", \"cols\": ");
    print_list(c.flows[k].cols, "%u");
    std::printf(", \"flows\": ");
    print_list(c.flows[k].flows, "%.17g");
    std::printf("}");
  }
  std::printf("]}\n");
  return 0;
}

int run(int argc, char** argv) {
  if (argc < 9) throw std::runtime_error("usage: case_driver CASE_DIR MODEL DT T_PER_FLOWMAP N STEPS SEED C0 [feed]");
  const mcst_host::CmaCase c = mcst_host::load_case(argv[1]);
  const char* model = argv[2];
  const double dt = mcst_host::time_step(c, std::atof(argv[3]));
  const uint32_t per_map = mcst_host::steps_per_flowmap(std::atof(argv[4]), c.n_maps(), dt);
  const uint64_t n_p = (uint64_t)std::atoll(argv[5]);
  const uint32_t steps = (uint32_t)std::atoi(argv[6]);
  const uint64_t seed = (uint64_t)std::atoll(argv[7]);
  const double c0 = std::atof(argv[8]);
  int model_id = 0, nv = 0, ns = 0, np = 0;
  check(mcst_model_lookup(model, &model_id, &nv, &ns, &np), "model");
  mcst_engine* e = nullptr;
  check(mcst_engine_create(model, 0, &e), "engine");
  int32_t r = -1;
  check(mcst_run_create(e, c.n_c(), (uint32_t)ns, dt, seed, &r), "run");
  check(mcst_host::define_flowmaps(e, r, c, per_map), "flowmaps");
  if (argc >= 13) {
    const uint32_t in = (uint32_t)std::atoi(argv[9]);
    const int64_t out = std::atoll(argv[10]);
    const uint32_t sp = 0;
    const double conc = std::atof(argv[12]);
    check(mcst_run_add_feed(e, r, MCST_FEED_CONSTANT, std::atof(argv[11]), 0.0, in, out, 1, &sp, &conc), "feed");
  }
  std::vector<double> C0((size_t)ns * c.n_c(), c0);
  check(mcst_run_set_concentrations(e, r, C0.data()), "liquid");
  check(mcst_run_init_particles(e, r, n_p, 4 * n_p + 1024, 1.0), "particles");
  check(mcst_engine_commit(e), "commit");
  check(mcst_engine_step(e, steps), "step");
  uint64_t cnt[MCST_C_COUNT];
  check(mcst_run_get_counters(e, r, cnt), "counters");
  uint64_t n = 0;
  check(mcst_run_get_n(e, r, &n), "n");
  std::vector<uint32_t> pos(n);
  if (n) check(mcst_run_get_particles(e, r, pos.data(), nullptr, nullptr, nullptr, nullptr), "particles");
  uint64_t pos_sum = 0;
  for (uint64_t i = 0; i < n; ++i) pos_sum += (uint64_t)pos[i] * (i % 7 + 1);
  std::vector<double> m((size_t)ns * c.n_c());
  check(mcst_run_get_mass(e, r, m.data()), "mass");
  std::printf("{\"n_c\": %u, \"n_maps\": %zu, \"dt\": %.17g, \"steps_per_map\": %u, \"counters\": ", c.n_c(),
              c.n_maps(), dt, per_map);
  print_list(std::vector<uint64_t>(cnt, cnt + MCST_C_COUNT), "%" PRIu64);
  std::printf(", \"pos_checksum\": %" PRIu64 ", \"mass\": ", pos_sum);
  print_list(m, "%.17g");
  std::printf("}\n");
  check(mcst_engine_destroy(&e), "destroy");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 3 && std::string(argv[1]) == "--dump") return dump(argc, argv);
    return run(argc, argv);
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "case_driver: %s\n", ex.what());
    return 1;
  }
}
