// udf.hpp -- what a model plugin hands the library (mcst_model_load).
//
// A plugin is a shared library built from udf_plugin.hip with a model struct
// of the contract in models.hpp (build.py --udf).  It exports one C function,
// mcst_udf_describe, which fills a UdfModel: the model's sizes and host
// functions, and the launchers of the step kernels instantiated for it.  The
// plugin's kernels read the library's device structures directly, so the
// library refuses a plugin whose structure sizes differ from its own (built
// against other headers).
#pragma once

#include <stdint.h>

#include "kernels.hpp"

namespace mcst {

constexpr uint32_t UDF_ABI_VERSION = 9;

struct UdfModel {
  uint32_t abi_version;
  uint32_t batch_bytes, run_bytes, map_bytes, state_bytes;  // sizeof BatchDev, RunDev, MapDev, RunState
  const char* name;
  int nv, ns, np, nl, nr;  // nr: record terms (record_terms<M>())
  bool can_die;
  void (*defaults)(double* p);
  void (*init)(const double* p, double* x, uint64_t seed, uint32_t slot);
  double (*mass)(const double* p, const double* x);
  ModelOps ops;
};

inline void udf_sizes(UdfModel& u) {
  u.batch_bytes = (uint32_t)sizeof(BatchDev);
  u.run_bytes = (uint32_t)sizeof(RunDev);
  u.map_bytes = (uint32_t)sizeof(MapDev);
  u.state_bytes = (uint32_t)sizeof(RunState);
}

}  // namespace mcst

// Returns 0 and fills *out when abi_version is the plugin's own.
extern "C" typedef int (*mcst_udf_describe_fn)(mcst::UdfModel* out, uint32_t abi_version);
