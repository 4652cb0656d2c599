// kernels.hpp -- launchers of the step kernels: the model-independent ones
// (kernels.hip) and the table of a model's own (step_kernels.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "mcst_common.hpp"

namespace mcst {

// Launch shape of the particle pass: block size (256/512/1024 threads per
// 1024-slot tile), grid, the LDS bytes of the per-block contribution
// accumulator (priv) or none (direct global atomics), whether some run
// records exit and division ages (probe: a 256-thread variant), and whether
// some run has a leaving flow (leave: the variant that reads and writes
// hydraulic ages and exit thresholds; without one the pass issues none of
// those loads).
struct ParticleLaunch {
  int block = 256;
  int grid = 1;
  bool priv = false;
  bool probe = false;
  bool win = false;  // the accumulator holds a window of compartments, not whole runs
  bool leave = true;
  size_t lds = 0;
};

// A model's kernels: transport (its record() per compartment), the particle
// pass (particle_kernel: the kernel of a launch shape, for occupancy queries;
// nullptr when the shape has no variant), resolve (its division();
// resolve_kernel for occupancy queries) and the export values (its mass()).  The same table comes from the library for the
// built-in models and from a plugin for a model loaded at run time.
struct ModelOps {
  void (*transport)(const BatchDev& b, int grid, hipStream_t s);
  void (*particle)(const BatchDev& b, const ParticleLaunch& L, hipStream_t s);
  const void* (*particle_kernel)(int block, bool priv, bool probe, bool win, bool leave);
  // resolve: grid blocks plus aux blocks first in the grid, which transport
  // the next step (ahead) and, with contribution rows, sum the rows first
  // (for the next step's transport, or into S)
  void (*resolve)(const BatchDev& b, int grid, uint32_t aux, bool ahead, hipStream_t s);
  const void* (*resolve_kernel)();  // for occupancy queries
  void (*properties)(const BatchDev& b, uint32_t run, uint64_t n, double* values, double* spatial, hipStream_t s);
};
const ModelOps* builtin_model_ops(int model_id);

void launch_removal(const BatchDev& b, uint32_t n_var, int grid, hipStream_t s);
void launch_layout(const BatchDev& b, hipStream_t s);
void launch_repartition(const BatchDev& b, uint32_t run, unsigned long long* hist, hipStream_t s);
void launch_clear_u64(unsigned long long* p, uint64_t n, hipStream_t s);

// Stable sort of one run's particles by position (sort.hip).  pos, status
// and the n_f64 FP64 fields (state variables and both ages) are the run's
// own ranges; keys, iota, perm (n u32 each), buf (8n bytes) and temp
// (temp_bytes, from sort_temp_bytes) are scratch.
struct SortArgs {
  uint64_t n = 0;
  uint32_t end_bit = 1;      // keys are below 2^end_bit
  uint32_t* pos = nullptr;
  uint8_t* status = nullptr;
  double* f64[MAX_VAR + 2] = {};
  uint32_t n_f64 = 0;
  uint32_t* u32[2] = {};
  uint32_t n_u32 = 0;
  // to_dst: every field is gathered once, straight into these arrays (the
  // sorted order at index 0..n), and the source arrays are left as they were
  bool to_dst = false;
  uint32_t* dst_pos = nullptr;
  uint8_t* dst_status = nullptr;
  double* dst_f64[MAX_VAR + 2] = {};
  uint32_t* dst_u32[2] = {};
  uint32_t *keys = nullptr, *iota = nullptr, *perm = nullptr;
  void* buf = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
};
hipError_t sort_temp_bytes(uint64_t n, uint32_t end_bit, size_t* bytes);
hipError_t sort_run_particles(const SortArgs& a, hipStream_t s);

}  // namespace mcst
