// kernels.hpp -- launchers of the step kernels (defined in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mcst_common.hpp"

namespace mcst {

void launch_transport(const BatchDev& b, int grid, hipStream_t s);
void launch_particle(int model_id, const BatchDev& b, int grid, hipStream_t s);
const void* particle_kernel_ptr(int model_id);
void launch_bookkeep(const BatchDev& b, hipStream_t s);
void launch_removal(const BatchDev& b, uint32_t n_var, int grid, hipStream_t s);
void launch_merge(const BatchDev& b, uint32_t n_var, int grid, hipStream_t s);
void launch_layout(const BatchDev& b, hipStream_t s);
void launch_repartition(const BatchDev& b, uint32_t run, unsigned long long* hist, hipStream_t s);

}  // namespace mcst
