// sort.hip -- stable reorder of one run's particles by compartment.
//
// The particle pass reads a per-compartment record for every particle and
// adds its contributions to a per-compartment accumulator.  With the slots of
// a compartment stored together, each block touches a small neighbourhood of
// compartments: the records stay in L1 and the contribution window of large
// networks covers the block's particles.  Moves and newborns (appended at the
// end) scatter that order over time; this pass restores it.  The order is a
// stable sort by position (slot order among equal positions), the same
// permutation the CPU oracle applies, so runs stay comparable step by step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hipcub/device/device_radix_sort.hpp>

#include "kernels.hpp"

namespace mcst {

namespace {

__global__ void __launch_bounds__(256) k_iota(uint64_t n, uint32_t* out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)i;
}

template <class T>
__global__ void __launch_bounds__(256) k_gather(uint64_t n, const uint32_t* __restrict__ perm, const T* __restrict__ src,
                                                T* __restrict__ dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

int grid_for(uint64_t n) { return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384)); }

}  // namespace

hipError_t sort_temp_bytes(uint64_t n, uint32_t end_bit, size_t* bytes) {
  *bytes = 0;
  return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                            (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, (int)end_bit);
}

hipError_t sort_run_particles(const SortArgs& a, hipStream_t s) {
  const uint64_t n = a.n;
  if (n < 2) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(256), 0, s, n, a.iota);
  size_t bytes = a.temp_bytes;
  hipError_t err = hipcub::DeviceRadixSort::SortPairs(a.temp, bytes, (const uint32_t*)a.pos, a.keys, (const uint32_t*)a.iota,
                                                      a.perm, (int)n, 0, (int)a.end_bit, s);
  if (err != hipSuccess) return err;
  // positions: the sorted keys; every other field through the permutation
  if ((err = hipMemcpyAsync(a.pos, a.keys, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s)) != hipSuccess) return err;
  hipLaunchKernelGGL(k_gather<uint8_t>, dim3(grid_for(n)), dim3(256), 0, s, n, (const uint32_t*)a.perm,
                     (const uint8_t*)a.status, (uint8_t*)a.buf);
  if ((err = hipMemcpyAsync(a.status, a.buf, n, hipMemcpyDeviceToDevice, s)) != hipSuccess) return err;
  for (uint32_t f = 0; f < a.n_f64; ++f) {
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(256), 0, s, n, (const uint32_t*)a.perm,
                       (const double*)a.f64[f], (double*)a.buf);
    if ((err = hipMemcpyAsync(a.f64[f], a.buf, n * sizeof(double), hipMemcpyDeviceToDevice, s)) != hipSuccess) return err;
  }
  return hipGetLastError();
}

}  // namespace mcst
