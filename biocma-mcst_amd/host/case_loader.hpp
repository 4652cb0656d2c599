// case_loader.hpp -- CMA case input on the C++ host: the flowmaps a run is
// built from, read from disk, and their period.
//
// The reference reads cases through rcmtool (a Rust library outside the
// reference tree, subprojects/rcmtool.wrap) from
// GlobalInitialiser::init_transitionner / host_init_state
// (apps/core/src/global_initaliser.cpp:200-244,521-568), and swaps flowmaps
// in its main loop with need_advance / advance
// (apps/core/src/host_specific.cpp:81-87,263-266).  The files are those of
// the reference's one shipped case (apps/api/tests/data/0d):
//   * a vector file (vofL.raw, liquid volumes): u32 count, then count f64
//     values, little-endian;
//   * a matrix file (flowL.raw, flows F(i, j) from i to j): u32 rows, u32
//     cols, then rows*cols f64 values, row-major (trailing bytes ignored);
//   * the case index (cma_case): a binary header followed by u32
//     length-prefixed relative paths of the flowmap files.
// Cases with several flowmaps sit in numbered subdirectories i_0, i_1, ...
// of the case directory (each with vofL.raw and flowL.raw, or a sparse
// flowL.coo: u64 nnz, u32 rows[nnz], u32 cols[nnz], f64 flows[nnz]); a
// directory holding the files itself is a single flowmap.
// On the device the flowmap of step k is (k / steps_per_map) mod n_maps, so
// the swap needs no host call: define_flowmaps hands every map over once.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mcst.h"

namespace mcst_host {

struct CooFlows {
  std::vector<uint32_t> rows, cols;
  std::vector<double> flows;
};

struct CmaCase {
  std::vector<std::vector<double>> volumes;  // [n_maps][n_c] liquid volumes
  std::vector<CooFlows> flows;               // [n_maps] off-diagonal flows F(i, j)
  uint32_t n_c() const { return volumes.empty() ? 0u : (uint32_t)volumes[0].size(); }
  size_t n_maps() const { return volumes.size(); }
};

// Readers of the files above; each throws std::runtime_error naming the
// file on a short or malformed file.
std::vector<double> read_raw_vector(const std::string& path);
std::vector<double> read_raw_matrix(const std::string& path, uint32_t& rows, uint32_t& cols);
CooFlows read_coo(const std::string& path);
// The relative paths listed in a case index: every u32 length followed by
// that many printable bytes ending in ".raw".
std::vector<std::string> read_case_index(const std::string& path);

// A case directory (one flowmap, or numbered i_k subdirectories).
CmaCase load_case(const std::string& directory);

// Smallest V_i / Q_out(i) over compartments with an outflow and over
// flowmaps (infinity when no compartment has one).
double min_residence_time(const CmaCase& c);
// get_time_step (global_initaliser.cpp:52-85): the user's dt if > 0 or for a
// 0-D case, else min residence time / 100; throws when neither exists.
double time_step(const CmaCase& c, double user_dt);
// compute_n_per_flowmap (global_initaliser.cpp:104-114).
uint32_t steps_per_flowmap(double t_per_flowmap, size_t n_maps, double dt);

// Hands every flowmap of the case to run `run` (mcst_run_add_flowmap, in
// order) and sets the flowmap period; returns the first error code.
int define_flowmaps(mcst_engine* e, int32_t run, const CmaCase& c, uint32_t steps_per_map);

}  // namespace mcst_host
