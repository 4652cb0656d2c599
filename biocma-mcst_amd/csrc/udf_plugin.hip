// udf_plugin.hip -- a model plugin: the step kernels instantiated for one
// model struct, loaded by the library at run time (mcst_model_load, or the
// model name "udf_model" with the plugin's path in BIOMC_LIB_UDF, as the
// reference's udf_model: apps/api/src/udf_handle.cpp:21-38).
//
// Built by `python biocma-mcst_amd/build.py --udf <header> --model <Struct>
// -o <plugin.so>`: MCST_UDF_HEADER names the header that defines the struct,
// MCST_UDF_MODEL the struct.  The struct follows the contract of models.hpp
// (NAME, NV, NS, NP, NL, CAN_DIE, defaults, init, record, update, division,
// mass); its ID member is not used: the library numbers loaded models itself.
// record() writes exactly NR terms (r[0..NR); NR is NS unless the struct
// declares it): the record row holds the move probability and the exit
// threshold right after them.
#include "step_kernels.hpp"
#include "udf.hpp"

#ifndef MCST_UDF_HEADER
#error "MCST_UDF_HEADER must name the header that defines the model struct"
#endif
#ifndef MCST_UDF_MODEL
#error "MCST_UDF_MODEL must name the model struct"
#endif

#include MCST_UDF_HEADER

namespace {

using Model = MCST_UDF_MODEL;

double model_mass(const double* p, const double* x) { return Model::mass(p, x); }

static_assert(Model::NV >= 1 && Model::NV <= mcst::MAX_VAR, "state variables: 1..MAX_VAR");
static_assert(Model::NS >= 1 && Model::NS <= mcst::MAX_SPECIES, "species: 1..MAX_SPECIES");
static_assert(Model::NP >= 0 && Model::NP <= mcst::MAX_PARAMS, "parameters: 0..MAX_PARAMS");
static_assert(Model::NL == 2 || Model::NL == 4, "slots per lane: 2 or 4");
static_assert(mcst::record_terms<Model>() >= 1 && mcst::record_terms<Model>() <= 14, "record terms: 1..14");

}  // namespace

extern "C" __attribute__((visibility("default"))) int mcst_udf_describe(mcst::UdfModel* out,
                                                                        uint32_t abi_version) {
  if (!out || abi_version != mcst::UDF_ABI_VERSION) return -1;
  mcst::UdfModel u{};
  u.abi_version = mcst::UDF_ABI_VERSION;
  mcst::udf_sizes(u);
  u.name = Model::NAME;
  u.nv = Model::NV;
  u.ns = Model::NS;
  u.np = Model::NP;
  u.nl = Model::NL;
  u.nr = mcst::record_terms<Model>();
  u.can_die = Model::CAN_DIE;
  u.defaults = &Model::defaults;
  u.init = &Model::init;
  u.mass = &model_mass;
  u.ops = mcst::model_ops<Model>();
  *out = u;
  return 0;
}
