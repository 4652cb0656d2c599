// engine.hip -- host side of libmcst: runs, flowmap tables, device upload,
// the per-step launch sequence and the C ABI declared in include/mcst.h.
//
// The engine is the MI355X counterpart of the reference's SimulationUnit +
// CycleFunctors + ParticlesContainer + ReactorDomain + ScalarSimulation for
// one batch of runs (apps/libs/simulation/public/simulation/simulation.hpp:44-239,
// apps/libs/simulation/public/simulation/kernels/kernels.hpp:25-224).  All
// state stays on the device between steps; a step needs no host round trip.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <thread>
#include <vector>

#include "../../include/mcst.h"
#include "kernels.hpp"
#include "models.hpp"
#include "udf.hpp"

namespace mcst {
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) return fail(MCST_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// ------------------------------------------------------------ model registry
struct ModelEntry {
  const char* name;
  int id, nv, ns, np, nl, nr;
  bool can_die;
  void (*defaults)(double*);
  void (*init)(const double*, double*, uint64_t, uint32_t);
  double (*mass)(const double*, const double*);
  const ModelOps* ops;
};

template <class M>
double host_mass(const double* p, const double* x) {
  return M::mass(p, x);
}

template <class M>
ModelEntry entry() {
  return ModelEntry{M::NAME,   M::ID,         M::NV, M::NS, M::NP, M::NL, record_terms<M>(), M::CAN_DIE,
                    &M::defaults, &M::init, &host_mass<M>, builtin_model_ops(M::ID)};
}

const std::vector<ModelEntry>& registry() {
  static const std::vector<ModelEntry> r = {entry<MonodMass>(), entry<DualMonod>(), entry<Monod>(),
                                            entry<FixedLength>(), entry<SimpleAcetate>(), entry<QuadMonod>()};
  return r;
}

// Models loaded from plugins (mcst_model_load), numbered from UDF_ID_BASE in
// load order.  Entries are never removed and plugins never unloaded, so
// pointers to entries stay valid for the life of the process.
constexpr int UDF_ID_BASE = 100;
constexpr const char* UDF_ALIAS = "udf_model";  // the reference's name for the plugin model

struct UdfRegistry {
  std::mutex mu;
  std::deque<ModelEntry> entries;
  std::deque<ModelOps> ops;
  std::deque<std::string> names, paths;
  const ModelEntry* alias = nullptr;  // the plugin named by BIOMC_LIB_UDF
};

UdfRegistry& udf_registry() {
  static UdfRegistry r;
  return r;
}

// Loads the plugin at `path` (registry lock held).  Loading the same path
// again returns the entry of the first load.
int load_plugin(UdfRegistry& u, const char* path, const ModelEntry** out) {
  if (!path || !*path) return fail(MCST_ERR_ARG, "empty plugin path");
  char* real = realpath(path, nullptr);  // one file under several names is one plugin
  const std::string key = real ? real : path;
  std::free(real);
  for (size_t k = 0; k < u.paths.size(); ++k)
    if (u.paths[k] == key) {
      *out = &u.entries[k];
      return MCST_OK;
    }
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    const char* why = dlerror();
    return fail(MCST_ERR_ARG, std::string("cannot load model plugin: ") + (why ? why : path));
  }
  auto describe = (mcst_udf_describe_fn)dlsym(h, "mcst_udf_describe");
  UdfModel d{};
  if (!describe || describe(&d, UDF_ABI_VERSION) != 0) {
    dlclose(h);
    return fail(MCST_ERR_ARG, std::string(path) + " is not a model plugin of this library version");
  }
  UdfModel ref{};
  udf_sizes(ref);
  if (d.batch_bytes != ref.batch_bytes || d.run_bytes != ref.run_bytes || d.map_bytes != ref.map_bytes ||
      d.state_bytes != ref.state_bytes) {
    dlclose(h);
    return fail(MCST_ERR_ARG, std::string(path) + " was built against other device structures: rebuild it");
  }
  std::string why;
  if (!d.name || !*d.name) why = "has no name";
  else if (d.nv < 1 || d.nv > MAX_VAR || d.ns < 1 || d.ns > MAX_SPECIES || d.np < 0 || d.np > MAX_PARAMS ||
           (d.nl != 2 && d.nl != 4) || d.nr < 1 || d.nr > 14)
    why = "has sizes out of range";
  else if (!d.defaults || !d.init || !d.mass || !d.ops.transport || !d.ops.particle || !d.ops.particle_kernel ||
           !d.ops.resolve || !d.ops.resolve_kernel || !d.ops.properties)
    why = "misses a function";
  else if (std::strcmp(d.name, UDF_ALIAS) == 0)
    why = std::string("may not be named ") + UDF_ALIAS;
  else {
    for (const auto& m : registry())
      if (std::strcmp(m.name, d.name) == 0) why = std::string("has the name of a built-in model (") + d.name + ")";
    for (const auto& n : u.names)
      if (n == d.name) why = std::string("has the name of a model already loaded (") + d.name + ")";
  }
  if (!why.empty()) {
    dlclose(h);
    return fail(MCST_ERR_ARG, std::string("model plugin ") + path + " " + why);
  }
  u.names.emplace_back(d.name);
  u.paths.emplace_back(key);
  u.ops.push_back(d.ops);
  u.entries.push_back(ModelEntry{u.names.back().c_str(), UDF_ID_BASE + (int)u.entries.size(), d.nv, d.ns, d.np,
                                 d.nl, d.nr, d.can_die, d.defaults, d.init, d.mass, &u.ops.back()});
  *out = &u.entries.back();
  return MCST_OK;
}

// The model of a name: built-in, loaded, or "udf_model" (the plugin named by
// the environment variable BIOMC_LIB_UDF, loaded on first use).  Sets the
// error message when there is none.
const ModelEntry* find_model(const char* name) {
  if (!name) {
    fail(MCST_ERR_ARG, "null model name");
    return nullptr;
  }
  for (const auto& m : registry())
    if (std::strcmp(m.name, name) == 0) return &m;
  UdfRegistry& u = udf_registry();
  std::lock_guard<std::mutex> lock(u.mu);
  for (const auto& m : u.entries)
    if (std::strcmp(m.name, name) == 0) return &m;
  if (std::strcmp(name, UDF_ALIAS) == 0) {
    if (u.alias) return u.alias;
    const char* path = std::getenv("BIOMC_LIB_UDF");
    if (!path || !*path) {
      fail(MCST_ERR_ARG, "udf_model: BIOMC_LIB_UDF names no model plugin");
      return nullptr;
    }
    const ModelEntry* m = nullptr;
    if (load_plugin(u, path, &m) != MCST_OK) return nullptr;
    u.alias = m;
    return m;
  }
  fail(MCST_ERR_ARG, std::string("unknown model: ") + name);
  return nullptr;
}

const ModelEntry* find_model(int id) {
  for (const auto& m : registry())
    if (m.id == id) return &m;
  UdfRegistry& u = udf_registry();
  std::lock_guard<std::mutex> lock(u.mu);
  if (id >= UDF_ID_BASE && (size_t)(id - UDF_ID_BASE) < u.entries.size()) return &u.entries[id - UDF_ID_BASE];
  return nullptr;
}

// FNV-1a of the model name: what a checkpoint records of its model (loaded
// models are numbered in load order, so their ids vary between processes).
uint32_t model_tag(const char* name) {
  uint32_t h = 2166136261u;
  for (const char* c = name; *c; ++c) h = (h ^ (uint8_t)*c) * 16777619u;
  return h;
}

// ------------------------------------------------------------ flowmap tables
struct HostMap {
  std::vector<double> V, inv_v, out_flow, p_move, q_exit, cdf, val;
  std::vector<uint32_t> nbr, col_ptr, src;
  uint32_t m = 1;
};

struct Edge {
  uint32_t i, j;
  uint64_t ord;
  double q;
};

// Transition data of one flowmap.  Off-diagonal flows F(i,j) (duplicates
// summed in input order, non-positive totals dropped) give:
//   out_flow[i] = sum_j F(i,j)                (diag of the transition matrix)
//   p_move[i]   = dt*out_flow[i] / V[i]        (probability_leaving.hpp:33-46)
//   cdf[i][k]   = sum_{k'<=k} F(i,nbr_k') / out_flow[i], neighbours ascending,
//                 rows padded with cdf 1 and the last neighbour
//   transport matrix by destination column: F(i,j) and -out_flow[j]
//                 (documentation/pages/03_cma.md:33-64)
int build_map(uint32_t n, const double* V, uint64_t nnz, const uint32_t* rows, const uint32_t* cols,
              const double* flows, double dt, HostMap& f) {
  std::vector<Edge> e;
  e.reserve(nnz);
  for (uint64_t t = 0; t < nnz; ++t) {
    if (rows[t] >= n || cols[t] >= n) return fail(MCST_ERR_ARG, "flowmap entry out of range");
    if (!std::isfinite(flows[t])) return fail(MCST_ERR_ARG, "flowmap entry is not finite");
    if (rows[t] == cols[t]) continue;
    e.push_back(Edge{rows[t], cols[t], t, flows[t]});
  }
  std::sort(e.begin(), e.end(), [](const Edge& a, const Edge& b) {
    if (a.i != b.i) return a.i < b.i;
    if (a.j != b.j) return a.j < b.j;
    return a.ord < b.ord;
  });
  std::vector<Edge> u;
  u.reserve(e.size());
  for (size_t t = 0; t < e.size();) {
    Edge acc = e[t];
    size_t k = t + 1;
    for (; k < e.size() && e[k].i == acc.i && e[k].j == acc.j; ++k) acc.q += e[k].q;
    if (acc.q > 0.0) u.push_back(acc);
    t = k;
  }
  std::vector<uint64_t> rs(n + 1, 0);
  for (const auto& x : u) rs[x.i + 1]++;
  uint32_t m = 1;
  for (uint32_t i = 0; i < n; ++i) {
    m = std::max<uint32_t>(m, (uint32_t)rs[i + 1]);
    rs[i + 1] += rs[i];
  }
  f.m = m;
  f.V.assign(V, V + n);
  f.inv_v.resize(n);
  f.out_flow.assign(n, 0.0);
  f.p_move.resize(n);
  f.q_exit.assign(n, 2.0);
  f.cdf.assign((size_t)n * m, 0.0);
  f.nbr.assign((size_t)n * m, 0);
  for (uint32_t i = 0; i < n; ++i) {
    if (!(V[i] > 0.0) || !std::isfinite(V[i])) return fail(MCST_ERR_ARG, "compartment volume must be positive");
    f.inv_v[i] = 1.0 / V[i];
    double acc = 0.0;
    for (uint64_t t = rs[i]; t < rs[i + 1]; ++t) acc += u[t].q;
    f.out_flow[i] = acc;
    f.p_move[i] = (dt * acc) / V[i];
    double run = 0.0;
    uint32_t c = 0;
    for (uint64_t t = rs[i]; t < rs[i + 1]; ++t, ++c) {
      run += u[t].q;
      f.cdf[(size_t)i * m + c] = run / acc;
      f.nbr[(size_t)i * m + c] = u[t].j;
    }
    const bool has = rs[i + 1] > rs[i];
    const uint32_t last = has ? u[rs[i + 1] - 1].j : i;
    for (; c < m; ++c) {
      f.cdf[(size_t)i * m + c] = has ? 1.0 : 0.0;
      f.nbr[(size_t)i * m + c] = last;
    }
  }
  std::vector<Edge> tr(u);
  for (uint32_t i = 0; i < n; ++i) tr.push_back(Edge{i, i, 0, -f.out_flow[i]});
  std::sort(tr.begin(), tr.end(), [](const Edge& a, const Edge& b) {
    if (a.j != b.j) return a.j < b.j;
    return a.i < b.i;
  });
  if (tr.size() >= 0xFFFFFFFFull) return fail(MCST_ERR_ARG, "flowmap too large");
  f.col_ptr.assign(n + 1, 0);
  f.src.resize(tr.size());
  f.val.resize(tr.size());
  for (size_t t = 0; t < tr.size(); ++t) {
    f.col_ptr[tr[t].j + 1]++;
    f.src[t] = tr[t].i;
    f.val[t] = tr[t].q;
  }
  for (uint32_t j = 0; j < n; ++j) f.col_ptr[j + 1] += f.col_ptr[j];
  return MCST_OK;
}

// Transition data of one flowmap from the tables the reference's
// hydrodynamic state hands over (ReactorDomain::update, domain.cpp:44-74;
// ScalarSimulation::set_transition, implScalar.cpp:316-320): neighbour and
// cumulative-probability rows [n][n_cols] used as given (rcmtool's
// neighbour order and padding), out-flows (p_move = dt*Q/V), and the
// transition matrix as COO triplets, taken by destination column, ascending
// source, duplicates summed in input order (as sparse_from_coo sums them).
int build_map_tables(uint32_t n, const double* V, uint32_t n_cols, const uint64_t* nbr, const double* cdf,
                     const double* out_flow, uint64_t nnz, const uint64_t* rows, const uint64_t* cols,
                     const double* vals, double dt, HostMap& f) {
  if (n_cols == 0) return fail(MCST_ERR_ARG, "neighbour rows need at least one column");
  for (uint32_t i = 0; i < n; ++i) {
    if (!(V[i] > 0.0) || !std::isfinite(V[i])) return fail(MCST_ERR_ARG, "compartment volume must be positive");
    if (!(out_flow[i] >= 0.0) || !std::isfinite(out_flow[i]))
      return fail(MCST_ERR_ARG, "out-flows must be finite and non-negative");
    for (uint32_t k = 0; k < n_cols; ++k) {
      const size_t o = (size_t)i * n_cols + k;
      if (nbr[o] >= n) return fail(MCST_ERR_ARG, "neighbour id out of range");
      if (!std::isfinite(cdf[o])) return fail(MCST_ERR_ARG, "cumulative probability is not finite");
      if (k && cdf[o] < cdf[o - 1]) return fail(MCST_ERR_ARG, "cumulative probability rows must not decrease");
    }
  }
  std::vector<Edge> e;
  e.reserve(nnz);
  for (uint64_t t = 0; t < nnz; ++t) {
    if (rows[t] >= n || cols[t] >= n) return fail(MCST_ERR_ARG, "transition entry out of range");
    if (!std::isfinite(vals[t])) return fail(MCST_ERR_ARG, "transition entry is not finite");
    e.push_back(Edge{(uint32_t)rows[t], (uint32_t)cols[t], t, vals[t]});
  }
  if (e.size() >= 0xFFFFFFFFull) return fail(MCST_ERR_ARG, "flowmap too large");
  std::sort(e.begin(), e.end(), [](const Edge& a, const Edge& b) {
    if (a.j != b.j) return a.j < b.j;
    if (a.i != b.i) return a.i < b.i;
    return a.ord < b.ord;
  });
  f.m = n_cols;
  f.V.assign(V, V + n);
  f.inv_v.resize(n);
  f.out_flow.assign(out_flow, out_flow + n);
  f.p_move.resize(n);
  f.q_exit.assign(n, 2.0);
  for (uint32_t i = 0; i < n; ++i) {
    f.inv_v[i] = 1.0 / V[i];
    f.p_move[i] = (dt * out_flow[i]) / V[i];
  }
  f.cdf.assign(cdf, cdf + (size_t)n * n_cols);
  f.nbr.resize((size_t)n * n_cols);
  for (size_t o = 0; o < f.nbr.size(); ++o) f.nbr[o] = (uint32_t)nbr[o];
  f.col_ptr.assign(n + 1, 0);
  f.src.clear();
  f.val.clear();
  for (size_t t = 0; t < e.size();) {
    Edge acc = e[t];
    size_t k = t + 1;
    for (; k < e.size() && e[k].i == acc.i && e[k].j == acc.j; ++k) acc.q += e[k].q;
    f.col_ptr[acc.j + 1]++;
    f.src.push_back(acc.i);
    f.val.push_back(acc.q);
    t = k;
  }
  for (uint32_t j = 0; j < n; ++j) f.col_ptr[j + 1] += f.col_ptr[j];
  if (f.src.empty()) {  // keep every device pointer valid
    f.src.push_back(0);
    f.val.push_back(0.0);
  }
  return MCST_OK;
}

// ------------------------------------------------------------ runs
struct HostFeed {
  int kind;
  bool gas;  // a feed of the gas phase: its output is a gas outlet only
  double f0, param;
  uint32_t in;
  int64_t out;
  std::vector<uint32_t> sp;
  std::vector<double> conc;
};

struct HostRun {
  uint32_t n_c = 0, n_s = 0;
  double dt = 0.0;
  uint64_t seed = 0;
  double params[MAX_PARAMS] = {};
  std::vector<HostMap> maps;
  uint32_t steps_per_map = 1;
  std::vector<HostFeed> feeds;
  std::vector<double> C0;
  std::vector<double> kla, cstar;  // gas-liquid transfer, [n_s][n_c] (empty: none)
  // two-phase runs: Henry constants, initial gas concentrations, the gas side
  // of each flowmap with its energy dissipation (empty: none), the kLa rule
  bool two_phase = false;
  double henry[MAX_SPECIES] = {};
  std::vector<double> Cg0;
  std::vector<HostMap> gmaps;
  std::vector<std::vector<double>> g_eps;
  int kla_mode = MCST_KLA_FIXED;
  double kla_fixed[MAX_SPECIES] = {};
  uint32_t kla_species = 0;
  double bubble_d = 5e-3;
  std::vector<uint32_t> vf;        // time-dependent feeds (indices into feeds), in feed order
  std::vector<uint32_t> vx_out;    // outlets whose first leaving feed is time-dependent ...
  std::vector<uint32_t> vx_feed;   // ... and that feed (index into vf)
  double* d_vf = nullptr;          // device: this step's flows of vf, then the outlets' thresholds
  double ratio = 0.01;
  uint64_t min_removal = 0;
  uint32_t div_cap = 0xFFFFFFFFu; // newborns per step at most (the reference's division buffer)
  uint32_t probe_cap = 0;
  uint32_t step_target = 0xFFFFFFFFu;
  uint32_t sort_interval = 0;    // reorder the particles by compartment every k steps (0: never)
  uint32_t grow_every = 0;       // capacity check every k steps (0: fixed capacity)
  double grow_factor = 2.5;      // ... new capacity = ceil(n * factor)
  uint32_t host_step = 0;        // the run's step count as the device will have it
  // initial particles
  bool has_particles = false;
  uint64_t n = 0, cap = 0;
  std::vector<uint32_t> pos;
  std::vector<uint8_t> status;
  std::vector<double> state;  // [nv][n]
  std::vector<double> ah, ad;
  double weight = 0.0;
  // state restored from a checkpoint (mcst_run_restore): liquid mass and
  // pending contributions, step, counters and probes
  bool restored = false;
  std::vector<double> r_m, r_S, r_exit, r_div, r_cg, r_mg;
  uint32_t r_step = 0;
  unsigned long long r_counters[C_COUNT] = {};
  // device-side placement
  int phase = 0;                 // 0 defined, 1 on the device, 2 retired
  uint64_t p_off = 0;            // first slot (a multiple of TILE)
  uint32_t tiles = 0;            // slot range in tiles: [p_off/TILE, p_off/TILE + tiles)
  uint32_t map_base = 0;
  std::vector<void*> dmem;       // the run's own device arrays
};

}  // namespace
}  // namespace mcst

struct mcst_engine {
  const mcst::ModelEntry* model = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<mcst::HostRun> runs;
  bool committed = false;
  // capacity reserved for runs admitted after commit (0: sized by the runs at commit)
  uint64_t reserve_slots = 0;
  uint32_t reserve_runs = 0;
  // device
  std::vector<void*> allocs;        // engine-wide arrays, freed at destroy
  mcst::BatchDev bd{};
  mcst::RunDev* d_runs = nullptr;
  mcst::RunState* d_rs = nullptr;
  uint32_t runs_cap = 0;            // entries of d_runs / d_rs / comp_off / tile_off
  uint32_t admitted = 0;            // runs [0, admitted) have been admitted
  std::vector<mcst::RunDev> h_runs;
  std::vector<std::pair<uint32_t, uint32_t>> free_tiles;  // free slot ranges of the pool, in tiles
  std::vector<mcst::MapDev> h_maps;
  mcst::MapDev* d_maps = nullptr;
  uint32_t maps_cap = 0;
  std::vector<uint32_t> comp_off;
  uint32_t* d_comp = nullptr;
  int grid_particle = 1, grid_transport = 1;
  mcst::ParticleLaunch plaunch;
  int tuned_block = 0;              // particle-pass shape chosen by mcst_engine_autotune (0: none)
  void* sort_mem = nullptr;         // scratch of the particle reorder, sized for sort_n slots
  uint64_t sort_n = 0;
  size_t sort_temp = 0;
  // a second set of the pool's particle arrays (alt_P slots, in allocs): the
  // reorder of an engine's only live run gathers into it and swaps it in
  uint32_t* alt_pos = nullptr;
  uint8_t* alt_status = nullptr;
  double* alt_state = nullptr;
  double* alt_age_hyd = nullptr;
  double* alt_age_div = nullptr;
  uint32_t* alt_clk_hyd = nullptr;
  uint32_t* alt_clk_div = nullptr;
  uint64_t alt_P = 0;
  // contribution rows of the particle pass (BatchDev::part), part_bytes
  // allocated; part_aux: the resolve kernel's blocks that sum them
  double* part_mem = nullptr;
  size_t part_bytes = 0;
  uint32_t part_aux = 0;
  int tuned_group = -1;
  int grid_move = 1;
  bool removal_possible = true;
  bool any_leave = true;            // some run on the device has a leaving flow (the pass's leave variant)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint32_t graph_steps = 0;
};

namespace mcst {
namespace {

int dev_alloc(mcst_engine* e, void** p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t err = hipMalloc(p, bytes);
  if (err != hipSuccess) return fail(MCST_ERR_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(err));
  e->allocs.push_back(*p);
  return MCST_OK;
}

template <class T>
int dev_upload(mcst_engine* e, T** p, const T* src, size_t count) {
  int rc = dev_alloc(e, (void**)p, count * sizeof(T));
  if (rc) return rc;
  if (count) HIP_TRY(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
  return MCST_OK;
}

// Allocation of one run's own arrays (freed when the run is retired).
int run_alloc(HostRun& h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t err = hipMalloc(p, bytes);
  if (err != hipSuccess) return fail(MCST_ERR_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(err));
  h.dmem.push_back(*p);
  return MCST_OK;
}

template <class T>
int run_upload(HostRun& h, T** p, const T* src, size_t count) {
  int rc = run_alloc(h, (void**)p, count * sizeof(T));
  if (rc) return rc;
  if (count) HIP_TRY(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
  return MCST_OK;
}

void free_run_memory(HostRun& h) {
  for (void* p : h.dmem) (void)hipFree(p);
  h.dmem.clear();
}

void drop_graph(mcst_engine* e) {
  if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->graph_exec = nullptr;
  e->graph = nullptr;
  e->graph_steps = 0;
}

void release(mcst_engine* e) {
  drop_graph(e);
  e->part_mem = nullptr;  // freed with allocs
  e->part_bytes = 0;
  e->part_aux = 0;
  e->alt_pos = nullptr;  // freed with allocs
  e->alt_status = nullptr;
  e->alt_state = e->alt_age_hyd = e->alt_age_div = nullptr;
  e->alt_clk_hyd = e->alt_clk_div = nullptr;
  e->alt_P = 0;
  if (e->sort_mem) (void)hipFree(e->sort_mem);
  e->sort_mem = nullptr;
  e->sort_n = 0;
  e->sort_temp = 0;
  for (auto& h : e->runs) free_run_memory(h);
  for (void* p : e->allocs) (void)hipFree(p);
  e->allocs.clear();
  e->d_maps = nullptr;
  e->maps_cap = 0;
}

// First-fit range of `n` tiles from the pool's free list; false if none.
bool take_tiles(mcst_engine* e, uint32_t n, uint32_t& start) {
  if (n == 0) {
    start = 0;
    return true;
  }
  for (size_t i = 0; i < e->free_tiles.size(); ++i) {
    auto& f = e->free_tiles[i];
    if (f.second < n) continue;
    start = f.first;
    f.first += n;
    f.second -= n;
    if (f.second == 0) e->free_tiles.erase(e->free_tiles.begin() + (long)i);
    return true;
  }
  return false;
}

void give_tiles(mcst_engine* e, uint32_t start, uint32_t n) {
  if (n == 0) return;
  auto& fl = e->free_tiles;
  fl.push_back({start, n});
  std::sort(fl.begin(), fl.end());
  std::vector<std::pair<uint32_t, uint32_t>> merged;
  for (const auto& f : fl) {
    if (!merged.empty() && merged.back().first + merged.back().second == f.first)
      merged.back().second += f.second;
    else
      merged.push_back(f);
  }
  fl.swap(merged);
}

HostRun* get_run(mcst_engine* e, int32_t run) {
  if (!e || run < 0 || (size_t)run >= e->runs.size()) return nullptr;
  return &e->runs[run];
}

// Feed sources F[s][c], sink[c] (constant feeds) and the per-map leave
// thresholds exp(-dt*Q/V) of the first leaving flow of each outlet
// (update_feed, simulation.model.cpp:72-122; find_flow, move_kernel.hpp:105-127).
void feed_tables(const HostRun& r, std::vector<double>& F, std::vector<double>& sink, bool gas = false) {
  F.assign((size_t)r.n_s * r.n_c, 0.0);
  sink.assign(r.n_c, 0.0);
  for (const auto& fd : r.feeds) {
    if (fd.gas != gas) continue;
    if (fd.kind != MCST_FEED_CONSTANT) continue;  // added by the transport kernel each step
    const double q = fd.f0;
    for (size_t v = 0; v < fd.sp.size(); ++v) F[(size_t)fd.sp[v] * r.n_c + fd.in] += q * fd.conc[v];
    if (fd.out >= 0) sink[(size_t)fd.out] += q;
  }
}

// feed_descriptor.cpp:18-28
double feed_flow(const HostFeed& fd, double t) {
  switch (fd.kind) {
    case MCST_FEED_LINEAR:
      return fd.f0 + t * fd.param;
    case MCST_FEED_EXPONENTIAL:
      return fd.f0 + std::exp(fd.param * t);
    default:
      return fd.f0;
  }
}

void exit_thresholds(const HostRun& r, HostMap& mp) {
  mp.q_exit.assign(r.n_c, 2.0);
  std::vector<char> seen(r.n_c, 0);
  for (const auto& fd : r.feeds) {
    if (fd.out < 0 || fd.gas) continue;
    const uint32_t o = (uint32_t)fd.out;
    if (seen[o]) continue;
    seen[o] = 1;
    if (fd.kind == MCST_FEED_CONSTANT) mp.q_exit[o] = std::exp(-(r.dt * fd.f0) / mp.V[o]);
    // a time-dependent first feed: its threshold is written before each step
  }
}

// Kinematic viscosity of water at 20 C as the reference's correlation gives
// it (impl_mtr.cpp:58-84 at temperature 20, :98-99), and the oxygen
// diffusion constant (impl_mtr.cpp:97).
constexpr double WATER_NU_20C = 1.0023e-6;
constexpr double O2_DIFFUSION = 1e-9;

// kLa [n_s][n_c] of flowmap k of a two-phase run: FixedKla rows, or for
// one species kL = 0.3 (eps nu)^(1/4) Sc^(-1/2) times the interfacial area
// a = 6 alpha / (d_b (1 - alpha)), alpha = V_gas / (V_liq + V_gas)
// (impl_mtr.cpp:21-50,90-134).
std::vector<double> gas_kla(const HostRun& h, size_t k) {
  const uint32_t n = h.n_c;
  std::vector<double> kla((size_t)h.n_s * n, 0.0);
  if (h.kla_mode == MCST_KLA_FIXED) {
    for (uint32_t s = 0; s < h.n_s; ++s)
      for (uint32_t j = 0; j < n; ++j) kla[(size_t)s * n + j] = h.kla_fixed[s];
    return kla;
  }
  const double nu = WATER_NU_20C;
  const double sc = nu / O2_DIFFUSION;
  const HostMap &f = h.maps[k], &g = h.gmaps[k];
  for (uint32_t j = 0; j < n; ++j) {
    const double kl = (0.3 * std::pow(h.g_eps[k][j] * nu, 0.25)) * std::pow(sc, -0.5);
    const double alpha = g.V[j] / (f.V[j] + g.V[j]);
    const double a = (6.0 * alpha) / (h.bubble_d * (1.0 - alpha));
    kla[(size_t)h.kla_species * n + j] = kl * a;
  }
  return kla;
}

int check_definable(const mcst_engine* e, const HostRun& h, uint32_t r) {
  if (h.maps.empty()) return fail(MCST_ERR_STATE, "run " + std::to_string(r) + " has no flowmap");
  if (!h.has_particles) return fail(MCST_ERR_STATE, "run " + std::to_string(r) + " has no particles");
  if (h.two_phase && h.gmaps.size() != h.maps.size())
    return fail(MCST_ERR_STATE, "two-phase run " + std::to_string(r) + " needs one gas flowmap per flowmap");
  if (h.two_phase && !h.kla.empty())
    return fail(MCST_ERR_STATE, "run " + std::to_string(r) + " has both a gas phase and a fixed gas transfer");
  if (h.two_phase && h.kla_mode == MCST_KLA_TURBULENCE)
    for (size_t k = 0; k < h.gmaps.size(); ++k)
      if (h.g_eps[k].empty())
        return fail(MCST_ERR_STATE, "turbulence kLa needs the energy dissipation of every gas flowmap");
  return MCST_OK;
}

// Launch shapes for the runs on the device: LDS contribution accumulator
// when the largest live run's [n_s][n_c] fits, then the block size with the
// most resident waves, up to 8 per CU; a grid of that many blocks per CU.  MCST_PARTICLE_BLOCK
// (256/512/1024), MCST_CONTRIB_LDS (0/1) and MCST_TILE_GROUP override.
int choose_launch(mcst_engine* e) {
  const ModelEntry& M = *e->model;
  BatchDev& b = e->bd;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, e->device));
  size_t lds_need = 0;
  uint32_t live = 0;
  bool probe = false;
  for (const auto& h : e->runs)
    if (h.phase == 1) {
      lds_need = std::max(lds_need, (size_t)h.n_s * h.n_c * sizeof(double));
      probe = probe || h.probe_cap > 0;
      ++live;
    }
  const size_t lds_cap = std::min<size_t>(std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin),
                                          96 * 1024);
  // the accumulator holds the whole [n_s][n_c] of a run when it fits, else a
  // window of compartments (particles outside it add to S directly)
  const bool windowed = lds_need > lds_cap;
  const size_t ns_b = sizeof(double) * std::max<uint32_t>(1, (uint32_t)M.ns);
  // a window is sized so that two blocks fit a CU's LDS (with 1 KB for the
  // kernel's own shared variables): on the 50 000-compartment lattice two
  // 512-thread blocks per CU with windows of 2496 compartments took 5.36 ms
  // per step against 5.59 ms for one block with 3072
  size_t win_cap = lds_cap;
  if (windowed && prop.maxSharedMemoryPerMultiProcessor > 2 * 1024)
    win_cap = std::min(lds_cap, (size_t)prop.maxSharedMemoryPerMultiProcessor / 2 - 1024);
  lds_need = std::min(lds_need, ((windowed ? win_cap : lds_cap) / ns_b) * ns_b);
  bool priv = lds_need > 0;
  if (const char* v = std::getenv("MCST_CONTRIB_LDS")) priv = priv && std::atoi(v) != 0;
  b.acc_w = priv ? (uint32_t)(lds_need / ns_b) : 0u;
  int forced_block = e->tuned_block;
  if (const char* v = std::getenv("MCST_PARTICLE_BLOCK")) forced_block = std::atoi(v);
  ParticleLaunch best;
  int best_waves = -1;
  for (int blk : {256, 512, 1024}) {
    if (!priv && blk != 256) continue;
    if (forced_block && blk != forced_block) continue;
    const void* fn = M.ops->particle_kernel(blk, priv, false, priv && windowed, e->any_leave);
    if (!fn) continue;
    const size_t lds = priv ? lds_need : 0;
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, blk, lds) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, fn) == hipSuccess && fa.localSizeBytes > 0 && !forced_block)
      per_cu = per_cu / 2;  // register spills: prefer a variant that fits
    // waves per CU count up to 8, and a grid holds at most 8 waves per CU
    // (but one block of 1024): the pass is bound by the misses a CU's vector
    // L1 keeps in flight, and more waves only share them (on the
    // 1e7-particle network 16 waves per CU took 0.204-0.214 ms against 0.167
    // ms for 8); on a tie the smaller block wins (more blocks, shorter
    // flushes).  With a compartment window, up to 16 (two 512-thread
    // blocks): its waves also wait on window flushes and global atomics of
    // the particles outside the window
    const int wcap = windowed ? 16 : 8;
    const int use = std::min(per_cu, std::max(1, wcap * 64 / blk));
    const int waves = std::min(use * blk / 64, wcap);
    if (per_cu > 0 && waves > best_waves) {
      best_waves = waves;
      best.block = blk;
      best.priv = priv;
      best.win = priv && windowed;
      best.lds = lds;
      best.grid = use * prop.multiProcessorCount;
    }
  }
  if (best_waves < 0) {  // no resident configuration with the accumulator: go direct
    best = ParticleLaunch{};
    b.acc_w = 0;
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, M.ops->particle_kernel(256, false, false, false, e->any_leave), 256, 0));
    best.grid = std::min(std::max(per_cu, 1), 2) * prop.multiProcessorCount;  // 8 waves per CU
  }
  best.grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)best.grid, b.T_max));
  best.probe = probe;
  best.leave = e->any_leave;
  if (best.probe && best.lds > 64 * 1024)
    (void)hipFuncSetAttribute(M.ops->particle_kernel(256, best.priv, true, best.win, true),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)best.lds);
  // one run whose accumulator fits: blocks take groups in turn, which keeps
  // the wave front of the whole grid inside a narrow window of the particle
  // arrays; a batch, or a window of compartments: one contiguous range per
  // block, so it meets few runs and few compartments
  b.tile_group = e->tuned_group >= 0 ? (uint32_t)e->tuned_group : (live <= 1 && !windowed ? 1u : 0u);
  if (const char* v = std::getenv("MCST_TILE_GROUP")) b.tile_group = (uint32_t)std::max(0, std::atoi(v));
  // contribution rows (MCST_CONTRIB_ROWS=1; one live run whose whole
  // accumulator is in LDS): each block of the pass stores its accumulator as
  // a row and the resolve kernel sums the rows in a fixed order, instead of
  // each block adding its accumulator to S with FP64 atomics.  On the
  // 1e7-particle network the pass takes 0.156 ms instead of 0.167 but the
  // resolve kernel 0.031 instead of 0.020 (the rows are written and read
  // back: 41 MB each way), the same step, so the atomics stay the default.
  b.part = nullptr;
  b.part_rows = b.part_w = b.part_run = 0;
  e->part_aux = 0;
  bool rows = false;
  if (const char* v = std::getenv("MCST_CONTRIB_ROWS")) rows = live == 1 && best.priv && !best.win && std::atoi(v) != 0;
  if (rows) {
    uint32_t r1 = 0;
    for (uint32_t r = 0; r < e->runs.size(); ++r)
      if (e->runs[r].phase == 1) r1 = r;
    const HostRun& h1 = e->runs[r1];
    const uint64_t w = (uint64_t)h1.n_s * h1.n_c;
    const size_t bytes = (size_t)best.grid * w * sizeof(double);
    if (bytes > e->part_bytes) {  // a larger row buffer
      HIP_TRY(hipStreamSynchronize(e->stream));
      drop_graph(e);
      if (e->part_mem) {
        (void)hipFree(e->part_mem);
        e->allocs.erase(std::remove(e->allocs.begin(), e->allocs.end(), (void*)e->part_mem), e->allocs.end());
        e->part_mem = nullptr;
        e->part_bytes = 0;
      }
      if (dev_alloc(e, (void**)&e->part_mem, bytes) == MCST_OK) e->part_bytes = bytes;
    }
    if (e->part_mem && bytes <= e->part_bytes) {
      const uint32_t ns = (uint32_t)std::max(1, M.ns);
      const uint32_t cpb = ns >= 16 ? 1u : 16u / ns;  // part_cpb<NS>() of the kernels
      b.part = e->part_mem;
      b.part_rows = (uint32_t)best.grid;
      b.part_w = (uint32_t)w;
      b.part_run = r1;
      e->part_aux = (h1.n_c + cpb - 1) / cpb;
    }
  }
  e->plaunch = best;
  e->grid_particle = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)prop.multiProcessorCount * 4, b.T_max));
  // one thread per compartment, plus the block that lays out the tiles
  e->grid_transport = (int)std::min<uint64_t>(((uint64_t)b.n_comp_total + 255) / 256, 8192) + 1;
  // resolve: enough blocks for groups of 32 sub-chunks, up to the blocks
  // that are resident at once (a second round of blocks would start after
  // the first had worked through their groups); the kernel picks 32 or 64
  // from the step's tile count
  int res_per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&res_per_cu, M.ops->resolve_kernel(), 256, 0));
  const uint64_t resident = (uint64_t)prop.multiProcessorCount * (uint64_t)std::max(1, res_per_cu);
  const uint64_t n_sc = (uint64_t)b.T_max * b.nsub;
  e->grid_move = (int)std::max<uint64_t>(1, std::min<uint64_t>(resident, (n_sc + 31) / 32));
  if (std::getenv("MCST_VERBOSE"))
    std::fprintf(stderr, "[mcst] particle pass: block=%d grid=%d lds=%zu priv=%d group=%u window=%u leave=%d rows=%d (lds cap %zu)\n",
                 best.block, best.grid, best.lds, (int)best.priv, b.tile_group, b.acc_w, (int)best.leave,
                 (int)(b.part != nullptr), lds_cap);
  return MCST_OK;
}

// Compartment offsets of the transport space (a run that is not on the
// device has none) and the batch-wide counts the kernels read.
int refresh_tables(mcst_engine* e) {
  BatchDev& b = e->bd;
  const uint32_t R = e->admitted;
  e->comp_off.assign(R + 1, 0);
  uint32_t n_comp = 0;
  for (uint32_t r = 0; r < R; ++r) {
    e->comp_off[r] = n_comp;
    if (e->runs[r].phase == 1) n_comp += e->runs[r].n_c;
  }
  e->comp_off[R] = n_comp;
  HIP_TRY(hipMemcpy(e->d_comp, e->comp_off.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (e->h_maps.size() > e->maps_cap) {  // the flowmap descriptor table grows
    const uint32_t cap = (uint32_t)std::max<size_t>(e->h_maps.size(), 2 * (size_t)e->maps_cap);
    MapDev* d = nullptr;
    int rc = dev_alloc(e, (void**)&d, cap * sizeof(MapDev));
    if (rc) return rc;
    e->d_maps = d;
    e->maps_cap = cap;
  }
  if (!e->h_maps.empty())
    HIP_TRY(hipMemcpy(e->d_maps, e->h_maps.data(), e->h_maps.size() * sizeof(MapDev), hipMemcpyHostToDevice));
  b.R = R;
  b.n_comp_total = n_comp;
  b.maps = e->d_maps;
  e->any_leave = false;
  for (uint32_t r = 0; r < R; ++r)
    if (e->runs[r].phase == 1 && e->h_runs[r].enable_leave) e->any_leave = true;
  e->removal_possible = e->model->can_die || e->any_leave;
  return choose_launch(e);
}

// Upload run r into its slot range and its own arrays.
int upload_run(mcst_engine* e, uint32_t r) {
  HostRun& h = e->runs[r];
  BatchDev& b = e->bd;
  const uint32_t nv = (uint32_t)e->model->nv;
  int rc;
  if (h.n) {
    HIP_TRY(hipMemcpy(b.pos + h.p_off, h.pos.data(), h.n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b.status + h.p_off, h.status.data(), h.n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b.age_hyd + h.p_off, h.ah.data(), h.n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b.age_div + h.p_off, h.ad.data(), h.n * sizeof(double), hipMemcpyHostToDevice));
    // the run's age clocks start at 0: every age is its base age
    HIP_TRY(hipMemset(b.clk_hyd + h.p_off, 0, h.n * sizeof(uint32_t)));
    HIP_TRY(hipMemset(b.clk_div + h.p_off, 0, h.n * sizeof(uint32_t)));
    for (uint32_t v = 0; v < nv; ++v)
      HIP_TRY(hipMemcpy(b.state + (uint64_t)v * b.P + h.p_off, h.state.data() + (size_t)v * h.n,
                        h.n * sizeof(double), hipMemcpyHostToDevice));
  }
  RunDev& d = e->h_runs[r];
  d = RunDev{};
  h.map_base = (uint32_t)e->h_maps.size();
  for (auto& mp : h.maps) {
    exit_thresholds(h, mp);
    MapDev md{};
    md.m = mp.m;
    if ((rc = run_upload(h, (double**)&md.p_move, mp.p_move.data(), mp.p_move.size()))) return rc;
    if ((rc = run_upload(h, (double**)&md.q_exit, mp.q_exit.data(), mp.q_exit.size()))) return rc;
    if ((rc = run_upload(h, (double**)&md.cdf, mp.cdf.data(), mp.cdf.size()))) return rc;
    if ((rc = run_upload(h, (uint32_t**)&md.nbr, mp.nbr.data(), mp.nbr.size()))) return rc;
    if (mp.m <= MOVE_ROW_M) {  // one 64-byte record per compartment for the movers
      std::vector<uint32_t> rows((size_t)h.n_c * 16, 0u);
      for (uint32_t i = 0; i < h.n_c; ++i) {
        uint32_t* w = rows.data() + (size_t)i * 16;
        for (uint32_t c = 0; c < 5; ++c) {
          const double v = c + 1 < mp.m ? mp.cdf[(size_t)i * mp.m + c] : 2.0;
          std::memcpy(w + 2 * c, &v, sizeof v);
        }
        for (uint32_t k = 0; k < MOVE_ROW_M; ++k) w[10 + k] = mp.nbr[(size_t)i * mp.m + (k < mp.m ? k : mp.m - 1)];
      }
      if ((rc = run_upload(h, (uint32_t**)&md.move_row, rows.data(), rows.size()))) return rc;
    }
    if ((rc = run_upload(h, (uint32_t**)&md.col_ptr, mp.col_ptr.data(), mp.col_ptr.size()))) return rc;
    if ((rc = run_upload(h, (uint32_t**)&md.src, mp.src.data(), mp.src.size()))) return rc;
    if ((rc = run_upload(h, (double**)&md.val, mp.val.data(), mp.val.size()))) return rc;
    if ((rc = run_upload(h, (double**)&md.inv_v, mp.inv_v.data(), mp.inv_v.size()))) return rc;
    if ((rc = run_upload(h, (double**)&md.vol, mp.V.data(), mp.V.size()))) return rc;
    if (h.two_phase) {
      const size_t k = e->h_maps.size() - h.map_base;
      const HostMap& g = h.gmaps[k];
      const std::vector<double> kla = gas_kla(h, k);
      if ((rc = run_upload(h, (uint32_t**)&md.g_col_ptr, g.col_ptr.data(), g.col_ptr.size()))) return rc;
      if ((rc = run_upload(h, (uint32_t**)&md.g_src, g.src.data(), g.src.size()))) return rc;
      if ((rc = run_upload(h, (double**)&md.g_val, g.val.data(), g.val.size()))) return rc;
      if ((rc = run_upload(h, (double**)&md.g_inv_v, g.inv_v.data(), g.inv_v.size()))) return rc;
      if ((rc = run_upload(h, (double**)&md.g_kla, kla.data(), kla.size()))) return rc;
    }
    e->h_maps.push_back(md);
  }
  const size_t nsc = (size_t)h.n_s * h.n_c;
  std::vector<double> C0 = h.C0.empty() ? std::vector<double>(nsc, 0.0) : h.C0;
  std::vector<double> m0(nsc);
  for (uint32_t sp = 0; sp < h.n_s; ++sp)
    for (uint32_t j = 0; j < h.n_c; ++j) m0[(size_t)sp * h.n_c + j] = C0[(size_t)sp * h.n_c + j] * h.maps[0].V[j];
  std::vector<double> F, sink;
  feed_tables(h, F, sink);
  const uint32_t rw = (uint32_t)rec_stride(e->model->nr);
  std::vector<double> rec((size_t)2 * h.n_c * rw, 0.0);  // two tables, by step parity
  std::vector<double> zeros(nsc, 0.0);
  if ((rc = run_upload(h, &d.C[0], C0.data(), nsc))) return rc;
  if ((rc = run_upload(h, &d.C[1], C0.data(), nsc))) return rc;
  if ((rc = run_upload(h, &d.m, h.restored ? h.r_m.data() : m0.data(), nsc))) return rc;
  if ((rc = run_upload(h, &d.S, h.restored ? h.r_S.data() : zeros.data(), nsc))) return rc;
  if ((rc = run_upload(h, (double**)&d.F, F.data(), nsc))) return rc;
  if ((rc = run_upload(h, (double**)&d.sink, sink.data(), sink.size()))) return rc;
  if (h.two_phase) {
    std::vector<double> Fg, sink_g, mg(nsc);
    feed_tables(h, Fg, sink_g, true);
    const std::vector<double>& cg = h.restored ? h.r_cg : h.Cg0;
    for (uint32_t sp = 0; sp < h.n_s; ++sp)
      for (uint32_t j = 0; j < h.n_c; ++j) mg[(size_t)sp * h.n_c + j] = cg[(size_t)sp * h.n_c + j] * h.gmaps[0].V[j];
    if ((rc = run_upload(h, &d.Cg[0], cg.data(), nsc))) return rc;
    if ((rc = run_upload(h, &d.Cg[1], cg.data(), nsc))) return rc;
    if ((rc = run_upload(h, &d.mg, h.restored ? h.r_mg.data() : mg.data(), nsc))) return rc;
    if ((rc = run_upload(h, (double**)&d.Fg, Fg.data(), nsc))) return rc;
    if ((rc = run_upload(h, (double**)&d.sink_g, sink_g.data(), sink_g.size()))) return rc;
    for (uint32_t sp = 0; sp < h.n_s; ++sp) d.henry[sp] = h.henry[sp];
  }
  h.vf.clear();
  h.vx_out.clear();
  h.vx_feed.clear();
  std::vector<int32_t> vf_of(h.feeds.size(), -1);
  for (uint32_t k = 0; k < h.feeds.size(); ++k)
    if (h.feeds[k].kind != MCST_FEED_CONSTANT) {
      vf_of[k] = (int32_t)h.vf.size();
      h.vf.push_back(k);
    }
  if (!h.vf.empty()) {
    // sources and sinks of the time-dependent feeds by compartment, in feed order
    std::vector<uint32_t> sp_ptr(h.n_c + 1, 0), sk_ptr(h.n_c + 1, 0);
    for (uint32_t q = 0; q < h.vf.size(); ++q) {
      const HostFeed& fd = h.feeds[h.vf[q]];
      sp_ptr[fd.in + 1] += (uint32_t)fd.sp.size();
      if (fd.out >= 0) sk_ptr[(size_t)fd.out + 1] += 1;
    }
    for (uint32_t j = 0; j < h.n_c; ++j) sp_ptr[j + 1] += sp_ptr[j], sk_ptr[j + 1] += sk_ptr[j];
    std::vector<uint32_t> src_k(sp_ptr[h.n_c]), src_s(sp_ptr[h.n_c]), sink_k(sk_ptr[h.n_c]);
    std::vector<double> src_c(sp_ptr[h.n_c]);
    std::vector<uint32_t> fill_s(sp_ptr.begin(), sp_ptr.end() - 1), fill_k(sk_ptr.begin(), sk_ptr.end() - 1);
    for (uint32_t q = 0; q < h.vf.size(); ++q) {
      const HostFeed& fd = h.feeds[h.vf[q]];
      const uint32_t phase = fd.gas ? VF_GAS : 0u;
      for (size_t v = 0; v < fd.sp.size(); ++v) {
        const uint32_t at = fill_s[fd.in]++;
        src_k[at] = q, src_s[at] = fd.sp[v] | phase, src_c[at] = fd.conc[v];
      }
      if (fd.out >= 0) sink_k[fill_k[(size_t)fd.out]++] = q | phase;
    }
    // outlets whose first leaving feed is time-dependent take this step's threshold
    std::vector<int32_t> vx_slot(h.n_c, -1);
    std::vector<char> seen(h.n_c, 0);
    for (uint32_t k = 0; k < h.feeds.size(); ++k) {
      const HostFeed& fd = h.feeds[k];
      if (fd.out < 0 || fd.gas || seen[(size_t)fd.out]) continue;
      seen[(size_t)fd.out] = 1;
      if (vf_of[k] >= 0) {
        vx_slot[(size_t)fd.out] = (int32_t)h.vx_out.size();
        h.vx_out.push_back((uint32_t)fd.out);
        h.vx_feed.push_back((uint32_t)vf_of[k]);
      }
    }
    auto up_u32 = [&](const uint32_t** dst, std::vector<uint32_t> v) {
      if (v.empty()) v.push_back(0u);  // keep every pointer valid
      return run_upload(h, (uint32_t**)dst, v.data(), v.size());
    };
    if (src_c.empty()) src_c.push_back(0.0);
    std::vector<double> zero_vf(h.vf.size() + h.vx_out.size(), 0.0);
    if ((rc = run_upload(h, &h.d_vf, zero_vf.data(), zero_vf.size()))) return rc;
    d.vf_q = h.d_vf;
    if ((rc = up_u32(&d.vf_src_ptr, sp_ptr))) return rc;
    if ((rc = up_u32(&d.vf_src_k, src_k))) return rc;
    if ((rc = up_u32(&d.vf_src_s, src_s))) return rc;
    if ((rc = run_upload(h, (double**)&d.vf_src_c, src_c.data(), src_c.size()))) return rc;
    if ((rc = up_u32(&d.vf_sink_ptr, sk_ptr))) return rc;
    if ((rc = up_u32(&d.vf_sink_k, sink_k))) return rc;
    if ((rc = run_upload(h, (int32_t**)&d.vx_slot, vx_slot.data(), vx_slot.size()))) return rc;
    d.n_vf = (uint32_t)h.vf.size();
  }
  if (!h.kla.empty()) {
    if ((rc = run_upload(h, (double**)&d.kla, h.kla.data(), nsc))) return rc;
    if ((rc = run_upload(h, (double**)&d.cstar, h.cstar.data(), nsc))) return rc;
  }
  if ((rc = run_upload(h, &d.rec, rec.data(), rec.size()))) return rc;
  if (h.probe_cap) {
    if ((rc = run_alloc(h, (void**)&d.probe_exit, (size_t)h.probe_cap * sizeof(double)))) return rc;
    if ((rc = run_alloc(h, (void**)&d.probe_div, (size_t)h.probe_cap * sizeof(double)))) return rc;
    const size_t ke = std::min<size_t>(h.r_exit.size(), h.probe_cap), kd = std::min<size_t>(h.r_div.size(), h.probe_cap);
    if (ke) HIP_TRY(hipMemcpy(d.probe_exit, h.r_exit.data(), ke * sizeof(double), hipMemcpyHostToDevice));
    if (kd) HIP_TRY(hipMemcpy(d.probe_div, h.r_div.data(), kd * sizeof(double), hipMemcpyHostToDevice));
  }
  d.p_off = h.p_off;
  d.seed = h.seed;
  d.dt = h.dt;
  d.weight = h.weight;
  d.ratio = h.ratio;
  d.min_removal = h.min_removal;
  std::memcpy(d.params, h.params, sizeof(d.params));
  d.cap = (uint32_t)h.cap;
  d.n_c = h.n_c;
  d.n_s = h.n_s;
  d.rec_w = rw;
  d.enable_move = h.n_c > 1 ? 1u : 0u;
  d.enable_leave = 0;
  for (const auto& fd : h.feeds)
    if (fd.out >= 0 && !fd.gas) d.enable_leave = 1;
  d.n_maps = (uint32_t)h.maps.size();
  d.steps_per_map = h.steps_per_map;
  d.map_base = h.map_base;
  d.div_cap = h.div_cap;
  d.comp_off = 0;
  d.probe_cap = h.probe_cap;
  d.step_target = h.step_target;
  RunState st;
  std::memset(&st, 0, sizeof(st));
  st.n = (uint32_t)h.n;
  uint64_t inactive = 0;
  for (uint64_t q = 0; q < h.n; ++q) inactive += h.status[q] != IDLE;
  st.inactive = (uint32_t)inactive;
  h.host_step = h.restored ? h.r_step : 0u;
  if (h.restored) {
    std::memcpy(st.counters, h.r_counters, sizeof(st.counters));
    st.step = h.r_step;
    st.probe_exit_n = std::min<uint64_t>(h.r_exit.size(), h.probe_cap);
    st.probe_div_n = std::min<uint64_t>(h.r_div.size(), h.probe_cap);
  }
  st.counters[C_N] = h.n;
  st.counters[C_INACTIVE] = inactive;
  HIP_TRY(hipMemcpy(e->d_runs + r, &d, sizeof(RunDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->d_rs + r, &st, sizeof(RunState), hipMemcpyHostToDevice));
  // host copies of the initial particles are no longer needed
  std::vector<uint32_t>().swap(h.pos);
  std::vector<uint8_t>().swap(h.status);
  std::vector<double>().swap(h.state);
  std::vector<double>().swap(h.ah);
  std::vector<double>().swap(h.ad);
  std::vector<double>().swap(h.r_m);
  std::vector<double>().swap(h.r_S);
  std::vector<double>().swap(h.r_exit);
  std::vector<double>().swap(h.r_div);
  std::vector<double>().swap(h.r_cg);
  std::vector<double>().swap(h.r_mg);
  h.phase = 1;
  return MCST_OK;
}

// Admit the defined runs in id order while the pool has room: they take
// part from the next step on.  The stream is idle during the update.
int admit(mcst_engine* e, uint32_t* n_admitted) {
  HIP_TRY(hipStreamSynchronize(e->stream));
  uint32_t k = 0;
  int rc = MCST_OK;
  while (e->admitted < e->runs.size()) {
    const uint32_t r = e->admitted;
    HostRun& h = e->runs[r];
    if ((rc = check_definable(e, h, r))) break;
    if (r >= e->runs_cap) {
      rc = fail(MCST_ERR_MEMORY, "run table full");
      break;
    }
    const uint32_t tiles = (uint32_t)((h.cap + TILE - 1) / TILE);
    uint32_t start = 0;
    if (!take_tiles(e, tiles, start)) {
      rc = fail(MCST_ERR_MEMORY, "particle pool full");
      break;
    }
    h.p_off = (uint64_t)start * TILE;
    h.tiles = tiles;
    if ((rc = upload_run(e, r))) {
      give_tiles(e, start, tiles);
      free_run_memory(h);
      break;
    }
    ++e->admitted;
    ++k;
  }
  if (n_admitted) *n_admitted = k;
  if (k) {
    drop_graph(e);
    const int rc2 = refresh_tables(e);
    if (rc2) return rc2;
  }
  return rc;
}

// Entries of the request block counts of a pool of T tiles (blocks of 2^6,
// 2^12 and 2^18 sub-chunks; the resolve kernel clears entry n_sc >> 18 too).
struct RequestLevels {
  uint64_t n1, n2, n3;
};
RequestLevels request_levels(uint64_t T, uint32_t nsub) {
  const uint64_t n_sc = T * nsub;
  return RequestLevels{(n_sc + 63) >> 6, (n_sc + 4095) >> 12, (n_sc >> 18) + 1};
}

int commit(mcst_engine* e) {
  const ModelEntry& M = *e->model;
  const uint32_t R0 = (uint32_t)e->runs.size();
  if (R0 == 0 && e->reserve_runs == 0) return fail(MCST_ERR_STATE, "no run defined");
  uint64_t tiles = 0;
  for (uint32_t r = 0; r < R0; ++r) tiles += (e->runs[r].cap + TILE - 1) / TILE;
  tiles = std::max<uint64_t>(tiles, (e->reserve_slots + TILE - 1) / TILE);
  if (tiles == 0) tiles = 1;
  const uint64_t P = tiles * TILE;
  if (P >= 0xFFFFFFFFull) return fail(MCST_ERR_ARG, "more than 2^32-1 particle slots in one engine");
  const uint64_t T_max = tiles;
  const uint32_t nv = (uint32_t)M.nv;
  BatchDev& b = e->bd;
  b = BatchDev{};
  b.P = P;
  b.T_max = (uint32_t)T_max;
  b.R = 0;
  b.n_comp_total = 0;
  b.nsub = TILE / (64u * (uint32_t)M.nl);
  b.ns_max = std::max<uint32_t>(1, (uint32_t)M.ns);  // one model per engine; thread s == 0 also writes the map fields
  b.stages = STAGE_ALL;
  e->runs_cap = std::max(R0, e->reserve_runs);
  const uint32_t RC = e->runs_cap;
  int rc;
  if ((rc = dev_alloc(e, (void**)&b.pos, P * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.status, P))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.state, P * nv * sizeof(double)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.age_hyd, P * sizeof(double)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.age_div, P * sizeof(double)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.clk_hyd, P * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.clk_div, P * sizeof(uint32_t)))) return rc;
  uint32_t** tile_arrays[] = {&b.tile_gap, &b.tile_fill, &b.tile_gap_off, &b.tile_fill_off};
  for (uint32_t** a : tile_arrays)
    if ((rc = dev_alloc(e, (void**)a, T_max * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.fill_src, P * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.tile_off, (RC + 1) * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.move_rec, T_max * TILE * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.div_rec, T_max * TILE * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.div_pos, T_max * TILE * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.div_stage, T_max * b.nsub * DIV_STAGE * nv * sizeof(double)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.sub_moves, T_max * b.nsub * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.sub_divs, T_max * b.nsub * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&e->d_runs, (size_t)RC * sizeof(RunDev)))) return rc;
  if ((rc = dev_alloc(e, (void**)&e->d_rs, (size_t)RC * sizeof(RunState)))) return rc;
  if ((rc = dev_alloc(e, (void**)&e->d_comp, (size_t)(RC + 1) * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.rm_tile_off, (RC + 1) * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.ticket, (1 + RESOLVE_TICKETS) * sizeof(uint32_t)))) return rc;
  const RequestLevels lv = request_levels(T_max, b.nsub);
  if ((rc = dev_alloc(e, (void**)&b.div_l1, lv.n1 * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.div_l2, lv.n2 * sizeof(uint32_t)))) return rc;
  if ((rc = dev_alloc(e, (void**)&b.div_l3, lv.n3 * sizeof(uint32_t)))) return rc;
  HIP_TRY(hipMemset(b.div_l1, 0, lv.n1 * sizeof(uint32_t)));
  HIP_TRY(hipMemset(b.div_l2, 0, lv.n2 * sizeof(uint32_t)));
  HIP_TRY(hipMemset(b.div_l3, 0, lv.n3 * sizeof(uint32_t)));
  HIP_TRY(hipMemset(b.ticket, 0, (1 + RESOLVE_TICKETS) * sizeof(uint32_t)));
  HIP_TRY(hipMemset(b.rm_tile_off, 0, (RC + 1) * sizeof(uint32_t)));
  HIP_TRY(hipMemset(b.tile_off, 0, (RC + 1) * sizeof(uint32_t)));
  HIP_TRY(hipMemset(b.sub_moves, 0, T_max * b.nsub * sizeof(uint32_t)));
  HIP_TRY(hipMemset(b.sub_divs, 0, T_max * b.nsub * sizeof(uint32_t)));
  HIP_TRY(hipMemset(e->d_rs, 0, (size_t)RC * sizeof(RunState)));
  b.runs = e->d_runs;
  b.rs = e->d_rs;
  b.comp_off = e->d_comp;
  e->h_runs.assign(RC, RunDev{});
  e->free_tiles.assign(1, {0u, (uint32_t)T_max});
  e->admitted = 0;
  // with reserved capacity, runs that do not fit yet stay defined (admit them later)
  rc = admit(e, nullptr);
  if (rc && !(rc == MCST_ERR_MEMORY && e->reserve_runs)) return rc;
  if (e->admitted < R0 && !e->reserve_runs) return fail(MCST_ERR_MEMORY, "the runs do not fit the engine");
  if (R0 == 0 && (rc = refresh_tables(e))) return rc;
  launch_layout(b, e->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->committed = true;
  return MCST_OK;
}

// Take a finished run off the device: its slots and arrays return to the
// pool, and it stops taking part in the batch.  The id is not reused.
int retire(mcst_engine* e, uint32_t r) {
  HIP_TRY(hipStreamSynchronize(e->stream));
  HostRun& h = e->runs[r];
  RunDev& d = e->h_runs[r];
  d.step_target = 0;  // inactive from now on: no tiles, no transport
  HIP_TRY(hipMemcpy(e->d_runs + r, &d, sizeof(RunDev), hipMemcpyHostToDevice));
  RunState st;
  HIP_TRY(hipMemcpy(&st, e->d_rs + r, sizeof(RunState), hipMemcpyDeviceToHost));
  st.n = 0;
  st.active = 0;
  st.compact = 0;
  HIP_TRY(hipMemcpy(e->d_rs + r, &st, sizeof(RunState), hipMemcpyHostToDevice));
  give_tiles(e, (uint32_t)(h.p_off / TILE), h.tiles);
  h.phase = 2;
  drop_graph(e);
  int rc = refresh_tables(e);
  free_run_memory(h);
  return rc;
}

// The kernels of one step of the batch.  With `ev`, records ev[0] before
// the transport and ev[k] after the k-th stage (transport, particle pass,
// resolve with the bookkeeping, removal) for mcst_engine_profile.
// `transported`: the step's transport (and tile layout) already ran, inside
// the previous step's resolve kernel.  `ahead`: run the next step's transport
// inside this step's resolve kernel.
void launch_step(mcst_engine* e, uint32_t stages, const hipEvent_t* ev = nullptr, bool transported = false,
                 bool ahead = false) {
  if (e->bd.R == 0) return;  // no run admitted yet
  BatchDev b = e->bd;
  b.stages = stages;
  const uint32_t nv = (uint32_t)e->model->nv;
  const ModelOps& ops = *e->model->ops;
  auto mark = [&](int k) {
    if (ev) (void)hipEventRecord(ev[k], e->stream);
  };
  mark(0);
  if (!transported) ops.transport(b, e->grid_transport, e->stream);
  mark(1);
  ops.particle(b, e->plaunch, e->stream);
  mark(2);
  // aux blocks of the resolve kernel: with contribution rows, the blocks
  // that sum them (and transport the next step with the sums when ahead);
  // else, when ahead, the next step's transport
  const uint32_t aux = b.part ? e->part_aux : ahead ? (uint32_t)(e->grid_transport - 1) : 0u;
  ops.resolve(b, e->grid_move, aux, ahead, e->stream);
  mark(3);
  if (e->removal_possible) launch_removal(b, nv, e->grid_particle, e->stream);
  mark(4);
}

// Whether consecutive steps of a call may run each next step's transport
// inside the resolve kernel: not when the host writes flows before every
// step (time-dependent feeds) or may re-lay the runs between steps (growth),
// and only for whole steps.
bool can_transport_ahead(const mcst_engine* e, uint32_t stages);

int read_state(mcst_engine* e, int32_t run, RunState& s) {
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(&s, e->d_rs + run, sizeof(RunState), hipMemcpyDeviceToHost));
  return MCST_OK;
}

// The hydraulic and division ages of slots [o, o + n) of a run in state s
// (read_state, the stream idle), from the device's base ages and age-clock
// origins (clock_age, the kernels' expression): an Idle slot ages with its
// run's clocks, another keeps the ages it had when it stopped being Idle.
// The base ages are read into the outputs and converted there; an output
// that is null or not 8-byte aligned goes through a host buffer.
int read_ages(mcst_engine* e, uint64_t o, uint64_t n, const RunState& s, double dt, void* age_hyd, void* age_div) {
  const BatchDev& b = e->bd;
  std::vector<uint8_t> st(n);
  std::vector<uint32_t> ch(n), cd(n);
  std::vector<double> th, td;
  auto place = [&](void* out, std::vector<double>& tmp) -> double* {
    if (out && (uintptr_t)out % alignof(double) == 0) return (double*)out;
    tmp.resize(n);
    return tmp.data();
  };
  double* ah = place(age_hyd, th);
  double* ad = place(age_div, td);
  HIP_TRY(hipMemcpy(st.data(), b.status + o, n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(ah, b.age_hyd + o, n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(ad, b.age_div + o, n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(ch.data(), b.clk_hyd + o, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(cd.data(), b.clk_div + o, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (uint64_t q = 0; q < n; ++q) {
    if (st[q] == IDLE) {
      ah[q] = clock_age(ah[q], s.clock_hyd, ch[q], dt);
      ad[q] = div_age(ad[q], s.clock_div, cd[q], dt);
    } else {
      ad[q] = div_age(ad[q], ch[q], cd[q], dt);
    }
  }
  if (age_hyd && ah != age_hyd) std::memcpy(age_hyd, ah, n * sizeof(double));
  if (age_div && ad != age_div) std::memcpy(age_div, ad, n * sizeof(double));
  return MCST_OK;
}

// Frees the second set of particle arrays (the stream idle).
void release_alt(mcst_engine* e) {
  void* arrays[] = {e->alt_pos, e->alt_status, e->alt_state, e->alt_age_hyd,
                    e->alt_age_div, e->alt_clk_hyd, e->alt_clk_div};
  for (void* p : arrays) {
    if (!p) continue;
    (void)hipFree(p);
    e->allocs.erase(std::remove(e->allocs.begin(), e->allocs.end(), p), e->allocs.end());
  }
  e->alt_pos = nullptr;
  e->alt_status = nullptr;
  e->alt_state = e->alt_age_hyd = e->alt_age_div = nullptr;
  e->alt_clk_hyd = e->alt_clk_div = nullptr;
  e->alt_P = 0;
}

// The second set of particle arrays, sized like the pool's; false (and no
// set) when the device cannot hold it.
bool ensure_alt(mcst_engine* e) {
  const uint64_t P = e->bd.P, nv = (uint64_t)e->model->nv;
  if (e->alt_P == P) return true;
  release_alt(e);
  if (dev_alloc(e, (void**)&e->alt_pos, P * sizeof(uint32_t)) || dev_alloc(e, (void**)&e->alt_status, P) ||
      dev_alloc(e, (void**)&e->alt_state, P * nv * sizeof(double)) ||
      dev_alloc(e, (void**)&e->alt_age_hyd, P * sizeof(double)) ||
      dev_alloc(e, (void**)&e->alt_age_div, P * sizeof(double)) ||
      dev_alloc(e, (void**)&e->alt_clk_hyd, P * sizeof(uint32_t)) ||
      dev_alloc(e, (void**)&e->alt_clk_div, P * sizeof(uint32_t))) {
    release_alt(e);
    return false;
  }
  e->alt_P = P;
  return true;
}

// Reorders run r's particles by compartment, stable in slot order
// (sort.hip).  Reads the particle count, so it waits for the stream.  The
// engine's only live run is gathered once into the second set of arrays,
// which then becomes the pool's (a captured graph, which holds the old
// pointers, is dropped); a run of a batch is gathered into scratch and
// copied back.
int sort_run(mcst_engine* e, uint32_t r) {
  RunState st;
  int rc = read_state(e, (int32_t)r, st);
  if (rc) return rc;
  const HostRun& h = e->runs[r];
  const uint64_t n = st.n;
  if (n < 2 || h.n_c < 2) return MCST_OK;
  uint32_t end_bit = 1;
  while (end_bit < 32 && ((uint64_t)1 << end_bit) < h.n_c) ++end_bit;
  size_t need = 0;
  HIP_TRY(sort_temp_bytes(n, end_bit, &need));
  if (n > e->sort_n || need > e->sort_temp) {  // grow the scratch (with headroom for later calls)
    const uint64_t cap = std::max<uint64_t>(n + n / 4, e->sort_n);
    size_t temp = 0;
    HIP_TRY(sort_temp_bytes(cap, 32, &temp));  // the widest keys: enough for every run
    temp = std::max(temp, need);
    HIP_TRY(hipStreamSynchronize(e->stream));  // the old scratch may still be in use
    if (e->sort_mem) HIP_TRY(hipFree(e->sort_mem));
    e->sort_mem = nullptr;
    e->sort_n = 0;
    e->sort_temp = 0;
    HIP_TRY(hipMalloc(&e->sort_mem, cap * (3 * sizeof(uint32_t) + sizeof(double)) + temp + 256));
    e->sort_n = cap;
    e->sort_temp = temp;
  }
  BatchDev& b = e->bd;
  const uint64_t o = h.p_off, cap = e->sort_n;
  char* m = (char*)e->sort_mem;
  SortArgs a;
  a.n = n;
  a.end_bit = end_bit;
  a.pos = b.pos + o;
  a.status = b.status + o;
  const uint32_t nv = (uint32_t)e->model->nv;
  for (uint32_t v = 0; v < nv; ++v) a.f64[v] = b.state + (uint64_t)v * b.P + o;
  a.f64[nv] = b.age_hyd + o;
  a.f64[nv + 1] = b.age_div + o;
  a.n_f64 = nv + 2;
  a.u32[0] = b.clk_hyd + o;
  a.u32[1] = b.clk_div + o;
  a.n_u32 = 2;
  a.keys = (uint32_t*)m;
  a.iota = a.keys + cap;
  a.perm = a.iota + cap;
  a.buf = (void*)(((uintptr_t)(a.perm + cap) + 255) & ~(uintptr_t)255);
  a.temp = (char*)a.buf + cap * sizeof(double);
  a.temp_bytes = e->sort_temp;
  uint32_t live = 0;
  for (const auto& x : e->runs) live += x.phase == 1 ? 1u : 0u;
  const bool swap = live == 1 && ensure_alt(e);
  if (swap) {
    a.to_dst = true;
    a.dst_pos = e->alt_pos + o;
    a.dst_status = e->alt_status + o;
    for (uint32_t v = 0; v < nv; ++v) a.dst_f64[v] = e->alt_state + (uint64_t)v * b.P + o;
    a.dst_f64[nv] = e->alt_age_hyd + o;
    a.dst_f64[nv + 1] = e->alt_age_div + o;
    a.dst_u32[0] = e->alt_clk_hyd + o;
    a.dst_u32[1] = e->alt_clk_div + o;
  }
  HIP_TRY(sort_run_particles(a, e->stream));
  if (swap) {  // later launches read the sorted set; the old one is the next reorder's target
    std::swap(b.pos, e->alt_pos);
    std::swap(b.status, e->alt_status);
    std::swap(b.state, e->alt_state);
    std::swap(b.age_hyd, e->alt_age_hyd);
    std::swap(b.age_div, e->alt_age_div);
    std::swap(b.clk_hyd, e->alt_clk_hyd);
    std::swap(b.clk_div, e->alt_clk_div);
    drop_graph(e);
  }
  return MCST_OK;
}

// Grows the particle pool to at least `min_tiles` tiles: the particle arrays
// keep their slots (the state's field stride becomes the new pool size),
// the per-step lists and staging rows are reallocated empty.  Every new
// array is allocated before any old one is touched, so a failed allocation
// leaves the engine as it was.
int grow_pool(mcst_engine* e, uint64_t min_tiles) {
  BatchDev& b = e->bd;
  HIP_TRY(hipStreamSynchronize(e->stream));
  release_alt(e);  // sized for the old pool
  const uint64_t T0 = b.T_max, P0 = b.P;
  const uint64_t T1 = std::max<uint64_t>(min_tiles, 2 * T0), P1 = T1 * TILE;
  if (P1 >= 0xFFFFFFFFull) return fail(MCST_ERR_MEMORY, "more than 2^32-1 particle slots in one engine");
  const uint64_t nv = (uint64_t)e->model->nv;
  const RequestLevels lv = request_levels(T1, b.nsub);
  struct Move {
    void** slot;     // the BatchDev field
    size_t bytes;    // new size
    size_t keep;     // bytes copied from the old array (per field for the state)
    uint64_t fields; // > 0: re-strided [fields][P] array
    void* fresh;
  };
  std::vector<Move> mv = {
      {(void**)&b.pos, P1 * 4, P0 * 4, 0, nullptr},
      {(void**)&b.status, P1, P0, 0, nullptr},
      {(void**)&b.age_hyd, P1 * 8, P0 * 8, 0, nullptr},
      {(void**)&b.age_div, P1 * 8, P0 * 8, 0, nullptr},
      {(void**)&b.clk_hyd, P1 * 4, P0 * 4, 0, nullptr},
      {(void**)&b.clk_div, P1 * 4, P0 * 4, 0, nullptr},
      {(void**)&b.state, P1 * nv * 8, P0 * 8, nv, nullptr},
      {(void**)&b.div_l1, lv.n1 * 4, 0, 0, nullptr},
      {(void**)&b.div_l2, lv.n2 * 4, 0, 0, nullptr},
      {(void**)&b.div_l3, lv.n3 * 4, 0, 0, nullptr},
      {(void**)&b.tile_gap, T1 * 4, 0, 0, nullptr},
      {(void**)&b.tile_fill, T1 * 4, 0, 0, nullptr},
      {(void**)&b.tile_gap_off, T1 * 4, 0, 0, nullptr},
      {(void**)&b.tile_fill_off, T1 * 4, 0, 0, nullptr},
      {(void**)&b.fill_src, P1 * 4, 0, 0, nullptr},
      {(void**)&b.move_rec, P1 * 4, 0, 0, nullptr},
      {(void**)&b.div_rec, P1 * 4, 0, 0, nullptr},
      {(void**)&b.div_pos, P1 * 4, 0, 0, nullptr},
      {(void**)&b.div_stage, T1 * b.nsub * DIV_STAGE * nv * 8, 0, 0, nullptr},
      {(void**)&b.sub_moves, T1 * b.nsub * 4, 0, 0, nullptr},
      {(void**)&b.sub_divs, T1 * b.nsub * 4, 0, 0, nullptr},
  };
  for (auto& m : mv) {
    const hipError_t err = hipMalloc(&m.fresh, m.bytes);
    if (err != hipSuccess) {
      for (auto& q : mv)
        if (q.fresh) (void)hipFree(q.fresh);
      return fail(MCST_ERR_MEMORY, std::string("growing the particle pool: ") + hipGetErrorString(err));
    }
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  for (auto& m : mv) {
    HIP_TRY(hipMemsetAsync(m.fresh, 0, m.bytes, e->stream));
    if (m.fields) {
      for (uint64_t v = 0; v < m.fields; ++v)
        HIP_TRY(hipMemcpyAsync((char*)m.fresh + v * P1 * 8, (char*)*m.slot + v * P0 * 8, m.keep,
                               hipMemcpyDeviceToDevice, e->stream));
    } else if (m.keep) {
      HIP_TRY(hipMemcpyAsync(m.fresh, *m.slot, m.keep, hipMemcpyDeviceToDevice, e->stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  for (auto& m : mv) {
    for (auto& a : e->allocs)
      if (a == *m.slot) a = m.fresh;
    (void)hipFree(*m.slot);
    *m.slot = m.fresh;
  }
  b.P = P1;
  b.T_max = (uint32_t)T1;
  give_tiles(e, (uint32_t)T0, (uint32_t)(T1 - T0));
  return MCST_OK;
}

// Takes `n` free tiles starting exactly at tile `at`; false if they are not free.
bool take_tiles_at(mcst_engine* e, uint32_t at, uint32_t n) {
  if (n == 0) return true;
  for (size_t i = 0; i < e->free_tiles.size(); ++i) {
    auto& f = e->free_tiles[i];
    if (f.first != at || f.second < n) continue;
    f.first += n;
    f.second -= n;
    if (f.second == 0) e->free_tiles.erase(e->free_tiles.begin() + (long)i);
    return true;
  }
  return false;
}

// Gives run r (n particles) capacity new_cap: in place when the tiles after
// its range are free, else in a new range (taken from the pool, grown if
// needed) to which its particles are copied.
int regrow_run(mcst_engine* e, uint32_t r, uint64_t n, uint64_t new_cap) {
  HostRun& h = e->runs[r];
  const uint32_t start0 = (uint32_t)(h.p_off / TILE), tiles0 = h.tiles;
  const uint32_t tiles1 = (uint32_t)((new_cap + TILE - 1) / TILE);
  uint32_t start1 = start0;
  if (tiles1 > tiles0 && !take_tiles_at(e, start0 + tiles0, tiles1 - tiles0)) {
    if (!take_tiles(e, tiles1, start1)) {
      int rc = grow_pool(e, (uint64_t)e->bd.T_max + tiles1);
      if (rc) return rc;
      if (!take_tiles(e, tiles1, start1)) return fail(MCST_ERR_MEMORY, "particle pool full after growing it");
    }
    BatchDev& b = e->bd;
    const uint64_t o0 = h.p_off, o1 = (uint64_t)start1 * TILE;
    HIP_TRY(hipMemcpyAsync(b.pos + o1, b.pos + o0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(b.status + o1, b.status + o0, n, hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(b.age_hyd + o1, b.age_hyd + o0, n * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(b.age_div + o1, b.age_div + o0, n * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(b.clk_hyd + o1, b.clk_hyd + o0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(b.clk_div + o1, b.clk_div + o0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, e->stream));
    for (uint32_t v = 0; v < (uint32_t)e->model->nv; ++v)
      HIP_TRY(hipMemcpyAsync(b.state + (uint64_t)v * b.P + o1, b.state + (uint64_t)v * b.P + o0, n * sizeof(double),
                             hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    give_tiles(e, start0, tiles0);
  }
  h.p_off = (uint64_t)start1 * TILE;
  h.tiles = std::max(tiles0, tiles1);
  h.cap = new_cap;
  RunDev& d = e->h_runs[r];
  d.p_off = h.p_off;
  d.cap = (uint32_t)new_cap;
  HIP_TRY(hipMemcpy(e->d_runs + r, &d, sizeof(RunDev), hipMemcpyHostToDevice));
  return MCST_OK;
}

// Container growth (merge_buffer / _resize, particles_container.hpp:575-643,
// whose allocation factor is 2.5): before every grow_every-th step, a run
// with fewer free slots than particles gets capacity ceil(n * grow_factor).
// The checks fall on fixed steps, so results do not depend on timing (the
// division budget depends on the free capacity); the oracle applies the
// same rule.
int grow_runs(mcst_engine* e) {
  bool any = false;
  for (auto& h : e->runs)
    if (h.phase == 1 && h.grow_every && h.host_step < h.step_target && h.host_step % h.grow_every == 0) any = true;
  if (!any) return MCST_OK;
  HIP_TRY(hipStreamSynchronize(e->stream));
  bool grown = false;
  for (uint32_t r = 0; r < e->runs.size(); ++r) {
    HostRun& h = e->runs[r];
    if (h.phase != 1 || !h.grow_every || h.host_step >= h.step_target || h.host_step % h.grow_every) continue;
    RunState st;
    HIP_TRY(hipMemcpy(&st, e->d_rs + r, sizeof(RunState), hipMemcpyDeviceToHost));
    const uint64_t n = st.n;
    if (h.cap - n >= n) continue;
    const uint64_t new_cap = (uint64_t)std::ceil((double)n * h.grow_factor);
    if (new_cap <= h.cap) continue;
    if (new_cap >= 0xFFFFFFFFull) return fail(MCST_ERR_MEMORY, "run capacity above 2^32-1 slots");
    const int rc = regrow_run(e, r, n, new_cap);
    if (rc) return rc;
    grown = true;
  }
  if (grown) {
    drop_graph(e);
    return refresh_tables(e);
  }
  return MCST_OK;
}

bool has_growth(const mcst_engine* e) {
  for (const auto& h : e->runs)
    if (h.phase == 1 && h.grow_every) return true;
  return false;
}

// The host's mirror of the device step counters: k steps of the batch
// advance every live run that has not reached its step target.
void count_steps(mcst_engine* e, uint64_t k) {
  for (auto& h : e->runs)
    if (h.phase == 1) h.host_step = (uint32_t)std::min<uint64_t>(h.step_target, h.host_step + k);
}

// Flows of the time-dependent feeds for the step each live run is about to
// take (t = step * dt, as update_feed), and the exit thresholds of the
// outlets they lead, computed on the host like the constant feeds'.
int prepare_feeds(mcst_engine* e) {
  bool synced = false;
  std::vector<double> v;
  for (auto& h : e->runs) {
    if (h.phase != 1 || h.vf.empty() || h.host_step >= h.step_target) continue;
    if (!synced) {  // the previous step's transport has read the last values
      HIP_TRY(hipStreamSynchronize(e->stream));
      synced = true;
    }
    const double t = (double)h.host_step * h.dt;
    v.assign(h.vf.size() + h.vx_out.size(), 0.0);
    for (size_t q = 0; q < h.vf.size(); ++q) v[q] = feed_flow(h.feeds[h.vf[q]], t);
    const HostMap& mp = h.maps[(h.host_step / h.steps_per_map) % h.maps.size()];
    for (size_t x = 0; x < h.vx_out.size(); ++x) {
      const uint32_t o = h.vx_out[x];
      v[h.vf.size() + x] = std::exp(-(h.dt * v[h.vx_feed[x]]) / mp.V[o]);
    }
    HIP_TRY(hipMemcpy(h.d_vf, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return MCST_OK;
}

bool has_time_feeds(const mcst_engine* e) {
  for (const auto& h : e->runs)
    if (h.phase == 1 && !h.vf.empty()) return true;
  return false;
}

bool can_transport_ahead(const mcst_engine* e, uint32_t stages) {
  return stages == STAGE_ALL && !has_time_feeds(e) && !has_growth(e);
}

// One step of the batch, then the reorders that fall due (transported,
// ahead: as launch_step).
int step_once(mcst_engine* e, uint32_t stages, const hipEvent_t* ev = nullptr, bool transported = false,
              bool ahead = false) {
  const int rcg = grow_runs(e);
  if (rcg) return rcg;
  const int rcf = prepare_feeds(e);
  if (rcf) return rcf;
  std::vector<uint32_t> before(e->runs.size());
  for (size_t r = 0; r < e->runs.size(); ++r) before[r] = e->runs[r].host_step;
  launch_step(e, stages, ev, transported, ahead);
  count_steps(e, 1);
  for (size_t r = 0; r < e->runs.size(); ++r) {
    const HostRun& h = e->runs[r];
    if (h.phase == 1 && h.sort_interval && h.host_step != before[r] && h.host_step % h.sort_interval == 0) {
      const int rc = sort_run(e, (uint32_t)r);
      if (rc) return rc;
    }
  }
  return MCST_OK;
}

}  // namespace
}  // namespace mcst

using namespace mcst;

#define CHECK_ENGINE(e)                                                     \
  do {                                                                      \
    if (!(e)) return fail(MCST_ERR_ARG, "null engine");                     \
  } while (0)
#define CHECK_DEFINE(e, r, h)                                               \
  CHECK_ENGINE(e);                                                          \
  HostRun* h = get_run(e, r);                                               \
  if (!h) return fail(MCST_ERR_ARG, "bad run id");                          \
  if (h->phase != 0) return fail(MCST_ERR_STATE, "run already on the device")
#define CHECK_READY(e, r)                                                   \
  CHECK_ENGINE(e);                                                          \
  if (!(e)->committed) return fail(MCST_ERR_STATE, "engine not committed"); \
  if (!get_run(e, r)) return fail(MCST_ERR_ARG, "bad run id");              \
  if (get_run(e, r)->phase != 1) return fail(MCST_ERR_STATE, "run not on the device")

extern "C" {

int mcst_version(int* major, int* minor, int* patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return MCST_OK;
}

const char* mcst_last_error(void) { return g_err.c_str(); }

int mcst_model_load(const char* path, int* model_id) {
  UdfRegistry& u = udf_registry();
  std::lock_guard<std::mutex> lock(u.mu);
  const ModelEntry* m = nullptr;
  const int rc = load_plugin(u, path, &m);
  if (rc) return rc;
  if (model_id) *model_id = m->id;
  return MCST_OK;
}

int mcst_model_lookup(const char* name, int* model_id, int* n_var, int* n_species, int* n_params) {
  const ModelEntry* m = find_model(name);
  if (!m) return MCST_ERR_ARG;  // find_model set the message
  if (model_id) *model_id = m->id;
  if (n_var) *n_var = m->nv;
  if (n_species) *n_species = m->ns;
  if (n_params) *n_params = m->np;
  return MCST_OK;
}

int mcst_model_default_params(int model_id, double* params, int n) {
  const ModelEntry* m = find_model(model_id);
  if (!m || !params || n < m->np) return fail(MCST_ERR_ARG, "bad model id or parameter buffer");
  double p[MAX_PARAMS] = {};
  m->defaults(p);
  for (int i = 0; i < m->np; ++i) params[i] = p[i];
  return MCST_OK;
}

int mcst_engine_create(const char* model_name, int device, mcst_engine** out) {
  if (!out) return fail(MCST_ERR_ARG, "null output handle");
  *out = nullptr;
  const ModelEntry* m = find_model(model_name);
  if (!m) return MCST_ERR_ARG;  // find_model set the message
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(MCST_ERR_DEVICE, "no such HIP device");
  HIP_TRY(hipSetDevice(device));
  mcst_engine* e = new (std::nothrow) mcst_engine();
  if (!e) return fail(MCST_ERR_MEMORY, "out of host memory");
  e->model = m;
  e->device = device;
  hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (err != hipSuccess) {
    delete e;
    return fail(MCST_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(err));
  }
  *out = e;
  return MCST_OK;
}

int mcst_engine_destroy(mcst_engine** pe) {
  if (!pe || !*pe) return MCST_OK;
  mcst_engine* e = *pe;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  release(e);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  *pe = nullptr;
  return MCST_OK;
}

int mcst_run_create(mcst_engine* e, uint32_t n_compartments, uint32_t n_species, double dt, uint64_t seed,
                    int32_t* run_id) {
  CHECK_ENGINE(e);
  if (e->committed && e->reserve_runs == 0)
    return fail(MCST_ERR_STATE, "engine already committed without reserved capacity");
  if (n_compartments == 0) return fail(MCST_ERR_ARG, "a run needs at least one compartment");
  if ((int)n_species != e->model->ns)
    return fail(MCST_ERR_ARG, "model " + std::string(e->model->name) + " needs " + std::to_string(e->model->ns) +
                                  " species");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(MCST_ERR_ARG, "dt must be positive");
  HostRun h;
  h.n_c = n_compartments;
  h.n_s = n_species;
  h.dt = dt;
  h.seed = seed;
  e->model->defaults(h.params);
  e->runs.push_back(std::move(h));
  if (run_id) *run_id = (int32_t)e->runs.size() - 1;
  return MCST_OK;
}

int mcst_run_set_model_params(mcst_engine* e, int32_t run, const double* params, int n) {
  CHECK_DEFINE(e, run, h);
  if (n < 0 || n > MAX_PARAMS || (n && !params)) return fail(MCST_ERR_ARG, "bad parameter vector");
  for (int i = 0; i < n; ++i) h->params[i] = params[i];
  return MCST_OK;
}

int mcst_run_add_flowmap(mcst_engine* e, int32_t run, const double* volumes, uint64_t nnz, const uint32_t* rows,
                         const uint32_t* cols, const double* flows) {
  CHECK_DEFINE(e, run, h);
  if (!volumes || (nnz && (!rows || !cols || !flows))) return fail(MCST_ERR_ARG, "null flowmap array");
  HostMap mp;
  const int rc = build_map(h->n_c, volumes, nnz, rows, cols, flows, h->dt, mp);
  if (rc) return rc;
  h->maps.push_back(std::move(mp));
  return MCST_OK;
}

int mcst_run_add_flowmap_tables(mcst_engine* e, int32_t run, const double* volumes, uint32_t n_cols,
                                const uint64_t* neighbors, const double* cumulative, const double* out_flows,
                                uint64_t nnz, const uint64_t* rows, const uint64_t* cols, const double* values) {
  CHECK_DEFINE(e, run, h);
  if (!volumes || !neighbors || !cumulative || !out_flows || (nnz && (!rows || !cols || !values)))
    return fail(MCST_ERR_ARG, "null flowmap array");
  HostMap mp;
  const int rc = build_map_tables(h->n_c, volumes, n_cols, neighbors, cumulative, out_flows, nnz, rows, cols, values,
                                  h->dt, mp);
  if (rc) return rc;
  h->maps.push_back(std::move(mp));
  return MCST_OK;
}

int mcst_run_set_steps_per_flowmap(mcst_engine* e, int32_t run, uint32_t steps) {
  CHECK_DEFINE(e, run, h);
  if (steps == 0) return fail(MCST_ERR_ARG, "steps per flowmap must be positive");
  h->steps_per_map = steps;
  return MCST_OK;
}

static int add_feed(mcst_engine* e, int32_t run, bool gas, int kind, double f0, double param, uint32_t input,
                    int64_t output, uint32_t n_values, const uint32_t* species, const double* conc) {
  CHECK_DEFINE(e, run, h);
  if (kind < MCST_FEED_CONSTANT || kind > MCST_FEED_EXPONENTIAL) return fail(MCST_ERR_ARG, "bad feed kind");
  if (input >= h->n_c || output >= (int64_t)h->n_c) return fail(MCST_ERR_ARG, "feed position out of range");
  if (f0 < 0.0) return fail(MCST_ERR_ARG, "negative feed flow");  // feed_descriptor.cpp:53-57
  if (n_values > MAX_SPECIES || (n_values && (!species || !conc))) return fail(MCST_ERR_ARG, "bad feed values");
  HostFeed fd{kind, gas, f0, param, input, output < 0 ? -1 : output, {}, {}};
  for (uint32_t v = 0; v < n_values; ++v) {
    if (species[v] >= h->n_s) return fail(MCST_ERR_ARG, "feed species out of range");
    if (!(conc[v] > 0.0)) return fail(MCST_ERR_ARG, "feed concentration must be positive");
    fd.sp.push_back(species[v]);
    fd.conc.push_back(conc[v]);
  }
  h->feeds.push_back(std::move(fd));
  return MCST_OK;
}

int mcst_run_add_feed(mcst_engine* e, int32_t run, int kind, double f0, double param, uint32_t input,
                      int64_t output, uint32_t n_values, const uint32_t* species, const double* conc) {
  return add_feed(e, run, false, kind, f0, param, input, output, n_values, species, conc);
}

int mcst_run_add_gas_feed(mcst_engine* e, int32_t run, int kind, double f0, double param, uint32_t input,
                          int64_t output, uint32_t n_values, const uint32_t* species, const double* conc) {
  return add_feed(e, run, true, kind, f0, param, input, output, n_values, species, conc);
}

int mcst_run_set_gas_phase(mcst_engine* e, int32_t run, const double* henry, const double* c_gas) {
  CHECK_DEFINE(e, run, h);
  if (!henry || !c_gas) return fail(MCST_ERR_ARG, "null gas phase array");
  const size_t nsc = (size_t)h->n_s * h->n_c;
  for (uint32_t sp = 0; sp < h->n_s; ++sp)
    if (!(henry[sp] >= 0.0) || !std::isfinite(henry[sp])) return fail(MCST_ERR_ARG, "Henry constants must be finite, >= 0");
  for (size_t i = 0; i < nsc; ++i)
    if (!(c_gas[i] >= 0.0) || !std::isfinite(c_gas[i])) return fail(MCST_ERR_ARG, "negative initial gas concentration");
  h->two_phase = true;
  for (uint32_t sp = 0; sp < h->n_s; ++sp) h->henry[sp] = henry[sp];
  h->Cg0.assign(c_gas, c_gas + nsc);
  return MCST_OK;
}

int mcst_run_set_kla(mcst_engine* e, int32_t run, int mode, const double* kla, uint32_t species,
                     double bubble_diameter) {
  CHECK_DEFINE(e, run, h);
  if (mode == MCST_KLA_FIXED) {
    if (!kla) return fail(MCST_ERR_ARG, "null kLa values");
    for (uint32_t sp = 0; sp < h->n_s; ++sp)
      if (!(kla[sp] >= 0.0) || !std::isfinite(kla[sp])) return fail(MCST_ERR_ARG, "kLa must be finite, >= 0");
    for (uint32_t sp = 0; sp < h->n_s; ++sp) h->kla_fixed[sp] = kla[sp];
  } else if (mode == MCST_KLA_TURBULENCE) {
    if (species >= h->n_s) return fail(MCST_ERR_ARG, "kLa species out of range");
    if (!(bubble_diameter > 0.0) || !std::isfinite(bubble_diameter))
      return fail(MCST_ERR_ARG, "bubble diameter must be positive");
    h->kla_species = species;
    h->bubble_d = bubble_diameter;
  } else {
    return fail(MCST_ERR_ARG, "unknown kLa rule");
  }
  h->kla_mode = mode;
  return MCST_OK;
}

int mcst_run_add_gas_flowmap(mcst_engine* e, int32_t run, const double* gas_volumes, uint64_t nnz,
                             const uint32_t* rows, const uint32_t* cols, const double* flows,
                             const double* energy_dissipation) {
  CHECK_DEFINE(e, run, h);
  if (!gas_volumes || (nnz && (!rows || !cols || !flows))) return fail(MCST_ERR_ARG, "null gas flowmap array");
  HostMap g;
  const int rc = build_map(h->n_c, gas_volumes, nnz, rows, cols, flows, h->dt, g);
  if (rc) return rc;
  std::vector<double> eps;
  if (energy_dissipation) {
    for (uint32_t j = 0; j < h->n_c; ++j)
      if (!(energy_dissipation[j] >= 0.0) || !std::isfinite(energy_dissipation[j]))
        return fail(MCST_ERR_ARG, "energy dissipation must be finite, >= 0");
    eps.assign(energy_dissipation, energy_dissipation + h->n_c);
  }
  h->gmaps.push_back(std::move(g));
  h->g_eps.push_back(std::move(eps));
  return MCST_OK;
}

int mcst_run_set_concentrations(mcst_engine* e, int32_t run, const double* C) {
  CHECK_DEFINE(e, run, h);
  h->restored = false;  // a checkpoint restored earlier no longer applies
  if (!C) return fail(MCST_ERR_ARG, "null concentrations");
  const size_t nsc = (size_t)h->n_s * h->n_c;
  for (size_t i = 0; i < nsc; ++i)
    if (!(C[i] >= 0.0)) return fail(MCST_ERR_ARG, "negative initial concentration");
  h->C0.assign(C, C + nsc);
  return MCST_OK;
}

int mcst_run_set_removal(mcst_engine* e, int32_t run, double ratio, uint64_t min_removal) {
  CHECK_DEFINE(e, run, h);
  if (!(ratio >= 0.0)) return fail(MCST_ERR_ARG, "removal ratio must be >= 0");
  h->ratio = ratio;
  h->min_removal = min_removal;
  return MCST_OK;
}

int mcst_run_set_gas_transfer(mcst_engine* e, int32_t run, const double* kla, const double* c_star) {
  CHECK_DEFINE(e, run, h);
  if (!kla || !c_star) return fail(MCST_ERR_ARG, "null gas-liquid transfer arrays");
  const size_t nsc = (size_t)h->n_s * h->n_c;
  for (size_t i = 0; i < nsc; ++i)
    if (!(kla[i] >= 0.0) || !std::isfinite(kla[i]) || !std::isfinite(c_star[i]))
      return fail(MCST_ERR_ARG, "gas-liquid transfer coefficients must be finite and non-negative");
  h->kla.assign(kla, kla + nsc);
  h->cstar.assign(c_star, c_star + nsc);
  return MCST_OK;
}

int mcst_run_set_sort_interval(mcst_engine* e, int32_t run, uint32_t steps) {
  CHECK_ENGINE(e);
  HostRun* h = get_run(e, run);
  if (!h) return fail(MCST_ERR_ARG, "bad run id");
  h->sort_interval = steps;
  return MCST_OK;
}

int mcst_run_sort_particles(mcst_engine* e, int32_t run) {
  CHECK_READY(e, run);
  HIP_TRY(hipSetDevice(e->device));
  return sort_run(e, (uint32_t)run);
}

int mcst_run_set_growth(mcst_engine* e, int32_t run, uint32_t every, double factor) {
  CHECK_ENGINE(e);
  HostRun* h = get_run(e, run);
  if (!h) return fail(MCST_ERR_ARG, "bad run id");
  if (h->phase == 2) return fail(MCST_ERR_STATE, "run retired");
  if (every && !(factor > 1.0 && factor <= 16.0)) return fail(MCST_ERR_ARG, "growth factor must be in (1, 16]");
  h->grow_every = every;
  if (every) h->grow_factor = factor;
  return MCST_OK;
}

int mcst_run_set_division_cap(mcst_engine* e, int32_t run, uint32_t cap) {
  CHECK_DEFINE(e, run, h);
  if (cap == 0) return fail(MCST_ERR_ARG, "division cap must be at least 1");
  h->div_cap = cap;
  return MCST_OK;
}

int mcst_run_set_probes(mcst_engine* e, int32_t run, uint32_t cap) {
  CHECK_DEFINE(e, run, h);
  h->probe_cap = cap;
  return MCST_OK;
}

int mcst_run_set_step_target(mcst_engine* e, int32_t run, uint32_t steps) {
  CHECK_DEFINE(e, run, h);
  h->step_target = steps;
  return MCST_OK;
}

// Stores the drawn particles in compartment order: a stable counting sort by
// position (draw order among equal positions).  Slots of one compartment
// then sit together, so the blocks of the particle pass each read the
// records of a small neighbourhood of compartments that stays in L1, and the
// order decays only as the particles diffuse through the network.
static void order_by_compartment(HostRun& h, uint32_t nv) {
  const uint64_t n = h.n;
  if (n < 2 || h.n_c < 2) return;
  const uint32_t n_c = h.n_c;
  // a stable counting sort in T chunks: per-chunk histograms, offsets in
  // (compartment, chunk) order, then each chunk places its particles in order
  const unsigned T = n < 1000000 ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint64_t chunk = (n + T - 1) / T;
  std::vector<uint64_t> off((size_t)T * n_c, 0);
  std::vector<uint64_t> dst(n);
  auto par = [&](auto&& fn) {
    if (T == 1) {
      fn(0u);
      return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t) th.emplace_back(fn, t);
    for (auto& x : th) x.join();
  };
  par([&](unsigned t) {
    const uint64_t a = t * chunk, z = std::min<uint64_t>(n, a + chunk);
    uint64_t* hist = off.data() + (size_t)t * n_c;
    for (uint64_t p = a; p < z; ++p) ++hist[h.pos[p]];
  });
  uint64_t run = 0;
  for (uint32_t c = 0; c < n_c; ++c)
    for (unsigned t = 0; t < T; ++t) {
      uint64_t& o = off[(size_t)t * n_c + c];
      const uint64_t k = o;
      o = run;
      run += k;
    }
  par([&](unsigned t) {
    const uint64_t a = t * chunk, z = std::min<uint64_t>(n, a + chunk);
    uint64_t* next = off.data() + (size_t)t * n_c;
    for (uint64_t p = a; p < z; ++p) dst[p] = next[h.pos[p]]++;
  });
  std::vector<uint32_t> pos(n);
  par([&](unsigned t) {
    const uint64_t a = t * chunk, z = std::min<uint64_t>(n, a + chunk);
    for (uint64_t p = a; p < z; ++p) pos[dst[p]] = h.pos[p];
  });
  h.pos.swap(pos);
  std::vector<double> x(n);
  for (uint32_t v = 0; v < nv; ++v) {
    double* col = h.state.data() + (size_t)v * n;
    par([&](unsigned t) {
      const uint64_t a = t * chunk, z = std::min<uint64_t>(n, a + chunk);
      for (uint64_t p = a; p < z; ++p) x[dst[p]] = col[p];
    });
    par([&](unsigned t) {
      const uint64_t a = t * chunk, z = std::min<uint64_t>(n, a + chunk);
      std::copy(x.begin() + (long)a, x.begin() + (long)z, col + a);
    });
  }
}

int mcst_run_init_particles(mcst_engine* e, int32_t run, uint64_t n, uint64_t capacity, double x0) {
  CHECK_DEFINE(e, run, h);
  h->restored = false;  // a checkpoint restored earlier no longer applies
  if (h->maps.empty()) return fail(MCST_ERR_STATE, "add a flowmap before the particles");
  if (capacity < n || capacity >= 0xFFFFFFFFull) return fail(MCST_ERR_ARG, "capacity must be in [n, 2^32)");
  const ModelEntry& M = *e->model;
  const uint32_t nv = (uint32_t)M.nv;
  h->n = n;
  h->cap = capacity;
  h->pos.assign(n, 0);
  h->status.assign(n, IDLE);
  h->state.assign((size_t)nv * n, 0.0);
  h->ah.assign(n, 0.0);
  h->ad.assign(n, 0.0);
  std::vector<double> mass(n);
  const uint32_t n_c = h->n_c;
  auto work = [&](uint64_t a, uint64_t z) {
    double x[MAX_VAR];
    for (uint64_t p = a; p < z; ++p) {
      const U4 w = draw4(h->seed, (uint32_t)p, 0u, TAG_INIT, 1u);
      h->pos[p] = (uint32_t)(((uint64_t)w.x * n_c) >> 32);
      M.init(h->params, x, h->seed, (uint32_t)p);
      for (uint32_t v = 0; v < nv; ++v) h->state[(size_t)v * n + p] = x[v];
      mass[p] = M.mass(h->params, x);
    }
  };
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 100000 || hw == 1) {
    work(0, n);
  } else {
    std::vector<std::thread> th;
    const uint64_t chunk = (n + hw - 1) / hw;
    for (unsigned t = 0; t < hw; ++t) {
      const uint64_t a = t * chunk, z = std::min<uint64_t>(n, a + chunk);
      if (a < z) th.emplace_back(work, a, z);
    }
    for (auto& t : th) t.join();
  }
  double total = 0.0;  // summed in draw order (the reduction of mc_init_first)
  for (uint64_t p = 0; p < n; ++p) total += mass[p];
  double vt = 0.0;
  for (uint32_t j = 0; j < n_c; ++j) vt += h->maps[0].V[j];
  h->weight = n ? (x0 * vt) / total : 0.0;
  order_by_compartment(*h, nv);
  h->has_particles = true;
  return MCST_OK;
}

int mcst_run_set_particles(mcst_engine* e, int32_t run, uint64_t n, uint64_t capacity, const uint32_t* pos,
                           const uint8_t* status, const double* state, const double* age_hyd,
                           const double* age_div, double weight) {
  CHECK_DEFINE(e, run, h);
  h->restored = false;  // a checkpoint restored earlier no longer applies
  if (capacity < n || capacity >= 0xFFFFFFFFull) return fail(MCST_ERR_ARG, "capacity must be in [n, 2^32)");
  if (n && (!pos || !status || !state || !age_hyd || !age_div)) return fail(MCST_ERR_ARG, "null particle array");
  const uint32_t nv = (uint32_t)e->model->nv;
  for (uint64_t p = 0; p < n; ++p) {
    if (pos[p] >= h->n_c) return fail(MCST_ERR_ARG, "particle position out of range");
    if (status[p] > DEAD) return fail(MCST_ERR_ARG, "bad particle status");
  }
  h->n = n;
  h->cap = capacity;
  h->pos.assign(pos, pos + n);
  h->status.assign(status, status + n);
  h->state.assign(state, state + (size_t)nv * n);
  h->ah.assign(age_hyd, age_hyd + n);
  h->ad.assign(age_div, age_div + n);
  h->weight = weight;
  h->has_particles = true;
  return MCST_OK;
}

int mcst_engine_commit(mcst_engine* e) {
  CHECK_ENGINE(e);
  if (e->committed) return fail(MCST_ERR_STATE, "engine already committed");
  HIP_TRY(hipSetDevice(e->device));
  const int rc = commit(e);
  if (rc) release(e);
  return rc;
}

int mcst_engine_reserve(mcst_engine* e, uint64_t slots, uint32_t runs) {
  CHECK_ENGINE(e);
  if (e->committed) return fail(MCST_ERR_STATE, "engine already committed");
  if (slots >= 0xFFFFFFFFull || runs > (1u << 20)) return fail(MCST_ERR_ARG, "reservation too large");
  e->reserve_slots = slots;
  e->reserve_runs = runs;
  return MCST_OK;
}

int mcst_engine_admit(mcst_engine* e, uint32_t* n_admitted) {
  CHECK_ENGINE(e);
  if (n_admitted) *n_admitted = 0;
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  HIP_TRY(hipSetDevice(e->device));
  return admit(e, n_admitted);
}

int mcst_engine_retire(mcst_engine* e, int32_t run) {
  CHECK_ENGINE(e);
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  HostRun* h = get_run(e, run);
  if (!h) return fail(MCST_ERR_ARG, "bad run id");
  if (h->phase != 1) return fail(MCST_ERR_STATE, "run not on the device");
  HIP_TRY(hipSetDevice(e->device));
  return retire(e, (uint32_t)run);
}

int mcst_engine_step(mcst_engine* e, uint32_t n_steps) {
  CHECK_ENGINE(e);
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  HIP_TRY(hipSetDevice(e->device));
  // every step but the last transports the next one inside its resolve
  // kernel; the last leaves the liquid as a step ends, so what the host
  // reads between calls is the same either way
  const bool ahead = can_transport_ahead(e, STAGE_ALL);
  for (uint32_t k = 0; k < n_steps; ++k) {
    const int rc = step_once(e, STAGE_ALL, nullptr, ahead && k > 0, ahead && k + 1 < n_steps);
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  return MCST_OK;
}

int mcst_engine_step_stages(mcst_engine* e, uint32_t stages) {
  CHECK_ENGINE(e);
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  if (stages & ~STAGE_ALL) return fail(MCST_ERR_ARG, "unknown stage bits");
  HIP_TRY(hipSetDevice(e->device));
  const int rc = step_once(e, stages);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return MCST_OK;
}

int mcst_engine_capture(mcst_engine* e, uint32_t n_steps) {
  CHECK_ENGINE(e);
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  if (n_steps == 0) return fail(MCST_ERR_ARG, "capture at least one step");
  if (has_time_feeds(e))
    return fail(MCST_ERR_STATE, "a run has time-dependent feeds: their flows are set by the host before each step");
  if (has_growth(e)) return fail(MCST_ERR_STATE, "a run grows its capacity: the host checks it between steps");
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (e->graph_exec) HIP_TRY(hipGraphExecDestroy(e->graph_exec));
  if (e->graph) HIP_TRY(hipGraphDestroy(e->graph));
  e->graph_exec = nullptr;
  e->graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  const bool ahead = can_transport_ahead(e, STAGE_ALL);
  for (uint32_t k = 0; k < n_steps; ++k) launch_step(e, STAGE_ALL, nullptr, ahead && k > 0, ahead && k + 1 < n_steps);
  HIP_TRY(hipStreamEndCapture(e->stream, &e->graph));
  HIP_TRY(hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
  e->graph_steps = n_steps;
  return MCST_OK;
}

int mcst_engine_replay(mcst_engine* e, uint32_t times) {
  CHECK_ENGINE(e);
  if (!e->graph_exec) return fail(MCST_ERR_STATE, "no captured graph");
  HIP_TRY(hipSetDevice(e->device));
  for (uint32_t k = 0; k < times; ++k) HIP_TRY(hipGraphLaunch(e->graph_exec, e->stream));
  count_steps(e, (uint64_t)times * e->graph_steps);
  return MCST_OK;
}

// Times the particle pass over `steps` real steps for each launch shape
// (block size x tile grouping) and keeps the fastest, or the untuned shape
// when none beats it by 2 %.  Results do not depend
// on the shape (every event is decided per slot and per tile; only the order
// of the FP64 contribution atomics changes), so the trial steps are ordinary
// steps of the runs.
int mcst_engine_autotune(mcst_engine* e, uint32_t steps, int32_t* block, int32_t* group, double* particle_ms) {
  CHECK_ENGINE(e);
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  if (steps == 0) return fail(MCST_ERR_ARG, "tune over at least one step");
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipStreamSynchronize(e->stream));
  drop_graph(e);
  // the shape chosen without tuning: kept unless another is 2 % faster (a
  // few trial steps carry that much timing noise)
  e->tuned_block = 0;
  e->tuned_group = -1;
  int rc = choose_launch(e);
  if (rc) return rc;
  const int def_block = e->plaunch.block, def_group = (int)e->bd.tile_group;
  double best = 1e300, def_ms = -1.0;
  int best_block = 0, best_group = -1;
  for (int blk : {256, 512, 1024}) {
    e->tuned_block = blk;
    e->tuned_group = -1;
    if ((rc = choose_launch(e))) break;
    if (e->plaunch.block != blk || (e->plaunch.probe && blk != 256)) continue;  // not resident / same kernel
    for (int g : {0, 1, 2, 4, 8}) {
      e->bd.tile_group = (uint32_t)g;
      double tot = 0.0, per[6] = {};
      if ((rc = mcst_engine_profile(e, steps, &tot, per))) break;
      if (blk == def_block && g == def_group) def_ms = per[1];
      if (per[1] < best) {
        best = per[1];
        best_block = blk;
        best_group = g;
      }
    }
    if (rc) break;
  }
  if (def_ms >= 0.0 && !(best < 0.98 * def_ms)) {
    best = def_ms;
    best_block = def_block;
    best_group = def_group;
  }
  e->tuned_block = best_block;
  e->tuned_group = best_group;
  const int rc2 = choose_launch(e);
  if (rc) return rc;
  if (rc2) return rc2;
  if (block) *block = best_block;
  if (group) *group = best_group;
  if (particle_ms) *particle_ms = best;
  return MCST_OK;
}

int mcst_engine_synchronize(mcst_engine* e) {
  CHECK_ENGINE(e);
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return MCST_OK;
}

int mcst_engine_profile(mcst_engine* e, uint32_t n_steps, double* total_ms, double* per_kernel_ms) {
  CHECK_ENGINE(e);
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  HIP_TRY(hipSetDevice(e->device));
  constexpr int K = 4;
  // every event created here is destroyed on every return path
  struct Events {
    std::vector<hipEvent_t> v;
    ~Events() {
      for (auto& x : v)
        if (x) (void)hipEventDestroy(x);
    }
  } events;
  events.v.assign((size_t)n_steps * (K + 1) + 2, nullptr);
  for (auto& x : events.v) HIP_TRY(hipEventCreate(&x));
  hipEvent_t* ev = events.v.data();
  hipEvent_t t0 = events.v[events.v.size() - 2], t1 = events.v.back();
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipEventRecord(t0, e->stream));
  // the same steps as mcst_engine_step (capacity checks, time-dependent
  // flows, reorders), with events between the kernels
  const uint32_t steps = e->bd.R ? n_steps : 0;
  for (uint32_t s = 0; s < steps; ++s) {
    const int rc = step_once(e, STAGE_ALL, &ev[(size_t)s * (K + 1)]);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(t1, e->stream));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventSynchronize(t1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
  if (total_ms) *total_ms = ms;
  if (per_kernel_ms) {
    for (int k = 0; k < K; ++k) {
      double acc = 0.0;
      for (uint32_t s = 0; s < steps; ++s) {
        float d = 0.f;
        HIP_TRY(hipEventElapsedTime(&d, ev[(size_t)s * (K + 1) + k], ev[(size_t)s * (K + 1) + k + 1]));
        acc += d;
      }
      per_kernel_ms[k] = steps ? acc / steps : 0.0;
    }
  }
  return MCST_OK;
}

int mcst_run_get_counters(mcst_engine* e, int32_t run, uint64_t* counters) {
  CHECK_READY(e, run);
  if (!counters) return fail(MCST_ERR_ARG, "null counter buffer");
  RunState s;
  int rc = read_state(e, run, s);
  if (rc) return rc;
  for (int i = 0; i < C_COUNT; ++i) counters[i] = s.counters[i];
  counters[C_STEPS] = s.step;
  counters[C_N] = s.n;
  counters[C_INACTIVE] = s.inactive;
  return MCST_OK;
}

int mcst_engine_get_steps(mcst_engine* e, uint32_t* steps, uint32_t n_runs) {
  CHECK_ENGINE(e);
  if (!e->committed) return fail(MCST_ERR_STATE, "engine not committed");
  if (!steps) return fail(MCST_ERR_ARG, "null output");
  if (n_runs > e->runs.size()) return fail(MCST_ERR_ARG, "more runs asked for than defined");
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipStreamSynchronize(e->stream));
  const uint32_t on_dev = std::min<uint32_t>(n_runs, e->runs_cap);
  std::vector<RunState> st(on_dev);
  if (on_dev) HIP_TRY(hipMemcpy(st.data(), e->d_rs, on_dev * sizeof(RunState), hipMemcpyDeviceToHost));
  for (uint32_t r = 0; r < n_runs; ++r) steps[r] = (r < on_dev && e->runs[r].phase == 1) ? st[r].step : 0u;
  return MCST_OK;
}

int mcst_run_get_n(mcst_engine* e, int32_t run, uint64_t* n) {
  CHECK_READY(e, run);
  RunState s;
  int rc = read_state(e, run, s);
  if (rc) return rc;
  if (n) *n = s.n;
  return MCST_OK;
}

int mcst_run_get_weight(mcst_engine* e, int32_t run, double* w) {
  CHECK_ENGINE(e);
  HostRun* h = get_run(e, run);
  if (!h) return fail(MCST_ERR_ARG, "bad run id");
  if (w) *w = h->weight;
  return MCST_OK;
}

int mcst_run_get_particles(mcst_engine* e, int32_t run, uint32_t* pos, uint8_t* status, double* state,
                           double* age_hyd, double* age_div) {
  CHECK_READY(e, run);
  RunState s;
  int rc = read_state(e, run, s);
  if (rc) return rc;
  const HostRun& h = e->runs[run];
  const BatchDev& b = e->bd;
  const uint64_t n = s.n;
  if (n == 0) return MCST_OK;
  if (pos) HIP_TRY(hipMemcpy(pos, b.pos + h.p_off, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (status) HIP_TRY(hipMemcpy(status, b.status + h.p_off, n, hipMemcpyDeviceToHost));
  if (age_hyd || age_div) {
    rc = read_ages(e, h.p_off, n, s, h.dt, age_hyd, age_div);
    if (rc) return rc;
  }
  if (state)
    for (int v = 0; v < e->model->nv; ++v)
      HIP_TRY(hipMemcpy(state + (size_t)v * n, b.state + (uint64_t)v * b.P + h.p_off, n * sizeof(double),
                        hipMemcpyDeviceToHost));
  return MCST_OK;
}

static int get_liquid(mcst_engine* e, int32_t run, double* out, int which) {
  CHECK_READY(e, run);
  if (!out) return fail(MCST_ERR_ARG, "null output");
  RunState s;
  int rc = read_state(e, run, s);
  if (rc) return rc;
  const RunDev& d = e->h_runs[run];
  const size_t nsc = (size_t)d.n_s * d.n_c;
  const double* src = which == 0 ? d.C[s.step & 1u] : (which == 1 ? d.m : which == 2 ? d.S :
                      which == 3 ? d.Cg[s.step & 1u] : d.mg);
  if (!src) return fail(MCST_ERR_STATE, "run has no gas phase");
  HIP_TRY(hipMemcpy(out, src, nsc * sizeof(double), hipMemcpyDeviceToHost));
  return MCST_OK;
}

int mcst_run_get_concentrations(mcst_engine* e, int32_t run, double* C) { return get_liquid(e, run, C, 0); }
int mcst_run_get_mass(mcst_engine* e, int32_t run, double* m) { return get_liquid(e, run, m, 1); }
int mcst_run_get_contributions(mcst_engine* e, int32_t run, double* S) { return get_liquid(e, run, S, 2); }
int mcst_run_get_gas_concentrations(mcst_engine* e, int32_t run, double* C) { return get_liquid(e, run, C, 3); }
int mcst_run_get_gas_mass(mcst_engine* e, int32_t run, double* m) { return get_liquid(e, run, m, 4); }

int mcst_run_get_repartition(mcst_engine* e, int32_t run, uint64_t* hist) {
  CHECK_READY(e, run);
  if (!hist) return fail(MCST_ERR_ARG, "null output");
  const uint32_t n_c = e->h_runs[run].n_c;
  unsigned long long* d = nullptr;
  HIP_TRY(hipMalloc(&d, n_c * sizeof(unsigned long long)));
  hipError_t err = hipMemsetAsync(d, 0, n_c * sizeof(unsigned long long), e->stream);
  if (err == hipSuccess) {
    launch_repartition(e->bd, (uint32_t)run, d, e->stream);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err == hipSuccess) err = hipMemcpy(hist, d, n_c * sizeof(uint64_t), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (err != hipSuccess) return fail(MCST_ERR_DEVICE, std::string("repartition: ") + hipGetErrorString(err));
  return MCST_OK;
}

// ---------------------------------------------------------------- allocation-free repartition
// The device scratch of mcst_run_get_repartition_noalloc: one counter per
// compartment, in a file-scope device array that each GPU the library runs
// on holds until the process ends (8 MiB).  Host code uses the address
// hipGetSymbolAddress gives for the current device, never the array's name
// (which on the host is only a stand-in for the symbol).
namespace {
constexpr uint32_t NOALLOC_MAX_COMPARTMENTS = 1u << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace
static __device__ __align__(256) unsigned long long scratch[NOALLOC_MAX_COMPARTMENTS];

namespace {
struct ScratchTable {
  std::mutex mu;
  unsigned long long* ptr[NOALLOC_MAX_DEVICES] = {};
};

// The device address of `scratch` on `device`, looked up once per device
// (on the first call on it) with that device current.
int scratch_on_device(int device, unsigned long long** out) {
  static ScratchTable t;
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(MCST_ERR_DEVICE, "device id above the scratch table");
  std::lock_guard<std::mutex> lock(t.mu);
  if (!t.ptr[device]) {
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p)
      return fail(MCST_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
    t.ptr[device] = (unsigned long long*)p;
  }
  *out = t.ptr[device];
  return MCST_OK;
}
}  // namespace

int mcst_run_get_repartition_noalloc(mcst_engine* e, int32_t run, uint64_t* hist) {
  CHECK_READY(e, run);
  if (!hist) return fail(MCST_ERR_ARG, "null output");
  const uint32_t n_c = e->h_runs[run].n_c;
  if (n_c > NOALLOC_MAX_COMPARTMENTS)
    return fail(MCST_ERR_ARG, "run has " + std::to_string(n_c) + " compartments; the allocation-free repartition holds " +
                                  std::to_string(NOALLOC_MAX_COMPARTMENTS));
  HIP_TRY(hipSetDevice(e->device));
  unsigned long long* d = nullptr;
  const int rc = scratch_on_device(e->device, &d);
  if (rc) return rc;
  // the array keeps the last call's counts: clear the run's n_c counters
  launch_clear_u64(d, n_c, e->stream);
  launch_repartition(e->bd, (uint32_t)run, d, e->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(hist, d, n_c * sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return MCST_OK;
}

int mcst_run_get_properties(mcst_engine* e, int32_t run, double* values, double* spatial) {
  CHECK_READY(e, run);
  if (!values && !spatial) return fail(MCST_ERR_ARG, "null outputs");
  RunState s;
  int rc = read_state(e, run, s);
  if (rc) return rc;
  const RunDev& d = e->h_runs[run];
  const uint64_t rows = (uint64_t)e->model->nv + 1, n = s.n;
  // the per-particle values take device memory only when they are asked for
  const size_t vb = values ? rows * n * sizeof(double) : 0, sb = rows * d.n_c * sizeof(double);
  char* buf = nullptr;
  HIP_TRY(hipMalloc((void**)&buf, vb + sb + 8));
  double* dv = values ? (double*)buf : nullptr;
  double* ds = (double*)(buf + vb);
  hipError_t err = hipMemsetAsync(ds, 0, sb, e->stream);
  if (err == hipSuccess && n) {
    e->model->ops->properties(e->bd, (uint32_t)run, n, dv, ds, e->stream);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err == hipSuccess && values && vb) err = hipMemcpy(values, dv, vb, hipMemcpyDeviceToHost);
  if (err == hipSuccess && spatial) err = hipMemcpy(spatial, ds, sb, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (err != hipSuccess) return fail(MCST_ERR_DEVICE, std::string("properties: ") + hipGetErrorString(err));
  return MCST_OK;
}

static int get_probe(mcst_engine* e, int32_t run, double* out, uint64_t max, uint64_t* n, bool exit_ages) {
  CHECK_READY(e, run);
  RunState s;
  int rc = read_state(e, run, s);
  if (rc) return rc;
  const RunDev& d = e->h_runs[run];
  const uint64_t have = std::min<uint64_t>(exit_ages ? s.probe_exit_n : s.probe_div_n, d.probe_cap);
  if (n) *n = have;
  const uint64_t k = std::min(have, max);
  if (out && k) HIP_TRY(hipMemcpy(out, exit_ages ? d.probe_exit : d.probe_div, k * sizeof(double), hipMemcpyDeviceToHost));
  return MCST_OK;
}

int mcst_run_get_exit_ages(mcst_engine* e, int32_t run, double* out, uint64_t max, uint64_t* n) {
  return get_probe(e, run, out, max, n, true);
}
int mcst_run_get_division_ages(mcst_engine* e, int32_t run, double* out, uint64_t max, uint64_t* n) {
  return get_probe(e, run, out, max, n, false);
}

// ---------------------------------------------------------------- checkpoints
// A checkpoint is one flat little-endian buffer: the header below, then
// pos u32[n], status u8[n] padded to 8 bytes, state f64[n_var][n],
// age_hyd f64[n], age_div f64[n], C, m, S f64[n_s][n_c] each, the
// recorded exit and division ages, and for a two-phase run (flags bit 0)
// the gas concentrations and gas mass f64[n_s][n_c] each.  It holds everything a step carries to
// the next (random draws are keyed by step, so none is stored); the network,
// feeds and model parameters are the run's definition and are not stored.
// The header's `model` is the model name's tag (model_tag).
namespace {
struct CkptHeader {
  char magic[8];
  uint32_t version, model, n_c, n_s, nv, step, flags, pad;
  uint64_t n, cap, seed, n_exit, n_div;
  double dt, weight;
  unsigned long long counters[C_COUNT];
};
constexpr char CKPT_MAGIC[8] = {'M', 'C', 'S', 'T', 'C', 'K', 'P', 'T'};
constexpr uint32_t CKPT_VERSION = 3;
constexpr uint32_t CKPT_TWO_PHASE = 1u;

uint64_t ckpt_bytes(uint64_t n, uint32_t nv, uint64_t nsc, uint64_t n_exit, uint64_t n_div, bool gas) {
  return sizeof(CkptHeader) + 4 * n + ((n + 7) / 8) * 8 +
         8 * ((uint64_t)nv * n + 2 * n + (gas ? 5 : 3) * nsc + n_exit + n_div);
}
}  // namespace

int mcst_run_checkpoint(mcst_engine* e, int32_t run, void* buf, uint64_t size, uint64_t* needed) {
  CHECK_READY(e, run);
  RunState s;
  int rc = read_state(e, run, s);
  if (rc) return rc;
  const RunDev& d = e->h_runs[run];
  const uint32_t nv = (uint32_t)e->model->nv;
  const uint64_t n = s.n, nsc = (uint64_t)d.n_s * d.n_c;
  const uint64_t n_exit = std::min<uint64_t>(s.probe_exit_n, d.probe_cap);
  const uint64_t n_div = std::min<uint64_t>(s.probe_div_n, d.probe_cap);
  const bool gas = d.Cg[0] != nullptr;
  const uint64_t bytes = ckpt_bytes(n, nv, nsc, n_exit, n_div, gas);
  if (needed) *needed = bytes;
  if (!buf) return MCST_OK;
  if (size < bytes) return fail(MCST_ERR_ARG, "checkpoint buffer too small (" + std::to_string(bytes) + " bytes needed)");
  CkptHeader hd{};
  std::memcpy(hd.magic, CKPT_MAGIC, 8);
  hd.version = CKPT_VERSION;
  hd.model = model_tag(e->model->name);
  hd.n_c = d.n_c;
  hd.n_s = d.n_s;
  hd.nv = nv;
  hd.step = s.step;
  hd.flags = gas ? CKPT_TWO_PHASE : 0u;
  hd.n = n;
  hd.cap = d.cap;
  hd.seed = d.seed;
  hd.n_exit = n_exit;
  hd.n_div = n_div;
  hd.dt = d.dt;
  hd.weight = d.weight;
  std::memcpy(hd.counters, s.counters, sizeof(hd.counters));
  char* p = (char*)buf;
  std::memcpy(p, &hd, sizeof(hd));
  p += sizeof(hd);
  const BatchDev& b = e->bd;
  const uint64_t o = d.p_off;
  auto down = [&](const void* src, uint64_t nb) -> hipError_t {
    hipError_t err = nb ? hipMemcpy(p, src, nb, hipMemcpyDeviceToHost) : hipSuccess;
    p += nb;
    return err;
  };
  HIP_TRY(down(b.pos + o, 4 * n));
  HIP_TRY(down(b.status + o, n));
  std::memset(p, 0, ((n + 7) / 8) * 8 - n);
  p += ((n + 7) / 8) * 8 - n;
  for (uint32_t v = 0; v < nv; ++v) HIP_TRY(down(b.state + (uint64_t)v * b.P + o, 8 * n));
  if ((rc = read_ages(e, o, n, s, d.dt, (double*)p, (double*)(p + 8 * n)))) return rc;
  p += 16 * n;
  HIP_TRY(down(d.C[s.step & 1u], 8 * nsc));
  HIP_TRY(down(d.m, 8 * nsc));
  HIP_TRY(down(d.S, 8 * nsc));
  HIP_TRY(down(d.probe_exit, 8 * n_exit));
  HIP_TRY(down(d.probe_div, 8 * n_div));
  if (gas) {
    HIP_TRY(down(d.Cg[s.step & 1u], 8 * nsc));
    HIP_TRY(down(d.mg, 8 * nsc));
  }
  return MCST_OK;
}

int mcst_run_restore(mcst_engine* e, int32_t run, const void* buf, uint64_t size) {
  CHECK_DEFINE(e, run, h);
  if (!buf || size < sizeof(CkptHeader)) return fail(MCST_ERR_ARG, "checkpoint buffer too small");
  CkptHeader hd;
  std::memcpy(&hd, buf, sizeof(hd));
  if (std::memcmp(hd.magic, CKPT_MAGIC, 8) != 0 || hd.version != CKPT_VERSION)
    return fail(MCST_ERR_ARG, "not a checkpoint of this library version");
  const uint32_t nv = (uint32_t)e->model->nv;
  if (hd.model != model_tag(e->model->name) || hd.nv != nv) return fail(MCST_ERR_ARG, "checkpoint of another model");
  if (hd.n_c != h->n_c || hd.n_s != h->n_s || hd.dt != h->dt || hd.seed != h->seed)
    return fail(MCST_ERR_ARG, "checkpoint of a run with another compartment count, species count, dt or seed");
  const uint64_t n = hd.n, nsc = (uint64_t)hd.n_s * hd.n_c;
  if (hd.cap < n || hd.cap >= 0xFFFFFFFFull) return fail(MCST_ERR_ARG, "bad particle counts in the checkpoint");
  const bool gas = (hd.flags & CKPT_TWO_PHASE) != 0;
  if (gas != h->two_phase)
    return fail(MCST_ERR_ARG, gas ? "checkpoint of a two-phase run: define the gas phase first"
                                  : "checkpoint of a run without a gas phase");
  if (size < ckpt_bytes(n, nv, nsc, hd.n_exit, hd.n_div, gas)) return fail(MCST_ERR_ARG, "truncated checkpoint");
  const char* p = (const char*)buf + sizeof(hd);
  auto take = [&](auto& vec, uint64_t count) {
    using T = typename std::decay_t<decltype(vec)>::value_type;
    vec.resize(count);
    if (count) std::memcpy(vec.data(), p, count * sizeof(T));
    p += count * sizeof(T);
  };
  std::vector<uint32_t> pos;
  std::vector<uint8_t> status;
  take(pos, n);
  take(status, n);
  p += ((n + 7) / 8) * 8 - n;
  for (uint64_t q = 0; q < n; ++q) {
    if (pos[q] >= h->n_c) return fail(MCST_ERR_ARG, "particle position out of range in the checkpoint");
    if (status[q] > DEAD) return fail(MCST_ERR_ARG, "bad particle status in the checkpoint");
  }
  h->pos.swap(pos);
  h->status.swap(status);
  take(h->state, (uint64_t)nv * n);
  take(h->ah, n);
  take(h->ad, n);
  take(h->C0, nsc);
  take(h->r_m, nsc);
  take(h->r_S, nsc);
  take(h->r_exit, hd.n_exit);
  take(h->r_div, hd.n_div);
  if (gas) {
    take(h->r_cg, nsc);
    take(h->r_mg, nsc);
  }
  h->n = n;
  h->cap = hd.cap;
  h->weight = hd.weight;
  h->has_particles = true;
  h->restored = true;
  h->r_step = hd.step;
  std::memcpy(h->r_counters, hd.counters, sizeof(hd.counters));
  return MCST_OK;
}

}  // extern "C"
