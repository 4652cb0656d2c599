// case_loader.cpp -- CMA case input on the C++ host (see case_loader.hpp).
#include "case_loader.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <set>
#include <stdexcept>

namespace mcst_host {
namespace {

std::vector<char> slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error(path + ": cannot be opened");
  return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

// little-endian scalars (the files' byte order; this host is little-endian too)
template <class T>
T get(const std::vector<char>& b, size_t at) {
  T v;
  std::memcpy(&v, b.data() + at, sizeof(T));
  return v;
}

bool is_dir(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool is_file(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string join(const std::string& a, const std::string& b) { return a.empty() || a.back() == '/' ? a + b : a + "/" + b; }

void read_map(const std::string& d, std::vector<double>& V, CooFlows& F) {
  V = read_raw_vector(join(d, "vofL.raw"));
  const size_t n = V.size();
  if (is_file(join(d, "flowL.coo"))) {
    F = read_coo(join(d, "flowL.coo"));
  } else {
    uint32_t rows = 0, cols = 0;
    const std::vector<double> M = read_raw_matrix(join(d, "flowL.raw"), rows, cols);
    if (rows != n || cols != n)
      throw std::runtime_error(d + ": flow matrix " + std::to_string(rows) + "x" + std::to_string(cols) + " for " +
                               std::to_string(n) + " compartments");
    F = CooFlows{};
    for (uint32_t i = 0; i < rows; ++i)  // row-major nonzeros; the diagonal carries no exchange
      for (uint32_t j = 0; j < cols; ++j) {
        const double q = M[(size_t)i * cols + j];
        if (q != 0.0 && i != j) {
          F.rows.push_back(i);
          F.cols.push_back(j);
          F.flows.push_back(q);
        }
      }
  }
  for (size_t t = 0; t < F.flows.size(); ++t)
    if (F.rows[t] >= n || F.cols[t] >= n)
      throw std::runtime_error(d + ": flow entry outside the " + std::to_string(n) + " compartments");
}

}  // namespace

std::vector<double> read_raw_vector(const std::string& path) {
  const std::vector<char> b = slurp(path);
  if (b.size() < 4) throw std::runtime_error(path + ": shorter than its count");
  const uint32_t n = get<uint32_t>(b, 0);
  if (b.size() < 4 + 8 * (size_t)n)
    throw std::runtime_error(path + ": " + std::to_string(n) + " values announced, " +
                             std::to_string((b.size() - 4) / 8) + " present");
  std::vector<double> v(n);
  if (n) std::memcpy(v.data(), b.data() + 4, 8 * (size_t)n);
  return v;
}

std::vector<double> read_raw_matrix(const std::string& path, uint32_t& rows, uint32_t& cols) {
  const std::vector<char> b = slurp(path);
  if (b.size() < 8) throw std::runtime_error(path + ": shorter than its shape");
  rows = get<uint32_t>(b, 0);
  cols = get<uint32_t>(b, 4);
  const size_t count = (size_t)rows * cols;
  if (b.size() < 8 + 8 * count)
    throw std::runtime_error(path + ": " + std::to_string(rows) + "x" + std::to_string(cols) +
                             " values announced, " + std::to_string((b.size() - 8) / 8) + " present");
  std::vector<double> m(count);
  if (count) std::memcpy(m.data(), b.data() + 8, 8 * count);
  return m;
}

CooFlows read_coo(const std::string& path) {
  const std::vector<char> b = slurp(path);
  if (b.size() < 8) throw std::runtime_error(path + ": shorter than its entry count");
  const uint64_t nnz = get<uint64_t>(b, 0);
  if (nnz > (b.size() - 8) / 16 || b.size() < 8 + nnz * 16)
    throw std::runtime_error(path + ": " + std::to_string(nnz) + " entries announced, file too short");
  CooFlows c;
  c.rows.resize(nnz);
  c.cols.resize(nnz);
  c.flows.resize(nnz);
  if (nnz) {
    std::memcpy(c.rows.data(), b.data() + 8, 4 * nnz);
    std::memcpy(c.cols.data(), b.data() + 8 + 4 * nnz, 4 * nnz);
    std::memcpy(c.flows.data(), b.data() + 8 + 8 * nnz, 8 * nnz);
  }
  return c;
}

std::vector<std::string> read_case_index(const std::string& path) {
  const std::vector<char> b = slurp(path);
  std::vector<std::string> out;
  size_t i = 0;
  while (i + 4 <= b.size()) {
    const uint32_t L = get<uint32_t>(b, i);
    if (L > 4 && L <= 4096 && i + 4 + L <= b.size()) {
      const std::string s(b.data() + i + 4, L);
      bool printable = true;
      for (unsigned char ch : s) printable = printable && ch >= 32 && ch < 127;
      if (printable && s.size() >= 4 && s.compare(s.size() - 4, 4, ".raw") == 0) {
        out.push_back(s);
        i += 4 + L;
        continue;
      }
    }
    ++i;
  }
  return out;
}

CmaCase load_case(const std::string& directory) {
  std::vector<std::string> subs;
  for (size_t k = 0; is_dir(join(directory, "i_" + std::to_string(k))); ++k)
    subs.push_back(join(directory, "i_" + std::to_string(k)));
  if (subs.empty()) {
    const std::string idx = join(directory, "cma_case");
    if (is_file(idx)) {
      std::set<std::string> listed;
      for (const auto& p : read_case_index(idx)) listed.insert(p.substr(p.find_last_of('/') + 1));
      if (!listed.count("vofL.raw")) throw std::runtime_error(idx + ": lists no liquid volumes");
    }
    subs.push_back(directory);
  }
  CmaCase c;
  for (const auto& d : subs) {
    c.volumes.emplace_back();
    c.flows.emplace_back();
    read_map(d, c.volumes.back(), c.flows.back());
  }
  for (const auto& v : c.volumes)
    if (v.size() != c.volumes[0].size())
      throw std::runtime_error(directory + ": flowmaps with different compartment counts");
  if (c.n_c() == 0) throw std::runtime_error(directory + ": no compartment");
  return c;
}

double min_residence_time(const CmaCase& c) {
  double best = std::numeric_limits<double>::infinity();
  for (size_t k = 0; k < c.n_maps(); ++k) {
    const std::vector<double>& V = c.volumes[k];
    std::vector<double> qout(V.size(), 0.0);
    for (size_t t = 0; t < c.flows[k].flows.size(); ++t) qout[c.flows[k].rows[t]] += c.flows[k].flows[t];
    for (size_t i = 0; i < V.size(); ++i)
      if (qout[i] > 0.0) best = std::min(best, V[i] / qout[i]);
  }
  return best;
}

double time_step(const CmaCase& c, double user_dt) {
  if (user_dt > 0.0 || c.n_c() == 1) return user_dt;
  const double tau = min_residence_time(c);
  if (!std::isfinite(tau) || tau == 0.0)
    throw std::runtime_error("no time step given and none can be estimated from the flowmaps");
  return tau / 100.0;
}

uint32_t steps_per_flowmap(double t_per_flowmap, size_t n_maps, double dt) {
  if (t_per_flowmap == 0.0 || n_maps == 1) return 1;
  return (uint32_t)(t_per_flowmap / dt) + 1u;
}

int define_flowmaps(mcst_engine* e, int32_t run, const CmaCase& c, uint32_t steps_per_map) {
  for (size_t k = 0; k < c.n_maps(); ++k) {
    const CooFlows& f = c.flows[k];
    const int rc = mcst_run_add_flowmap(e, run, c.volumes[k].data(), f.flows.size(), f.rows.data(), f.cols.data(),
                                        f.flows.data());
    if (rc != MCST_OK) return rc;
  }
  return mcst_run_set_steps_per_flowmap(e, run, steps_per_map);
}

}  // namespace mcst_host
