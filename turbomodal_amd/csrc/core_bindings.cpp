// pybind11 module `_core`: the reference's host object model (Material, Mesh,
// NodeSet, ConstraintGroup, FluidConfig, SolverConfig/Status/Progress,
// ModalResult, StationaryFrameResult, CyclicSymmetrySolver) with the same
// Python signatures as src/python_bindings.cpp:53-410.  All numerics go
// through the C ABI of libtmhip.so (include/tmhip.h); this file only does the
// integer/host work the reference also does on the host: gmsh parsing
// (src/mesh.cpp:30-206), boundary identification and matching
// (src/mesh.cpp:208-351), constrained-DOF construction
// (src/cyclic_solver.cpp:107-193) and DOF classification (214-238).
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <unordered_map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tmhip.h"

namespace py = pybind11;

namespace tmcore {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
struct Material {
  double E, nu, rho;
  double T_ref = 293.15;
  double E_slope = 0.0;
  Material(double E_, double nu_, double rho_) : E(E_), nu(nu_), rho(rho_) { validate(); }
  Material(double E_, double nu_, double rho_, double T, double s)
      : E(E_), nu(nu_), rho(rho_), T_ref(T), E_slope(s) { validate(); }
  void validate() const {  // src/material.cpp:15-22
    if (E <= 0.0) throw std::invalid_argument("Young's modulus E must be positive");
    if (nu <= 0.0 || nu >= 0.5) throw std::invalid_argument("Poisson's ratio nu must be in (0, 0.5)");
    if (rho <= 0.0) throw std::invalid_argument("Density rho must be positive");
  }
  Material at_temperature(double T) const {  // src/material.cpp:33-38
    const double Ea = E + E_slope * (T - T_ref);
    if (Ea <= 0.0) throw std::runtime_error("Temperature-adjusted E is non-positive");
    return Material(Ea, nu, rho, T_ref, E_slope);
  }
  py::array_t<double> constitutive_matrix() const {
    py::array_t<double> D({6, 6});
    auto d = D.mutable_unchecked<2>();
    const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) d(i, j) = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) d(i, j) = c * nu;
    for (int i = 0; i < 3; ++i) d(i, i) = c * (1.0 - nu);
    for (int i = 3; i < 6; ++i) d(i, i) = c * (1.0 - 2.0 * nu) / 2.0;
    return D;
  }
};

struct NodeSet {
  std::string name;
  std::vector<int> node_ids;
};

enum class BCType { FIXED, DISPLACEMENT, FRICTIONLESS, ELASTIC_SUPPORT, CYLINDRICAL };

struct ConstraintGroup {
  std::string name;
  std::vector<int> node_ids;
  BCType type = BCType::FIXED;
  std::array<bool, 3> constrained_components = {true, true, true};
  std::array<double, 3> surface_normal = {0, 0, 0};
  std::array<double, 3> spring_stiffness = {0, 0, 0};
  std::array<double, 3> cylinder_axis = {0, 0, 1};
  std::array<double, 3> cylinder_origin = {0, 0, 0};
};

// ---------------------------------------------------------------------------
struct Mesh {
  std::vector<double> nodes;   // N x 3
  std::vector<int> elements;   // E x 10
  std::vector<NodeSet> node_sets;
  std::vector<int> left_boundary, right_boundary, free_boundary;
  std::vector<std::pair<int, int>> matched_pairs;
  int num_sectors = 0;
  int rotation_axis = 2;

  int num_nodes() const { return static_cast<int>(nodes.size() / 3); }
  int num_elements() const { return static_cast<int>(elements.size() / 10); }
  int num_dof() const { return 3 * num_nodes(); }

  const NodeSet* find_node_set(const std::string& name) const {
    for (const auto& ns : node_sets)
      if (ns.name == name) return &ns;
    return nullptr;
  }

  void planes(int& c1, int& c2, int& cax) const {
    if (rotation_axis == 0) { c1 = 1; c2 = 2; cax = 0; }
    else if (rotation_axis == 1) { c1 = 0; c2 = 2; cax = 1; }
    else { c1 = 0; c2 = 1; cax = 2; }
  }

  // gmsh MSH 2.x ASCII: nodes, TET10 volume elements (mid-edge 8/9 order
  // detected geometrically), physical groups -> node sets.
  void load_from_gmsh(const std::string& filename) {
    std::ifstream f(filename);
    if (!f.is_open()) throw std::runtime_error("Cannot open mesh file: " + filename);
    std::map<int, std::string> phys;
    std::vector<double> xyz;
    std::map<int, int> g2l;
    std::vector<int> tets;
    std::map<std::string, std::set<int>> sets;
    std::string line;
    auto strip = [](std::string& s) {
      while (!s.empty() && (s.back() == '\r' || s.back() == '\n' || s.back() == ' ')) s.pop_back();
    };
    auto count_of = [](int type) {
      switch (type) {
        case 1: return 2; case 2: return 3; case 3: return 4; case 4: return 4; case 5: return 8;
        case 8: return 3; case 9: return 6; case 11: return 10; case 15: return 1;
        default: return -1;
      }
    };
    while (std::getline(f, line)) {
      strip(line);
      if (line == "$MeshFormat") {
        std::getline(f, line);
        std::istringstream is(line);
        double ver = 0;
        int ft = 0, ds = 0;
        is >> ver >> ft >> ds;
        if (ver < 2.0 || ver >= 3.0)
          throw std::runtime_error("Unsupported Gmsh format version: " + std::to_string(ver) +
                                   ". Only v2.x supported.");
        if (ft != 0) throw std::runtime_error("Only ASCII Gmsh files supported (file-type must be 0).");
        std::getline(f, line);
      } else if (line == "$PhysicalNames") {
        std::getline(f, line);
        const int n = std::stoi(line);
        for (int i = 0; i < n; ++i) {
          std::getline(f, line);
          std::istringstream is(line);
          int dim, tag;
          std::string name;
          is >> dim >> tag >> name;
          if (!name.empty() && name.front() == '"') name.erase(0, 1);
          if (!name.empty() && name.back() == '"') name.pop_back();
          phys[tag] = name;
        }
        std::getline(f, line);
      } else if (line == "$Nodes") {
        std::getline(f, line);
        const int n = std::stoi(line);
        xyz.reserve(3 * (size_t)n);
        for (int i = 0; i < n; ++i) {
          std::getline(f, line);
          std::istringstream is(line);
          int id;
          double x, y, z;
          is >> id >> x >> y >> z;
          g2l[id] = static_cast<int>(xyz.size() / 3);
          xyz.push_back(x);
          xyz.push_back(y);
          xyz.push_back(z);
        }
        std::getline(f, line);
      } else if (line == "$Elements") {
        std::getline(f, line);
        const int n = std::stoi(line);
        for (int i = 0; i < n; ++i) {
          std::getline(f, line);
          std::istringstream is(line);
          int id, type, ntag;
          is >> id >> type >> ntag;
          std::vector<int> tags(ntag);
          for (int t = 0; t < ntag; ++t) is >> tags[t];
          const int nn = count_of(type);
          if (nn < 0) throw std::runtime_error("Unknown Gmsh element type: " + std::to_string(type));
          std::vector<int> en(nn);
          for (int q = 0; q < nn; ++q) {
            int g;
            is >> g;
            auto it = g2l.find(g);
            if (it == g2l.end()) throw std::runtime_error("Element references unknown node ID: " + std::to_string(g));
            en[q] = it->second;
          }
          if (type == 11) {
            auto p = [&](int k, int c) { return xyz[3 * (size_t)en[k] + c]; };
            double d13 = 0, d23 = 0;
            for (int c = 0; c < 3; ++c) {
              const double m13 = 0.5 * (p(1, c) + p(3, c)), m23 = 0.5 * (p(2, c) + p(3, c));
              d13 += (p(8, c) - m13) * (p(8, c) - m13);
              d23 += (p(8, c) - m23) * (p(8, c) - m23);
            }
            if (d23 < d13) std::swap(en[8], en[9]);
            tets.insert(tets.end(), en.begin(), en.end());
          }
          if (ntag > 0 && tags[0] > 0 && phys.count(tags[0]))
            sets[phys[tags[0]]].insert(en.begin(), en.end());
        }
        std::getline(f, line);
      }
    }
    nodes = std::move(xyz);
    elements = std::move(tets);
    node_sets.clear();
    for (auto& [name, ids] : sets) node_sets.push_back({name, std::vector<int>(ids.begin(), ids.end())});
  }

  void load_from_arrays(py::array_t<double, py::array::c_style | py::array::forcecast> coords,
                        py::array_t<int, py::array::c_style | py::array::forcecast> conn,
                        const std::vector<NodeSet>& sets, int ns, int axis) {
    if (coords.ndim() != 2 || coords.shape(1) != 3)
      throw std::runtime_error("node_coords must have 3 columns (x, y, z)");
    if (conn.ndim() != 2 || conn.shape(1) != 10)
      throw std::runtime_error("element_connectivity must have 10 columns (TET10)");
    if (axis < 0 || axis > 2) throw std::runtime_error("rotation_axis must be 0 (X), 1 (Y), or 2 (Z)");
    nodes.assign(coords.data(), coords.data() + coords.size());
    elements.assign(conn.data(), conn.data() + conn.size());
    node_sets = sets;
    num_sectors = ns;
    rotation_axis = axis;
    identify_cyclic_boundaries(1e-6);
    match_boundary_nodes();
  }

  void identify_cyclic_boundaries(double /*tolerance*/) {
    left_boundary.clear();
    right_boundary.clear();
    free_boundary.clear();
    const NodeSet* l = find_node_set("left_boundary");
    const NodeSet* r = find_node_set("right_boundary");
    if (!l || !r)
      throw std::runtime_error(
          "Mesh must have physical groups named 'left_boundary' and 'right_boundary' for cyclic symmetry.");
    left_boundary = l->node_ids;
    right_boundary = r->node_ids;
    if (const NodeSet* fb = find_node_set("free_boundary")) free_boundary = fb->node_ids;
  }

  // Greedy nearest-partner matching after rotating right nodes by -alpha;
  // tolerance max(1e-6, 1e-4 r_left); unmatched nodes move to free_boundary.
  void match_boundary_nodes() {
    if (num_sectors <= 0) throw std::runtime_error("num_sectors must be set before matching boundary nodes.");
    if (left_boundary.empty() || right_boundary.empty())
      throw std::runtime_error("Call identify_cyclic_boundaries() before match_boundary_nodes().");
    const double alpha = 2.0 * kPi / num_sectors;
    const double ca = std::cos(-alpha), sa = std::sin(-alpha);
    int c1, c2, cax;
    planes(c1, c2, cax);
    const size_t R = right_boundary.size();
    std::vector<std::array<double, 3>> rp(R);
    for (size_t j = 0; j < R; ++j) {
      const int id = right_boundary[j];
      const double r1 = nodes[3 * (size_t)id + c1], r2 = nodes[3 * (size_t)id + c2];
      rp[j] = {ca * r1 - sa * r2, sa * r1 + ca * r2, nodes[3 * (size_t)id + cax]};
    }
    std::vector<char> used(R, 0);
    matched_pairs.clear();
    std::vector<int> nl, nr, fl;
    // The reference pairs each left node, in order, with the nearest unused
    // right node and keeps the pair when it lies within the tolerance
    // (src/mesh.cpp:289-333).  A pair is only kept within the tolerance, so a
    // uniform grid of cell size = the largest tolerance, searched over the 27
    // cells around the left node, finds the same node (ties: the first index)
    // in O(L + R) instead of O(L R).
    auto tol_of = [&](double l1, double l2) { return std::max(1e-6, 1e-4 * std::sqrt(l1 * l1 + l2 * l2)); };
    double cell = 1e-6;
    for (int lid : left_boundary)
      cell = std::max(cell, tol_of(nodes[3 * (size_t)lid + c1], nodes[3 * (size_t)lid + c2]));
    auto cell_of = [&](double v) { return static_cast<long long>(std::floor(v / cell)); };
    auto key = [](long long i, long long j, long long k) {
      return ((i & 0x1fffff) << 42) | ((j & 0x1fffff) << 21) | (k & 0x1fffff);
    };
    std::unordered_map<long long, std::vector<int>> grid;
    grid.reserve(2 * R);
    for (size_t j = 0; j < R; ++j) grid[key(cell_of(rp[j][0]), cell_of(rp[j][1]), cell_of(rp[j][2]))].push_back((int)j);
    for (int lid : left_boundary) {
      const double l1 = nodes[3 * (size_t)lid + c1], l2 = nodes[3 * (size_t)lid + c2],
                   la = nodes[3 * (size_t)lid + cax];
      const long long g1 = cell_of(l1), g2 = cell_of(l2), g3 = cell_of(la);
      double best = std::numeric_limits<double>::max();
      long best_j = -1;
      for (long long d1c = -1; d1c <= 1; ++d1c)
        for (long long d2c = -1; d2c <= 1; ++d2c)
          for (long long d3c = -1; d3c <= 1; ++d3c) {
            auto it = grid.find(key(g1 + d1c, g2 + d2c, g3 + d3c));
            if (it == grid.end()) continue;
            for (int j : it->second) {
              if (used[j]) continue;
              const double d1 = l1 - rp[j][0], d2 = l2 - rp[j][1], d3 = la - rp[j][2];
              const double d = std::sqrt(d1 * d1 + d2 * d2 + d3 * d3);
              if (d < best || (d == best && j < best_j)) {
                best = d;
                best_j = j;
              }
            }
          }
      if (best_j < 0 || best > tol_of(l1, l2)) {
        fl.push_back(lid);
        continue;
      }
      used[best_j] = 1;
      nl.push_back(lid);
      nr.push_back(right_boundary[best_j]);
      matched_pairs.push_back({lid, right_boundary[best_j]});
    }
    std::set<int> fs(free_boundary.begin(), free_boundary.end());
    fs.insert(fl.begin(), fl.end());
    for (size_t j = 0; j < R; ++j)
      if (!used[j]) fs.insert(right_boundary[j]);
    left_boundary = nl;
    right_boundary = nr;
    free_boundary.assign(fs.begin(), fs.end());
  }
};

// ---------------------------------------------------------------------------
struct FluidConfig {
  enum class Type { NONE, GAS_AIC, KWAK_ANALYTICAL, POTENTIAL_FLOW_BEM, LIQUID_ACOUSTIC_FEM };
  Type type = Type::NONE;
  double fluid_density = 0.0, disk_radius = 0.0, disk_thickness = 0.0, speed_of_sound = 0.0;
};

struct SolverConfig {
  int nev = 20, ncv = 0;
  double shift = 0.0, tolerance = 1e-8;
  int max_iterations = 1000;
};
struct SolverStatus {
  bool converged = false;
  int num_converged = 0, iterations = 0;
  double max_residual = 0.0;
  std::string message;
};
struct SolverProgress {
  int completed = 0, total = 0, harmonic_k = 0;
  bool converged = true;
  int iterations = 0, num_converged = 0, num_modes = 0;
  double max_residual = 0.0, min_freq_hz = 0.0, max_freq_hz = 0.0, elapsed_s = 0.0;
};

struct ModalResult {
  int harmonic_index = 0;
  double rpm = 0.0;
  std::vector<double> frequencies;
  py::array_t<std::complex<double>> mode_shapes = py::array_t<std::complex<double>>(std::vector<py::ssize_t>{0, 0});
  std::vector<int> whirl_direction;
  bool converged = true;
  // engine diagnostics (not in the reference's struct)
  std::vector<double> modal_mass;    // N * Re(phi^H M_sector phi), kg
  std::vector<double> residuals;     // ||K x - lambda M x|| / ||K x|| per unique eigenpair
  std::vector<double> eigenvalues;   // lambda per unique eigenpair
  int restarts = 0;
  double solve_seconds = 0.0;

  std::vector<double> wave_propagation_velocity(double radius) const {
    std::vector<double> v(frequencies.size(), 0.0);
    for (size_t i = 0; i < frequencies.size(); ++i)
      if (harmonic_index > 0) v[i] = 2.0 * kPi * frequencies[i] * radius / harmonic_index;
    return v;
  }
};

struct StationaryFrameResult {
  std::vector<double> frequencies;
  std::vector<int> whirl_direction;
};

// ---------------------------------------------------------------------------
// engine handles
// ---------------------------------------------------------------------------
// the context's memory reserve outside solve calls (tm_set_memory_reserve's default)
constexpr double kDefaultReserve = 0.15;

class Engine {
 public:
  static Engine& get() {
    static Engine e;
    return e;
  }
  tm_context* ctx() {
    std::lock_guard<std::mutex> g(mu_);
    if (!ctx_) {
      tm_context* c = nullptr;
      int rc = tm_context_create(-1, &c);
      if (rc != TM_OK) throw std::runtime_error(std::string("HIP engine unavailable: ") + tm_last_error(nullptr));
      ctx_ = c;
    }
    return ctx_;
  }
  std::mutex& mutex() { return mu_; }
  static void check(tm_context* c, int rc) {
    if (rc == TM_OK) return;
    const std::string msg = tm_last_error(c);
    if (rc == TM_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
  }

 private:
  std::mutex mu_;
  tm_context* ctx_ = nullptr;
};

// ---------------------------------------------------------------------------
class CyclicSymmetrySolver {
 public:
  CyclicSymmetrySolver(const Mesh& mesh, const Material& mat, const FluidConfig& fluid, bool apply_hub)
      : mesh_(mesh), mat_(mat), fluid_(fluid) {
    if (apply_hub) {
      if (const NodeSet* hub = mesh_.find_node_set("hub_constraint")) {
        ConstraintGroup cg;
        cg.name = "hub_constraint";
        cg.node_ids = hub->node_ids;
        cg.type = BCType::FIXED;
        constraints_.push_back(cg);
        has_constraints_ = true;
      }
    }
    check_fluid();
  }
  CyclicSymmetrySolver(const Mesh& mesh, const Material& mat, const std::vector<ConstraintGroup>& cons,
                       const FluidConfig& fluid)
      : mesh_(mesh), mat_(mat), fluid_(fluid), constraints_(cons), has_constraints_(!cons.empty()) {
    check_fluid();
  }
  ~CyclicSymmetrySolver() {
    if (op_) tm_operator_destroy(op_);
    if (emesh_) tm_mesh_destroy(emesh_);
  }

  // build_constrained_dofs (src/cyclic_solver.cpp:107-193)
  std::vector<int> build_constrained_dofs() const {
    std::set<int> s;
    for (const auto& cg : constraints_)
      for (int node : cg.node_ids) {
        switch (cg.type) {
          case BCType::FIXED:
            for (int c = 0; c < 3; ++c) s.insert(3 * node + c);
            break;
          case BCType::DISPLACEMENT:
            for (int c = 0; c < 3; ++c)
              if (cg.constrained_components[c]) s.insert(3 * node + c);
            break;
          case BCType::FRICTIONLESS: {
            const auto& n = cg.surface_normal;
            const double nm = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            int best = 0;
            double bv = nm > 0 ? std::fabs(n[0] / nm) : 0.0;
            for (int c = 1; c < 3; ++c) {
              const double v = nm > 0 ? std::fabs(n[c] / nm) : 0.0;
              if (v > bv) { bv = v; best = c; }
            }
            s.insert(3 * node + best);
            break;
          }
          case BCType::ELASTIC_SUPPORT:
            break;
          case BCType::CYLINDRICAL: {
            const auto& ax = cg.cylinder_axis;
            const double an = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
            const double a[3] = {ax[0] / an, ax[1] / an, ax[2] / an};
            double p[3];
            for (int c = 0; c < 3; ++c) p[c] = mesh_.nodes[3 * (size_t)node + c] - cg.cylinder_origin[c];
            const double pa = p[0] * a[0] + p[1] * a[1] + p[2] * a[2];
            double rv[3] = {p[0] - pa * a[0], p[1] - pa * a[1], p[2] - pa * a[2]};
            const double r = std::sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
            if (r < 1e-12) {
              for (int c = 0; c < 3; ++c)
                if (cg.constrained_components[c]) s.insert(3 * node + c);
              break;
            }
            for (double& x : rv) x /= r;
            double t[3] = {a[1] * rv[2] - a[2] * rv[1], a[2] * rv[0] - a[0] * rv[2], a[0] * rv[1] - a[1] * rv[0]};
            const double tn = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            for (double& x : t) x /= tn;
            const double* dirs[3] = {rv, t, a};
            for (int c = 0; c < 3; ++c) {
              if (!cg.constrained_components[c]) continue;
              int best = 0;
              double bv = std::fabs(dirs[c][0]);
              for (int k = 1; k < 3; ++k)
                if (std::fabs(dirs[c][k]) > bv) { bv = std::fabs(dirs[c][k]); best = k; }
              s.insert(3 * node + best);
            }
            break;
          }
        }
      }
    return std::vector<int>(s.begin(), s.end());
  }

  std::vector<ModalResult> solve_at_rpm(double rpm, int num_modes, const std::vector<int>& hi, int max_threads,
                                        bool coriolis, double min_freq, py::object progress_cb,
                                        bool allow_condensation, double memory_reserve_fraction) {
    (void)max_threads;
    (void)allow_condensation;  // the reference's Guyan branch is not part of this engine
    auto res = solve_points({rpm}, num_modes, hi, coriolis, min_freq, progress_cb, memory_reserve_fraction);
    return res.empty() ? std::vector<ModalResult>{} : res[0];
  }

  std::vector<std::vector<ModalResult>> solve_rpm_sweep(const std::vector<double>& rpms, int num_modes,
                                                        const std::vector<int>& hi, int max_threads,
                                                        bool coriolis, double min_freq,
                                                        double memory_reserve_fraction) {
    (void)max_threads;
    return solve_points(rpms, num_modes, hi, coriolis, min_freq, py::none(), memory_reserve_fraction);
  }

  // engine statistics of the cached mesh (symbolic analysis)
  py::dict engine_info() {
    ensure_engine();
    tm_mesh_info mi{};
    Engine::check(Engine::get().ctx(), tm_mesh_get_info(emesh_, &mi));
    py::dict d;
    d["n_free"] = mi.n_free;
    d["n_dof_system"] = mi.n_dof_system;
    d["nnz_blocks"] = mi.nnz_blocks;
    d["nnz_L"] = mi.nnz_L;
    d["factor_flops"] = mi.factor_flops;
    d["n_supernodes"] = mi.n_supernodes;
    d["tree_levels"] = mi.tree_levels;
    d["max_front"] = mi.max_front;
    d["front_bytes_fp64"] = mi.front_bytes_fp64;
    return d;
  }

  // compute_stationary_frame (src/cyclic_solver.cpp:374-454)
  static StationaryFrameResult compute_stationary_frame(const ModalResult& r, int num_sectors) {
    StationaryFrameResult sf;
    const int k = r.harmonic_index, max_k = num_sectors / 2;
    const double omega = r.rpm * 2.0 * kPi / 60.0;
    const int nm = static_cast<int>(r.frequencies.size());
    bool cor = false;
    for (int m = 0; m < nm && m < (int)r.whirl_direction.size(); ++m)
      if (r.whirl_direction[m] != 0) cor = true;
    auto sort_pairs = [](std::vector<double>& f, std::vector<int>& w) {
      std::vector<int> idx(f.size());
      for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return f[a] < f[b]; });
      std::vector<double> f2;
      std::vector<int> w2;
      for (int i : idx) { f2.push_back(f[i]); w2.push_back(w[i]); }
      f = f2;
      w = w2;
    };
    if (cor && std::fabs(omega) > 0.0 && k > 0) {
      const double kw = k * std::fabs(omega) / (2.0 * kPi);
      for (int m = 0; m < nm; ++m) {
        const int w = r.whirl_direction[m];
        sf.frequencies.push_back(std::fabs(r.frequencies[m] + w * kw));
        sf.whirl_direction.push_back(w);
      }
      sort_pairs(sf.frequencies, sf.whirl_direction);
    } else if (std::fabs(omega) > 0.0 && k > 0 && k < max_k) {
      const double kw = k * std::fabs(omega) / (2.0 * kPi);
      for (int m = 0; m < nm; ++m) {
        sf.frequencies.push_back(r.frequencies[m] + kw);
        sf.whirl_direction.push_back(1);
        sf.frequencies.push_back(std::fabs(r.frequencies[m] - kw));
        sf.whirl_direction.push_back(-1);
      }
      sort_pairs(sf.frequencies, sf.whirl_direction);
    } else {
      sf.frequencies = r.frequencies;
      sf.whirl_direction.assign(nm, 0);
    }
    return sf;
  }

 private:
  const Mesh& mesh_;
  Material mat_;
  FluidConfig fluid_;
  std::vector<ConstraintGroup> constraints_;
  bool has_constraints_ = false;
  tm_mesh* emesh_ = nullptr;
  tm_operator* op_ = nullptr;

 public:
  // eigensolver of the engine: TM_METHOD_BLOCK (default) or TM_METHOD_LANCZOS
  int method_ = TM_METHOD_BLOCK;
  // factor precision: TM_PREC_FP64 (default) or TM_PREC_MIXED (complex-FP32 factor,
  // FP64 refinement of every shift-invert solve; block method only)
  int precision_ = TM_PREC_FP64;

 private:

  void check_fluid() const {
    if (fluid_.type != FluidConfig::Type::NONE && fluid_.fluid_density > 0.0)
      throw std::runtime_error("fluid coupling is not part of the MI355X cyclic solve (dry structures only)");
  }

  // build_spring_stiffness (src/cyclic_solver.cpp:193-212): per DOF, the sum of the
  // positive spring stiffnesses of the ELASTIC_SUPPORT groups that hold its node
  std::vector<double> spring_stiffness_full(bool& any) const {
    std::vector<double> k(3 * (size_t)mesh_.num_nodes(), 0.0);
    any = false;
    for (const auto& cg : constraints_) {
      if (cg.type != BCType::ELASTIC_SUPPORT) continue;
      for (int node : cg.node_ids)
        for (int c = 0; c < 3; ++c)
          if (cg.spring_stiffness[c] > 0.0) {
            k.at(3 * (size_t)node + c) += cg.spring_stiffness[c];
            any = true;
          }
    }
    return k;
  }

  void ensure_engine() {
    if (op_) return;
    tm_context* ctx = Engine::get().ctx();
    const int nn = mesh_.num_nodes();
    if (mesh_.num_sectors <= 0) throw std::runtime_error("mesh has no sector count");
    std::vector<int32_t> cls(nn, TM_NODE_INTERIOR), partner(nn, -1);
    // classify_dofs: right takes precedence over left (src/cyclic_solver.cpp:214-238)
    for (int l : mesh_.left_boundary) cls.at(l) = TM_NODE_LEFT;
    for (int r : mesh_.right_boundary) cls.at(r) = TM_NODE_RIGHT;
    for (auto& pr : mesh_.matched_pairs) partner.at(pr.second) = pr.first;
    std::vector<uint8_t> cm(3 * (size_t)nn, 0);
    for (int d : build_constrained_dofs())
      if (d >= 0 && d < 3 * nn) cm[d] = 1;
    std::vector<int32_t> el(mesh_.elements.begin(), mesh_.elements.end());
    tm_mesh* m = nullptr;
    {
      py::gil_scoped_release rel;
      std::lock_guard<std::mutex> g(Engine::get().mutex());
      Engine::check(ctx, tm_mesh_create(ctx, mesh_.nodes.data(), nn, el.data(), mesh_.num_elements(), cls.data(),
                                        partner.data(), cm.data(), mesh_.num_sectors, mesh_.rotation_axis, &m));
    }
    emesh_ = m;
    tm_operator* op = nullptr;
    {
      py::gil_scoped_release rel;
      std::lock_guard<std::mutex> g(Engine::get().mutex());
      Engine::check(ctx, tm_operator_create(ctx, emesh_, mat_.E, mat_.nu, mat_.rho, has_constraints_ ? 1 : 0, &op));
    }
    op_ = op;
    bool springs = false;
    const std::vector<double> k = spring_stiffness_full(springs);
    if (springs) {
      py::gil_scoped_release rel;
      std::lock_guard<std::mutex> g(Engine::get().mutex());
      Engine::check(ctx, tm_operator_add_springs(op_, k.data()));
    }
  }

  std::vector<std::vector<ModalResult>> solve_points(const std::vector<double>& rpms, int num_modes,
                                                     const std::vector<int>& hi, bool coriolis, double min_freq,
                                                     py::object progress_cb, double memory_reserve_fraction) {
    ensure_engine();
    tm_context* ctx = Engine::get().ctx();
    if (!(memory_reserve_fraction >= 0.0 && memory_reserve_fraction <= 1.0))
      throw py::value_error("memory_reserve_fraction must lie in [0, 1]");
    const int max_k = mesh_.num_sectors / 2;
    std::vector<int> ks;
    if (hi.empty())
      for (int k = 0; k <= max_k; ++k) ks.push_back(k);
    else
      for (int k : hi)
        if (k >= 0 && k <= max_k) ks.push_back(k);
    const int nn = mesh_.num_nodes();
    const size_t np = rpms.size() * ks.size();
    std::vector<tm_problem> probs(np);
    std::vector<tm_result> res(np);
    std::vector<std::vector<double>> freq(np), ev(np), rsd(np), mass(np);
    // mode shapes land directly in the numpy arrays handed back (no staging copy)
    std::vector<py::array_t<std::complex<double>>> shapes;
    shapes.reserve(np);
    std::vector<std::vector<int32_t>> whirl(np);
    size_t q = 0;
    for (double rpm : rpms)
      for (int k : ks) {
        tm_problem& p = probs[q];
        p.op = op_;
        p.rpm = rpm;
        p.harmonic = k;
        p.num_modes = num_modes;
        p.precision = precision_;
        p.conv_policy = TM_CONV_TIGHT;
        p.shift = 1.0;
        p.tol_rel = 1e-12;
        p.max_restarts = 1000;
        p.method = method_;
        p.coriolis = coriolis ? 1 : 0;
        // rpm continuation between the sweep's adjacent rpm points
        p.warm_start = rpms.size() > 1 ? 1 : 0;
        p.min_frequency = min_freq;
        freq[q].assign(num_modes, 0.0);
        ev[q].assign(num_modes, 0.0);
        rsd[q].assign(num_modes, 0.0);
        mass[q].assign(num_modes, 0.0);
        whirl[q].assign(num_modes, 0);
        shapes.emplace_back(std::vector<py::ssize_t>{3 * (py::ssize_t)nn, num_modes});
        tm_result& r = res[q];
        r.frequencies = freq[q].data();
        r.mode_shapes = reinterpret_cast<double*>(shapes[q].mutable_data());
        r.whirl = whirl[q].data();
        r.eigenvalues = ev[q].data();
        r.residuals = rsd[q].data();
        r.modal_mass = mass[q].data();
        ++q;
      }
    auto t0 = std::chrono::steady_clock::now();
    {
      py::gil_scoped_release rel;
      std::lock_guard<std::mutex> g(Engine::get().mutex());
      // device memory kept out of batch sizing, as the reference keeps host RAM out of
      // its concurrency cap (src/cyclic_solver.cpp:1257-1272); set under the engine lock
      // for this call only, so concurrent callers do not see each other's value
      Engine::check(ctx, tm_set_memory_reserve(ctx, memory_reserve_fraction));
      const int rc = tm_solve_batch(ctx, probs.data(), static_cast<int32_t>(np), res.data());
      tm_set_memory_reserve(ctx, kDefaultReserve);
      Engine::check(ctx, rc);
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<std::vector<ModalResult>> out(rpms.size());
    q = 0;
    int done = 0;
    for (size_t ir = 0; ir < rpms.size(); ++ir)
      for (size_t ik = 0; ik < ks.size(); ++ik, ++q) {
        const tm_result& r = res[q];
        ++done;
        if (!progress_cb.is_none() && ir == 0) {
          SolverProgress pg;
          pg.completed = done;
          pg.total = static_cast<int>(ks.size());
          pg.harmonic_k = ks[ik];
          pg.converged = r.status == TM_OK && r.converged;
          pg.iterations = r.restarts;
          pg.num_converged = r.n_unique;
          pg.num_modes = r.n_modes;
          pg.elapsed_s = elapsed;
          if (r.n_modes > 0) {
            pg.min_freq_hz = *std::min_element(freq[q].begin(), freq[q].begin() + r.n_modes);
            pg.max_freq_hz = *std::max_element(freq[q].begin(), freq[q].begin() + r.n_modes);
            pg.max_residual = *std::max_element(rsd[q].begin(), rsd[q].begin() + std::max(1, r.n_unique));
          }
          progress_cb(pg);
        }
        if (r.status != TM_OK || r.n_modes <= 0) continue;  // dropped, as the reference does
        ModalResult m;
        m.harmonic_index = ks[ik];
        m.rpm = rpms[ir];
        m.frequencies.assign(freq[q].begin(), freq[q].begin() + r.n_modes);
        m.whirl_direction.assign(whirl[q].begin(), whirl[q].begin() + r.n_modes);
        m.modal_mass.assign(mass[q].begin(), mass[q].begin() + r.n_modes);
        m.eigenvalues.assign(ev[q].begin(), ev[q].begin() + r.n_unique);
        m.residuals.assign(rsd[q].begin(), rsd[q].begin() + r.n_unique);
        m.converged = r.converged != 0;
        m.restarts = r.restarts;
        m.solve_seconds = r.seconds;
        if (r.n_modes == num_modes) {
          m.mode_shapes = shapes[q];
        } else {  // fewer modes: the engine wrote [3 nn][n_modes] at the front of the buffer
          py::array_t<std::complex<double>> a(std::vector<py::ssize_t>{3 * (py::ssize_t)nn, r.n_modes});
          std::memcpy(a.mutable_data(), shapes[q].data(), sizeof(std::complex<double>) * 3 * (size_t)nn * r.n_modes);
          m.mode_shapes = a;
        }
        out[ir].push_back(std::move(m));
      }
    return out;
  }
};

}  // namespace tmcore

using namespace tmcore;

PYBIND11_MODULE(_core, m) {
  m.doc() = "Cyclic symmetry modal solve on MI355X (gfx950 HIP engine behind a C ABI)";

  py::class_<Material>(m, "Material")
      .def(py::init<double, double, double>(), py::arg("E"), py::arg("nu"), py::arg("rho"))
      .def(py::init<double, double, double, double, double>(), py::arg("E"), py::arg("nu"), py::arg("rho"),
           py::arg("T_ref"), py::arg("E_slope"))
      .def_readwrite("E", &Material::E)
      .def_readwrite("nu", &Material::nu)
      .def_readwrite("rho", &Material::rho)
      .def_readwrite("T_ref", &Material::T_ref)
      .def_readwrite("E_slope", &Material::E_slope)
      .def("at_temperature", &Material::at_temperature, py::arg("T"))
      .def("validate", &Material::validate)
      .def("constitutive_matrix", &Material::constitutive_matrix)
      .def("__repr__", [](const Material& a) {
        return "<Material E=" + std::to_string(a.E) + " nu=" + std::to_string(a.nu) +
               " rho=" + std::to_string(a.rho) + ">";
      });

  py::class_<NodeSet>(m, "NodeSet")
      .def(py::init<>())
      .def_readwrite("name", &NodeSet::name)
      .def_readwrite("node_ids", &NodeSet::node_ids)
      .def("__repr__", [](const NodeSet& ns) {
        return "<NodeSet '" + ns.name + "' with " + std::to_string(ns.node_ids.size()) + " nodes>";
      });

  py::enum_<BCType>(m, "BCType")
      .value("FIXED", BCType::FIXED)
      .value("DISPLACEMENT", BCType::DISPLACEMENT)
      .value("FRICTIONLESS", BCType::FRICTIONLESS)
      .value("ELASTIC_SUPPORT", BCType::ELASTIC_SUPPORT)
      .value("CYLINDRICAL", BCType::CYLINDRICAL);

  py::class_<ConstraintGroup>(m, "ConstraintGroup")
      .def(py::init<>())
      .def_readwrite("name", &ConstraintGroup::name)
      .def_readwrite("node_ids", &ConstraintGroup::node_ids)
      .def_readwrite("type", &ConstraintGroup::type)
      .def_readwrite("constrained_components", &ConstraintGroup::constrained_components)
      .def_readwrite("surface_normal", &ConstraintGroup::surface_normal)
      .def_readwrite("spring_stiffness", &ConstraintGroup::spring_stiffness)
      .def_readwrite("cylinder_axis", &ConstraintGroup::cylinder_axis)
      .def_readwrite("cylinder_origin", &ConstraintGroup::cylinder_origin);

  py::class_<Mesh>(m, "Mesh")
      .def(py::init<>())
      .def_property(
          "nodes",
          [](const Mesh& a) {
            py::array_t<double> r(std::vector<py::ssize_t>{a.num_nodes(), 3});
            std::copy(a.nodes.begin(), a.nodes.end(), r.mutable_data());
            return r;
          },
          [](Mesh& a, py::array_t<double, py::array::c_style | py::array::forcecast> v) {
            if (v.ndim() != 2 || v.shape(1) != 3) throw std::runtime_error("nodes must be (N, 3)");
            a.nodes.assign(v.data(), v.data() + v.size());
          })
      .def_property(
          "elements",
          [](const Mesh& a) {
            py::array_t<int> r(std::vector<py::ssize_t>{a.num_elements(), 10});
            std::copy(a.elements.begin(), a.elements.end(), r.mutable_data());
            return r;
          },
          [](Mesh& a, py::array_t<int, py::array::c_style | py::array::forcecast> v) {
            if (v.ndim() != 2 || v.shape(1) != 10) throw std::runtime_error("elements must be (E, 10)");
            a.elements.assign(v.data(), v.data() + v.size());
          })
      .def_readwrite("node_sets", &Mesh::node_sets)
      .def_readwrite("left_boundary", &Mesh::left_boundary)
      .def_readwrite("right_boundary", &Mesh::right_boundary)
      .def_readwrite("matched_pairs", &Mesh::matched_pairs)
      .def_readwrite("free_boundary", &Mesh::free_boundary)
      .def_readwrite("num_sectors", &Mesh::num_sectors)
      .def_readwrite("rotation_axis", &Mesh::rotation_axis)
      .def("load_from_gmsh", &Mesh::load_from_gmsh, py::arg("filename"), py::call_guard<py::gil_scoped_release>())
      .def("load_from_arrays", &Mesh::load_from_arrays, py::arg("node_coords"), py::arg("element_connectivity"),
           py::arg("node_sets"), py::arg("num_sectors"), py::arg("rotation_axis") = 2)
      .def("identify_cyclic_boundaries", &Mesh::identify_cyclic_boundaries, py::arg("tolerance") = 1e-6)
      .def("match_boundary_nodes", &Mesh::match_boundary_nodes)
      .def("find_node_set", &Mesh::find_node_set, py::arg("name"), py::return_value_policy::reference_internal)
      .def("num_nodes", &Mesh::num_nodes)
      .def("num_elements", &Mesh::num_elements)
      .def("num_dof", &Mesh::num_dof)
      .def("__repr__", [](const Mesh& a) {
        return "<Mesh " + std::to_string(a.num_nodes()) + " nodes, " + std::to_string(a.num_elements()) +
               " elements, " + std::to_string(a.num_sectors) + " sectors>";
      });

  py::class_<SolverConfig>(m, "SolverConfig")
      .def(py::init<>())
      .def_readwrite("nev", &SolverConfig::nev)
      .def_readwrite("ncv", &SolverConfig::ncv)
      .def_readwrite("shift", &SolverConfig::shift)
      .def_readwrite("tolerance", &SolverConfig::tolerance)
      .def_readwrite("max_iterations", &SolverConfig::max_iterations);

  py::class_<SolverStatus>(m, "SolverStatus")
      .def(py::init<>())
      .def_readonly("converged", &SolverStatus::converged)
      .def_readonly("num_converged", &SolverStatus::num_converged)
      .def_readonly("iterations", &SolverStatus::iterations)
      .def_readonly("max_residual", &SolverStatus::max_residual)
      .def_readonly("message", &SolverStatus::message);

  py::class_<SolverProgress>(m, "SolverProgress")
      .def(py::init<>())
      .def_readonly("completed", &SolverProgress::completed)
      .def_readonly("total", &SolverProgress::total)
      .def_readonly("harmonic_k", &SolverProgress::harmonic_k)
      .def_readonly("converged", &SolverProgress::converged)
      .def_readonly("iterations", &SolverProgress::iterations)
      .def_readonly("num_converged", &SolverProgress::num_converged)
      .def_readonly("num_modes", &SolverProgress::num_modes)
      .def_readonly("max_residual", &SolverProgress::max_residual)
      .def_readonly("min_freq_hz", &SolverProgress::min_freq_hz)
      .def_readonly("max_freq_hz", &SolverProgress::max_freq_hz)
      .def_readonly("elapsed_s", &SolverProgress::elapsed_s);

  py::class_<ModalResult>(m, "ModalResult")
      .def(py::init<>())
      .def_readwrite("harmonic_index", &ModalResult::harmonic_index)
      .def_readwrite("rpm", &ModalResult::rpm)
      .def_property(
          "frequencies",
          [](const ModalResult& r) { return py::array_t<double>(r.frequencies.size(), r.frequencies.data()); },
          [](ModalResult& r, std::vector<double> v) { r.frequencies = std::move(v); })
      .def_readwrite("mode_shapes", &ModalResult::mode_shapes)
      .def_property(
          "whirl_direction",
          [](const ModalResult& r) { return py::array_t<int>(r.whirl_direction.size(), r.whirl_direction.data()); },
          [](ModalResult& r, std::vector<int> v) { r.whirl_direction = std::move(v); })
      .def_readwrite("converged", &ModalResult::converged)
      .def_property_readonly("modal_mass",
                             [](const ModalResult& r) { return py::array_t<double>(r.modal_mass.size(), r.modal_mass.data()); })
      .def_property_readonly("residuals",
                             [](const ModalResult& r) { return py::array_t<double>(r.residuals.size(), r.residuals.data()); })
      .def_property_readonly("eigenvalues",
                             [](const ModalResult& r) { return py::array_t<double>(r.eigenvalues.size(), r.eigenvalues.data()); })
      .def_readonly("restarts", &ModalResult::restarts)
      .def_readonly("solve_seconds", &ModalResult::solve_seconds)
      .def("wave_propagation_velocity", &ModalResult::wave_propagation_velocity, py::arg("radius"))
      .def("__repr__", [](const ModalResult& r) {
        std::string s = "<ModalResult ND=" + std::to_string(r.harmonic_index) + " rpm=" + std::to_string(r.rpm) +
                        " n_modes=" + std::to_string(r.frequencies.size());
        if (!r.frequencies.empty()) {
          s += " freqs=[";
          const int n = std::min<int>(static_cast<int>(r.frequencies.size()), 6);
          for (int i = 0; i < n; ++i) {
            if (i) s += ", ";
            char buf[64];
            std::snprintf(buf, sizeof(buf), "%.1f", r.frequencies[i]);
            s += buf;
            if ((int)r.whirl_direction.size() > i && r.whirl_direction[i] != 0)
              s += r.whirl_direction[i] > 0 ? " FW" : " BW";
          }
          if ((int)r.frequencies.size() > n) s += ", ...";
          s += "] Hz";
        }
        return s + ">";
      });

  py::class_<StationaryFrameResult>(m, "StationaryFrameResult")
      .def(py::init<>())
      .def_property(
          "frequencies",
          [](const StationaryFrameResult& r) { return py::array_t<double>(r.frequencies.size(), r.frequencies.data()); },
          [](StationaryFrameResult& r, std::vector<double> v) { r.frequencies = std::move(v); })
      .def_property(
          "whirl_direction",
          [](const StationaryFrameResult& r) {
            return py::array_t<int>(r.whirl_direction.size(), r.whirl_direction.data());
          },
          [](StationaryFrameResult& r, std::vector<int> v) { r.whirl_direction = std::move(v); });

  py::enum_<FluidConfig::Type>(m, "FluidType")
      .value("NONE", FluidConfig::Type::NONE)
      .value("GAS_AIC", FluidConfig::Type::GAS_AIC)
      .value("LIQUID_ANALYTICAL", FluidConfig::Type::KWAK_ANALYTICAL)
      .value("KWAK_ANALYTICAL", FluidConfig::Type::KWAK_ANALYTICAL)
      .value("POTENTIAL_FLOW_BEM", FluidConfig::Type::POTENTIAL_FLOW_BEM)
      .value("LIQUID_ACOUSTIC_FEM", FluidConfig::Type::LIQUID_ACOUSTIC_FEM);

  py::class_<FluidConfig>(m, "FluidConfig")
      .def(py::init<>())
      .def_readwrite("type", &FluidConfig::type)
      .def_readwrite("fluid_density", &FluidConfig::fluid_density)
      .def_readwrite("disk_radius", &FluidConfig::disk_radius)
      .def_readwrite("disk_thickness", &FluidConfig::disk_thickness)
      .def_readwrite("speed_of_sound", &FluidConfig::speed_of_sound);

  // engine-wide timing of every kernel group with HIP events on the engine stream
  m.def("engine_set_profiling", [](bool on) {
    tm_context* c = Engine::get().ctx();
    std::lock_guard<std::mutex> g(Engine::get().mutex());
    Engine::check(c, tm_set_profiling(c, on ? 1 : 0));
  });
  m.def("engine_reset_kernel_stats", []() {
    tm_context* c = Engine::get().ctx();
    std::lock_guard<std::mutex> g(Engine::get().mutex());
    Engine::check(c, tm_reset_kernel_stats(c));
  });
  m.def("engine_kernel_stats", []() {
    tm_context* c = Engine::get().ctx();
    std::lock_guard<std::mutex> g(Engine::get().mutex());
    int n = tm_kernel_stats(c, nullptr, 0);
    if (n < 0) Engine::check(c, n);
    std::vector<tm_kernel_stat> v(std::max(n, 1));
    tm_kernel_stats(c, v.data(), n);
    py::dict d;
    for (int i = 0; i < n; ++i) {
      py::dict e;
      e["launches"] = v[i].launches;
      e["total_ms"] = v[i].total_ms;
      e["alg_bytes"] = v[i].alg_bytes;
      e["alg_flops"] = v[i].alg_flops;
      d[py::str(v[i].name)] = e;
    }
    return d;
  });
  m.def("engine_free_bytes", []() { return tm_device_free_bytes(Engine::get().ctx()); });

  py::class_<CyclicSymmetrySolver>(m, "CyclicSymmetrySolver")
      .def(py::init<const Mesh&, const Material&, const FluidConfig&, bool>(), py::arg("mesh"), py::arg("material"),
           py::arg("fluid") = FluidConfig(), py::arg("apply_hub_constraint") = true, py::keep_alive<1, 2>())
      .def(py::init<const Mesh&, const Material&, const std::vector<ConstraintGroup>&, const FluidConfig&>(),
           py::arg("mesh"), py::arg("material"), py::arg("constraints"), py::arg("fluid") = FluidConfig(),
           py::keep_alive<1, 2>())
      .def("solve_at_rpm", &CyclicSymmetrySolver::solve_at_rpm, py::arg("rpm"), py::arg("num_modes_per_harmonic"),
           py::arg("harmonic_indices") = std::vector<int>{}, py::arg("max_threads") = 0,
           py::arg("include_coriolis") = false, py::arg("min_frequency") = 0.0, py::arg("progress_cb") = py::none(),
           py::arg("allow_condensation") = false, py::arg("memory_reserve_fraction") = 0.2)
      .def("solve_rpm_sweep", &CyclicSymmetrySolver::solve_rpm_sweep, py::arg("rpm_values"),
           py::arg("num_modes_per_harmonic"), py::arg("harmonic_indices") = std::vector<int>{},
           py::arg("max_threads") = 0, py::arg("include_coriolis") = false, py::arg("min_frequency") = 0.0,
           py::arg("memory_reserve_fraction") = 0.2)
      .def("engine_info", &CyclicSymmetrySolver::engine_info)
      .def_property(
          "engine_method",
          [](const CyclicSymmetrySolver& s) { return std::string(s.method_ == TM_METHOD_BLOCK ? "block" : "lanczos"); },
          [](CyclicSymmetrySolver& s, const std::string& v) {
            if (v == "block") s.method_ = TM_METHOD_BLOCK;
            else if (v == "lanczos") s.method_ = TM_METHOD_LANCZOS;
            else throw std::invalid_argument("engine_method: 'block' or 'lanczos'");
          })
      .def_property(
          "engine_precision",
          [](const CyclicSymmetrySolver& s) { return std::string(s.precision_ == TM_PREC_MIXED ? "mixed" : "fp64"); },
          [](CyclicSymmetrySolver& s, const std::string& v) {
            if (v == "fp64") s.precision_ = TM_PREC_FP64;
            else if (v == "mixed") s.precision_ = TM_PREC_MIXED;
            else throw std::invalid_argument("engine_precision: 'fp64' or 'mixed'");
          })
      .def("build_constrained_dofs", &CyclicSymmetrySolver::build_constrained_dofs)
      .def_static("compute_stationary_frame", &CyclicSymmetrySolver::compute_stationary_frame,
                  py::arg("rotating_result"), py::arg("num_sectors"));
}
