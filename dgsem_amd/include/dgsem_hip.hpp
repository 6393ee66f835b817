// dgsem_hip.hpp -- header-only C++ mirror of the reference's solver
// interface over the C ABI (include/dgsem_amd.h).
//
// A user of RetrO598/DGSEM composes a solver from template arguments,
//   StructuredSolver<Equations, Basis, VolumeFlux, SurfaceFlux, Mesh, BoundarySet>
// (src/solver/structured_solver.hpp:19-58) and drives it with
//   SSPRK3<T, Solver, Mesh, Solution>::step / TimeIntegrator::solve
// (src/time_integrator/SSPRK3.hpp:145-183, integrator_base.hpp:21-40).
// The same spellings compile against this header; the template arguments are
// mapped onto a dgsem_config and every RHS / stage runs in the HIP kernels.
// Device memory for the state comes from hipMalloc; nothing here needs
// Kokkos.  Errors raise std::runtime_error carrying dgsem_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/dgsem_amd.h"

namespace dgsem_hip {

// --- equations (src/equations/*): value_type, NDIMS, NVARS, get_gamma -------
namespace equations {
template <class T>
struct CompressibleEuler2D {
  using value_type = T;
  static constexpr std::size_t NDIMS = 2, NVARS = 4;
  explicit CompressibleEuler2D(T gamma = T(1.4)) : gamma_(gamma) {}
  T get_gamma() const { return gamma_; }
  T gamma_;
};
template <class T>
struct CompressibleEuler3D {
  using value_type = T;
  static constexpr std::size_t NDIMS = 3, NVARS = 5;
  explicit CompressibleEuler3D(T gamma = T(1.4)) : gamma_(gamma) {}
  T get_gamma() const { return gamma_; }
  T gamma_;
};
template <class T>
struct CompressibleNavierStokes3D {
  using value_type = T;
  static constexpr std::size_t NDIMS = 3, NVARS = 5, NGRAD_VARS = 5;
  explicit CompressibleNavierStokes3D(T gamma = T(1.4), T mu = T(6.25e-4), T prandtl = T(0.72))
      : gamma_(gamma), mu_(mu), prandtl_(prandtl) {}
  T get_gamma() const { return gamma_; }
  T get_mu() const { return mu_; }
  T get_prandtl() const { return prandtl_; }
  T gamma_, mu_, prandtl_;
};
template <class T>
struct CompressibleNavierStokes2D {
  using value_type = T;
  static constexpr std::size_t NDIMS = 2, NVARS = 4, NGRAD_VARS = 4;
  explicit CompressibleNavierStokes2D(T gamma = T(1.4), T mu = T(5.0e-3), T prandtl = T(0.73))
      : gamma_(gamma), mu_(mu), prandtl_(prandtl) {}
  T get_gamma() const { return gamma_; }
  T get_mu() const { return mu_; }
  T get_prandtl() const { return prandtl_; }
  T gamma_, mu_, prandtl_;
};
template <class E>
struct is_navier_stokes : std::false_type {};
template <class T>
struct is_navier_stokes<CompressibleNavierStokes2D<T>> : std::true_type {};
template <class T>
struct is_navier_stokes<CompressibleNavierStokes3D<T>> : std::true_type {};
}  // namespace equations

// --- basis (polynomial_basis/basis.hpp) --------------------------------------
namespace Basis {
template <class T, std::size_t Polydeg>
struct LobattoLegendreBasis {
  static constexpr std::size_t NNodes = Polydeg + 1;
  static constexpr int polydeg = (int)Polydeg;
};
}  // namespace Basis

// --- flux tags (base/numerical_flux.hpp) -------------------------------------
template <class Eq>
struct ChandrashekarFlux {
  static constexpr int kind = DGSEM_FLUX_CHANDRASHEKAR;
};
template <class Eq>
struct LaxFriedrichsFlux {
  static constexpr int kind = DGSEM_FLUX_LLF;
};
template <class Eq>
struct HLLCFlux {
  static constexpr int kind = DGSEM_FLUX_HLLC;
};

// --- indicator and volume-integral policies (volume_flux.hpp:686-879) --------
template <class Basis, class Eq>
struct HGIndicator {};

template <class Basis, class Eq, template <class> class NumericFlux>
struct VolumeIntegralSplitForm {
  static constexpr int policy = DGSEM_VOLUME_SPLIT_FORM;
  static constexpr int volume_flux = NumericFlux<Eq>::kind;
  static constexpr int fv_flux = DGSEM_FLUX_LLF;
};

struct VolumeIntegralShockCapturingBase {};

template <class Basis, class Eq, template <class> class NumericFlux, template <class> class FvFlux,
          class Indicator>
struct VolumeIntegralShockCapturingHG : VolumeIntegralShockCapturingBase {
  static constexpr int policy = DGSEM_VOLUME_SHOCK_CAPTURING;
  static constexpr int volume_flux = NumericFlux<Eq>::kind;
  static constexpr int fv_flux = FvFlux<Eq>::kind;
};

// --- mesh and boundaries (mesh/structured_mesh.hpp, boundary_condition) ------
template <class T, std::size_t NDIM>
class StructuredMesh {
 public:
  StructuredMesh(const std::array<T, NDIM>& left, const std::array<T, NDIM>& right,
                 const std::array<std::size_t, NDIM>& n_cells)
      : left_(left), right_(right), n_cells_(n_cells) {}
  std::array<std::size_t, NDIM> get_num_cells() const { return n_cells_; }
  std::array<T, NDIM> get_domain_left() const { return left_; }
  std::array<T, NDIM> get_domain_right() const { return right_; }
  std::size_t get_nelem() const {
    std::size_t e = 1;
    for (auto c : n_cells_) e *= c;
    return e;
  }

 private:
  std::array<T, NDIM> left_, right_;
  std::array<std::size_t, NDIM> n_cells_;
};

// Boundary conditions (src/boundary_condition/*.hpp).  Faces are ordered
// -x, +x, -y, +y[, -z, +z] as in BoundarySet (boundary_set.hpp:175-240).
struct PeriodicBC {
  static constexpr int kind = DGSEM_BC_PERIODIC;
};
struct SlipWallBC {  // slip_wall_boundary.hpp:10-66
  static constexpr int kind = DGSEM_BC_SLIP_WALL;
};
struct OutflowBC {  // outflow_boundary.hpp:10-110
  static constexpr int kind = DGSEM_BC_OUTFLOW;
};
struct NSNoSlipAdiabaticWallBC {  // navier_stokes_wall_boundary.hpp:39-146
  static constexpr int kind = DGSEM_BC_NS_NOSLIP_WALL;
};
struct NSSlipAdiabaticWallBC {  // navier_stokes_wall_boundary.hpp:148-217
  static constexpr int kind = DGSEM_BC_NS_SLIP_WALL;
};
// DirichletBC<Func> (dirichlet_boundary.hpp:11-196): func is either a state
// indexable by variable (a constant outer state) or a callable
// func(coord, t) -> state, evaluated here on the host at the boundary nodes
// and uploaded whenever the solver is asked for a new time.
template <class Func>
struct DirichletBC {
  static constexpr int kind = DGSEM_BC_DIRICHLET;
  Func func;
  explicit DirichletBC(Func func_) : func(func_) {}
};
// MixedDirichletSlipWallBC<Func, Predicate> (dirichlet_boundary.hpp:198-275):
// func(coord, t) where use_dirichlet(coord, t), the slip wall elsewhere.
template <class Func, class Predicate>
struct MixedDirichletSlipWallBC {
  static constexpr int kind = DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL;
  Func func;
  Predicate use_dirichlet;
  MixedDirichletSlipWallBC(Func func_, Predicate use_dirichlet_)
      : func(func_), use_dirichlet(use_dirichlet_) {}
};

template <class... BCs>
struct BoundarySet {
  static constexpr std::size_t NFACES = sizeof...(BCs);
  static constexpr std::array<int, sizeof...(BCs)> kinds{BCs::kind...};
  std::tuple<BCs...> bcs;
  explicit BoundarySet(BCs... b) : bcs(b...) {}
  BoundarySet() = default;
};

inline void check(int rc, const dgsem_handle* h, const char* what) {
  if (rc != DGSEM_OK)
    throw std::runtime_error(std::string(what) + ": " + (h ? dgsem_last_error(h) : "error"));
}

// --- solution: caller-owned device arrays (solution_container.hpp:157-291) ---
template <class Mesh, class Basis, class Eq>
class Solution {
 public:
  using value_type = typename Eq::value_type;
  static constexpr std::size_t NDIMS = Eq::NDIMS, NVARS = Eq::NVARS;

  explicit Solution(const Mesh& mesh) {
    std::size_t nn = 1;
    for (std::size_t d = 0; d < NDIMS; ++d) nn *= Basis::NNodes;
    size_ = mesh.get_nelem() * nn * NVARS;
    u_ = alloc();
    du_ = alloc();
  }
  Solution clone_shape() const {
    Solution s;
    s.size_ = size_;
    s.u_ = s.alloc();
    s.du_ = s.alloc();
    return s;
  }
  value_type* state() { return u_.get(); }
  const value_type* state() const { return u_.get(); }
  value_type* rhs() { return du_.get(); }
  const value_type* rhs() const { return du_.get(); }
  std::size_t size() const { return size_; }

 private:
  Solution() = default;
  struct Free {
    void operator()(value_type* p) const { (void)hipFree(p); }
  };
  std::shared_ptr<value_type> alloc() const {
    void* p = nullptr;
    if (hipMalloc(&p, size_ * sizeof(value_type)) != hipSuccess) throw std::bad_alloc();
    return std::shared_ptr<value_type>(static_cast<value_type*>(p), Free{});
  }
  std::size_t size_ = 0;
  std::shared_ptr<value_type> u_, du_;
};

// --- the solver (structured_solver.hpp:19-125) --------------------------------
template <class Equations, class BasisT, class VolumeFlux, class SurfaceFlux, class Mesh,
          class BoundarySetType>
class StructuredSolver {
 public:
  using value_type = typename Equations::value_type;
  using EquationType = Equations;
  static constexpr std::size_t NDIMS = Equations::NDIMS, NVARS = Equations::NVARS;
  static_assert(std::is_same_v<value_type, double>, "the GPU path computes in FP64");
  static_assert(BoundarySetType::NFACES == 2 * NDIMS,
                "StructuredSolver requires one boundary condition per mesh face");
  using solution = Solution<Mesh, BasisT, Equations>;

  StructuredSolver(const Equations& eq, const Mesh& mesh, const BoundarySetType& bcs,
                   hipStream_t stream = nullptr)
      : eq_(eq), mesh_(mesh), bcs_(bcs), stream_(stream) {
    dgsem_config c;
    dgsem_config_default(&c, (int)NDIMS, BasisT::polydeg);
    const auto nc = mesh.get_num_cells();
    const auto L = mesh.get_domain_left();
    const auto R = mesh.get_domain_right();
    for (std::size_t d = 0; d < NDIMS; ++d) {
      c.cells[d] = (int)nc[d];
      c.domain_left[d] = L[d];
      c.domain_right[d] = R[d];
    }
    for (std::size_t f = 0; f < 2 * NDIMS; ++f) c.face_boundary[f] = BoundarySetType::kinds[f];
    constant_dirichlet_states(c, std::make_index_sequence<2 * NDIMS>{});
    c.gamma = eq.get_gamma();
    if constexpr (equations::is_navier_stokes<Equations>::value) {
      c.equation = DGSEM_NAVIER_STOKES;
      c.mu = eq.get_mu();
      c.prandtl = eq.get_prandtl();
    }
    c.volume_integral = VolumeFlux::policy;
    c.volume_flux = VolumeFlux::volume_flux;
    c.fv_flux = VolumeFlux::fv_flux;
    c.surface_flux = SurfaceFlux::kind;
    dgsem_handle* h = nullptr;
    const int rc = dgsem_create(&c, stream, &h);
    h_.reset(h);
    check(rc, h, "dgsem_create");
    int64_t E = 0;
    int nn = 0, nv = 0;
    check(dgsem_sizes(h, &E, &nn, &nv), h, "dgsem_sizes");
    n_elements_ = (std::size_t)E;
    n_dofs_ = (std::size_t)nn;
    prepare_boundary_nodes(std::make_index_sequence<2 * NDIMS>{});
    update_boundary_values(value_type(0));
  }

  // Re-evaluates the time-dependent boundary states (callable Dirichlet and
  // mixed faces) at time t and uploads them; a no-op for constant data or an
  // unchanged time.  Synchronises the solver's stream.
  void update_boundary_values(value_type t) {
    if (!has_timed_faces_ || (values_time_set_ && t == values_time_)) return;
    evaluate_faces(t, std::make_index_sequence<2 * NDIMS>{});
    values_time_ = t;
    values_time_set_ = true;
  }

  // A curvilinear mesh: the element geometry of `mapping` (any type with
  // eval(std::array<T, NDIMS> reference_coordinate) -> physical coordinate,
  // like LinearMapping, mapping.hpp:31-50) instead of the mesh's linear map,
  // as StructuredElementInitializer<T, Basis, Mapping, NDIMS> builds it
  // (data_container.hpp:729-768).  The mapping runs on the host; the metric
  // terms are computed on the GPU.
  template <class Mapping>
    requires requires(const Mapping& m, std::array<value_type, NDIMS> x) { m.eval(x); }
  StructuredSolver(const Equations& eq, const Mesh& mesh, const Mapping& mapping,
                   const BoundarySetType& bcs, hipStream_t stream = nullptr)
      : StructuredSolver(eq, mesh, bcs, stream) {
    std::vector<double> x(n_elements_ * n_dofs_ * NDIMS);
    check(dgsem_reference_coordinates(h_.get(), x.data()), h_.get(), "reference_coordinates");
    for (std::size_t i = 0; i < n_elements_ * n_dofs_; ++i) {
      std::array<value_type, NDIMS> r;
      for (std::size_t d = 0; d < NDIMS; ++d) r[d] = x[i * NDIMS + d];
      const auto p = mapping.eval(r);
      for (std::size_t d = 0; d < NDIMS; ++d) x[i * NDIMS + d] = p[d];
    }
    void* dev = nullptr;
    if (hipMalloc(&dev, x.size() * sizeof(double)) != hipSuccess) throw std::bad_alloc();
    std::unique_ptr<double, FreeDevice> owned(static_cast<double*>(dev));
    if (hipMemcpy(dev, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("hipMemcpy to device failed");
    check(dgsem_set_node_coordinates(h_.get(), owned.get()), h_.get(), "set_node_coordinates");
    if (has_timed_faces_) {  // boundary nodes moved with the mapping
      values_time_set_ = false;
      prepare_boundary_nodes(std::make_index_sequence<2 * NDIMS>{});
      update_boundary_values(value_type(0));
    }
  }

  // set_indicator_parameters, structured_solver.hpp:118-123
  void set_indicator_parameters(value_type amax, value_type amin, bool smooth) {
    check(dgsem_set_indicator_parameters(h_.get(), amax, amin, smooth ? 1 : 0), h_.get(),
          "set_indicator_parameters");
  }
  std::size_t get_ndofs() const { return n_dofs_; }

  // initialize(initial_condition, sol), structured_solver.hpp:145-155: the
  // initial-condition functor is evaluated at the node coordinates (on the
  // host here) and packed into the device layout.
  template <class IC>
  void initialize(const IC& ic, solution& sol) {
    std::vector<double> x(n_elements_ * n_dofs_ * NDIMS);
    check(dgsem_node_coordinates(h_.get(), x.data()), h_.get(), "node_coordinates");
    std::vector<double> u(n_elements_ * n_dofs_ * NVARS);
    for (std::size_t i = 0; i < n_elements_ * n_dofs_; ++i) {
      std::array<double, NDIMS> xc;
      for (std::size_t d = 0; d < NDIMS; ++d) xc[d] = x[i * NDIMS + d];
      const auto s = ic(xc);
      for (std::size_t v = 0; v < NVARS; ++v) u[i * NVARS + v] = s[v];
    }
    upload_reference_layout(u, sol);
  }

  // Host state in the reference's LayoutRight (e, dof, var) <-> device.
  void upload_reference_layout(const std::vector<double>& u, solution& sol) {
    double* tmp = sol.rhs();  // scratch for the packed copy
    if (hipMemcpy(tmp, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("hipMemcpy to device failed");
    check(dgsem_pack_state(h_.get(), tmp, sol.state()), h_.get(), "pack_state");
    check(dgsem_synchronize(h_.get()), h_.get(), "synchronize");
  }
  std::vector<double> download_reference_layout(solution& sol) {
    std::vector<double> u(sol.size());
    check(dgsem_unpack_state(h_.get(), sol.state(), sol.rhs()), h_.get(), "unpack_state");
    check(dgsem_synchronize(h_.get()), h_.get(), "synchronize");
    if (hipMemcpy(u.data(), sol.rhs(), u.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("hipMemcpy to host failed");
    return u;
  }

  // calc_rhs(sol, time), structured_solver.hpp:463-480
  void calc_rhs(solution& sol, value_type time) {
    update_boundary_values(time);
    check(dgsem_rhs(h_.get(), sol.state(), sol.rhs(), time), h_.get(), "calc_rhs");
  }

  // one fused SSPRK3 stage (RHS at `time` + parallel_ma2/ma3)
  void ssprk3_stage(int stage, value_type dt, const value_type* u_base, const value_type* u_k,
                    value_type* u_out, value_type time = value_type(0)) {
    update_boundary_values(time);
    check(dgsem_ssprk3_stage(h_.get(), stage, dt, u_base, u_k, u_out), h_.get(), "ssprk3_stage");
  }
  void synchronize() { check(dgsem_synchronize(h_.get()), h_.get(), "synchronize"); }
  dgsem_handle* handle() { return h_.get(); }

 private:
  struct Destroy {
    void operator()(dgsem_handle* h) const { dgsem_destroy(h); }
  };
  struct FreeDevice {
    void operator()(double* p) const { (void)hipFree(p); }
  };

  template <class BC>
  static constexpr bool timed_dirichlet() {
    if constexpr (BC::kind == DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL) {
      return true;
    } else if constexpr (BC::kind == DGSEM_BC_DIRICHLET) {
      return std::is_invocable_v<decltype(std::declval<BC>().func),
                                 const std::array<value_type, NDIMS>&, value_type>;
    } else {
      return false;
    }
  }

  template <std::size_t... F>
  void constant_dirichlet_states(dgsem_config& c, std::index_sequence<F...>) {
    (constant_state<F>(c), ...);
  }
  template <std::size_t F>
  void constant_state(dgsem_config& c) {
    using BC = std::tuple_element_t<F, decltype(bcs_.bcs)>;
    if constexpr (BC::kind == DGSEM_BC_DIRICHLET && !timed_dirichlet<BC>()) {
      const auto& s = std::get<F>(bcs_.bcs).func;
      for (std::size_t v = 0; v < NVARS; ++v) c.face_state[F][v] = s[v];
    }
  }

  // Boundary node coordinates of every face with time-dependent data, in
  // the order the ABI reads per-node states: (element position along the
  // face) * (N+1) + (node position along the face).
  template <std::size_t... F>
  void prepare_boundary_nodes(std::index_sequence<F...>) {
    (prepare_face<F>(), ...);
  }
  template <std::size_t F>
  void prepare_face() {
    using BC = std::tuple_element_t<F, decltype(bcs_.bcs)>;
    if constexpr (timed_dirichlet<BC>()) {
      static_assert(NDIMS == 2, "non-periodic boundaries are built for 2D meshes");
      has_timed_faces_ = true;
      std::vector<double> x(n_elements_ * n_dofs_ * NDIMS);
      check(dgsem_node_coordinates(h_.get(), x.data()), h_.get(), "node_coordinates");
      const auto nc = mesh_.get_num_cells();
      const std::size_t n = BasisT::NNodes, dir = F / 2, upper = F % 2;
      const std::size_t count = dir == 0 ? nc[1] : nc[0];
      auto& coords = face_coords_[F];
      coords.clear();
      for (std::size_t idx = 0; idx < count; ++idx)
        for (std::size_t q = 0; q < n; ++q) {
          const std::size_t ie = dir == 0 ? (upper ? nc[0] - 1 : 0) : idx;
          const std::size_t je = dir == 0 ? idx : (upper ? nc[1] - 1 : 0);
          const std::size_t i = dir == 0 ? (upper ? n - 1 : 0) : q;
          const std::size_t j = dir == 0 ? q : (upper ? n - 1 : 0);
          const std::size_t node = (ie * nc[1] + je) * n_dofs_ + j * n + i;
          coords.push_back({x[node * NDIMS], x[node * NDIMS + 1]});
        }
      const std::size_t width = BC::kind == DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL ? NVARS + 1 : NVARS;
      face_host_[F].assign(coords.size() * width, 0.0);
      void* p = nullptr;
      if (hipMalloc(&p, face_host_[F].size() * sizeof(double)) != hipSuccess) throw std::bad_alloc();
      face_dev_[F].reset(static_cast<double*>(p));
      check(dgsem_set_boundary_values(h_.get(), (int)F, face_dev_[F].get()), h_.get(),
            "set_boundary_values");
    }
  }

  template <std::size_t... F>
  void evaluate_faces(value_type t, std::index_sequence<F...>) {
    (evaluate_face<F>(t), ...);
  }
  template <std::size_t F>
  void evaluate_face(value_type t) {
    using BC = std::tuple_element_t<F, decltype(bcs_.bcs)>;
    if constexpr (timed_dirichlet<BC>()) {
      const auto& bc = std::get<F>(bcs_.bcs);
      const bool mixed = BC::kind == DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL;
      const std::size_t width = mixed ? NVARS + 1 : NVARS;
      auto& host = face_host_[F];
      // the previous upload may still be read by queued kernels
      if (hipStreamSynchronize(stream_) != hipSuccess) throw std::runtime_error("hipStreamSynchronize");
      for (std::size_t k = 0; k < face_coords_[F].size(); ++k) {
        const auto& xc = face_coords_[F][k];
        const auto s = bc.func(xc, t);
        for (std::size_t v = 0; v < NVARS; ++v) host[k * width + v] = s[v];
        if constexpr (BC::kind == DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL)
          host[k * width + NVARS] = bc.use_dirichlet(xc, t) ? 1.0 : 0.0;
      }
      if (hipMemcpyAsync(face_dev_[F].get(), host.data(), host.size() * sizeof(double),
                         hipMemcpyHostToDevice, stream_) != hipSuccess ||
          hipStreamSynchronize(stream_) != hipSuccess)
        throw std::runtime_error("boundary state upload failed");
    }
  }

  Equations eq_;
  Mesh mesh_;
  BoundarySetType bcs_;
  hipStream_t stream_ = nullptr;
  std::unique_ptr<dgsem_handle, Destroy> h_;
  std::size_t n_elements_ = 0, n_dofs_ = 0;
  bool has_timed_faces_ = false, values_time_set_ = false;
  value_type values_time_ = value_type(0);
  std::array<std::vector<std::array<double, NDIMS>>, 2 * NDIMS> face_coords_;
  std::array<std::vector<double>, 2 * NDIMS> face_host_;
  std::array<std::unique_ptr<double, FreeDevice>, 2 * NDIMS> face_dev_;
};

// --- time integration (SSPRK3.hpp:145-183, integrator_base.hpp:21-40) --------
template <class T, class Solver, class Mesh, class SolutionT>
class SSPRK3 {
 public:
  SSPRK3(const SolutionT& sol, const Mesh& mesh, T final_time = T(0))
      : tmp1_(sol.clone_shape()), tmp2_(sol.clone_shape()), mesh_(mesh), final_time_(final_time) {}

  // u1 = u + dt L(u); u2 = 3/4 u + 1/4 u1 + dt/4 L(u1); u = 1/3 u + 2/3 u2 + 2dt/3 L(u2),
  // with L evaluated at t, t + dt and t + dt/2 (SSPRK3.hpp:159-181)
  void step(Solver& solver, SolutionT& sol, T dt) {
    solver.ssprk3_stage(1, dt, nullptr, sol.state(), tmp1_.state(), time_);
    solver.ssprk3_stage(2, dt, sol.state(), tmp1_.state(), tmp2_.state(), time_ + dt);
    solver.ssprk3_stage(3, dt, sol.state(), tmp2_.state(), sol.state(), time_ + T(0.5) * dt);
  }
  void solve(Solver& solver, SolutionT& sol, T dt) {
    while (time_ < final_time_) {
      step(solver, sol, dt);
      time_ += dt;
      ++iter_;
    }
    solver.synchronize();
  }
  T get_time() const { return time_; }

 private:
  SolutionT tmp1_, tmp2_;
  const Mesh& mesh_;
  T time_ = T(0), final_time_;
  int iter_ = 0;
};

}  // namespace dgsem_hip
