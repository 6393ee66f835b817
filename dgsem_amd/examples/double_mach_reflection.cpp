// Double Mach reflection, composed the way the reference's example is
// (examples/Euler2D/double_mach_reflection/double_mach_reflection.cpp) but
// against dgsem_hip.hpp: a time-dependent Dirichlet inflow on the left and
// top, outflow on the right, and the mixed Dirichlet/slip-wall bottom
// (inflow for x < 1/6, reflecting wall beyond).  N = 4, shock capturing.
// Usage: double_mach_reflection nx ny steps dt out_prefix
// writes out_prefix.u0.bin and out_prefix.u.bin, the initial and final
// states in the reference layout (e, dof, var) as raw FP64.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numbers>
#include <string>

#include "../include/dgsem_hip.hpp"

namespace {

// primitive (rho, vx, vy, p) -> conservative
std::array<double, 4> prim_to_cons(double rho, double vx, double vy, double p, double gamma) {
  return {rho, rho * vx, rho * vy, p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)};
}

// post-shock state left of the oblique Mach-10 shock, quiescent gas right of it
std::array<double, 4> double_mach_state(const std::array<double, 2>& x, double t, double gamma) {
  const double shock = 1.0 / 6.0 + (x[1] + 20.0 * t) / std::sqrt(3.0);
  if (x[0] < shock) {
    const double phi = std::numbers::pi / 6.0;
    return prim_to_cons(8.0, 8.25 * std::cos(phi), -8.25 * std::sin(phi), 116.5, gamma);
  }
  return prim_to_cons(1.4, 0.0, 0.0, 1.0, gamma);
}

struct BoundaryState {
  double gamma;
  std::array<double, 4> operator()(const std::array<double, 2>& x, double t) const {
    return double_mach_state(x, t, gamma);
  }
};
struct BottomInflowRegion {
  bool operator()(const std::array<double, 2>& x, double /*t*/) const { return x[0] < 1.0 / 6.0; }
};
struct Initial {
  double gamma;
  std::array<double, 4> operator()(const std::array<double, 2>& x) const {
    return double_mach_state(x, 0.0, gamma);
  }
};

bool write(const std::string& path, const std::vector<double>& u) {
  FILE* f = std::fopen(path.c_str(), "wb");
  const bool ok = f && std::fwrite(u.data(), sizeof(double), u.size(), f) == u.size();
  if (f) std::fclose(f);
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  using namespace dgsem_hip;
  using value_type = double;
  using Eq = equations::CompressibleEuler2D<value_type>;
  using MyBasis = Basis::LobattoLegendreBasis<value_type, 4>;
  using SurfaceFlux = LaxFriedrichsFlux<Eq>;
  using VolumeFlux = VolumeIntegralShockCapturingHG<MyBasis, Eq, ChandrashekarFlux,
                                                    LaxFriedrichsFlux, HGIndicator<MyBasis, Eq>>;
  using Mesh = StructuredMesh<value_type, 2>;

  if (argc < 6) {
    std::fprintf(stderr, "usage: %s nx ny steps dt out_prefix\n", argv[0]);
    return 2;
  }
  const std::size_t nx = std::strtoull(argv[1], nullptr, 10), ny = std::strtoull(argv[2], nullptr, 10);
  const int steps = std::atoi(argv[3]);
  const double dt = std::strtod(argv[4], nullptr);
  const std::string prefix = argv[5];
  constexpr double gamma = 1.4;
  try {
    const BoundaryState inflow{gamma};
    auto boundaries = BoundarySet(DirichletBC(inflow), OutflowBC{},
                                  MixedDirichletSlipWallBC(inflow, BottomInflowRegion{}),
                                  DirichletBC(inflow));
    using Solver = StructuredSolver<Eq, MyBasis, VolumeFlux, SurfaceFlux, Mesh, decltype(boundaries)>;
    using Sol = Solver::solution;
    Mesh mesh({0.0, 0.0}, {4.0, 1.0}, {nx, ny});
    Solver solver(Eq{gamma}, mesh, boundaries);
    solver.set_indicator_parameters(0.5, 0.001, false);
    Sol sol(mesh);
    solver.initialize(Initial{gamma}, sol);
    if (!write(prefix + ".u0.bin", solver.download_reference_layout(sol))) return 1;
    // solve() steps while t < final time: (steps - 1/2) dt gives exactly `steps`
    SSPRK3<value_type, Solver, Mesh, Sol> rk(sol, mesh, (steps - 0.5) * dt);
    rk.solve(solver, sol, dt);
    if (!write(prefix + ".u.bin", solver.download_reference_layout(sol))) return 1;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
