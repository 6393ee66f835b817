// Two-dimensional Riemann problem, composed the way the reference's example
// is (examples/Euler2D/riemann2d/riemann2d.cpp) but against dgsem_hip.hpp:
// four constant quadrants meeting at (0.8, 0.8) on the unit square, outflow
// on every face, N = 5, shock capturing with alpha_max = 0.5,
// alpha_min = 0.001, dt = 0.2 min(dx, dy) / ((2 (N+1) - 1) 2).
// Usage: riemann2d nx ny steps out_prefix
// writes out_prefix.u0.bin and out_prefix.u.bin, the initial and final
// states in the reference layout (e, dof, var) as raw FP64.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../include/dgsem_hip.hpp"

namespace {

std::array<double, 4> prim_to_cons(double rho, double vx, double vy, double p, double gamma) {
  return {rho, rho * vx, rho * vy, p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)};
}

struct Quadrants {
  double gamma;
  std::array<double, 4> operator()(const std::array<double, 2>& c) const {
    const bool right = c[0] >= 0.8, top = c[1] >= 0.8;
    if (right && top) return prim_to_cons(1.5, 0.0, 0.0, 1.5, gamma);
    if (top) return prim_to_cons(0.5323, 1.206, 0.0, 0.3, gamma);
    if (!right) return prim_to_cons(0.138, 1.206, 1.206, 0.029, gamma);
    return prim_to_cons(0.5323, 0.0, 1.206, 0.3, gamma);
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
  using MyBasis = Basis::LobattoLegendreBasis<value_type, 5>;
  using SurfaceFlux = LaxFriedrichsFlux<Eq>;
  using VolumeFlux = VolumeIntegralShockCapturingHG<MyBasis, Eq, ChandrashekarFlux,
                                                    LaxFriedrichsFlux, HGIndicator<MyBasis, Eq>>;
  using Mesh = StructuredMesh<value_type, 2>;

  if (argc < 5) {
    std::fprintf(stderr, "usage: %s nx ny steps out_prefix\n", argv[0]);
    return 2;
  }
  const std::size_t nx = std::strtoull(argv[1], nullptr, 10), ny = std::strtoull(argv[2], nullptr, 10);
  const int steps = std::atoi(argv[3]);
  const std::string prefix = argv[4];
  constexpr double gamma = 1.4;
  try {
    auto boundaries = BoundarySet(OutflowBC{}, OutflowBC{}, OutflowBC{}, OutflowBC{});
    using Solver = StructuredSolver<Eq, MyBasis, VolumeFlux, SurfaceFlux, Mesh, decltype(boundaries)>;
    using Sol = Solver::solution;
    Mesh mesh({0.0, 0.0}, {1.0, 1.0}, {nx, ny});
    Solver solver(Eq{gamma}, mesh, boundaries);
    solver.set_indicator_parameters(0.5, 0.001, false);
    Sol sol(mesh);
    solver.initialize(Quadrants{gamma}, sol);
    if (!write(prefix + ".u0.bin", solver.download_reference_layout(sol))) return 1;
    const double dx = 1.0 / nx, dy = 1.0 / ny;
    const double dt = 0.2 * std::min(dx, dy) / ((2.0 * MyBasis::NNodes - 1.0) * 2.0);
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
