// Astrophysical jet, composed the way the reference's example is
// (examples/Euler2D/astro_jet/astro_jet.cpp) but against dgsem_hip.hpp:
// Dirichlet faces left and right, periodic in y, N = 3, shock capturing with
// alpha_max = 0.3, alpha_min = 1e-4.  The left face's state turns into the
// Mach ~2350 jet (rho 5, vx 800, p 0.4127) on |y| < 0.05 once t > 0, so the RHS at a
// step's first stage sees the ambient state and the later stages the jet.
// Usage: astro_jet nx ny steps dt out_prefix
// writes out_prefix.u0.bin and out_prefix.u.bin, the initial and final
// states in the reference layout (e, dof, var) as raw FP64.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../include/dgsem_hip.hpp"

namespace {

std::array<double, 4> prim_to_cons(double rho, double vx, double vy, double p, double gamma) {
  return {rho, rho * vx, rho * vy, p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)};
}

// ambient gas at rest; the jet on the left face for |y| < 0.05 and t > 0
struct JetState {
  double gamma;
  std::array<double, 4> operator()(const std::array<double, 2>& x, double t) const {
    if (t > 0.0 && std::abs(x[0] + 0.5) < 1e-8 && std::abs(x[1]) < 0.05)
      return prim_to_cons(5.0, 800.0, 0.0, 0.4127, gamma);
    return prim_to_cons(0.5, 0.0, 0.0, 0.4127, gamma);
  }
};
struct Initial {
  double gamma;
  std::array<double, 4> operator()(const std::array<double, 2>&) const {
    return prim_to_cons(0.5, 0.0, 0.0, 0.4127, gamma);
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
  using MyBasis = Basis::LobattoLegendreBasis<value_type, 3>;
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
    const JetState inflow{gamma};
    auto boundaries = BoundarySet(DirichletBC(inflow), DirichletBC(inflow), PeriodicBC{}, PeriodicBC{});
    using Solver = StructuredSolver<Eq, MyBasis, VolumeFlux, SurfaceFlux, Mesh, decltype(boundaries)>;
    using Sol = Solver::solution;
    Mesh mesh({-0.5, -0.5}, {0.5, 0.5}, {nx, ny});
    Solver solver(Eq{gamma}, mesh, boundaries);
    solver.set_indicator_parameters(0.3, 0.0001, false);
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
