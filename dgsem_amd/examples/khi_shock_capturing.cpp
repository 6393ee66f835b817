// Kelvin-Helmholtz instability with subcell shock capturing, written the way
// the reference's example is (examples/Euler2D/kelvin_helmholtz_hg/
// kelvin_helmholtz_hg.cpp:34-126) but against dgsem_hip.hpp: the same
// template composition, initial condition and SSPRK3 loop, running in the
// HIP kernels.  Usage: khi_shock_capturing nx ny N_steps dt alpha_max out.bin
// writes the final state in the reference layout (e, dof, var) as raw FP64.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numbers>

#include "../include/dgsem_hip.hpp"

template <class T>
struct KelvinHelmholtzInitial {
  std::array<T, 4> operator()(const std::array<T, 2>& x) const {
    const T slope = 15.0;
    const T B = std::tanh(slope * x[1] + 7.5) - std::tanh(slope * x[1] - 7.5);
    const T rho = 0.5 + 0.75 * B;
    const T v1 = 0.5 * (B - 1.0);
    const T v2 = 0.1 * std::sin(2.0 * std::numbers::pi_v<T> * x[0]);
    const T p = 1.0;
    return {rho, rho * v1, rho * v2, p / (1.4 - 1.0) + 0.5 * rho * (v1 * v1 + v2 * v2)};
  }
};

int main(int argc, char** argv) {
  using namespace dgsem_hip;
  using value_type = double;
  using Eq = equations::CompressibleEuler2D<value_type>;
  using MyBasis = Basis::LobattoLegendreBasis<value_type, 3>;
  using SurfaceFlux = LaxFriedrichsFlux<Eq>;
  using VolumeFlux = VolumeIntegralShockCapturingHG<MyBasis, Eq, ChandrashekarFlux,
                                                    LaxFriedrichsFlux, HGIndicator<MyBasis, Eq>>;
  using Mesh = StructuredMesh<value_type, 2>;
  using Boundaries = BoundarySet<PeriodicBC, PeriodicBC, PeriodicBC, PeriodicBC>;
  using Solver = StructuredSolver<Eq, MyBasis, VolumeFlux, SurfaceFlux, Mesh, Boundaries>;
  using Sol = Solver::solution;

  if (argc < 7) {
    std::fprintf(stderr, "usage: %s nx ny steps dt alpha_max out.bin\n", argv[0]);
    return 2;
  }
  const std::size_t nx = std::strtoull(argv[1], nullptr, 10), ny = std::strtoull(argv[2], nullptr, 10);
  const int steps = std::atoi(argv[3]);
  const double dt = std::strtod(argv[4], nullptr), amax = std::strtod(argv[5], nullptr);
  try {
    Mesh mesh({-1.0, -1.0}, {1.0, 1.0}, {nx, ny});
    Solver solver(Eq{1.4}, mesh, Boundaries(PeriodicBC{}, PeriodicBC{}, PeriodicBC{}, PeriodicBC{}));
    solver.set_indicator_parameters(amax, 0.001, false);
    Sol sol(mesh);
    solver.initialize(KelvinHelmholtzInitial<value_type>{}, sol);
    SSPRK3<value_type, Solver, Mesh, Sol> rk(sol, mesh, steps * dt);
    for (int s = 0; s < steps; ++s) rk.step(solver, sol, dt);
    const auto u = solver.download_reference_layout(sol);
    FILE* f = std::fopen(argv[6], "wb");
    if (!f || std::fwrite(u.data(), sizeof(double), u.size(), f) != u.size()) {
      std::fprintf(stderr, "cannot write %s\n", argv[6]);
      return 1;
    }
    std::fclose(f);
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
