// Daru-Tenaud viscous shock tube, composed the way the reference's example
// is (examples/NavierStokes2D/viscous_shock_tube/viscous_shock_tube.cpp) but
// against dgsem_hip.hpp: compressible Navier-Stokes (mu = 1e-3, Pr = 0.73)
// in [0, 1] x [0, 0.5] with no-slip adiabatic walls on three sides and a
// slip adiabatic wall on top, shock capturing with alpha_max = 1 and
// alpha_min = 0.1, N = 3.
// Usage: viscous_shock_tube nx ny steps dt out_prefix
// writes out_prefix.u0.bin and out_prefix.u.bin in the reference layout.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../include/dgsem_hip.hpp"

namespace {

struct DaruTenaudInitial {
  double gamma;
  std::array<double, 4> operator()(const std::array<double, 2>& x) const {
    const double rho = x[0] <= 0.5 ? 120.0 : 1.2;
    const double p = rho / gamma;
    return {rho, 0.0, 0.0, p / (gamma - 1.0)};
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
  using Eq = equations::CompressibleNavierStokes2D<value_type>;
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
  constexpr double gamma = 1.4, prandtl = 0.73, mu = 1.0e-3;
  try {
    auto boundaries = BoundarySet(NSNoSlipAdiabaticWallBC{}, NSNoSlipAdiabaticWallBC{},
                                  NSNoSlipAdiabaticWallBC{}, NSSlipAdiabaticWallBC{});
    using Solver = StructuredSolver<Eq, MyBasis, VolumeFlux, SurfaceFlux, Mesh, decltype(boundaries)>;
    using Sol = Solver::solution;
    Mesh mesh({0.0, 0.0}, {1.0, 0.5}, {nx, ny});
    Solver solver(Eq(gamma, mu, prandtl), mesh, boundaries);
    solver.set_indicator_parameters(1.0, 0.1, false);
    Sol sol(mesh);
    solver.initialize(DaruTenaudInitial{gamma}, sol);
    if (!write(prefix + ".u0.bin", solver.download_reference_layout(sol))) return 1;
    SSPRK3<value_type, Solver, Mesh, Sol> rk(sol, mesh, steps * dt);
    for (int s = 0; s < steps; ++s) rk.step(solver, sol, dt);
    solver.synchronize();
    if (!write(prefix + ".u.bin", solver.download_reference_layout(sol))) return 1;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
