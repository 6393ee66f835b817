// Inviscid Taylor-Green vortex driven the way the reference's drivers are
// (examples/Euler3D/inviscid_tgv/inviscid_tgv.cpp): an AbstractInitial
// functor, the solver built with an element container and periodic flags,
// and TimeIntegrator::solve with the printing, VTU, NaN-stopping and
// volume-average observers attached -- all spelled as in namespace DGSEM,
// all running on the GPU through dgsem_hip.hpp.
// Usage: taylor_green_observers cells N steps out_prefix
// writes <out_prefix>_step<k>.vtu and <out_prefix>_averages.csv.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numbers>
#include <string>

#include "../include/dgsem_hip.hpp"

template <class T>
struct TaylorGreen : DGSEM::AbstractInitial<TaylorGreen<T>, DGSEM::equations::CompressibleEuler3D<T>> {
  KOKKOS_INLINE_FUNCTION
  std::array<T, 5> operator()(std::array<T, 3> x) const {
    const T rho = 1.0;
    const T u = std::sin(x[0]) * std::cos(x[1]) * std::cos(x[2]);
    const T v = -std::cos(x[0]) * std::sin(x[1]) * std::cos(x[2]);
    const T p = 100.0 / 1.4 + rho / 16.0 * (std::cos(2 * x[0]) + std::cos(2 * x[1])) *
                                  (std::cos(2 * x[2]) + 2.0);
    return DGSEM::utils::prim_to_cons(std::array<T, 5>{rho, u, v, 0.0, p}, T(1.4));
  }
};

template <std::size_t Order>
int run(std::size_t cells, int steps, const std::string& prefix) {
  using T = double;
  using Eq = DGSEM::equations::CompressibleEuler3D<T>;
  using B = DGSEM::Basis::LobattoLegendreBasis<T, Order>;
  using Volume = DGSEM::VolumeIntegralSplitForm<B, Eq, DGSEM::ChandrashekarFlux>;
  using Surface = DGSEM::LaxFriedrichsFlux<Eq>;
  using Mesh = DGSEM::StructuredMesh<T, 3>;
  DGSEM::BasisGuard<B> guard;
  auto bcs = DGSEM::BoundarySet(DGSEM::PeriodicBC{}, DGSEM::PeriodicBC{}, DGSEM::PeriodicBC{},
                                DGSEM::PeriodicBC{}, DGSEM::PeriodicBC{}, DGSEM::PeriodicBC{});
  using Solver = DGSEM::StructuredSolver<Eq, B, Volume, Surface, Mesh, decltype(bcs)>;
  using Sol = DGSEM::Solution<Mesh, B, Eq>;

  const T L = 2.0 * std::numbers::pi;
  const std::array<std::size_t, 3> n_cells{cells, cells, cells};
  Mesh mesh({0.0, 0.0, 0.0}, {L, L, L}, n_cells);
  DGSEM::StructuredElementContainer<T, 3> container;
  Solver solver(Eq{1.4}, mesh, container, bcs, {true, true, true});
  Sol sol(mesh);
  solver.initialize(TaylorGreen<T>{}, sol);

  // dt = CFL dx / (3 (2N + 1) lambda_max), CFL 0.5
  const T lam = 1.0 + std::sqrt(1.4 * (100.0 / 1.4 + 0.25));
  const T dt = 0.5 * (L / (T)cells) / (3.0 * (2 * Order + 1) * lam);
  DGSEM::SSPRK3<T, Solver, Mesh, Sol> integrator(sol, mesh, steps * dt * (1.0 - 1e-12));

  using Checker = DGSEM::AnalyzerWrapper<B, Eq, DGSEM::DivergenceChecker<T, Eq::NVARS>>;
  using Averages = DGSEM::AnalyzerWrapper<B, Eq, DGSEM::VolumeAverageEuler<Eq>>;
  Checker checker;
  Averages averages;
  integrator.add_observer(std::make_unique<DGSEM::PrintObserver>(5));
  integrator.add_observer(DGSEM::make_analysis_observer<B, Eq>(
      DGSEM::PointwiseAnalysisTag{}, checker, sol, n_cells, DGSEM::StopOnNaN<Eq>()));
  integrator.add_observer(
      std::make_unique<DGSEM::VTUOutputObserver<T, B, Sol, decltype(container.node_coordinates), Eq>>(
          prefix, sol, container.node_coordinates, n_cells, 10));
  integrator.add_observer(DGSEM::make_analysis_observer<B, Eq>(
      DGSEM::VolumeWeightedAnalysisTag{}, averages, sol, container, n_cells,
      DGSEM::VolumeAverageCsvWriter<Eq>(prefix + "_averages.csv")));
  integrator.solve(solver, sol, dt);
  std::printf("dt %.17g\n", dt);
  return 0;
}

int main(int argc, char** argv) {
  DGSEM::KokkosSession session(argc, argv);
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s cells N steps out_prefix\n", argv[0]);
    return 2;
  }
  const std::size_t cells = std::strtoull(argv[1], nullptr, 10);
  const int order = std::atoi(argv[2]), steps = std::atoi(argv[3]);
  try {
    switch (order) {
      case 2: return run<2>(cells, steps, argv[4]);
      case 3: return run<3>(cells, steps, argv[4]);
      case 4: return run<4>(cells, steps, argv[4]);
      default: std::fprintf(stderr, "N must be 2, 3 or 4\n"); return 2;
    }
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
}
