// The plugin traits of dgsem_amd/include/dgsem_hip.hpp (the reference's
// equation and numerical-flux members) behind extern "C" functions, so that
// tests/test_traits_cpu.py can hold each member against the oracle.  Host
// code only: no kernel is launched.
#include <array>
#include <cstddef>

#include "../../dgsem_amd/include/dgsem_hip.hpp"

namespace {

template <std::size_t D>
using St = std::array<double, D + 2>;
template <std::size_t D>
using Nv = std::array<double, D>;

template <std::size_t D>
St<D> load(const double* p) {
  St<D> s;
  for (std::size_t i = 0; i < D + 2; ++i) s[i] = p[i];
  return s;
}
template <std::size_t D>
Nv<D> load_n(const double* p) {
  Nv<D> s;
  for (std::size_t i = 0; i < D; ++i) s[i] = p[i];
  return s;
}
template <std::size_t N>
void store(const std::array<double, N>& s, double* p) {
  for (std::size_t i = 0; i < N; ++i) p[i] = s[i];
}

template <std::size_t D>
using Euler = std::conditional_t<D == 2, DGSEM::equations::CompressibleEuler2D<double>,
                                 DGSEM::equations::CompressibleEuler3D<double>>;
template <std::size_t D>
using NS = std::conditional_t<D == 2, DGSEM::equations::CompressibleNavierStokes2D<double>,
                              DGSEM::equations::CompressibleNavierStokes3D<double>>;

// kind: 0 flux(u, dim), 1 flux(u, normal), 2 LLF(dim), 3 LLF(normal),
// 4 Chandrashekar(normal), 5 HLLC(normal), 6 Central(dim), 7 Central(normal, 2D)
template <std::size_t D>
void eval(int kind, double gamma, const double* ul, const double* ur, const double* n, int d,
          double* f) {
  const Euler<D> eq(gamma);
  const auto a = load<D>(ul), b = load<D>(ur);
  const auto nv = load_n<D>(n);
  switch (kind) {
    case 0: store(eq.flux(a, (std::size_t)d), f); break;
    case 1: store(eq.flux(a, nv), f); break;
    case 2: store(DGSEM::LaxFriedrichsFlux<Euler<D>>::numerical_flux(eq, a, b, (std::size_t)d), f); break;
    case 3: store(DGSEM::LaxFriedrichsFlux<Euler<D>>::numerical_flux(eq, a, b, nv), f); break;
    case 4: store(DGSEM::ChandrashekarFlux<Euler<D>>::numerical_flux(eq, a, b, nv), f); break;
    case 5: store(DGSEM::HLLCFlux<Euler<D>>::numerical_flux(eq, a, b, nv), f); break;
    case 6: store(DGSEM::CentralFlux<Euler<D>>::numerical_flux(eq, a, b, (std::size_t)d), f); break;
    case 7:
      if constexpr (D == 2) store(DGSEM::CentralFlux<Euler<D>>::numerical_flux(eq, a, b, nv), f);
      break;
    default: break;
  }
}

}  // namespace

extern "C" {

void tr_eval(int dim, int kind, double gamma, const double* ul, const double* ur, const double* n,
             int d, double* f) {
  if (dim == 2)
    eval<2>(kind, gamma, ul, ur, n, d, f);
  else
    eval<3>(kind, gamma, ul, ur, n, d, f);
}

// get_wave_speed(u_ll, u_rr, dim) (normal == nullptr) or (u_ll, u_rr, normal)
double tr_wave_speed(int dim, double gamma, const double* ul, const double* ur, const double* n,
                     int d) {
  if (dim == 2) {
    const Euler<2> eq(gamma);
    return n ? eq.get_wave_speed(load<2>(ul), load<2>(ur), load_n<2>(n))
             : eq.get_wave_speed(load<2>(ul), load<2>(ur), (std::size_t)d);
  }
  const Euler<3> eq(gamma);
  return n ? eq.get_wave_speed(load<3>(ul), load<3>(ur), load_n<3>(n))
           : eq.get_wave_speed(load<3>(ul), load<3>(ur), (std::size_t)d);
}

void tr_gradient_variables(int dim, double gamma, double mu, double prandtl, const double* u,
                           double* q) {
  if (dim == 2)
    store(NS<2>(gamma, mu, prandtl).gradient_variables(load<2>(u)), q);
  else
    store(NS<3>(gamma, mu, prandtl).gradient_variables(load<3>(u)), q);
}

// grad: [dim][dim + 2] gradient of the gradient variables
void tr_viscous_flux(int dim, double gamma, double mu, double prandtl, const double* q,
                     const double* grad, int d, double* f) {
  if (dim == 2) {
    std::array<std::array<double, 4>, 2> g;
    for (int c = 0; c < 2; ++c)
      for (int v = 0; v < 4; ++v) g[c][v] = grad[c * 4 + v];
    store(NS<2>(gamma, mu, prandtl).viscous_flux(load<2>(q), g, (std::size_t)d), f);
  } else {
    std::array<std::array<double, 5>, 3> g;
    for (int c = 0; c < 3; ++c)
      for (int v = 0; v < 5; ++v) g[c][v] = grad[c * 5 + v];
    store(NS<3>(gamma, mu, prandtl).viscous_flux(load<3>(q), g, (std::size_t)d), f);
  }
}

void tr_prim_to_cons(int dim, double gamma, const double* prim, double* cons) {
  if (dim == 2)
    store(DGSEM::utils::prim_to_cons(load<2>(prim), gamma), cons);
  else
    store(DGSEM::utils::prim_to_cons(load<3>(prim), gamma), cons);
}

}  // extern "C"
