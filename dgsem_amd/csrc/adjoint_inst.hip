// Instantiations of the tangent / adjoint kernels (adjoint_impl.hpp) for one
// (dim, N), compiled once per pair with -DDGSEM_INST_DIM and -DDGSEM_INST_N.
#include "adjoint_impl.hpp"

#if !defined(DGSEM_INST_DIM) || !defined(DGSEM_INST_N)
#error "DGSEM_INST_DIM and DGSEM_INST_N must be defined"
#endif

namespace dgsem {

template <>
hipError_t launch_jvp_dn<DGSEM_INST_DIM, DGSEM_INST_N>(bool nodal, const AdParams& a,
                                                       hipStream_t s) {
  // Navier-Stokes terms when the viscous scratch is set; the shock-capturing
  // blend when a frozen blending factor is
  const bool ns = a.dvisc != nullptr, sc = a.k.alpha_in != nullptr;
  if (sc)
    return nodal ? detail::launch_jvp_t<DGSEM_INST_DIM, DGSEM_INST_N, true, true>(ns, a, s)
                 : detail::launch_jvp_t<DGSEM_INST_DIM, DGSEM_INST_N, false, true>(ns, a, s);
  return nodal ? detail::launch_jvp_t<DGSEM_INST_DIM, DGSEM_INST_N, true, false>(ns, a, s)
               : detail::launch_jvp_t<DGSEM_INST_DIM, DGSEM_INST_N, false, false>(ns, a, s);
}

template <>
hipError_t launch_vjp_dn<DGSEM_INST_DIM, DGSEM_INST_N>(bool nodal, const AdParams& a,
                                                       hipStream_t s) {
  // Navier-Stokes terms when the viscous scratch is set; the shock-capturing
  // blend when a frozen blending factor is
  const bool ns = a.dvisc != nullptr, sc = a.k.alpha_in != nullptr;
  if (sc)
    return nodal ? detail::launch_vjp_t<DGSEM_INST_DIM, DGSEM_INST_N, true, true>(ns, a, s)
                 : detail::launch_vjp_t<DGSEM_INST_DIM, DGSEM_INST_N, false, true>(ns, a, s);
  return nodal ? detail::launch_vjp_t<DGSEM_INST_DIM, DGSEM_INST_N, true, false>(ns, a, s)
               : detail::launch_vjp_t<DGSEM_INST_DIM, DGSEM_INST_N, false, false>(ns, a, s);
}

template <>
hipError_t launch_gradient_dn<DGSEM_INST_DIM, DGSEM_INST_N>(bool nodal, const KParams& p,
                                                            hipStream_t s) {
  return nodal ? detail::launch_gradient_t<DGSEM_INST_DIM, DGSEM_INST_N, true>(p, s)
               : detail::launch_gradient_t<DGSEM_INST_DIM, DGSEM_INST_N, false>(p, s);
}

}  // namespace dgsem
