// One translation unit per (dimension, polynomial degree): compiled several
// times by dgsem_amd/build.py with -DDGSEM_INST_DIM / -DDGSEM_INST_N so the
// kernel instantiations build in parallel.
#include "parabolic_impl.hpp"
#include "rhs_kernel_impl.hpp"

#if !defined(DGSEM_INST_DIM) || !defined(DGSEM_INST_N)
#error "DGSEM_INST_DIM and DGSEM_INST_N must be defined"
#endif

namespace dgsem {
template <>
hipError_t launch_rhs_dn<DGSEM_INST_DIM, DGSEM_INST_N>(bool sc, bool nodal, bool bcs,
                                                        const KParams& p, hipStream_t s) {
  return detail::launch_dn<DGSEM_INST_DIM, DGSEM_INST_N>(sc, nodal, bcs, p, s);
}
template <>
hipError_t launch_parabolic_dn<DGSEM_INST_DIM, DGSEM_INST_N>(bool sc, bool nodal, bool bcs,
                                                              const KParams& p, hipStream_t s) {
#if DGSEM_INST_DIM == 2
  if (bcs)
    return nodal ? detail::launch_parabolic_t<DGSEM_INST_DIM, DGSEM_INST_N, true, true>(sc, p, s)
                 : detail::launch_parabolic_t<DGSEM_INST_DIM, DGSEM_INST_N, false, true>(sc, p, s);
#else
  if (bcs) return hipErrorInvalidValue;  // non-periodic boundaries are built for 2D
#endif
  return nodal ? detail::launch_parabolic_t<DGSEM_INST_DIM, DGSEM_INST_N, true, false>(sc, p, s)
               : detail::launch_parabolic_t<DGSEM_INST_DIM, DGSEM_INST_N, false, false>(sc, p, s);
}
}  // namespace dgsem
