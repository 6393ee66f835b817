// Instantiation of the batched ensemble kernel for one workgroup size,
// compiled once per size with -DDGSEM_BATCH_BLOCK=<threads>.
#include "batch_impl.hpp"

#ifndef DGSEM_BATCH_BLOCK
#error "DGSEM_BATCH_BLOCK must be defined"
#endif

namespace dgsem {

template <>
hipError_t launch_batch_bs<DGSEM_BATCH_BLOCK>(const BatchLaunch& b, hipStream_t stream) {
  return detail::launch_batch_block<DGSEM_BATCH_BLOCK>(b, stream);
}

}  // namespace dgsem
