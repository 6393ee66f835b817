// Dispatch of the fused RHS kernel over (dim, N) and the layout-conversion
// kernel.  The kernel body is in rhs_kernel_impl.hpp; its instantiations are
// compiled per (dim, N) from rhs_inst.hip.
#include <hip/hip_runtime.h>

#include "batch_impl.hpp"
#include "kernels.hpp"

namespace dgsem {
namespace {

__global__ void pack_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t E,
                            int nn, int nv, int to_internal) {
  const int64_t total = E * nn * nv;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    // i enumerates the internal layout [e][v][dof]
    const int64_t e = i / ((int64_t)nn * nv);
    const int r = (int)(i % ((int64_t)nn * nv));
    const int v = r / nn, dof = r % nn;
    const int64_t j = (e * nn + dof) * nv + v;  // reference layout [e][dof][v]
    if (to_internal)
      dst[i] = src[j];
    else
      dst[j] = src[i];
  }
}

// Column-major (LayoutLeft) view u(ie, je[, ke], dof, var): ie fastest.  The
// thread index enumerates that layout so that its side is contiguous; the
// element's internal index is (ie * ny + je) * nz + ke.
__global__ void pack_left_kernel(const double* __restrict__ src, double* __restrict__ dst, int nx,
                                 int ny, int nz, int nn, int nv, int to_internal) {
  const int64_t E = (int64_t)nx * ny * nz;
  const int64_t total = E * nn * nv;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = i;
    const int ie = (int)(r % nx);
    r /= nx;
    const int je = (int)(r % ny);
    r /= ny;
    const int ke = (int)(r % nz);
    r /= nz;
    const int dof = (int)(r % nn);
    const int v = (int)(r / nn);
    const int64_t e = ((int64_t)ie * ny + je) * nz + ke;
    const int64_t j = (e * nv + v) * nn + dof;  // internal layout [e][v][dof]
    if (to_internal)
      dst[j] = src[i];
    else
      dst[i] = src[j];
  }
}

}  // namespace

hipError_t launch_rhs(int dim, int N, bool sc, bool nodal, bool bcs, const KParams& p,
                      hipStream_t s) {
  if (dim == 2) {
    switch (N) {
      case 1: return launch_rhs_dn<2, 1>(sc, nodal, bcs, p, s);
      case 2: return launch_rhs_dn<2, 2>(sc, nodal, bcs, p, s);
      case 3: return launch_rhs_dn<2, 3>(sc, nodal, bcs, p, s);
      case 4: return launch_rhs_dn<2, 4>(sc, nodal, bcs, p, s);
      case 5: return launch_rhs_dn<2, 5>(sc, nodal, bcs, p, s);
      case 7: return launch_rhs_dn<2, 7>(sc, nodal, bcs, p, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (dim == 3) {
    switch (N) {
      case 2: return launch_rhs_dn<3, 2>(sc, nodal, bcs, p, s);
      case 3: return launch_rhs_dn<3, 3>(sc, nodal, bcs, p, s);
      case 4: return launch_rhs_dn<3, 4>(sc, nodal, bcs, p, s);
      case 5: return launch_rhs_dn<3, 5>(sc, nodal, bcs, p, s);
      case 7: return launch_rhs_dn<3, 7>(sc, nodal, bcs, p, s);
      default: return hipErrorInvalidValue;
    }
  }
  return hipErrorInvalidValue;
}
hipError_t launch_parabolic(int dim, int N, bool sc, bool nodal, bool bcs, const KParams& p,
                            hipStream_t s) {
  if (dim == 2) {
    switch (N) {
      case 1: return launch_parabolic_dn<2, 1>(sc, nodal, bcs, p, s);
      case 2: return launch_parabolic_dn<2, 2>(sc, nodal, bcs, p, s);
      case 3: return launch_parabolic_dn<2, 3>(sc, nodal, bcs, p, s);
      case 4: return launch_parabolic_dn<2, 4>(sc, nodal, bcs, p, s);
      case 5: return launch_parabolic_dn<2, 5>(sc, nodal, bcs, p, s);
      case 7: return launch_parabolic_dn<2, 7>(sc, nodal, bcs, p, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (dim == 3) {
    switch (N) {
      case 2: return launch_parabolic_dn<3, 2>(sc, nodal, bcs, p, s);
      case 3: return launch_parabolic_dn<3, 3>(sc, nodal, bcs, p, s);
      case 4: return launch_parabolic_dn<3, 4>(sc, nodal, bcs, p, s);
      case 5: return launch_parabolic_dn<3, 5>(sc, nodal, bcs, p, s);
      case 7: return launch_parabolic_dn<3, 7>(sc, nodal, bcs, p, s);
      default: return hipErrorInvalidValue;
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_pack(const double* src, double* dst, int64_t E, int nn, int nv,
                       bool to_internal, hipStream_t stream) {
  const int64_t total = E * nn * nv;
  if (total <= 0) return hipSuccess;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, dst, E, nn,
                     nv, to_internal ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_pack_left(const double* src, double* dst, const int cells[3], int nn, int nv,
                            bool to_internal, hipStream_t stream) {
  const int64_t total = (int64_t)cells[0] * cells[1] * cells[2] * nn * nv;
  if (total <= 0) return hipSuccess;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(pack_left_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, dst,
                     cells[0], cells[1], cells[2], nn, nv, to_internal ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_batch(int block, const BatchLaunch& b, hipStream_t stream) {
  switch (block) {
    case 128: return launch_batch_bs<128>(b, stream);
    case 256: return launch_batch_bs<256>(b, stream);
    default: return hipErrorInvalidValue;
  }
}

namespace {

template <int BLOCK>
int slots_for(int N) {
  switch (N) {
    case 1: return BLOCK / detail::Shape<2, 1>::LINES;
    case 2: return BLOCK / detail::Shape<2, 2>::LINES;
    case 3: return BLOCK / detail::Shape<2, 3>::LINES;
    case 4: return BLOCK / detail::Shape<2, 4>::LINES;
    case 5: return BLOCK / detail::Shape<2, 5>::LINES;
    default: return 0;
  }
}

template <int BLOCK>
size_t lds_for(int N) {
  switch (N) {
    case 1: return sizeof(double) * detail::batch_lds_doubles<1, BLOCK>();
    case 2: return sizeof(double) * detail::batch_lds_doubles<2, BLOCK>();
    case 3: return sizeof(double) * detail::batch_lds_doubles<3, BLOCK>();
    case 4: return sizeof(double) * detail::batch_lds_doubles<4, BLOCK>();
    case 5: return sizeof(double) * detail::batch_lds_doubles<5, BLOCK>();
    default: return 0;
  }
}

}  // namespace

int batch_slots(int block, int N) {
  if (block == 128) return slots_for<128>(N);
  if (block == 256) return slots_for<256>(N);
  return 0;
}

size_t batch_lds_bytes(int block, int N) {
  if (block == 128) return lds_for<128>(N);
  if (block == 256) return lds_for<256>(N);
  return 0;
}

}  // namespace dgsem
