// On-device analysis of a state: the reductions of the reference's
// analyzers (VolumeAverageEuler, volume_average_euler.hpp:30-53, with the
// volume weights of VolumeWeightedAnalyzerFunctor,
// volume_weighted_analyzer_functor.hpp:108-134; DivergenceChecker,
// divergence_checker.hpp:6-27) plus the largest signal speed for a CFL time
// step, in one pass over the state (include/dgsem_amd.h, dgsem_analyze).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/dgsem_amd.h"
#include "handle.hpp"
#include "kernels.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kOut = 5;  // volume, kinetic energy, entropy, non-finite count, max speed

// The 1D Gauss-Lobatto weights, passed by value: a node's tensor-product
// weight is formed in the kernel as w_i w_j (w_k), the same products in the
// same order as the host table, so it needs no device buffer.
struct Weights1D {
  double w[dgsem::kMaxN1];
};

// Partial sums of the analysis blocks for dgsem_analyze_noalloc: one set of
// kOut values per block of the grid-stride reduction, every slot written by
// its block before the final kernel reads it, so nothing needs clearing.
// Each GPU holds its own copy for the life of the process; host code reaches
// it only through the address hipGetSymbolAddress returns for that device.
static __device__ __align__(256) double scratch[kMaxBlocks * kOut];

// Each block reduces a fixed grid-stride set of nodes in a fixed tree order,
// so the sums do not depend on scheduling.
__global__ void __launch_bounds__(kBlock)
    analyze_kernel(int dim, int n, int nn, Weights1D w1, int nodal, double gamma, dgsem::KParams p,
                   double* __restrict__ partial) {
  __shared__ double s[kOut][kBlock];
  const int nv = dim + 2;
  const int64_t total = p.E * nn;
  double vol = 0.0, ke = 0.0, ent = 0.0, bad = 0.0, vmax = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kBlock) {
    const int64_t e = i / nn;
    const int node = (int)(i % nn);
    const double* u = p.u + e * nv * nn + node;
    double q[5];
    bool finite = true;
    for (int v = 0; v < nv; ++v) {
      q[v] = u[v * nn];
      if (!isfinite(q[v])) finite = false;
    }
    if (!finite) {
      bad += 1.0;
      continue;
    }
    const double ij = nodal ? p.invJ[e * nn + node] : p.invJ[e];
    // node = (k n + j) n + i
    const int ix = node % n, jx = (node / n) % n, kx = node / (n * n);
    double wq = w1.w[ix] * w1.w[jx];
    if (dim == 3) wq = wq * w1.w[kx];
    const double w = wq / fabs(ij);
    const double rho = q[0];
    double m2 = 0.0;
    for (int k = 0; k < dim; ++k) m2 += q[k + 1] * q[k + 1];
    // prim = cons_to_prim: v = m / rho, p = (gamma - 1)(E - rho |v|^2 / 2)
    double v2 = 0.0, vabs = 0.0;
    for (int k = 0; k < dim; ++k) {
      const double vk = q[k + 1] / rho;
      v2 += vk * vk;
      vabs = fmax(vabs, fabs(vk));
    }
    const double pr = (gamma - 1.0) * (q[dim + 1] - 0.5 * rho * v2);
    const double sp = log(pr) - gamma * log(rho);
    vol += w;
    ke += w * (0.5 * m2 / rho);
    ent += w * (-rho * sp / (gamma - 1.0));
    vmax = fmax(vmax, vabs + sqrt(gamma * pr / rho));
  }
  s[0][threadIdx.x] = vol;
  s[1][threadIdx.x] = ke;
  s[2][threadIdx.x] = ent;
  s[3][threadIdx.x] = bad;
  s[4][threadIdx.x] = vmax;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w /= 2) {
    if ((int)threadIdx.x < w) {
      for (int k = 0; k < 4; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
      s[4][threadIdx.x] = fmax(s[4][threadIdx.x], s[4][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < kOut) partial[blockIdx.x * kOut + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(kBlock)
    analyze_final_kernel(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
  __shared__ double s[kOut][kBlock];
  double a[kOut] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nblocks; i += kBlock) {
    for (int k = 0; k < 4; ++k) a[k] += partial[i * kOut + k];
    a[4] = fmax(a[4], partial[i * kOut + 4]);
  }
  for (int k = 0; k < kOut; ++k) s[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w /= 2) {
    if ((int)threadIdx.x < w) {
      for (int k = 0; k < 4; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
      s[4][threadIdx.x] = fmax(s[4][threadIdx.x], s[4][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < kOut) out[threadIdx.x] = s[threadIdx.x][0];
}

// Enqueues the two analysis kernels with the block partials in `partial`
// (kMaxBlocks * kOut doubles).
int launch_analyze(dgsem_handle* h, const double* u, double* partial, double* out) {
  dgsem::KParams p = h->base;
  p.u = u;
  Weights1D w1{};
  for (int i = 0; i < h->basis.n; ++i) w1.w[i] = h->basis.weights[i];
  const int64_t total = h->E * h->nn;
  int64_t nb = (total + kBlock - 1) / kBlock;
  nb = nb < 1 ? 1 : (nb > kMaxBlocks ? kMaxBlocks : nb);
  const bool nodal = h->cfg.geometry == DGSEM_GEOMETRY_NODAL;
  hipLaunchKernelGGL(analyze_kernel, dim3((unsigned)nb), dim3(kBlock), 0, h->stream, h->cfg.dim,
                     h->basis.n, h->nn, w1, nodal ? 1 : 0, h->cfg.gamma, p, partial);
  hipLaunchKernelGGL(analyze_final_kernel, dim3(1), dim3(kBlock), 0, h->stream, partial, (int)nb,
                     out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    h->err = std::string("analyze launch: ") + hipGetErrorString(e);
    return DGSEM_E_HIP;
  }
  return DGSEM_OK;
}

// Device address of `scratch` on each device, looked up once per device
// with that device current.
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

int scratch_address(dgsem_handle* h, double** ptr) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    h->err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return DGSEM_E_HIP;
  }
  if (dev != h->device) {
    h->err = "dgsem_analyze_noalloc: the handle's device must be the current device";
    return DGSEM_E_INVALID;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    h->err = "dgsem_analyze_noalloc: device index out of range";
    return DGSEM_E_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* a = nullptr;
    e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !a) {
      h->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return DGSEM_E_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(a);
  }
  *ptr = g_scratch_ptr[dev];
  return DGSEM_OK;
}

}  // namespace

extern "C" {

int dgsem_analyze(dgsem_handle* h, const double* u, double* out) {
  if (!h || !h->d_Ja || !u || !out) return DGSEM_E_INVALID;
  if (!h->d_an_partial &&
      hipMalloc((void**)&h->d_an_partial, (size_t)kMaxBlocks * kOut * sizeof(double)) != hipSuccess) {
    h->err = "hipMalloc failed";
    return DGSEM_E_ALLOC;
  }
  return launch_analyze(h, u, h->d_an_partial, out);
}

int dgsem_analyze_noalloc(dgsem_handle* h, const double* u, double* out) {
  if (!h || !h->d_Ja || !u || !out) return DGSEM_E_INVALID;
  if (h->basis.n > dgsem::kMaxN1) {
    h->err = "dgsem_analyze_noalloc: polynomial degree beyond the built maximum";
    return DGSEM_E_INVALID;
  }
  double* partial = nullptr;
  const int rc = scratch_address(h, &partial);
  return rc != DGSEM_OK ? rc : launch_analyze(h, u, partial, out);
}

}  // extern "C"
