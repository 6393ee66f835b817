// Element geometry on the GPU from node coordinates of any mapping: the
// covariant Jacobian matrix by the derivative matrix, the contravariant
// vectors (2D adjugate, 3D curl form), the inverse Jacobian and the subcell
// normals of the FV blending.  Reference: StructuredContainerInitializer::
// calc_jacobian_matrix / calc_contravariant_vectors / calc_inverse_jacobian
// (data_container.hpp:223-312 2D, :416-693 3D) and SubcellNormalVectors::
// initialize_host (global_def.hpp:175-251 2D, :299-439 3D), all host-serial
// there.  One workgroup per element stages its coordinates, Jacobian and
// metric terms in LDS.  Floating-point contraction is off and every sum runs
// in the reference's order, so the results equal the host construction
// bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dgsem_amd.h"
#include "handle.hpp"

namespace {

template <int DIM>
__device__ __forceinline__ int dof_of(const int* ix, int n) {
  return DIM == 2 ? ix[1] * n + ix[0] : (ix[2] * n + ix[1]) * n + ix[0];
}

// X_in [E][nn][DIM] -> jac [E][nn][DIM][DIM], ja_ref [E][nn][DIM][DIM],
// ja_dev [E][DIM][DIM][nn], invj [E][nn], sn [E][DIM][FN][n-1][DIM].
template <int DIM>
__global__ void __launch_bounds__(256) geometry_kernel(
    const double* __restrict__ X_in, const double* __restrict__ D, const double* __restrict__ w,
    int n, double* __restrict__ jac, double* __restrict__ ja_ref, double* __restrict__ ja_dev,
    double* __restrict__ invj, double* __restrict__ sn) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  const int nn = DIM == 2 ? n * n : n * n * n;
  const int FN = DIM == 2 ? n : n * n;
  double* X = lds;                  // [nn][DIM]
  double* JM = X + nn * DIM;        // [nn][DIM][DIM]
  double* JA = JM + nn * DIM * DIM;  // [nn][DIM][DIM]
  const int64_t e = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < nn * DIM; i += nt) X[i] = X_in[e * nn * DIM + i];
  __syncthreads();

  // covariant Jacobian: JM[dof][r][c] = sum_m D[ix_r][m] X(node with ix_r = m)[c]
  for (int dof = tid; dof < nn; dof += nt) {
    const int ix0[3] = {dof % n, (dof / n) % n, DIM == 3 ? dof / (n * n) : 0};
    for (int r = 0; r < DIM; ++r) {
      double acc[3] = {0.0, 0.0, 0.0};
      int ix[3] = {ix0[0], ix0[1], ix0[2]};
      const int self = ix0[r];
      for (int m = 0; m < n; ++m) {
        ix[r] = m;
        const int dm = dof_of<DIM>(ix, n);
        for (int c = 0; c < DIM; ++c) acc[c] += D[self * n + m] * X[dm * DIM + c];
      }
      for (int c = 0; c < DIM; ++c) {
        JM[(dof * DIM + r) * DIM + c] = acc[c];
        jac[((e * nn + dof) * DIM + r) * DIM + c] = acc[c];
      }
    }
  }
  __syncthreads();

  // contravariant vectors and 1/J
  for (int dof = tid; dof < nn; dof += nt) {
    const double* J = JM + dof * DIM * DIM;
    double* A = JA + dof * DIM * DIM;
    if constexpr (DIM == 2) {
      A[0] = J[3];   // Ja^1 = ( y_eta, -x_eta)
      A[1] = -J[2];
      A[2] = -J[1];  // Ja^2 = (-y_xi,  x_xi)
      A[3] = J[0];
      invj[e * nn + dof] = 1.0 / (J[0] * J[3] - J[2] * J[1]);
    } else {
      const int ix0[3] = {dof % n, (dof / n) % n, dof / (n * n)};
      for (int c = 0; c < 3; ++c) {
        const int m = (c + 1) % 3, l = (c + 2) % 3;
        // (A)_tA with Jacobian row rowA minus (B)_tB with row rowB, the
        // second sum accumulated into the same running value
        auto curl = [&](int tA, int rowA, int tB, int rowB) {
          int ix[3] = {ix0[0], ix0[1], ix0[2]};
          double r = 0.0;
          const int sA = ix0[tA];
          for (int ii = 0; ii < n; ++ii) {
            ix[tA] = ii;
            const int q = dof_of<3>(ix, n);
            r += 0.5 * D[sA * n + ii] *
                 (X[q * 3 + m] * JM[(q * 3 + rowA) * 3 + l] - X[q * 3 + l] * JM[(q * 3 + rowA) * 3 + m]);
          }
          ix[tA] = sA;
          const int sB = ix0[tB];
          for (int ii = 0; ii < n; ++ii) {
            ix[tB] = ii;
            const int q = dof_of<3>(ix, n);
            r -= 0.5 * D[sB * n + ii] *
                 (X[q * 3 + m] * JM[(q * 3 + rowB) * 3 + l] - X[q * 3 + l] * JM[(q * 3 + rowB) * 3 + m]);
          }
          return r;
        };
        A[0 * 3 + c] = curl(1, 2, 2, 1);
        A[1 * 3 + c] = curl(2, 0, 0, 2);
        A[2 * 3 + c] = curl(0, 1, 1, 0);
      }
      const double j11 = J[0], j12 = J[1], j13 = J[2];
      const double j21 = J[3], j22 = J[4], j23 = J[5];
      const double j31 = J[6], j32 = J[7], j33 = J[8];
      const double det =
          j11 * (j22 * j33 - j23 * j32) - j12 * (j21 * j33 - j23 * j31) + j13 * (j21 * j32 - j22 * j31);
      invj[e * nn + dof] = 1.0 / det;
    }
    for (int r = 0; r < DIM; ++r)
      for (int c = 0; c < DIM; ++c) {
        ja_ref[((e * nn + dof) * DIM + r) * DIM + c] = A[r * DIM + c];
        ja_dev[((e * DIM + r) * DIM + c) * nn + dof] = A[r * DIM + c];
      }
  }
  __syncthreads();

  // subcell normals: n_0 = Ja(0) + w_0 sum_m D(0,m) Ja(m),
  // n_f = n_{f-1} + w_f sum_m D(f,m) Ja(m), per line of direction r
  for (int t = tid; t < DIM * FN; t += nt) {
    const int r = t / FN, line = t % FN;
    int ix[3] = {0, 0, 0};
    int o = 0;
    for (int k = 0; k < DIM; ++k)
      if (k != r) ix[k] = (o++ == 0) ? line % n : line / n;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int f = 0; f < n - 1; ++f) {
      if (f == 0) {
        ix[r] = 0;
        const int d0 = dof_of<DIM>(ix, n);
        for (int c = 0; c < DIM; ++c) acc[c] = JA[(d0 * DIM + r) * DIM + c];
      }
      for (int m = 0; m < n; ++m) {
        ix[r] = m;
        const int dm = dof_of<DIM>(ix, n);
        const double wD = w[f] * D[f * n + m];
        for (int c = 0; c < DIM; ++c) acc[c] += wD * JA[(dm * DIM + r) * DIM + c];
      }
      double* out = sn + ((((e * DIM + r) * FN) + line) * (n - 1) + f) * DIM;
      for (int c = 0; c < DIM; ++c) out[c] = acc[c];
    }
  }
}

int gfail(dgsem_handle* h, int code, const std::string& msg) {
  h->err = msg;
  return code;
}

template <class T>
struct DeviceBuf {
  T* p = nullptr;
  ~DeviceBuf() { hipFree(p); }
  hipError_t alloc(size_t count) { return hipMalloc((void**)&p, count * sizeof(T)); }
};

}  // namespace

extern "C" int dgsem_reference_coordinates(const dgsem_handle* h, double* out) {
  if (!h || !h->d_Ja || !out) return DGSEM_E_INVALID;
  // calc_node_coordinates, data_container.hpp:193-221 (2D), :371-414 (3D):
  // the position of every node in the reference cube [-1, 1]^dim
  const int d = h->cfg.dim, n = h->basis.n, nn = h->nn;
  const int cells[3] = {h->cfg.cells[0], h->cfg.cells[1], d == 3 ? h->cfg.cells[2] : 1};
  double delta[3];
  for (int k = 0; k < d; ++k) delta[k] = 2.0 / (double)cells[k];
  for (int ie = 0; ie < cells[0]; ++ie)
    for (int je = 0; je < cells[1]; ++je)
      for (int ke = 0; ke < cells[2]; ++ke) {
        const int64_t e = ((int64_t)ie * cells[1] + je) * cells[2] + ke;
        const int eid[3] = {ie, je, ke};
        for (int dof = 0; dof < nn; ++dof) {
          const int nd[3] = {dof % n, (dof / n) % n, dof / (n * n)};
          for (int k = 0; k < d; ++k) {
            const double off = -1.0 + (double)eid[k] * delta[k] + 0.5 * delta[k];
            out[(e * nn + dof) * d + k] = off + 0.5 * delta[k] * h->basis.nodes[nd[k]];
          }
        }
      }
  return DGSEM_OK;
}

extern "C" int dgsem_set_node_coordinates(dgsem_handle* h, const double* coords) {
  if (!h || !h->d_Ja || !coords) return DGSEM_E_INVALID;
  if (h->cfg.geometry != DGSEM_GEOMETRY_NODAL)
    return gfail(h, DGSEM_E_UNSUPPORTED, "mapped coordinates need the nodal geometry mode");
  const int d = h->cfg.dim, n = h->basis.n, nn = h->nn;
  const int FN = d == 2 ? n : n * n;
  const int64_t E = h->E;
  // a mapped mesh is no longer a box: the kernels read the per-node metrics
  if (h->d_ja_box) {
    (void)hipStreamSynchronize(h->stream);  // queued launches may still read the table
    hipFree(h->d_ja_box);
    h->d_ja_box = nullptr;
    h->base.ja_box = nullptr;
  }
  const size_t nco = (size_t)E * nn * d, nmat = (size_t)E * nn * d * d, nsn = (size_t)E * d * FN * (n - 1) * d;
  DeviceBuf<double> dD, dw, djac, djaref, dsn;
  if (dD.alloc(n * n) || dw.alloc(n) || djac.alloc(nmat) || djaref.alloc(nmat) || dsn.alloc(nsn))
    return gfail(h, DGSEM_E_ALLOC, "geometry: device allocation failed");
  hipStream_t s = h->stream;
  hipError_t err = hipMemcpyAsync(dD.p, h->basis.D.data(), n * n * sizeof(double), hipMemcpyHostToDevice, s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(dw.p, h->basis.weights.data(), n * sizeof(double), hipMemcpyHostToDevice, s);
  if (err != hipSuccess) return gfail(h, DGSEM_E_HIP, "geometry: upload failed");
  const size_t lds = (size_t)nn * (d + 2 * d * d) * sizeof(double);
  if (d == 2) {
    hipLaunchKernelGGL(geometry_kernel<2>, dim3((unsigned)E), dim3(256), lds, s, coords, dD.p, dw.p,
                       n, djac.p, djaref.p, h->d_Ja, h->d_invJ, dsn.p);
  } else {
    err = hipFuncSetAttribute((const void*)geometry_kernel<3>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return gfail(h, DGSEM_E_HIP, "geometry: LDS attribute");
    hipLaunchKernelGGL(geometry_kernel<3>, dim3((unsigned)E), dim3(256), lds, s, coords, dD.p, dw.p,
                       n, djac.p, djaref.p, h->d_Ja, h->d_invJ, dsn.p);
  }
  err = hipGetLastError();
  if (err != hipSuccess) return gfail(h, DGSEM_E_HIP, "geometry: kernel launch failed");
  // the device subcell normals exist (nsn doubles) only for shock capturing
  if (h->cfg.volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING)
    err = hipMemcpyAsync(h->d_sn, dsn.p, nsn * sizeof(double), hipMemcpyDeviceToDevice, s);
  // host mirrors for the getters and the VTU writer
  auto& g = h->geo;
  if (err == hipSuccess)
    err = hipMemcpyAsync(g.coords.data(), coords, nco * sizeof(double), hipMemcpyDeviceToHost, s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(g.jacobian.data(), djac.p, nmat * sizeof(double), hipMemcpyDeviceToHost, s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(g.contravariant.data(), djaref.p, nmat * sizeof(double), hipMemcpyDeviceToHost, s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(g.inverse_jacobian.data(), h->d_invJ, (size_t)E * nn * sizeof(double),
                         hipMemcpyDeviceToHost, s);
  if (err == hipSuccess) {
    g.subcell_normals.resize(nsn);
    err = hipMemcpyAsync(g.subcell_normals.data(), dsn.p, nsn * sizeof(double), hipMemcpyDeviceToHost, s);
  }
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err != hipSuccess) return gfail(h, DGSEM_E_HIP, "geometry: copy failed");
  return DGSEM_OK;
}
