// C ABI of the tangent and discrete adjoint of the right-hand side and of
// the SSPRK3 step, and of the kinetic-energy objective (include/dgsem_amd.h,
// dgsem_rhs_jvp ... dgsem_kinetic_energy_grad).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dgsem_amd.h"
#include "handle.hpp"
#include "kernels.hpp"

namespace dgsem {

hipError_t launch_jvp(int dim, int N, bool nodal, bool ns, const AdParams& a, hipStream_t s) {
  (void)ns;
  if (dim == 2) {
    switch (N) {
      case 1: return launch_jvp_dn<2, 1>(nodal, a, s);
      case 2: return launch_jvp_dn<2, 2>(nodal, a, s);
      case 3: return launch_jvp_dn<2, 3>(nodal, a, s);
      case 4: return launch_jvp_dn<2, 4>(nodal, a, s);
      case 5: return launch_jvp_dn<2, 5>(nodal, a, s);
      case 7: return launch_jvp_dn<2, 7>(nodal, a, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (N) {
    case 2: return launch_jvp_dn<3, 2>(nodal, a, s);
    case 3: return launch_jvp_dn<3, 3>(nodal, a, s);
    case 4: return launch_jvp_dn<3, 4>(nodal, a, s);
    case 5: return launch_jvp_dn<3, 5>(nodal, a, s);
    case 7: return launch_jvp_dn<3, 7>(nodal, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_vjp(int dim, int N, bool nodal, bool ns, const AdParams& a, hipStream_t s) {
  (void)ns;
  if (dim == 2) {
    switch (N) {
      case 1: return launch_vjp_dn<2, 1>(nodal, a, s);
      case 2: return launch_vjp_dn<2, 2>(nodal, a, s);
      case 3: return launch_vjp_dn<2, 3>(nodal, a, s);
      case 4: return launch_vjp_dn<2, 4>(nodal, a, s);
      case 5: return launch_vjp_dn<2, 5>(nodal, a, s);
      case 7: return launch_vjp_dn<2, 7>(nodal, a, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (N) {
    case 2: return launch_vjp_dn<3, 2>(nodal, a, s);
    case 3: return launch_vjp_dn<3, 3>(nodal, a, s);
    case 4: return launch_vjp_dn<3, 4>(nodal, a, s);
    case 5: return launch_vjp_dn<3, 5>(nodal, a, s);
    case 7: return launch_vjp_dn<3, 7>(nodal, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_gradient(int dim, int N, bool nodal, const KParams& p, hipStream_t s) {
  if (dim == 2) {
    switch (N) {
      case 1: return launch_gradient_dn<2, 1>(nodal, p, s);
      case 2: return launch_gradient_dn<2, 2>(nodal, p, s);
      case 3: return launch_gradient_dn<2, 3>(nodal, p, s);
      case 4: return launch_gradient_dn<2, 4>(nodal, p, s);
      case 5: return launch_gradient_dn<2, 5>(nodal, p, s);
      case 7: return launch_gradient_dn<2, 7>(nodal, p, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (N) {
    case 2: return launch_gradient_dn<3, 2>(nodal, p, s);
    case 3: return launch_gradient_dn<3, 3>(nodal, p, s);
    case 4: return launch_gradient_dn<3, 4>(nodal, p, s);
    case 5: return launch_gradient_dn<3, 5>(nodal, p, s);
    case 7: return launch_gradient_dn<3, 7>(nodal, p, s);
    default: return hipErrorInvalidValue;
  }
}

namespace {

constexpr int kKeBlock = 256;
constexpr int64_t kKeMaxBlocks = 2048;

// Kinetic energy sum over nodes of w_q / |1/J| |m|^2 / (2 rho): each block
// sums a fixed grid-stride set of nodes and reduces in a fixed tree, so the
// result does not depend on scheduling.
__global__ void __launch_bounds__(kKeBlock)
    ke_kernel(int dim, int nn, const double* __restrict__ wq, int nodal, KParams p,
              double* __restrict__ partial) {
  __shared__ double s[kKeBlock];
  const int nv = dim + 2;
  const int64_t total = p.E * nn;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kKeBlock + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kKeBlock) {
    const int64_t e = i / nn;
    const int node = (int)(i % nn);
    const double* u = p.u + e * nv * nn + node;
    double m2 = 0.0;
    for (int k = 0; k < dim; ++k) m2 += u[(k + 1) * nn] * u[(k + 1) * nn];
    const double ij = nodal ? p.invJ[e * nn + node] : p.invJ[e];
    acc += wq[node] / fabs(ij) * (0.5 * m2 / u[0]);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kKeBlock / 2; w > 0; w /= 2) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kKeBlock)
    ke_final_kernel(const double* __restrict__ partial, int nblocks, double* __restrict__ result) {
  __shared__ double s[kKeBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += kKeBlock) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kKeBlock / 2; w > 0; w /= 2) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = s[0];
}

// grad += scale * dKE/du: d/drho = -w |m|^2 / (2 rho^2), d/dm = w m / rho.
__global__ void __launch_bounds__(kKeBlock)
    ke_grad_kernel(int dim, int nn, const double* __restrict__ wq, int nodal, KParams p,
                   double scale, double* __restrict__ grad) {
  const int nv = dim + 2;
  const int64_t total = p.E * nn;
  for (int64_t i = blockIdx.x * (int64_t)kKeBlock + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kKeBlock) {
    const int64_t e = i / nn;
    const int node = (int)(i % nn);
    const int64_t base = e * nv * nn + node;
    const double* u = p.u + base;
    const double ij = nodal ? p.invJ[e * nn + node] : p.invJ[e];
    const double w = scale * wq[node] / fabs(ij);
    const double ir = 1.0 / u[0];
    double m2 = 0.0;
    for (int k = 0; k < dim; ++k) m2 += u[(k + 1) * nn] * u[(k + 1) * nn];
    grad[base] += -0.5 * w * m2 * ir * ir;
    for (int k = 0; k < dim; ++k) grad[base + (k + 1) * nn] += w * u[(k + 1) * nn] * ir;
  }
}

int64_t ke_blocks(int64_t total) {
  const int64_t b = (total + kKeBlock - 1) / kKeBlock;
  return b < 1 ? 1 : (b > kKeMaxBlocks ? kKeMaxBlocks : b);
}

}  // namespace

int64_t ke_partials(int64_t E) { (void)E; return kKeMaxBlocks; }

hipError_t launch_kinetic_energy(int dim, int nn, const double* wq, bool nodal, const KParams& p,
                                 double* partial, double* result, hipStream_t s) {
  const int64_t nb = ke_blocks(p.E * nn);
  hipLaunchKernelGGL(ke_kernel, dim3((unsigned)nb), dim3(kKeBlock), 0, s, dim, nn, wq, nodal ? 1 : 0,
                     p, partial);
  hipLaunchKernelGGL(ke_final_kernel, dim3(1), dim3(kKeBlock), 0, s, partial, (int)nb, result);
  return hipGetLastError();
}

hipError_t launch_kinetic_energy_grad(int dim, int nn, const double* wq, bool nodal,
                                      const KParams& p, double scale, double* grad, hipStream_t s) {
  const int64_t nb = ke_blocks(p.E * nn);
  hipLaunchKernelGGL(ke_grad_kernel, dim3((unsigned)nb), dim3(kKeBlock), 0, s, dim, nn, wq,
                     nodal ? 1 : 0, p, scale, grad);
  return hipGetLastError();
}

}  // namespace dgsem

namespace {

int afail(dgsem_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

int ahip(dgsem_handle* h, hipError_t e, const char* where) {
  return afail(h, DGSEM_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

int alloc_once(dgsem_handle* h, double** p, size_t count) {
  if (*p) return DGSEM_OK;
  if (hipMalloc((void**)p, count * sizeof(double)) != hipSuccess)
    return afail(h, DGSEM_E_ALLOC, "hipMalloc failed");
  return DGSEM_OK;
}

// Scratch and the configuration check of the differentiated path.
int ad_prepare(dgsem_handle* h) {
  const auto& c = h->cfg;
  const bool sc = c.volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING;
  if (dgsem::uses_central_flux(c))
    return afail(h, DGSEM_E_UNSUPPORTED, "CentralFlux is not differentiated");
  const size_t cell = (size_t)h->E * h->nvars * h->nn;
  if (sc && !h->base.alpha_in) {
    // the indicator's alpha at the linearisation state, from a forward RHS
    int rc;
    if ((rc = alloc_once(h, &h->d_ad_alpha, (size_t)h->E)) || (rc = alloc_once(h, &h->d_ad_out, cell)))
      return rc;
  }
  if (c.equation == DGSEM_NAVIER_STOKES) {
    int rc;
    // the forward gradient, which the node-parallel forward path never stores
    if (!h->d_grad) {
      if ((rc = alloc_once(h, &h->d_grad, cell * c.dim))) return rc;
      h->base.grad = h->base.grad_out = h->d_grad;
    }
    if ((rc = alloc_once(h, &h->d_ad_dgrad, cell * c.dim)) ||
        (rc = alloc_once(h, &h->d_ad_dvisc, cell)) || (rc = alloc_once(h, &h->d_ad_qbar, cell)))
      return rc;
  }
  return DGSEM_OK;
}

dgsem::AdParams ad_params(dgsem_handle* h, const double* u) {
  dgsem::AdParams a{};
  a.k = h->base;
  a.k.u = u;
  a.k.grad = h->d_grad;
  a.k.grad_out = h->d_grad;
  a.dgrad = h->d_ad_dgrad;
  a.dvisc = h->cfg.equation == DGSEM_NAVIER_STOKES ? h->d_ad_dvisc : nullptr;
  a.qbar = h->d_ad_qbar;
  return a;
}

// out = a1 x1 + a2 x2 + b v + c Op(u) v with Op the tangent or adjoint.
int apply(dgsem_handle* h, bool adjoint, const double* u, const double* v, const double* x1,
          double a1, const double* x2, double a2, double b, double c, double* out) {
  int rc = ad_prepare(h);
  if (rc) return rc;
  const bool nodal = h->cfg.geometry == DGSEM_GEOMETRY_NODAL;
  const bool ns = h->cfg.equation == DGSEM_NAVIER_STOKES;
  dgsem::AdParams a = ad_params(h, u);
  if (h->cfg.volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING && !a.k.alpha_in) {
    // Freeze the blending factor the forward evaluation at u computes: a
    // forward RHS of u into scratch records it (the same kernel on the same
    // bits as the forward pass, so the same alpha).
    dgsem::KParams f = h->base;
    f.u = u;
    f.out = h->d_ad_out;
    f.alpha = h->d_ad_alpha;
    f.mode = dgsem::kModeRhs;
    if ((rc = dgsem::launch_params(h, f))) return rc;
    a.k.alpha_in = h->d_ad_alpha;
  }
  a.v = v;
  a.x1 = x1;
  a.a1 = a1;
  a.x2 = x2;
  a.a2 = a2;
  a.b = b;
  a.c = c;
  a.out = out;
  hipError_t e;
  if (ns) {
    e = dgsem::launch_gradient(h->cfg.dim, h->cfg.polydeg, nodal, a.k, h->stream);
    if (e != hipSuccess) return ahip(h, e, "gradient kernel launch");
  }
  e = adjoint ? dgsem::launch_vjp(h->cfg.dim, h->cfg.polydeg, nodal, ns, a, h->stream)
              : dgsem::launch_jvp(h->cfg.dim, h->cfg.polydeg, nodal, ns, a, h->stream);
  return e == hipSuccess ? DGSEM_OK : ahip(h, e, adjoint ? "adjoint kernel launch" : "tangent kernel launch");
}

bool distinct(const void* a, const void* b) { return a && b && a != b; }

// The positivity floor in the step derivatives.  A stage's result is the
// stage combination pre = c1 u_base + c2 u_k + c3 R(u_k) (stage 1:
// c1 u_k + c2 R(u_k)) clamped as the forward kernel's epilogue clamps it:
// rho' = max(rho, floor), then E' = floor / (gamma - 1) + kin(rho', m) where
// the pressure would fall below the floor.  This applies the clamp's
// derivative at pre to w in place: the tangent (adjoint = false) or its
// transpose.  pre is formed with the forward's expressions from R(u_k), so
// the clamps are the forward's.
template <int DIM>
__global__ void floor_clamp_kernel(int64_t nodes, int nn, const double* __restrict__ u_base,
                                   const double* __restrict__ u_k, const double* __restrict__ r,
                                   double c1, double c2, double c3, double floor_value, double gamma,
                                   bool adjoint, double* __restrict__ w) {
  constexpr int NV = DIM + 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  const int64_t g = (i / nn) * NV * nn + i % nn;
  double s[NV], x[NV];
  for (int v = 0; v < NV; ++v) {
    const double uk = u_k[g + v * nn];
    s[v] = u_base ? c1 * u_base[g + v * nn] + c2 * uk + c3 * r[g + v * nn] : c1 * uk + c2 * r[g + v * nn];
    x[v] = w[g + v * nn];
  }
  const bool rho_clamped = s[0] < floor_value;
  const double rho = rho_clamped ? floor_value : s[0];
  double m2 = 0.0;
  for (int k = 0; k < DIM; ++k) m2 += s[k + 1] * s[k + 1];
  const double kin = 0.5 * m2 / rho;
  const bool e_clamped = (gamma - 1.0) * (s[NV - 1] - kin) < floor_value;
  if (!adjoint) {
    const double drho = rho_clamped ? 0.0 : x[0];
    if (e_clamped) {
      double mdm = 0.0;
      for (int k = 0; k < DIM; ++k) mdm += s[k + 1] * x[k + 1];
      x[NV - 1] = mdm / rho - kin / rho * drho;
    }
    x[0] = drho;
  } else {
    double rho_b = x[0];
    if (e_clamped) {
      const double kin_b = x[NV - 1];
      for (int k = 0; k < DIM; ++k) x[k + 1] += kin_b * s[k + 1] / rho;
      rho_b -= kin_b * kin / rho;
      x[NV - 1] = 0.0;
    }
    x[0] = rho_clamped ? 0.0 : rho_b;
  }
  for (int v = 0; v < NV; ++v) w[g + v * nn] = x[v];
}

// Applies the derivative of stage `stage`'s clamp (SSPRK3 coefficients as
// stage_params forms them) to w in place; u_base is u0 for stages 2 and 3.
int floor_stage(dgsem_handle* h, bool adjoint, int stage, double dt, const double* u_base,
                const double* u_k, double* w) {
  const size_t cell = (size_t)h->E * h->nvars * h->nn;
  int rc;
  if ((rc = alloc_once(h, &h->d_fl_r, cell)) || (rc = alloc_once(h, &h->d_ad_alpha, (size_t)h->E)))
    return rc;
  // R(u_k), by the forward kernel with the blending factor the forward step
  // used at u_k (the indicator's, or the frozen array)
  dgsem::KParams f = h->base;
  f.u = u_k;
  f.out = h->d_fl_r;
  f.alpha = h->d_ad_alpha;
  f.mode = dgsem::kModeRhs;
  if ((rc = dgsem::launch_params(h, f))) return rc;
  dgsem::KParams sp;
  if ((rc = dgsem::stage_params(h, stage, dt, stage == 1 ? nullptr : u_base, u_k, w, &sp))) return rc;
  const int64_t nodes = h->E * (int64_t)h->nn;
  const unsigned grid = (unsigned)((nodes + 255) / 256);
  if (h->cfg.dim == 2)
    hipLaunchKernelGGL(floor_clamp_kernel<2>, dim3(grid), dim3(256), 0, h->stream, nodes, h->nn,
                       sp.u_base, u_k, h->d_fl_r, sp.c1, sp.c2, sp.c3, h->cfg.positivity_floor,
                       h->cfg.gamma, adjoint, w);
  else
    hipLaunchKernelGGL(floor_clamp_kernel<3>, dim3(grid), dim3(256), 0, h->stream, nodes, h->nn,
                       sp.u_base, u_k, h->d_fl_r, sp.c1, sp.c2, sp.c3, h->cfg.positivity_floor,
                       h->cfg.gamma, adjoint, w);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DGSEM_OK : ahip(h, e, "floor clamp kernel launch");
}

}  // namespace

namespace dgsem {

int prepare_weights(dgsem_handle* h) {
  if (h->d_wq) return DGSEM_OK;
  const auto& w = h->basis.weights;
  const int n = h->basis.n, d = h->cfg.dim;
  std::vector<double> wq(h->nn);
  for (int dof = 0; dof < h->nn; ++dof) {
    const int i = dof % n, j = (dof / n) % n, k = dof / (n * n);
    double x = w[i] * w[j];
    if (d == 3) x = x * w[k];
    wq[dof] = x;
  }
  if (hipMalloc((void**)&h->d_wq, wq.size() * sizeof(double)) != hipSuccess)
    return afail(h, DGSEM_E_ALLOC, "hipMalloc failed");
  hipError_t e = hipMemcpy(h->d_wq, wq.data(), wq.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) return ahip(h, e, "hipMemcpy");
  if (hipMalloc((void**)&h->d_ke_partial, dgsem::ke_partials(h->E) * sizeof(double)) != hipSuccess)
    return afail(h, DGSEM_E_ALLOC, "hipMalloc failed");
  return DGSEM_OK;
}

}  // namespace dgsem

extern "C" {

int dgsem_rhs_jvp(dgsem_handle* h, const double* u, const double* du, double* out) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (!distinct(u, out) || !distinct(du, out))
    return afail(h, DGSEM_E_INVALID, "out must be a distinct device array");
  return apply(h, false, u, du, nullptr, 0.0, nullptr, 0.0, 0.0, 1.0, out);
}

int dgsem_rhs_vjp(dgsem_handle* h, const double* u, const double* w, double* out) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (!distinct(u, out) || !distinct(w, out))
    return afail(h, DGSEM_E_INVALID, "out must be a distinct device array");
  return apply(h, true, u, w, nullptr, 0.0, nullptr, 0.0, 0.0, 1.0, out);
}

int dgsem_ssprk3_step_jvp(dgsem_handle* h, double dt, const double* u0, const double* u1,
                          const double* u2, const double* du0, double* du_out, double* tmp1,
                          double* tmp2) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (!u0 || !u1 || !u2 || !du0 || !du_out || !tmp1 || !tmp2 || tmp1 == tmp2 || du_out == tmp1 ||
      du_out == tmp2 || du_out == du0)
    return afail(h, DGSEM_E_INVALID, "du0, du_out, tmp1, tmp2 must be distinct device arrays");
  const bool fl = h->cfg.positivity_floor > 0.0;
  int rc;
  // du1 = du0 + dt J(u0) du0 (then the clamp's derivative, with the floor)
  if ((rc = apply(h, false, u0, du0, nullptr, 0.0, nullptr, 0.0, 1.0, dt, tmp1))) return rc;
  if (fl && (rc = floor_stage(h, false, 1, dt, nullptr, u0, tmp1))) return rc;
  // du2 = 3/4 du0 + 1/4 du1 + dt/4 J(u1) du1
  if ((rc = apply(h, false, u1, tmp1, du0, 3.0 / 4.0, nullptr, 0.0, 1.0 / 4.0, 1.0 / 4.0 * dt, tmp2)))
    return rc;
  if (fl && (rc = floor_stage(h, false, 2, dt, u0, u1, tmp2))) return rc;
  // du3 = 1/3 du0 + 2/3 du2 + 2 dt/3 J(u2) du2
  if ((rc = apply(h, false, u2, tmp2, du0, 1.0 / 3.0, nullptr, 0.0, 2.0 / 3.0, 2.0 / 3.0 * dt, du_out)))
    return rc;
  return fl ? floor_stage(h, false, 3, dt, u0, u2, du_out) : DGSEM_OK;
}

int dgsem_ssprk3_step_vjp(dgsem_handle* h, double dt, const double* u0, const double* u1,
                          const double* u2, const double* ubar, double* ubar_out, double* tmp1,
                          double* tmp2) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (!u0 || !u1 || !u2 || !ubar || !ubar_out || !tmp1 || !tmp2 || tmp1 == tmp2 ||
      ubar == tmp1 || ubar == tmp2 || ubar_out == tmp1 || ubar_out == tmp2)
    return afail(h, DGSEM_E_INVALID, "ubar, tmp1, tmp2 must be distinct device arrays");
  int rc;
  // with the floor, each stage's clamp is transposed first: w3 (the last
  // clamp's) in scratch, since ubar_out may alias ubar
  const double* w3 = ubar;
  if (h->cfg.positivity_floor > 0.0) {
    if ((rc = alloc_once(h, &h->d_fl_w, (size_t)h->E * h->nvars * h->nn))) return rc;
    const hipError_t e = hipMemcpyAsync(h->d_fl_w, ubar, sizeof(double) * h->E * h->nvars * h->nn,
                                        hipMemcpyDeviceToDevice, h->stream);
    if (e != hipSuccess) return ahip(h, e, "copy");
    if ((rc = floor_stage(h, true, 3, dt, u0, u2, h->d_fl_w))) return rc;
    w3 = h->d_fl_w;
  }
  // ubar2 = 2/3 w3 + 2 dt/3 J(u2)^T w3
  if ((rc = apply(h, true, u2, w3, nullptr, 0.0, nullptr, 0.0, 2.0 / 3.0, 2.0 / 3.0 * dt, tmp1)))
    return rc;
  if (w3 != ubar && (rc = floor_stage(h, true, 2, dt, u0, u1, tmp1))) return rc;
  // ubar1 = 1/4 ubar2 + dt/4 J(u1)^T ubar2
  if ((rc = apply(h, true, u1, tmp1, nullptr, 0.0, nullptr, 0.0, 1.0 / 4.0, 1.0 / 4.0 * dt, tmp2)))
    return rc;
  if (w3 != ubar && (rc = floor_stage(h, true, 1, dt, nullptr, u0, tmp2))) return rc;
  // ubar0 = 1/3 w3 + 3/4 ubar2 + ubar1 + dt J(u0)^T ubar1  (ubar_out may alias ubar)
  return apply(h, true, u0, tmp2, w3, 1.0 / 3.0, tmp1, 3.0 / 4.0, 1.0, dt, ubar_out);
}

int dgsem_kinetic_energy(dgsem_handle* h, const double* u, double* result) {
  if (!h || !h->d_Ja || !u || !result) return DGSEM_E_INVALID;
  int rc = dgsem::prepare_weights(h);
  if (rc) return rc;
  dgsem::KParams p = h->base;
  p.u = u;
  hipError_t e = dgsem::launch_kinetic_energy(h->cfg.dim, h->nn, h->d_wq,
                                              h->cfg.geometry == DGSEM_GEOMETRY_NODAL, p,
                                              h->d_ke_partial, result, h->stream);
  return e == hipSuccess ? DGSEM_OK : ahip(h, e, "kinetic energy launch");
}

int dgsem_kinetic_energy_grad(dgsem_handle* h, const double* u, double scale, double* grad) {
  if (!h || !h->d_Ja || !u || !grad || u == grad) return DGSEM_E_INVALID;
  int rc = dgsem::prepare_weights(h);
  if (rc) return rc;
  dgsem::KParams p = h->base;
  p.u = u;
  hipError_t e = dgsem::launch_kinetic_energy_grad(h->cfg.dim, h->nn, h->d_wq,
                                                   h->cfg.geometry == DGSEM_GEOMETRY_NODAL, p,
                                                   scale, grad, h->stream);
  return e == hipSuccess ? DGSEM_OK : ahip(h, e, "kinetic energy gradient launch");
}

}  // extern "C"
