// BR1 parabolic terms of the compressible Navier-Stokes RHS on the GPU.
//
// Two kernels per RHS, both one workgroup per EPW elements with line threads
// as in the inviscid kernel:
//   * gradient_kernel: q = (rho, v, T) at the nodes; along each line the weak
//     derivative D_hat q, the BR1 face value q_hat = (q_L + q_R)/2 lifted
//     with l_hat; then the physical gradient sum_d Ja^d (.) and 1/J.
//     Reference: GradientVolumeIntegralFunctor, GradientPhysicalTransform-
//     Functor, GradientInterfaceFluxFunctor, GradientSurfaceIntegralFunctor,
//     GradientJacobianProjFunctor (parabolic_gradient.hpp:17-478 3D,
//     :480-818 2D).
//   * viscous_kernel: the contravariant viscous flux Ja^d . (-F_v) at the
//     nodes, its weak D_hat divergence along each line, and the interface
//     flux, which is the upper ("primary") element's (-F_v).n only
//     (parabolic_viscous.hpp:318-334 3D, :570-585 2D), lifted with l_hat.
//     The result (before -1/J) is added by the inviscid kernel.
//     Elements of the node-parallel RHS kernel (node_kernel_for) take
//     gradient_node_kernel instead, which stores the per-node viscous fluxes
//     only; that kernel adds their terms in the reference's order.
// Layouts: gradient [E][dim][var][dof]; viscous part [E][var][dof], or the
// per-node fluxes [E][dim][var][dof].
// BCS (2D): some face is not periodic; the neighbour table holds -1 there and
// the face values come from the boundary condition (wall_gradient_value and
// the viscous kernel's wall branch).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "physics.hpp"
#include "rhs_kernel_impl.hpp"

namespace dgsem {
namespace detail {

// grad_vars and viscous_flux_dir (the pointwise gradient variables and
// viscous flux) live in physics.hpp, shared with the host-side traits.

template <bool NODAL, int DIM, int NN>
__device__ __forceinline__ double metric(const KParams& p, int64_t e, int r, int c, int dof) {
  if constexpr (NODAL)
    return p.Ja[((e * DIM + r) * DIM + c) * NN + dof];
  else
    return p.Ja[(e * DIM + r) * DIM + c];
}

template <bool NODAL, int NN>
__device__ __forceinline__ double inv_jac(const KParams& p, int64_t e, int dof) {
  if constexpr (NODAL)
    return p.invJ[e * NN + dof];
  else
    return p.invJ[e];
}

// Gradient face value at a boundary node: the adiabatic walls take the inner
// gradient variables q with the velocity set to zero (no slip,
// navier_stokes_wall_boundary.hpp:81-114) or with its component along the
// normal nrm = sign(J) Ja^dir removed (slip, :163-197 and :28-36).
// Dirichlet faces take the outer state's variables (computed by the caller,
// dirichlet_boundary.hpp:122-162).  Kinds without a gradient hook
// (boundary_set.hpp:16-33) leave the face value at zero, as the reference's
// cleared face array does.
template <int DIM>
__device__ __forceinline__ void wall_gradient_value(int kind, const double* q, const double* nrm,
                                                    double* qw) {
  constexpr int NV = DIM + 2;
  static_assert(DIM == 2, "non-periodic boundaries are built for 2D meshes");
  if (kind == kBcNsNoSlipWall) {
    qw[0] = q[0];
    qw[1] = 0.0;
    qw[2] = 0.0;
    qw[3] = q[3];
  } else if (kind == kBcNsSlipWall) {
    const double nn = nrm[0] * nrm[0] + nrm[1] * nrm[1];
    const double vn = (q[1] * nrm[0] + q[2] * nrm[1]) / nn;
    qw[0] = q[0];
    qw[1] = q[1] - vn * nrm[0];
    qw[2] = q[2] - vn * nrm[1];
    qw[3] = q[3];
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) qw[v] = 0.0;
  }
}

template <int DIM, int N, int EPW, bool NODAL, bool BCS>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) gradient_kernel(const KParams p) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES;
  constexpr int BLOCK = EPW * L;
  __shared__ double s_q[EPW * NV * NN];
  __shared__ double s_g[EPW * DIM * NV * NN];  // per direction: D_hat q + lift
  const int tid = threadIdx.x;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e0 = (int64_t)blk * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    double s[NV], q[NV];
    for (int v = 0; v < NV; ++v) s[v] = p.u[((e0 + sl) * NV + v) * NN + node];
    grad_vars<DIM>(p.gamma, s, q);
    for (int v = 0; v < NV; ++v) s_q[(sl * NV + v) * NN + node] = q[v];
  }
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  if (slot < nvalid) {
    const int64_t e = e0 + slot;
    int dofs[n];
#pragma unroll
    for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
    const double* q = s_q + slot * NV * NN;
    double acc[n][NV];
#pragma unroll
    for (int k = 0; k < n; ++k)
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        double r = 0.0;
#pragma unroll
        for (int i = 0; i < n; ++i) r += p.Dhat[k][i] * q[v * NN + dofs[i]];
        acc[k][v] = r;
      }
    // BR1 face values q_hat = (q_lower + q_upper) / 2 on both line ends
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
      const int64_t nb = p.nbr[e * 2 * DIM + 2 * dir + side];
      if constexpr (BCS) {
        if (nb < 0) {
          const int dof_own = side == 0 ? dofs[0] : dofs[n - 1];
          const double sg = inv_jac<NODAL, NN>(p, e, dof_own) >= 0.0 ? 1.0 : -1.0;
          double qo[NV], nrm[DIM], qw[NV];
          const int face = 2 * dir + side;
          if (p.bc[face] == kBcDirichlet) {
            const int64_t along = dir == 0 ? e % p.cells[1] : e / p.cells[1];
            const int64_t bnode = along * n + line;
            const double* vals = p.bc_values[face];
            double so[NV];
            for (int v = 0; v < NV; ++v) so[v] = vals ? vals[bnode * NV + v] : p.bc_state[face][v];
            grad_vars<DIM>(p.gamma, so, qw);
          } else {
            for (int v = 0; v < NV; ++v) qo[v] = q[v * NN + dof_own];
            for (int c = 0; c < DIM; ++c) nrm[c] = sg * metric<NODAL, DIM, NN>(p, e, dir, c, dof_own);
            wall_gradient_value<DIM>(p.bc[face], qo, nrm, qw);
          }
          if (side == 0) {
            for (int v = 0; v < NV; ++v) acc[0][v] -= sg * p.lhat0 * qw[v];
          } else {
            for (int v = 0; v < NV; ++v) acc[n - 1][v] += sg * p.lhatN * qw[v];
          }
          continue;
        }
      }
      const int dof_nb = side == 0 ? dofs[n - 1] : dofs[0];
      double s[NV], qn[NV];
      for (int v = 0; v < NV; ++v) s[v] = p.u[(nb * NV + v) * NN + dof_nb];
      grad_vars<DIM>(p.gamma, s, qn);
      // the face term uses the element's own normal sign(J) Ja^dir at its node
      const int dof_own = side == 0 ? dofs[0] : dofs[n - 1];
      const double sg = inv_jac<NODAL, NN>(p, e, dof_own) >= 0.0 ? 1.0 : -1.0;
      if (side == 0) {
        for (int v = 0; v < NV; ++v) acc[0][v] -= sg * p.lhat0 * (0.5 * (qn[v] + q[v * NN + dofs[0]]));
      } else {
        for (int v = 0; v < NV; ++v) acc[n - 1][v] += sg * p.lhatN * (0.5 * (q[v * NN + dofs[n - 1]] + qn[v]));
      }
    }
#pragma unroll
    for (int k = 0; k < n; ++k)
#pragma unroll
      for (int v = 0; v < NV; ++v) s_g[((slot * DIM + dir) * NV + v) * NN + dofs[k]] = acc[k][v];
  }
  __syncthreads();
  // physical gradient: g_c = (sum_d Ja^d_c G_d) / J
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int64_t e = e0 + sl;
    const double ij = inv_jac<NODAL, NN>(p, e, node);
    for (int c = 0; c < DIM; ++c) {
      double ja[DIM];
      for (int d = 0; d < DIM; ++d) ja[d] = metric<NODAL, DIM, NN>(p, e, d, c, node);
      for (int v = 0; v < NV; ++v) {
        double g = 0.0;
        for (int d = 0; d < DIM; ++d) g += ja[d] * s_g[((sl * DIM + d) * NV + v) * NN + node];
        p.grad_out[((e * DIM + c) * NV + v) * NN + node] = g * ij;
      }
    }
  }
}

// Contravariant viscous flux Ja^d . (-F_v) of one node: out[d][v].
template <int DIM, int NN, bool NODAL>
__device__ __forceinline__ void contravariant_viscous(const KParams& p, int64_t e, int node,
                                                      const double* u, double out[DIM][DIM + 2]) {
  constexpr int NV = DIM + 2;
  double q[NV], g[DIM][NV], fv[DIM][NV];
  grad_vars<DIM>(p.gamma, u, q);
  for (int c = 0; c < DIM; ++c)
    for (int v = 0; v < NV; ++v) g[c][v] = p.grad[((e * DIM + c) * NV + v) * NN + node];
  for (int c = 0; c < DIM; ++c) viscous_flux_dir<DIM>(p.mu, p.kappa, q, g, c, fv[c]);
  for (int d = 0; d < DIM; ++d) {
    double ja[DIM];
    for (int c = 0; c < DIM; ++c) ja[c] = metric<NODAL, DIM, NN>(p, e, d, c, node);
    for (int v = 0; v < NV; ++v) {
      double r = 0.0;
      for (int c = 0; c < DIM; ++c) r += ja[c] * -fv[c][v];
      out[d][v] = r;
    }
  }
}

template <int DIM, int N, int EPW, bool NODAL, bool BCS>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) viscous_kernel(const KParams p) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES;
  constexpr int BLOCK = EPW * L;
  __shared__ double s_f[EPW * DIM * NV * NN];  // Ja^d . (-F_v) per node and d
  __shared__ double s_du[EPW * NV * NN];
  const int tid = threadIdx.x;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e0 = (int64_t)blk * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int64_t e = e0 + sl;
    double s[NV], out[DIM][NV];
    for (int v = 0; v < NV; ++v) s[v] = p.u[(e * NV + v) * NN + node];
    contravariant_viscous<DIM, NN, NODAL>(p, e, node, s, out);
    for (int d = 0; d < DIM; ++d)
      for (int v = 0; v < NV; ++v) s_f[((sl * DIM + d) * NV + v) * NN + node] = out[d][v];
  }
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  const bool valid = slot < nvalid;
  double acc[n][NV];
  int dofs[n];
#pragma unroll
  for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
  if (valid) {
    const int64_t e = e0 + slot;
    const double* f = s_f + (slot * DIM + dir) * NV * NN;
#pragma unroll
    for (int k = 0; k < n; ++k)
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        double r = 0.0;
#pragma unroll
        for (int i = 0; i < n; ++i) r += p.Dhat[k][i] * f[v * NN + dofs[i]];
        acc[k][v] = r;
      }
    // Boundary faces (BCS): the own node's (-F_v).(sign(J) Ja^dir) for the
    // momentum components only at a no-slip adiabatic wall
    // (navier_stokes_wall_boundary.hpp:116-145), for every variable at a
    // Dirichlet face (dirichlet_boundary.hpp:164-194); every other kind
    // leaves the face value at zero (:199-216, boundary_set.hpp:35-51).
    bool bnd_lo = false, bnd_hi = false;
    int vlo[2] = {0, 0}, vhi[2] = {0, 0};  // variables carried, per side
    if constexpr (BCS) {
      bnd_lo = p.nbr[e * 2 * DIM + 2 * dir] < 0;
      bnd_hi = p.nbr[e * 2 * DIM + 2 * dir + 1] < 0;
      for (int side = 0; side < 2; ++side) {
        const int kind = p.bc[2 * dir + side];
        vlo[side] = kind == kBcDirichlet ? 0 : 1;
        vhi[side] = kind == kBcDirichlet ? NV : (kind == kBcNsNoSlipWall ? NV - 1 : 0);
      }
    }
    // lower face: this element is primary, (-F_v).(sign(J) Ja^dir) at its node
    {
      const double sg = inv_jac<NODAL, NN>(p, e, dofs[0]) >= 0.0 ? 1.0 : -1.0;
      if (!bnd_lo) {
        for (int v = 0; v < NV; ++v) acc[0][v] -= sg * f[v * NN + dofs[0]] * p.lhat0;
      } else {
        for (int v = vlo[0]; v < vhi[0]; ++v) acc[0][v] -= sg * f[v * NN + dofs[0]] * p.lhat0;
      }
    }
    if (bnd_hi) {
      const double sg = inv_jac<NODAL, NN>(p, e, dofs[n - 1]) >= 0.0 ? 1.0 : -1.0;
      for (int v = vlo[1]; v < vhi[1]; ++v) acc[n - 1][v] += sg * f[v * NN + dofs[n - 1]] * p.lhatN;
    } else {
      // upper face: the upper neighbour is primary
      const int64_t nb = p.nbr[e * 2 * DIM + 2 * dir + 1];
      double s[NV], q[NV], g[DIM][NV], fv[DIM][NV];
      for (int v = 0; v < NV; ++v) s[v] = p.u[(nb * NV + v) * NN + dofs[0]];
      grad_vars<DIM>(p.gamma, s, q);
#pragma unroll
      for (int c = 0; c < DIM; ++c)
#pragma unroll
        for (int v = 0; v < NV; ++v) g[c][v] = p.grad[((nb * DIM + c) * NV + v) * NN + dofs[0]];
#pragma unroll
      for (int c = 0; c < DIM; ++c) viscous_flux_dir<DIM>(p.mu, p.kappa, q, g, c, fv[c]);
      double ja[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) ja[c] = metric<NODAL, DIM, NN>(p, nb, dir, c, dofs[0]);
      const double sg = inv_jac<NODAL, NN>(p, nb, dofs[0]) >= 0.0 ? 1.0 : -1.0;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        double r = 0.0;
#pragma unroll
        for (int c = 0; c < DIM; ++c) r += ja[c] * -fv[c][v];
        acc[n - 1][v] += sg * r * p.lhatN;
      }
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    if (valid && dir == d) {
#pragma unroll
      for (int k = 0; k < n; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const int idx = (slot * NV + v) * NN + dofs[k];
          s_du[idx] = d == 0 ? acc[k][v] : s_du[idx] + acc[k][v];
        }
    }
    __syncthreads();
  }
  for (int i = tid; i < EPW * NV * NN; i += BLOCK) {
    if (i / (NV * NN) >= nvalid) continue;
    p.visc_out[e0 * NV * NN + i] = s_du[i];
  }
}

// Viscous fluxes for the node-parallel RHS kernel: at node `node` of
// element e the contravariant flux Ja^d . (-F_v) of each direction d, into
// p.visc_out [E][d][var][dof], rounded as ViscousFluxFunctor and the viscous
// volume integral's fref (parabolic_viscous.hpp:39-66 and :103-163 3D,
// :367-386 and :412-447 2D): -F_v of each physical direction, then
// Ja^d_0 f_0 + Ja^d_1 f_1 (+ Ja^d_2 f_2) from left to right, with no fused
// multiply-adds.  The interface value (-F_v) . (sign(J) Ja^d) of a face node
// is sign(J) times this sum to the bit (scaling by a sign is exact), so the
// RHS kernel needs nothing more.
template <int DIM, int N, bool NODAL>
__device__ __forceinline__ void store_viscous_fluxes(const KParams& p, int64_t e, int node,
                                                     const double* q, const double g[DIM][DIM + 2]) {
  constexpr int NV = DIM + 2, NN = Shape<DIM, N>::NN;
  double fv[DIM][NV];
#pragma unroll
  for (int c = 0; c < DIM; ++c) viscous_flux_dir<DIM>(p.mu, p.kappa, q, g, c, fv[c]);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double ja[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) ja[c] = metric<NODAL, DIM, NN>(p, e, d, c, node);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
      double r = ja[0] * -fv[0][v];
#pragma unroll
      for (int c = 1; c < DIM; ++c) r = r + ja[c] * -fv[c][v];
      p.visc_out[((e * DIM + d) * NV + v) * NN + node] = r;
    }
  }
}

// BR1 gradient for the node-parallel RHS kernel's elements (periodic, no
// shock capturing, N >= 7), one workgroup per element and one thread per
// node, rounded as the reference's scalar functors round it: the weak
// derivative sum_m D_hat(x_d, m) q(m) of each direction d from left to right
// (GradientVolumeIntegralFunctor, parabolic_gradient.hpp:40-79 3D,
// :502-527 2D), the physical transform Ja^0_c r_0 + Ja^1_c r_1 (+ Ja^2_c r_2)
// (:121-150 / :567-584), the lifted face values (-l_0 or l_N) q_hat times
// sign(J) Ja^d_c, q_hat = (q_lower + q_upper) / 2, in direction order
// (:186-291 and :348-430 / :628-685 and :737-788), then 1/J (:462-474 /
// :806-815); no fused multiply-adds.  The derivative of T = p / rho cancels
// terms some thousand times its size at N = 7, so these roundings show in
// the viscous terms at the 1e-12 level.  Writes the node's viscous fluxes
// (store_viscous_fluxes), not the gradient.
template <int DIM, int N, bool NODAL>
__global__ void __launch_bounds__((Shape<DIM, N>::NN)) gradient_node_kernel(const KParams p) {
  constexpr int n = N + 1, NV = DIM + 2, NN = Shape<DIM, N>::NN;
  __shared__ double s_q[NV * NN];
  __shared__ double s_Dh[n * n];
  const int a = threadIdx.x;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x);
  const int idx[3] = {a % n, (a / n) % n, DIM == 3 ? a / (n * n) : 0};
  constexpr int stride[3] = {1, n, n * n};
  const int nz = DIM == 3 ? p.cells[2] : 1;
  double u[NV], q[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) u[v] = p.u[(e * NV + v) * NN + a];
  grad_vars<DIM>(p.gamma, u, q);
#pragma unroll
  for (int v = 0; v < NV; ++v) s_q[v * NN + a] = q[v];
  if (a < n * n) s_Dh[a] = p.Dhat[a / n][a % n];
  __syncthreads();
  double r[DIM][NV];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double t = 0.0;
#pragma unroll
      for (int m = 0; m < n; ++m) {
#pragma clang fp contract(off)
        t = t + s_Dh[idx[d] * n + m] * s_q[v * NN + a + (m - idx[d]) * stride[d]];
      }
      r[d][v] = t;
    }
  double g[DIM][NV];
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    double ja[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) ja[d] = metric<NODAL, DIM, NN>(p, e, d, c, a);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
      double t = ja[0] * r[0][v];
#pragma unroll
      for (int d = 1; d < DIM; ++d) t = t + ja[d] * r[d][v];
      g[c][v] = t;
    }
  }
  const double ij = inv_jac<NODAL, NN>(p, e, a);
  const double sg = ij >= 0.0 ? 1.0 : -1.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const bool lower = idx[d] == 0, upper = idx[d] == N;
    if (lower || upper) {
      int64_t lo, hi;
      face_neighbours<DIM, false>(p, p.cells[0], p.cells[1], nz, e, d, lo, hi);
      const int64_t ne = lower ? lo : hi;
      const int nn = lower ? a + N * stride[d] : a - N * stride[d];
      double un[NV], qn[NV], nrm[DIM];
#pragma unroll
      for (int v = 0; v < NV; ++v) un[v] = p.u[(ne * NV + v) * NN + nn];
      grad_vars<DIM>(p.gamma, un, qn);
#pragma unroll
      for (int c = 0; c < DIM; ++c) nrm[c] = sg * metric<NODAL, DIM, NN>(p, e, d, c, a);
      const double factor = lower ? -p.lhat0 : p.lhatN;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
        const double fq = factor * (0.5 * (qn[v] + q[v]));
#pragma unroll
        for (int c = 0; c < DIM; ++c) g[c][v] = g[c][v] + fq * nrm[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c)
#pragma unroll
    for (int v = 0; v < NV; ++v) g[c][v] = g[c][v] * ij;
  // the node's gradient is complete: its viscous fluxes follow here, so the
  // gradient itself never goes to memory
  store_viscous_fluxes<DIM, N, NODAL>(p, e, a, q, g);
}

template <int DIM, int N>
constexpr int parabolic_elems_per_block() {
  constexpr int e = 192 / Shape<DIM, N>::LINES;
  return e < 1 ? 1 : e;
}

template <int DIM, int N, bool NODAL, bool BCS>
hipError_t launch_parabolic_t(bool sc, const KParams& p, hipStream_t s) {
  constexpr int EPW = parabolic_elems_per_block<DIM, N>();
  constexpr int BLOCK = EPW * Shape<DIM, N>::LINES;
  const int64_t nb = (p.E + EPW - 1) / EPW;
  if (nb <= 0) return hipSuccess;
  if constexpr (node_kernel_for(N, false, BCS)) {
    if (!sc) {
      hipLaunchKernelGGL((gradient_node_kernel<DIM, N, NODAL>), dim3((unsigned)p.E),
                         dim3(Shape<DIM, N>::NN), 0, s, p);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((gradient_kernel<DIM, N, EPW, NODAL, BCS>), dim3((unsigned)nb), dim3(BLOCK), 0, s, p);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL((viscous_kernel<DIM, N, EPW, NODAL, BCS>), dim3((unsigned)nb), dim3(BLOCK), 0, s, p);
  return hipGetLastError();
}

}  // namespace detail
}  // namespace dgsem
