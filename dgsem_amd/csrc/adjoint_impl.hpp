// Tangent (JVP) and discrete adjoint (VJP) kernels of the DGSEM right-hand
// side L(u) = -(1/J) (A(u) + R(u)): A the split-form volume integral (under
// shock capturing blended with the subcell finite volumes, alpha frozen)
// plus the LLF or HLLC surface flux lifted with l_hat, R the BR1 viscous part
// (Navier-Stokes).  On 2D meshes a face may be a boundary face (neighbour
// entry -1): the outer state, the wall gradient value and the wall viscous
// flux are functions of the own element's nodes, differentiated with them.
// They differentiate exactly what rhs_kernel / gradient_kernel /
// viscous_kernel compute, with the pointwise derivatives of physics_ad.hpp.
//
// Tangent: three kernels (NS) or one (Euler):
//   jvp_grad_kernel   dg = d(grad q)             -> a.dgrad
//   jvp_visc_kernel   dR                         -> a.dvisc
//   jvp_rhs_kernel    dA, then -(1/J)(dA + dR) and the output combination
// Adjoint, with r = -(1/J) w the cotangent of A + R:
//   vjp_visc1_kernel  Rbar -> Hbar (transposed D_hat and interface lifting)
//                     -> Fbar -> (qbar direct, gbar) -> Gbar  -> a.qbar, a.dgrad
//   vjp_visc2_kernel  Gbar -> qbar (transposed D_hat and BR1 face averages,
//                     gathered from the neighbours) -> ubar_visc -> a.dvisc
//   vjp_rhs_kernel    pairs and faces -> primitive / conservative cotangents,
//                     plus ubar_visc, and the output combination
// Every kernel gathers what it needs from the neighbours (no atomics), so
// results do not depend on scheduling.  The Navier-Stokes forward gradient
// of the state must be in a.k.grad.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"
#include "parabolic_impl.hpp"
#include "physics_ad.hpp"

namespace dgsem {
namespace detail {

template <int DIM, int N>
constexpr int ad_elems_per_block() {
  constexpr int e = 192 / Shape<DIM, N>::LINES;
  return e < 1 ? 1 : e;
}

__device__ __forceinline__ double combine(const AdParams& a, int64_t g, double r) {
  double o = a.c * r;
  if (a.b != 0.0) o += a.b * a.v[g];
  if (a.x1) o += a.a1 * a.x1[g];
  if (a.x2) o += a.a2 * a.x2[g];
  return o;
}

template <int DIM, int NN>
__device__ __forceinline__ void load_state(const double* u, int64_t e, int node, double* s) {
  for (int v = 0; v < DIM + 2; ++v) s[v] = u[(e * (DIM + 2) + v) * NN + node];
}

template <bool NODAL, int DIM, int NN>
__device__ __forceinline__ double sign_j(const KParams& p, int64_t e, int node) {
  return inv_jac<NODAL, NN>(p, e, node) >= 0.0 ? 1.0 : -1.0;
}

// Normal of subcell interface f (between nodes f and f + 1) on line `line`
// of direction dir: the nodal subcell normals, or the element's metric row
// (affine), as the forward's subcell flux uses them.
template <int DIM, int N, bool NODAL>
__device__ __forceinline__ void subcell_normal(const KParams& p, int64_t e, int dir, int line, int f,
                                               double* nrm) {
  constexpr int FN = Shape<DIM, N>::FN;
  for (int c = 0; c < DIM; ++c)
    nrm[c] = NODAL ? p.sn[((((e * DIM + dir) * FN) + line) * N + f) * DIM + c]
                   : p.Ja[(e * DIM + dir) * DIM + c];
}

// Frozen blending factor of element e (shock capturing): the split-form
// weight coef = 1 - alpha and the subcell weight alpha of a troubled
// element, coef = 1 and no subcell terms otherwise
// (volume_flux.hpp:843-869 with the atol test of :772-773).
template <bool SC>
__device__ __forceinline__ bool frozen_blend(const KParams& p, int64_t e, double& coef, double& alpha) {
  coef = 1.0;
  alpha = 0.0;
  if constexpr (SC) {
    alpha = p.alpha_in[e];
    if (fabs(alpha - 0.0) > p.atol) {
      coef = 1.0 - alpha;
      return true;
    }
  }
  return false;
}

// Non-periodic faces (2D meshes; the neighbour table holds -1 there).  The
// outer state of face `face` (2 dir + upper) at boundary node bnode is the
// one the forward's boundary_state forms (rhs_kernel_impl.hpp), with the
// line's states read from memory.  As a map of the state it is one of
// three kinds: a slip wall (and the inviscid part of the Navier-Stokes
// walls) mirrors the momentum about the face, linear and symmetric in the
// momentum; outflow takes the mean of the line's n nodes; a Dirichlet state
// (constant or per node) does not depend on the state.
enum class BcMap { kConstant, kMirror, kMean };

template <int DIM>
__device__ __forceinline__ BcMap bc_map(const KParams& p, int face, int64_t bnode) {
  constexpr int NV = DIM + 2;
  const int kind = p.bc[face];
  if (kind == kBcMixedDirichletSlipWall && p.bc_values[face][bnode * (NV + 1) + NV] != 0.0)
    return BcMap::kConstant;
  if (kind == kBcSlipWall || kind == kBcNsNoSlipWall || kind == kBcNsSlipWall ||
      kind == kBcMixedDirichletSlipWall)
    return BcMap::kMirror;
  return kind == kBcOutflow ? BcMap::kMean : BcMap::kConstant;
}

// Boundary node of line `line` of element e on a face of direction dir
// (2D, elements numbered e = ie * ny + je), as the forward numbers it.
template <int DIM, int N>
__device__ __forceinline__ int64_t boundary_node(const KParams& p, int64_t e, int dir, int line) {
  const int64_t along = dir == 0 ? e % p.cells[1] : e / p.cells[1];
  return along * (N + 1) + line;
}

// The outer state; ue: the element's states in memory ([var][node]).
template <int DIM, int N, int NN>
__device__ void bc_outer_state(const KParams& p, int face, int64_t bnode, BcMap m, const double* ue,
                               const int* dofs, const double* uin, const double* nrm, double* uout) {
  constexpr int n = N + 1, NV = DIM + 2;
  if (m == BcMap::kMirror) {
    const double rho = uin[0];
    double v[DIM], vn = 0.0, nn = 0.0;
    for (int c = 0; c < DIM; ++c) v[c] = uin[1 + c] / rho;
    for (int c = 0; c < DIM; ++c) nn += nrm[c] * nrm[c];
    for (int c = 0; c < DIM; ++c) vn += v[c] * nrm[c];
    vn = vn / nn;
    uout[0] = rho;
    for (int c = 0; c < DIM; ++c) uout[1 + c] = rho * (v[c] - 2.0 * vn * nrm[c]);
    uout[NV - 1] = uin[NV - 1];
  } else if (m == BcMap::kMean) {
    for (int v = 0; v < NV; ++v) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += ue[v * NN + dofs[k]];
      uout[v] = acc / (double)n;
    }
  } else {
    const double* vals = p.bc_values[face];
    if (p.bc[face] == kBcMixedDirichletSlipWall) {
      for (int v = 0; v < NV; ++v) uout[v] = vals[bnode * (NV + 1) + v];
    } else {
      for (int v = 0; v < NV; ++v) uout[v] = vals ? vals[bnode * NV + v] : p.bc_state[face][v];
    }
  }
}

// Its tangent: dve the tangent's states of the element, duin the tangent at
// the boundary node.
template <int DIM, int N, int NN>
__device__ void bc_outer_tangent(BcMap m, const double* dve, const int* dofs, const double* duin,
                                 const double* nrm, double* duout) {
  constexpr int n = N + 1, NV = DIM + 2;
  if (m == BcMap::kMirror) {
    double mn = 0.0, nn = 0.0;
    for (int c = 0; c < DIM; ++c) {
      mn += duin[1 + c] * nrm[c];
      nn += nrm[c] * nrm[c];
    }
    mn = mn / nn;
    duout[0] = duin[0];
    for (int c = 0; c < DIM; ++c) duout[1 + c] = duin[1 + c] - 2.0 * mn * nrm[c];
    duout[NV - 1] = duin[NV - 1];
  } else if (m == BcMap::kMean) {
    for (int v = 0; v < NV; ++v) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += dve[v * NN + dofs[k]];
      duout[v] = acc / (double)n;
    }
  } else {
    for (int v = 0; v < NV; ++v) duout[v] = 0.0;
  }
}

// Navier-Stokes faces of bounded meshes (parabolic_impl.hpp).  The gradient's
// face value is the wall map of the own node's gradient variables
// (wall_gradient_value: no slip zeroes the velocity, slip removes its normal
// part; both linear in q and symmetric) or a Dirichlet state's, which is
// constant; this applies the map's derivative (which is its transpose) to x.
template <int DIM>
__device__ __forceinline__ void wall_gradient_linear(int kind, const double* x, const double* nrm,
                                                     double* y) {
  if (kind == kBcDirichlet) {
    for (int v = 0; v < DIM + 2; ++v) y[v] = 0.0;
  } else {
    wall_gradient_value<DIM>(kind, x, nrm, y);
  }
}

// The variables [lo, hi) of the own node's viscous flux a boundary face
// carries: all at a Dirichlet face, the momentum at a no-slip wall, none
// otherwise.
__device__ __forceinline__ void wall_viscous_range(int kind, int nv, int& lo, int& hi) {
  lo = kind == kBcDirichlet ? 0 : 1;
  hi = kind == kBcDirichlet ? nv : (kind == kBcNsNoSlipWall ? nv - 1 : 0);
}

// ---------------------------------------------------------------------------
// Tangent
// ---------------------------------------------------------------------------

template <int DIM, int N, int EPW, bool NODAL>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) jvp_grad_kernel(const AdParams a) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES;
  constexpr int BLOCK = EPW * L;
  const KParams& p = a.k;
  __shared__ double s_dq[EPW * NV * NN];
  __shared__ double s_g[EPW * DIM * NV * NN];
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    double s[NV], ds[NV], dq[NV];
    load_state<DIM, NN>(p.u, e0 + sl, node, s);
    load_state<DIM, NN>(a.v, e0 + sl, node, ds);
    grad_vars_jvp<DIM>(p.gamma, s, ds, dq);
    for (int v = 0; v < NV; ++v) s_dq[(sl * NV + v) * NN + node] = dq[v];
  }
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  if (slot < nvalid) {
    const int64_t e = e0 + slot;
    int dofs[n];
    for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
    const double* dq = s_dq + slot * NV * NN;
    double acc[n][NV];
    for (int k = 0; k < n; ++k)
      for (int v = 0; v < NV; ++v) {
        double r = 0.0;
        for (int i = 0; i < n; ++i) r += p.Dhat[k][i] * dq[v * NN + dofs[i]];
        acc[k][v] = r;
      }
    for (int side = 0; side < 2; ++side) {
      const int64_t nb = p.nbr[e * 2 * DIM + 2 * dir + side];
      if constexpr (DIM == 2) {
        if (nb < 0) {
          // a boundary face: the wall map of the own node's variables
          const int dof_own = side == 0 ? dofs[0] : dofs[n - 1];
          const double sg = sign_j<NODAL, DIM, NN>(p, e, dof_own);
          double x[NV], nrm[DIM], dqw[NV];
          for (int v = 0; v < NV; ++v) x[v] = dq[v * NN + dof_own];
          for (int c = 0; c < DIM; ++c) nrm[c] = sg * metric<NODAL, DIM, NN>(p, e, dir, c, dof_own);
          wall_gradient_linear<DIM>(p.bc[2 * dir + side], x, nrm, dqw);
          if (side == 0) {
            for (int v = 0; v < NV; ++v) acc[0][v] -= sg * p.lhat0 * dqw[v];
          } else {
            for (int v = 0; v < NV; ++v) acc[n - 1][v] += sg * p.lhatN * dqw[v];
          }
          continue;
        }
      }
      const int dof_nb = side == 0 ? dofs[n - 1] : dofs[0];
      double s[NV], ds[NV], dqn[NV];
      load_state<DIM, NN>(p.u, nb, dof_nb, s);
      load_state<DIM, NN>(a.v, nb, dof_nb, ds);
      grad_vars_jvp<DIM>(p.gamma, s, ds, dqn);
      const int dof_own = side == 0 ? dofs[0] : dofs[n - 1];
      const double sg = sign_j<NODAL, DIM, NN>(p, e, dof_own);
      if (side == 0) {
        for (int v = 0; v < NV; ++v) acc[0][v] -= sg * p.lhat0 * (0.5 * (dqn[v] + dq[v * NN + dofs[0]]));
      } else {
        for (int v = 0; v < NV; ++v) acc[n - 1][v] += sg * p.lhatN * (0.5 * (dq[v * NN + dofs[n - 1]] + dqn[v]));
      }
    }
    for (int k = 0; k < n; ++k)
      for (int v = 0; v < NV; ++v) s_g[((slot * DIM + dir) * NV + v) * NN + dofs[k]] = acc[k][v];
  }
  __syncthreads();
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int64_t e = e0 + sl;
    const double ij = inv_jac<NODAL, NN>(p, e, node);
    for (int c = 0; c < DIM; ++c)
      for (int v = 0; v < NV; ++v) {
        double g = 0.0;
        for (int d = 0; d < DIM; ++d)
          g += metric<NODAL, DIM, NN>(p, e, d, c, node) * s_g[((sl * DIM + d) * NV + v) * NN + node];
        a.dgrad[((e * DIM + c) * NV + v) * NN + node] = g * ij;
      }
  }
}

// Tangent of Ja^d . (-F_v) at one node: out[d][v].
template <int DIM, int NN, bool NODAL>
__device__ __forceinline__ void contravariant_viscous_jvp(const AdParams& a, int64_t e, int node,
                                                          double out[DIM][DIM + 2]) {
  constexpr int NV = DIM + 2;
  const KParams& p = a.k;
  double s[NV], ds[NV], q[NV], dq[NV], g[DIM][NV], dg[DIM][NV], df[DIM][NV];
  load_state<DIM, NN>(p.u, e, node, s);
  load_state<DIM, NN>(a.v, e, node, ds);
  grad_vars<DIM>(p.gamma, s, q);
  grad_vars_jvp<DIM>(p.gamma, s, ds, dq);
  for (int c = 0; c < DIM; ++c)
    for (int v = 0; v < NV; ++v) {
      g[c][v] = p.grad[((e * DIM + c) * NV + v) * NN + node];
      dg[c][v] = a.dgrad[((e * DIM + c) * NV + v) * NN + node];
    }
  viscous_flux_jvp<DIM>(p.mu, p.kappa, q, g, dq, dg, df);
  for (int d = 0; d < DIM; ++d)
    for (int v = 0; v < NV; ++v) {
      double r = 0.0;
      for (int c = 0; c < DIM; ++c) r += metric<NODAL, DIM, NN>(p, e, d, c, node) * -df[c][v];
      out[d][v] = r;
    }
}

template <int DIM, int N, int EPW, bool NODAL>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) jvp_visc_kernel(const AdParams a) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES;
  constexpr int BLOCK = EPW * L;
  const KParams& p = a.k;
  __shared__ double s_f[EPW * DIM * NV * NN];
  __shared__ double s_du[EPW * NV * NN];
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    double out[DIM][NV];
    contravariant_viscous_jvp<DIM, NN, NODAL>(a, e0 + sl, node, out);
    for (int d = 0; d < DIM; ++d)
      for (int v = 0; v < NV; ++v) s_f[((sl * DIM + d) * NV + v) * NN + node] = out[d][v];
  }
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  const bool valid = slot < nvalid;
  double acc[n][NV];
  int dofs[n];
  for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
  if (valid) {
    const int64_t e = e0 + slot;
    const double* f = s_f + (slot * DIM + dir) * NV * NN;
    for (int k = 0; k < n; ++k)
      for (int v = 0; v < NV; ++v) {
        double r = 0.0;
        for (int i = 0; i < n; ++i) r += p.Dhat[k][i] * f[v * NN + dofs[i]];
        acc[k][v] = r;
      }
    // lower face: this element is primary; at a boundary face only the
    // variables the condition carries
    int vlo = 0, vhi = NV;
    if constexpr (DIM == 2) {
      if (p.nbr[e * 2 * DIM + 2 * dir] < 0) wall_viscous_range(p.bc[2 * dir], NV, vlo, vhi);
    }
    {
      const double sg = sign_j<NODAL, DIM, NN>(p, e, dofs[0]);
      for (int v = vlo; v < vhi; ++v) acc[0][v] -= sg * f[v * NN + dofs[0]] * p.lhat0;
    }
    const int64_t nb = p.nbr[e * 2 * DIM + 2 * dir + 1];
    bool own_upper = false;  // a boundary upper face: the own node's flux
    if constexpr (DIM == 2) own_upper = nb < 0;
    if (own_upper) {
      int ulo, uhi;
      wall_viscous_range(p.bc[2 * dir + 1], NV, ulo, uhi);
      const double sg = sign_j<NODAL, DIM, NN>(p, e, dofs[n - 1]);
      for (int v = ulo; v < uhi; ++v) acc[n - 1][v] += sg * f[v * NN + dofs[n - 1]] * p.lhatN;
    } else {
      double out[DIM][NV];
      contravariant_viscous_jvp<DIM, NN, NODAL>(a, nb, dofs[0], out);
      const double sg = sign_j<NODAL, DIM, NN>(p, nb, dofs[0]);
      for (int v = 0; v < NV; ++v) acc[n - 1][v] += sg * out[dir][v] * p.lhatN;
    }
  }
  for (int d = 0; d < DIM; ++d) {
    if (valid && dir == d) {
      for (int k = 0; k < n; ++k)
        for (int v = 0; v < NV; ++v) {
          const int idx = (slot * NV + v) * NN + dofs[k];
          s_du[idx] = d == 0 ? acc[k][v] : s_du[idx] + acc[k][v];
        }
    }
    __syncthreads();
  }
  for (int i = tid; i < EPW * NV * NN; i += BLOCK) {
    if (i / (NV * NN) >= nvalid) continue;
    a.dvisc[e0 * NV * NN + i] = s_du[i];
  }
}

template <int DIM, int N, int EPW, bool NODAL, bool NS, bool SC, int SF, int FF>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) jvp_rhs_kernel(const AdParams a) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES, NP = S::NP;
  constexpr int NQ = DIM + 3;
  constexpr int BLOCK = EPW * L;
  const KParams& p = a.k;
  __shared__ double s_p[EPW * NP * NN];  // rho, v, beta, kin, p, c; later the sums
  __shared__ double s_dq[EPW * NQ * NN];
  double* const s_acc = s_p;
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    double s[NV], ds[NV];
    load_state<DIM, NN>(p.u, e0 + sl, node, s);
    load_state<DIM, NN>(a.v, e0 + sl, node, ds);
    const NodePrims<DIM> q = node_prims<DIM>(p.gamma, s);
    PrimTangent<DIM> dq;
    prims_jvp<DIM>(p.gamma, s, q, ds, dq);
    double* d = s_p + sl * NP * NN + node;
    d[0] = q.rho;
    for (int k = 0; k < DIM; ++k) d[(1 + k) * NN] = q.v[k];
    d[(DIM + 1) * NN] = q.beta;
    d[(DIM + 2) * NN] = q.kin;
    d[(DIM + 3) * NN] = q.p;
    d[(DIM + 4) * NN] = q.c;
    double* t = s_dq + sl * NQ * NN + node;
    t[0] = dq.rho;
    for (int k = 0; k < DIM; ++k) t[(1 + k) * NN] = dq.v[k];
    t[(DIM + 1) * NN] = dq.beta;
    t[(DIM + 2) * NN] = dq.kin;
  }
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  const bool valid = slot < nvalid;
  int dofs[n];
  for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
  double acc[n][NV];
  for (int k = 0; k < n; ++k)
    for (int v = 0; v < NV; ++v) acc[k][v] = 0.0;
  if (valid) {
    const int64_t e = e0 + slot;
    const double* pr = s_p + slot * NP * NN;
    const double* tq = s_dq + slot * NQ * NN;
    auto prims = [&](int dof) {
      NodePrims<DIM> q;
      q.rho = pr[dof];
      for (int c = 0; c < DIM; ++c) q.v[c] = pr[(1 + c) * NN + dof];
      q.beta = pr[(DIM + 1) * NN + dof];
      q.kin = pr[(DIM + 2) * NN + dof];
      q.p = pr[(DIM + 3) * NN + dof];
      q.c = pr[(DIM + 4) * NN + dof];
      return q;
    };
    auto tangent = [&](int dof) {
      PrimTangent<DIM> t;
      t.rho = tq[dof];
      for (int c = 0; c < DIM; ++c) t.v[c] = tq[(1 + c) * NN + dof];
      t.beta = tq[(DIM + 1) * NN + dof];
      t.kin = tq[(DIM + 2) * NN + dof];
      return t;
    };
    auto ja = [&](int64_t el, int dof, double* out) {
      for (int c = 0; c < DIM; ++c) out[c] = metric<NODAL, DIM, NN>(p, el, dir, c, dof);
    };
    double coef, alpha;
    const bool troubled = frozen_blend<SC>(p, e, coef, alpha);
#pragma unroll
    for (int ia = 0; ia < n - 1; ++ia) {
      const NodePrims<DIM> qa = prims(dofs[ia]);
      const PrimTangent<DIM> ta = tangent(dofs[ia]);
      double jaa[DIM];
      ja(e, dofs[ia], jaa);
#pragma unroll
      for (int ib = ia + 1; ib < n; ++ib) {
        __builtin_amdgcn_sched_barrier(0);
        const NodePrims<DIM> qb = prims(dofs[ib]);
        const PrimTangent<DIM> tb = tangent(dofs[ib]);
        double jab[DIM], nrm[DIM], df[NV];
        ja(e, dofs[ib], jab);
        for (int c = 0; c < DIM; ++c) nrm[c] = NODAL ? 0.5 * (jaa[c] + jab[c]) : jaa[c];
        chandrashekar_jvp<DIM>(p.inv_gm1, qa, qb, ta, tb, nrm, df);
        const double wa = coef * p.Ds[ia][ib], wb = coef * p.Ds[ib][ia];
        for (int v = 0; v < NV; ++v) {
          acc[ia][v] += wa * df[v];
          acc[ib][v] += wb * df[v];
        }
      }
    }
    // first-order subcell fluxes of a troubled element, weighted by the
    // frozen alpha as the forward's blend weights them.  With LLF and up to
    // N = 5 the loop is unrolled: a run-time index would put the accumulators
    // in scratch memory (at 3D N = 3 the tangent takes 667 instead of 757 us
    // with every element troubled at 32^3).  The HLLC tangent measured 1.5%
    // slower unrolled, and at N = 7 the unrolled fluxes spill more.
    if constexpr (SC) {
      if (troubled) {
        constexpr int FV_UNROLL = N <= 5 && FF != kFluxHLLC ? N : 1;
#pragma unroll FV_UNROLL
        for (int fi = 1; fi < n; ++fi) {
          double ul[NV], ur[NV], dul[NV], dur[NV], nrm[DIM], df[NV];
          load_state<DIM, NN>(p.u, e, dofs[fi - 1], ul);
          load_state<DIM, NN>(p.u, e, dofs[fi], ur);
          load_state<DIM, NN>(a.v, e, dofs[fi - 1], dul);
          load_state<DIM, NN>(a.v, e, dofs[fi], dur);
          subcell_normal<DIM, N, NODAL>(p, e, dir, line, fi - 1, nrm);
          if constexpr (FF == kFluxHLLC)
            hllc_jvp<DIM>(p.gamma, ul, ur, dul, dur, nrm, df);
          else
            llf_jvp<DIM>(p.gamma, ul, node_prims<DIM>(p.gamma, ul), ur, node_prims<DIM>(p.gamma, ur),
                         dul, dur, nrm, df);
          const double wl = alpha * p.inv_w[fi - 1], wr = alpha * p.inv_w[fi];
          for (int v = 0; v < NV; ++v) {
            acc[fi - 1][v] += wl * df[v];
            acc[fi][v] -= wr * df[v];
          }
        }
      }
    }
    // faces, both states and their tangents from memory
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
      const bool lower = side == 0;
      const int64_t nb = p.nbr[e * 2 * DIM + 2 * dir + side];
      double ul[NV], ur[NV], dul[NV], dur[NV], nrm[DIM], df[NV];
      bool boundary = false;  // non-periodic faces exist on 2D meshes only
      if constexpr (DIM == 2) {
        if (nb < 0) {
          // a non-periodic face: the own trace, the normal of the own
          // element at it, and the outer state on the far side
          boundary = true;
          const int od = lower ? dofs[0] : dofs[n - 1];
          double uo[NV], duo[NV], ub[NV], dub[NV];
          load_state<DIM, NN>(p.u, e, od, uo);
          load_state<DIM, NN>(a.v, e, od, duo);
          ja(e, od, nrm);
          const double sg = sign_j<NODAL, DIM, NN>(p, e, od);
          for (int c = 0; c < DIM; ++c) nrm[c] *= sg;
          const int face = 2 * dir + side;
          const int64_t bnode = boundary_node<DIM, N>(p, e, dir, line);
          const BcMap m = bc_map<DIM>(p, face, bnode);
          bc_outer_state<DIM, N, NN>(p, face, bnode, m, p.u + e * NV * NN, dofs, uo, nrm, ub);
          bc_outer_tangent<DIM, N, NN>(m, a.v + e * NV * NN, dofs, duo, nrm, dub);
          for (int v = 0; v < NV; ++v) {
            ul[v] = lower ? ub[v] : uo[v];
            dul[v] = lower ? dub[v] : duo[v];
            ur[v] = lower ? uo[v] : ub[v];
            dur[v] = lower ? duo[v] : dub[v];
          }
        }
      }
      if (!boundary) {
        // the face's lower element contributes its last node, the upper one its first
        const int64_t el = lower ? nb : e, er = lower ? e : nb;
        const int dl = dofs[n - 1], dr = dofs[0];
        load_state<DIM, NN>(p.u, el, dl, ul);
        load_state<DIM, NN>(p.u, er, dr, ur);
        load_state<DIM, NN>(a.v, el, dl, dul);
        load_state<DIM, NN>(a.v, er, dr, dur);
        // the normal of the face's upper element at its first node
        ja(er, dr, nrm);
        const double sg = sign_j<NODAL, DIM, NN>(p, er, dr);
        for (int c = 0; c < DIM; ++c) nrm[c] *= sg;
      }
      if constexpr (SF == kFluxHLLC)
        hllc_jvp<DIM>(p.gamma, ul, ur, dul, dur, nrm, df);
      else
        llf_jvp<DIM>(p.gamma, ul, node_prims<DIM>(p.gamma, ul), ur, node_prims<DIM>(p.gamma, ur), dul,
                     dur, nrm, df);
      if (lower) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[0][v] -= df[v] * p.lhat0;
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[n - 1][v] += df[v] * p.lhatN;
      }
    }
  }
  __syncthreads();  // s_acc aliases the primitives
  for (int d = 0; d < DIM; ++d) {
    if (valid && dir == d) {
      for (int k = 0; k < n; ++k)
        for (int v = 0; v < NV; ++v) {
          const int idx = (slot * NV + v) * NN + dofs[k];
          s_acc[idx] = d == 0 ? acc[k][v] : s_acc[idx] + acc[k][v];
        }
    }
    __syncthreads();
  }
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int64_t e = e0 + sl;
    const double fac = -inv_jac<NODAL, NN>(p, e, node);
    for (int v = 0; v < NV; ++v) {
      const int64_t g = (e * NV + v) * NN + node;
      double r = s_acc[(sl * NV + v) * NN + node];
      if constexpr (NS) r += a.dvisc[g];
      a.out[g] = combine(a, g, r * fac);
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint
// ---------------------------------------------------------------------------

// r = -(1/J) w of element e at node: the cotangent of A + R.
template <int DIM, int NN, bool NODAL>
__device__ __forceinline__ void load_r(const AdParams& a, int64_t e, int node, double* r) {
  const double f = -inv_jac<NODAL, NN>(a.k, e, node);
  for (int v = 0; v < DIM + 2; ++v) r[v] = f * a.v[(e * (DIM + 2) + v) * NN + node];
}

template <int DIM, int N, int EPW, bool NODAL>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) vjp_visc1_kernel(const AdParams a) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES;
  constexpr int BLOCK = EPW * L;
  const KParams& p = a.k;
  __shared__ double s_r[EPW * NV * NN];
  __shared__ double s_hb[EPW * DIM * NV * NN];
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    double r[NV];
    load_r<DIM, NN, NODAL>(a, e0 + sl, node, r);
    for (int v = 0; v < NV; ++v) s_r[(sl * NV + v) * NN + node] = r[v];
  }
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  if (slot < nvalid) {
    const int64_t e = e0 + slot;
    int dofs[n];
    for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
    const double* r = s_r + slot * NV * NN;
    // Hbar_k = sum_k' Dhat[k'][k] r_k'  (transpose of the weak divergence)
    double hb[n][NV];
    for (int k = 0; k < n; ++k)
      for (int v = 0; v < NV; ++v) {
        double h = 0.0;
        for (int i = 0; i < n; ++i) h += p.Dhat[i][k] * r[v * NN + dofs[i]];
        hb[k][v] = h;
      }
    // the first node's value is lifted by this element (-l_hat0, lower face)
    // and by the lower neighbour (+l_hatN, its upper face, whose primary
    // element this is)
    const int64_t lo = p.nbr[e * 2 * DIM + 2 * dir];
    bool bnd_lo = false, bnd_hi = false;
    if constexpr (DIM == 2) {
      bnd_lo = lo < 0;
      bnd_hi = p.nbr[e * 2 * DIM + 2 * dir + 1] < 0;
    }
    if (!bnd_lo) {
      double rl[NV];
      load_r<DIM, NN, NODAL>(a, lo, dofs[n - 1], rl);
      const double sg = sign_j<NODAL, DIM, NN>(p, e, dofs[0]);
      for (int v = 0; v < NV; ++v) hb[0][v] += sg * (p.lhatN * rl[v] - p.lhat0 * r[v * NN + dofs[0]]);
    } else {
      // a boundary lower face lifts the variables its condition carries
      int vlo, vhi;
      wall_viscous_range(p.bc[2 * dir], NV, vlo, vhi);
      const double sg = sign_j<NODAL, DIM, NN>(p, e, dofs[0]);
      for (int v = vlo; v < vhi; ++v) hb[0][v] -= sg * p.lhat0 * r[v * NN + dofs[0]];
    }
    if (bnd_hi) {
      // a boundary upper face lifts the own last node's flux
      int vlo, vhi;
      wall_viscous_range(p.bc[2 * dir + 1], NV, vlo, vhi);
      const double sg = sign_j<NODAL, DIM, NN>(p, e, dofs[n - 1]);
      for (int v = vlo; v < vhi; ++v) hb[n - 1][v] += sg * p.lhatN * r[v * NN + dofs[n - 1]];
    }
    for (int k = 0; k < n; ++k)
      for (int v = 0; v < NV; ++v) s_hb[((slot * DIM + dir) * NV + v) * NN + dofs[k]] = hb[k][v];
  }
  __syncthreads();
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int64_t e = e0 + sl;
    double s[NV], q[NV], g[DIM][NV], fb[DIM][NV], qb[NV], gb[DIM][NV];
    load_state<DIM, NN>(p.u, e, node, s);
    grad_vars<DIM>(p.gamma, s, q);
    for (int c = 0; c < DIM; ++c)
      for (int v = 0; v < NV; ++v) {
        g[c][v] = p.grad[((e * DIM + c) * NV + v) * NN + node];
        gb[c][v] = 0.0;
        double f = 0.0;
        for (int d = 0; d < DIM; ++d)
          f -= metric<NODAL, DIM, NN>(p, e, d, c, node) * s_hb[((sl * DIM + d) * NV + v) * NN + node];
        fb[c][v] = f;
      }
    for (int v = 0; v < NV; ++v) qb[v] = 0.0;
    viscous_flux_vjp<DIM>(p.mu, p.kappa, q, g, fb, qb, gb);
    for (int v = 0; v < NV; ++v) a.qbar[(e * NV + v) * NN + node] = qb[v];
    const double ij = inv_jac<NODAL, NN>(p, e, node);
    for (int d = 0; d < DIM; ++d)
      for (int v = 0; v < NV; ++v) {
        double G = 0.0;
        for (int c = 0; c < DIM; ++c) G += metric<NODAL, DIM, NN>(p, e, d, c, node) * gb[c][v];
        a.dgrad[((e * DIM + d) * NV + v) * NN + node] = ij * G;
      }
  }
}

template <int DIM, int N, int EPW, bool NODAL>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) vjp_visc2_kernel(const AdParams a) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES;
  constexpr int BLOCK = EPW * L;
  const KParams& p = a.k;
  __shared__ double s_gb[EPW * DIM * NV * NN];
  __shared__ double s_qb[EPW * NV * NN];
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < nvalid * DIM * NV * NN; i += BLOCK) s_gb[i] = a.dgrad[e0 * DIM * NV * NN + i];
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  const bool valid = slot < nvalid;
  int dofs[n];
  for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
  double qb[n][NV];
  if (valid) {
    const int64_t e = e0 + slot;
    const double* G = s_gb + (slot * DIM + dir) * NV * NN;
    for (int k = 0; k < n; ++k)
      for (int v = 0; v < NV; ++v) {
        double h = 0.0;
        for (int i = 0; i < n; ++i) h += p.Dhat[i][k] * G[v * NN + dofs[i]];
        qb[k][v] = h;
      }
    // BR1 averages q_hat = (q_L + q_R)/2 on this element's faces and on the
    // neighbours' faces that used this element's face values
    const int64_t lo = p.nbr[e * 2 * DIM + 2 * dir], hi = p.nbr[e * 2 * DIM + 2 * dir + 1];
    const double sg0 = sign_j<NODAL, DIM, NN>(p, e, dofs[0]);
    const double sgn = sign_j<NODAL, DIM, NN>(p, e, dofs[n - 1]);
    bool bnd_lo = false, bnd_hi = false;
    if constexpr (DIM == 2) {
      bnd_lo = lo < 0;
      bnd_hi = hi < 0;
    }
    if (!bnd_lo) {
      const double sg_lo = sign_j<NODAL, DIM, NN>(p, lo, dofs[n - 1]);
      for (int v = 0; v < NV; ++v) {
        const double g_lo = a.dgrad[((lo * DIM + dir) * NV + v) * NN + dofs[n - 1]];
        qb[0][v] += 0.5 * (sg_lo * p.lhatN * g_lo - sg0 * p.lhat0 * G[v * NN + dofs[0]]);
      }
    }
    if (!bnd_hi) {
      const double sg_hi = sign_j<NODAL, DIM, NN>(p, hi, dofs[0]);
      for (int v = 0; v < NV; ++v) {
        const double g_hi = a.dgrad[((hi * DIM + dir) * NV + v) * NN + dofs[0]];
        qb[n - 1][v] += 0.5 * (sgn * p.lhatN * G[v * NN + dofs[n - 1]] - sg_hi * p.lhat0 * g_hi);
      }
    }
    // a boundary face's value is the wall map of the own node's variables:
    // its transpose (the map itself) takes the face term's cotangent back
    if constexpr (DIM == 2) {
      for (int side = 0; side < 2; ++side) {
        if (side == 0 ? !bnd_lo : !bnd_hi) continue;
        const int k = side == 0 ? 0 : n - 1;
        const double sg = side == 0 ? sg0 : sgn;
        double x[NV], nrm[DIM], y[NV];
        for (int v = 0; v < NV; ++v)
          x[v] = side == 0 ? -sg * p.lhat0 * G[v * NN + dofs[0]] : sg * p.lhatN * G[v * NN + dofs[n - 1]];
        for (int c = 0; c < DIM; ++c) nrm[c] = sg * metric<NODAL, DIM, NN>(p, e, dir, c, dofs[k]);
        wall_gradient_linear<DIM>(p.bc[2 * dir + side], x, nrm, y);
        for (int v = 0; v < NV; ++v) qb[k][v] += y[v];
      }
    }
  }
  for (int d = 0; d < DIM; ++d) {
    if (valid && dir == d) {
      for (int k = 0; k < n; ++k)
        for (int v = 0; v < NV; ++v) {
          const int idx = (slot * NV + v) * NN + dofs[k];
          s_qb[idx] = d == 0 ? qb[k][v] : s_qb[idx] + qb[k][v];
        }
    }
    __syncthreads();
  }
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int64_t e = e0 + sl;
    double s[NV], q[NV], ub[NV];
    load_state<DIM, NN>(p.u, e, node, s);
    for (int v = 0; v < NV; ++v) {
      q[v] = s_qb[(sl * NV + v) * NN + node] + a.qbar[(e * NV + v) * NN + node];
      ub[v] = 0.0;
    }
    grad_vars_vjp<DIM>(p.gamma, s, q, ub);
    for (int v = 0; v < NV; ++v) a.dvisc[(e * NV + v) * NN + node] = ub[v];
  }
}

// Whether the adjoint kernel forms the subcell cotangents of all directions
// at once: in 3D up to N = 3, where the per-direction arrays keep the LDS of
// a block within 80 KiB, so the two blocks per CU its registers allow still
// fit.  (With HLLC subcells and every element troubled, 3D N = 3 at 32^3,
// the adjoint took 5.7 ms with one direction at a time.)
template <int DIM, int N, int EPW, bool SC>
constexpr bool par_subcell_adjoint() {
  using S = Shape<DIM, N>;
  constexpr int lds = 8 * EPW * S::NN * (S::NP + 2 * S::NV + DIM + 3 + DIM * S::NV);
  return SC && DIM == 3 && N <= 3 && lds <= 80 * 1024;
}

template <int DIM, int N, int EPW, bool NODAL, bool NS, bool SC, int SF, int FF>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES)) vjp_rhs_kernel(const AdParams a) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES, NP = S::NP;
  constexpr int NQ = DIM + 3;
  constexpr int BLOCK = EPW * L;
  const KParams& p = a.k;
  __shared__ double s_p[EPW * NP * NN];  // rho, v, beta, kin, p, c
  __shared__ double s_r[EPW * NV * NN];
  __shared__ double s_qb[EPW * NQ * NN];
  __shared__ double s_ub[EPW * NV * NN];
  // PAR_FV (shock capturing, where the LDS allows): the troubled elements'
  // subcell cotangents are formed by all directions' lines at once, into
  // one array per direction, instead of one direction at a time.
  constexpr bool PAR_FV = par_subcell_adjoint<DIM, N, EPW, SC>();
  __shared__ double s_fv[PAR_FV ? DIM * EPW * NV * NN : 1];
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    for (int v = 0; v < NV; ++v) s_ub[(sl * NV + v) * NN + node] = 0.0;
    if (sl >= nvalid) continue;
    double s[NV], r[NV];
    load_state<DIM, NN>(p.u, e0 + sl, node, s);
    const NodePrims<DIM> q = node_prims<DIM>(p.gamma, s);
    double* d = s_p + sl * NP * NN + node;
    d[0] = q.rho;
    for (int k = 0; k < DIM; ++k) d[(1 + k) * NN] = q.v[k];
    d[(DIM + 1) * NN] = q.beta;
    d[(DIM + 2) * NN] = q.kin;
    d[(DIM + 3) * NN] = q.p;
    d[(DIM + 4) * NN] = q.c;
    load_r<DIM, NN, NODAL>(a, e0 + sl, node, r);
    for (int v = 0; v < NV; ++v) s_r[(sl * NV + v) * NN + node] = r[v];
  }
  __syncthreads();

  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  const bool valid = slot < nvalid;
  int dofs[n];
  for (int k = 0; k < n; ++k) dofs[k] = line_dof<DIM, N>(dir, line, k);
  PrimTangent<DIM> qb[n];
  for (int k = 0; k < n; ++k) {
    qb[k].rho = qb[k].beta = qb[k].kin = 0.0;
    for (int c = 0; c < DIM; ++c) qb[k].v[c] = 0.0;
  }
  double ub_lo[NV], ub_hi[NV];
  for (int v = 0; v < NV; ++v) ub_lo[v] = ub_hi[v] = 0.0;
  // an outflow face's outer state is the mean of the line: its cotangent
  // goes to every node of the line
  double ub_line[NV];
  bool line_mean = false;
  for (int v = 0; v < NV; ++v) ub_line[v] = 0.0;
  double coef = 1.0, alpha = 0.0;
  const bool troubled = valid && frozen_blend<SC>(p, e0 + slot, coef, alpha);
  if (valid) {
    const int64_t e = e0 + slot;
    const double* pr = s_p + slot * NP * NN;
    const double* rr = s_r + slot * NV * NN;
    auto prims = [&](int dof) {
      NodePrims<DIM> q;
      q.rho = pr[dof];
      for (int c = 0; c < DIM; ++c) q.v[c] = pr[(1 + c) * NN + dof];
      q.beta = pr[(DIM + 1) * NN + dof];
      q.kin = pr[(DIM + 2) * NN + dof];
      q.p = pr[(DIM + 3) * NN + dof];
      q.c = pr[(DIM + 4) * NN + dof];
      return q;
    };
    auto ja = [&](int64_t el, int dof, double* out) {
      for (int c = 0; c < DIM; ++c) out[c] = metric<NODAL, DIM, NN>(p, el, dir, c, dof);
    };
#pragma unroll
    for (int ia = 0; ia < n - 1; ++ia) {
      const NodePrims<DIM> qa = prims(dofs[ia]);
      double jaa[DIM];
      ja(e, dofs[ia], jaa);
#pragma unroll
      for (int ib = ia + 1; ib < n; ++ib) {
        // one pair at a time: interleaving pairs would keep several
        // adjoint flux evaluations live at once
        __builtin_amdgcn_sched_barrier(0);
        const NodePrims<DIM> qbn = prims(dofs[ib]);
        double jab[DIM], nrm[DIM], g[NV];
        ja(e, dofs[ib], jab);
        for (int c = 0; c < DIM; ++c) nrm[c] = NODAL ? 0.5 * (jaa[c] + jab[c]) : jaa[c];
        const double wa = coef * p.Ds[ia][ib], wb = coef * p.Ds[ib][ia];
        for (int v = 0; v < NV; ++v) g[v] = wa * rr[v * NN + dofs[ia]] + wb * rr[v * NN + dofs[ib]];
        chandrashekar_vjp<DIM>(p.inv_gm1, qa, qbn, nrm, g, qb[ia], qb[ib]);
      }
    }
    // faces: this element's share of each face flux's adjoint
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
      const bool lower = side == 0;
      const int64_t nb = p.nbr[e * 2 * DIM + 2 * dir + side];
      if (DIM == 2 && nb < 0) {
        // a non-periodic face (2D meshes only): the flux enters this element
        // alone, and the outer state's cotangent maps back through the
        // boundary condition
        const int od = lower ? dofs[0] : dofs[n - 1];
        double uo[NV], ro[NV], uob[NV], ul[NV], ur[NV], nrm[DIM], g[NV], ubl[NV], ubr[NV];
        load_state<DIM, NN>(p.u, e, od, uo);
        load_r<DIM, NN, NODAL>(a, e, od, ro);
        ja(e, od, nrm);
        const double sg = sign_j<NODAL, DIM, NN>(p, e, od);
        for (int c = 0; c < DIM; ++c) nrm[c] *= sg;
        const int face = 2 * dir + side;
        const int64_t bnode = boundary_node<DIM, N>(p, e, dir, line);
        const BcMap m = bc_map<DIM>(p, face, bnode);
        bc_outer_state<DIM, N, NN>(p, face, bnode, m, p.u + e * NV * NN, dofs, uo, nrm, uob);
        for (int v = 0; v < NV; ++v) {
          g[v] = lower ? -p.lhat0 * ro[v] : p.lhatN * ro[v];
          ul[v] = lower ? uob[v] : uo[v];
          ur[v] = lower ? uo[v] : uob[v];
          ubl[v] = ubr[v] = 0.0;
        }
        if constexpr (SF == kFluxHLLC) {
          hllc_vjp<DIM>(p.gamma, ul, ur, nrm, g, true, true, ubl, ubr);
        } else {
          const NodePrims<DIM> ql = node_prims<DIM>(p.gamma, ul), qr = node_prims<DIM>(p.gamma, ur);
          llf_vjp_side<DIM>(p.gamma, ul, ql, ur, qr, nrm, g, true, ubl);
          llf_vjp_side<DIM>(p.gamma, ul, ql, ur, qr, nrm, g, false, ubr);
        }
        double* own = lower ? ub_lo : ub_hi;
        const double* outer = lower ? ubl : ubr;
        for (int v = 0; v < NV; ++v) own[v] = lower ? ubr[v] : ubl[v];
        if (m == BcMap::kMirror) {
          // the mirror is symmetric: its transpose is itself
          double mn = 0.0, nn = 0.0;
          for (int c = 0; c < DIM; ++c) {
            mn += outer[1 + c] * nrm[c];
            nn += nrm[c] * nrm[c];
          }
          mn = mn / nn;
          own[0] += outer[0];
          for (int c = 0; c < DIM; ++c) own[1 + c] += outer[1 + c] - 2.0 * mn * nrm[c];
          own[NV - 1] += outer[NV - 1];
        } else if (m == BcMap::kMean) {
          line_mean = true;
          for (int v = 0; v < NV; ++v) ub_line[v] += outer[v] / (double)n;
        }
        continue;
      }
      const int64_t el = lower ? nb : e, er = lower ? e : nb;
      const int dl = dofs[n - 1], dr = dofs[0];
      double ul[NV], ur[NV], nrm[DIM], g[NV], rl[NV], rrv[NV];
      load_state<DIM, NN>(p.u, el, dl, ul);
      load_state<DIM, NN>(p.u, er, dr, ur);
      load_r<DIM, NN, NODAL>(a, el, dl, rl);
      load_r<DIM, NN, NODAL>(a, er, dr, rrv);
      ja(er, dr, nrm);
      const double sg = sign_j<NODAL, DIM, NN>(p, er, dr);
      for (int c = 0; c < DIM; ++c) nrm[c] *= sg;
      // the face flux enters the upper element's first node with -l_hat0
      // and the lower element's last node with +l_hatN
      for (int v = 0; v < NV; ++v) g[v] = p.lhatN * rl[v] - p.lhat0 * rrv[v];
      double ubf[NV];
      for (int v = 0; v < NV; ++v) ubf[v] = 0.0;
      // this element is the face's right state at its lower face
      if constexpr (SF == kFluxHLLC)
        hllc_vjp<DIM>(p.gamma, ul, ur, nrm, g, !lower, lower, ubf, ubf);
      else
        llf_vjp_side<DIM>(p.gamma, ul, node_prims<DIM>(p.gamma, ul), ur, node_prims<DIM>(p.gamma, ur),
                          nrm, g, !lower, ubf);
      if (lower) {
        for (int v = 0; v < NV; ++v) ub_lo[v] = ubf[v];
      } else {
        for (int v = 0; v < NV; ++v) ub_hi[v] = ubf[v];
      }
    }
  }
  for (int d = 0; d < DIM; ++d) {
    if (valid && dir == d) {
      for (int k = 0; k < n; ++k) {
        double* t = s_qb + slot * NQ * NN + dofs[k];
        if (d == 0) {
          t[0] = qb[k].rho;
          for (int c = 0; c < DIM; ++c) t[(1 + c) * NN] = qb[k].v[c];
          t[(DIM + 1) * NN] = qb[k].beta;
          t[(DIM + 2) * NN] = qb[k].kin;
        } else {
          t[0] += qb[k].rho;
          for (int c = 0; c < DIM; ++c) t[(1 + c) * NN] += qb[k].v[c];
          t[(DIM + 1) * NN] += qb[k].beta;
          t[(DIM + 2) * NN] += qb[k].kin;
        }
      }
      for (int v = 0; v < NV; ++v) {
        s_ub[(slot * NV + v) * NN + dofs[0]] += ub_lo[v];
        s_ub[(slot * NV + v) * NN + dofs[n - 1]] += ub_hi[v];
      }
      if (line_mean) {
        for (int k = 0; k < n; ++k)
          for (int v = 0; v < NV; ++v) s_ub[(slot * NV + v) * NN + dofs[k]] += ub_line[v];
      }
      // subcell fluxes of a troubled element: flux fi enters node fi - 1
      // with alpha w_inv(fi - 1) and node fi with -alpha w_inv(fi); each line
      // owns its nodes within this direction, so the sums need no atomics
      if constexpr (SC && !PAR_FV) {
        if (troubled) {
          const int64_t e = e0 + slot;
          const double* rr = s_r + slot * NV * NN;
#pragma unroll 1
          for (int fi = 1; fi < n; ++fi) {
            double ul[NV], ur[NV], nrm[DIM], g[NV], ubl[NV], ubr[NV];
            load_state<DIM, NN>(p.u, e, dofs[fi - 1], ul);
            load_state<DIM, NN>(p.u, e, dofs[fi], ur);
            subcell_normal<DIM, N, NODAL>(p, e, dir, line, fi - 1, nrm);
            const double wl = alpha * p.inv_w[fi - 1], wr = alpha * p.inv_w[fi];
            for (int v = 0; v < NV; ++v) {
              g[v] = wl * rr[v * NN + dofs[fi - 1]] - wr * rr[v * NN + dofs[fi]];
              ubl[v] = ubr[v] = 0.0;
            }
            if constexpr (FF == kFluxHLLC) {
              hllc_vjp<DIM>(p.gamma, ul, ur, nrm, g, true, true, ubl, ubr);
            } else {
              const NodePrims<DIM> ql = node_prims<DIM>(p.gamma, ul), qr = node_prims<DIM>(p.gamma, ur);
              llf_vjp_side<DIM>(p.gamma, ul, ql, ur, qr, nrm, g, true, ubl);
              llf_vjp_side<DIM>(p.gamma, ul, ql, ur, qr, nrm, g, false, ubr);
            }
            for (int v = 0; v < NV; ++v) {
              s_ub[(slot * NV + v) * NN + dofs[fi - 1]] += ubl[v];
              s_ub[(slot * NV + v) * NN + dofs[fi]] += ubr[v];
            }
          }
        }
      }
    }
    __syncthreads();
  }
  // PAR_FV: the same subcell terms, all directions at once now that the
  // line's other sums are in LDS.  Each direction has its own array, where
  // the line owns its nodes: node fi - 1 gets its second term here, node fi
  // its first.
  if constexpr (PAR_FV) {
    if (troubled) {
      const int64_t e = e0 + slot;
      const double* rr = s_r + slot * NV * NN;
      double* fvp = s_fv + (dir * EPW + slot) * NV * NN;
#pragma unroll 1
      for (int fi = 1; fi < n; ++fi) {
        double ul[NV], ur[NV], nrm[DIM], g[NV], ubl[NV], ubr[NV];
        load_state<DIM, NN>(p.u, e, dofs[fi - 1], ul);
        load_state<DIM, NN>(p.u, e, dofs[fi], ur);
        subcell_normal<DIM, N, NODAL>(p, e, dir, line, fi - 1, nrm);
        const double wl = alpha * p.inv_w[fi - 1], wr = alpha * p.inv_w[fi];
        for (int v = 0; v < NV; ++v) {
          g[v] = wl * rr[v * NN + dofs[fi - 1]] - wr * rr[v * NN + dofs[fi]];
          ubl[v] = ubr[v] = 0.0;
        }
        if constexpr (FF == kFluxHLLC) {
          hllc_vjp<DIM>(p.gamma, ul, ur, nrm, g, true, true, ubl, ubr);
        } else {
          const NodePrims<DIM> ql = node_prims<DIM>(p.gamma, ul), qr = node_prims<DIM>(p.gamma, ur);
          llf_vjp_side<DIM>(p.gamma, ul, ql, ur, qr, nrm, g, true, ubl);
          llf_vjp_side<DIM>(p.gamma, ul, ql, ur, qr, nrm, g, false, ubr);
        }
        for (int v = 0; v < NV; ++v) {
          double* lo = fvp + v * NN + dofs[fi - 1];
          *lo = fi == 1 ? ubl[v] : *lo + ubl[v];
          fvp[v * NN + dofs[fi]] = ubr[v];
        }
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int64_t e = e0 + sl;
    const double* pr = s_p + sl * NP * NN + node;
    NodePrims<DIM> q;
    q.rho = pr[0];
    for (int c = 0; c < DIM; ++c) q.v[c] = pr[(1 + c) * NN];
    q.beta = pr[(DIM + 1) * NN];
    q.kin = pr[(DIM + 2) * NN];
    q.p = pr[(DIM + 3) * NN];
    q.c = pr[(DIM + 4) * NN];
    const double* t = s_qb + sl * NQ * NN + node;
    PrimTangent<DIM> qbar;
    qbar.rho = t[0];
    for (int c = 0; c < DIM; ++c) qbar.v[c] = t[(1 + c) * NN];
    qbar.beta = t[(DIM + 1) * NN];
    qbar.kin = t[(DIM + 2) * NN];
    double ub[NV];
    for (int v = 0; v < NV; ++v) ub[v] = s_ub[(sl * NV + v) * NN + node];
    if constexpr (PAR_FV) {
      double cf, al;
      if (frozen_blend<SC>(p, e, cf, al)) {
        for (int d = 0; d < DIM; ++d)
          for (int v = 0; v < NV; ++v) ub[v] += s_fv[((d * EPW + sl) * NV + v) * NN + node];
      }
    }
    prims_vjp<DIM>(p.gamma, q, qbar, 0.0, ub);
    for (int v = 0; v < NV; ++v) {
      const int64_t g = (e * NV + v) * NN + node;
      double r = ub[v];
      if constexpr (NS) r += a.dvisc[g];
      a.out[g] = combine(a, g, r);
    }
  }
}

// The Navier-Stokes derivatives run the viscous kernels first (they form
// a.dvisc, which the RHS kernels add); every kernel takes the flux kinds as
// template arguments.
template <int DIM, int N, bool NODAL, bool SC, bool ADJ>
hipError_t launch_ad_t(bool ns, const AdParams& a, hipStream_t s) {
  constexpr int EPW = ad_elems_per_block<DIM, N>();
  constexpr int BLOCK = EPW * Shape<DIM, N>::LINES;
  const int64_t nb = (a.k.E + EPW - 1) / EPW;
  if (nb <= 0) return hipSuccess;
  const dim3 grid((unsigned)nb), block(BLOCK);
  if (ns) {
    if constexpr (ADJ) {
      hipLaunchKernelGGL((vjp_visc1_kernel<DIM, N, EPW, NODAL>), grid, block, 0, s, a);
      hipLaunchKernelGGL((vjp_visc2_kernel<DIM, N, EPW, NODAL>), grid, block, 0, s, a);
    } else {
      hipLaunchKernelGGL((jvp_grad_kernel<DIM, N, EPW, NODAL>), grid, block, 0, s, a);
      hipLaunchKernelGGL((jvp_visc_kernel<DIM, N, EPW, NODAL>), grid, block, 0, s, a);
    }
  }
  const bool shllc = a.k.surface_flux == kFluxHLLC, fhllc = SC && a.k.fv_flux == kFluxHLLC;
  auto go = [&](auto nsv, auto sf, auto ff) {
    constexpr bool NSv = decltype(nsv)::value;
    constexpr int SFv = decltype(sf)::value, FFv = decltype(ff)::value;
    if constexpr (ADJ)
      hipLaunchKernelGGL((vjp_rhs_kernel<DIM, N, EPW, NODAL, NSv, SC, SFv, FFv>), grid, block, 0, s,
                         a);
    else
      hipLaunchKernelGGL((jvp_rhs_kernel<DIM, N, EPW, NODAL, NSv, SC, SFv, FFv>), grid, block, 0, s,
                         a);
    return hipGetLastError();
  };
  auto flux = [&](auto nsv) {
    using LLF = std::integral_constant<int, kFluxLLF>;
    using HLLC = std::integral_constant<int, kFluxHLLC>;
    if constexpr (SC) {
      if (shllc) return fhllc ? go(nsv, HLLC{}, HLLC{}) : go(nsv, HLLC{}, LLF{});
      return fhllc ? go(nsv, LLF{}, HLLC{}) : go(nsv, LLF{}, LLF{});
    } else {
      return shllc ? go(nsv, HLLC{}, LLF{}) : go(nsv, LLF{}, LLF{});
    }
  };
  return ns ? flux(std::true_type{}) : flux(std::false_type{});
}

template <int DIM, int N, bool NODAL, bool SC>
hipError_t launch_jvp_t(bool ns, const AdParams& a, hipStream_t s) {
  return launch_ad_t<DIM, N, NODAL, SC, false>(ns, a, s);
}

template <int DIM, int N, bool NODAL, bool SC>
hipError_t launch_vjp_t(bool ns, const AdParams& a, hipStream_t s) {
  return launch_ad_t<DIM, N, NODAL, SC, true>(ns, a, s);
}

template <int DIM, int N, bool NODAL>
hipError_t launch_gradient_t(const KParams& p, hipStream_t s) {
  constexpr int EPW = parabolic_elems_per_block<DIM, N>();
  constexpr int BLOCK = EPW * Shape<DIM, N>::LINES;
  const int64_t nb = (p.E + EPW - 1) / EPW;
  if (nb <= 0) return hipSuccess;
  // 2D meshes may have boundary faces (a neighbour entry of -1): the kernel's
  // boundary branch forms their face values, as the derivative kernels do
  hipLaunchKernelGGL((gradient_kernel<DIM, N, EPW, NODAL, DIM == 2>), dim3((unsigned)nb), dim3(BLOCK), 0, s, p);
  return hipGetLastError();
}

}  // namespace detail
}  // namespace dgsem
