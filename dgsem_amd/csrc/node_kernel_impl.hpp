// Node-parallel variant of the fused RHS / SSPRK3 stage kernel for
// split-form elements of high polynomial degree, 2D and 3D (periodic
// meshes, no shock capturing).
//
// The line-thread kernel (rhs_kernel) keeps all N+1 node sums of a line in
// registers; at N = 7 in 3D that is 40 accumulators plus the line's metric
// rows, which takes the kernel past 400 registers and down to one wave per
// SIMD.
// Here one workgroup owns one element and one thread owns one node: it
// walks the lines through its node (three in 3D, two in 2D) and sums
// D_split(x, m) f(u_x, u_m)
// over the partners m, evaluating every pair flux itself.  Each pair is thus
// evaluated twice, once from either end (the flux is symmetric in its two
// nodes to the bit, so both ends agree and conservation is exact), but a
// thread holds only its node's five sums and one partner, so the kernel
// fits four waves per SIMD with two elements per CU, and there is no
// cross-thread accumulation: the node's result goes straight from registers
// to memory.
//
// At N = 7 the split-form sums cancel terms some thousand times larger than
// their result, so the RHS carries the last bits of every term and the
// order of the sums at the 1e-12 level.  This kernel therefore takes each
// node's terms in the reference's order and rounds every operation of the
// pair flux, the face flux and the sums as its scalar code does (true
// quotients, no fused multiply-adds).  Against a scalar restatement of the
// reference's algorithm this kernel stays within a few 1e-13 on the hardest
// small case (TGV, 3x2x4 elements), where the line kernel's faster
// arithmetic is at 2e-12.
//
// Reference kernels replaced: the same as rhs_kernel's split-form path
// (volume_flux.hpp:221-410, surface_flux.hpp:45-271,
// surface_integral.hpp:35-154, jacobian.hpp:22-105, SSPRK3.hpp:12-143).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "physics.hpp"

namespace dgsem {
namespace detail {

template <int DIM, int N, bool SC, bool BCS>
constexpr bool use_node_kernel() {
  return node_kernel_for(N, SC, BCS);
}

// The functions below round every operation as the reference's scalar code
// does (numerical_flux.hpp:520-594, compressible_euler3D.hpp:26-166,
// math_utils.hpp:8-24): true quotients and no fused multiply-adds.  At high
// degree the split-form sums cancel terms some thousand times larger than
// their result, so the last bits of each pair flux and the order of the sums
// show in the RHS at the 1e-12 level; with the reference's roundings and its
// summation order the kernel reproduces its results.

// a / b rounded to nearest for positive normal operands, in fewer
// instructions than the IEEE division sequence (which also scales for
// denormal and overflowing operands): the hardware reciprocal estimate with
// two Newton steps, the quotient, and one correction by its exact residual
// (Markstein's method).  Equal to a / b on 2e8 random operand pairs.
__device__ __forceinline__ double div_rn(double a, double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  const double q = a * y;
  return fma(fma(-b, q, a), y, q);
}

// Sums below run left to right from their first term, as the reference's
// written-out expressions do ((a + b) + c), in 2D and 3D alike.

// rho, v, beta = rho / (2 p), |v|^2 / 2, p and c of a conservative state.
template <int DIM>
__device__ __forceinline__ NodePrims<DIM> node_prims_exact(double g, const double* u) {
#pragma clang fp contract(off)
  NodePrims<DIM> q;
  q.rho = u[0];
#pragma unroll
  for (int k = 0; k < DIM; ++k) q.v[k] = u[k + 1] / u[0];
  double v2 = q.v[0] * q.v[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) v2 = v2 + q.v[k] * q.v[k];
  q.kin = 0.5 * v2;
  q.p = (g - 1.0) * (u[DIM + 1] - 0.5 * u[0] * v2);
  q.beta = 0.5 * u[0] / q.p;
  q.c = sqrt(g * q.p / u[0]);
  return q;
}

// Chandrashekar flux, symmetric in (a, b) to the bit: the series test uses
// ((x - y) / (x + y))^2 (whose rounding cannot reach the polynomial, which
// is 2 to within 1e-4), and the log branch orders its arguments by value.
template <int DIM>
__device__ __forceinline__ void flux_chandrashekar_exact(double inv_gm1, const NodePrims<DIM>& a,
                                                         const NodePrims<DIM>& b, const double* n,
                                                         double* f) {
#pragma clang fp contract(off)
  double vna = a.v[0] * n[0], vnb = b.v[0] * n[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) {
    vna = vna + a.v[k] * n[k];
    vnb = vnb + b.v[k] * n[k];
  }
  const double sr = a.rho + b.rho, sb = a.beta + b.beta;
  const double xr = a.rho - b.rho, xb = a.beta - b.beta;
  const double r1 = recip((sr * sr) * (sb * sb));
  const double f2r = (xr * xr) * ((sb * sb) * r1), f2b = (xb * xb) * ((sr * sr) * r1);
  double rho_mean, beta_mean;
  if (f2r < 1.0e-4) {
    rho_mean = div_rn(sr, 2.0 + f2r * (2.0 / 3.0 + f2r * (2.0 / 5.0 + f2r * (2.0 / 7.0))));
  } else {
    const double lo = dmin(a.rho, b.rho), hi = dmax(a.rho, b.rho);
    rho_mean = (hi - lo) / log(hi / lo);
  }
  if (f2b < 1.0e-4) {
    beta_mean = div_rn(sb, 2.0 + f2b * (2.0 / 3.0 + f2b * (2.0 / 5.0 + f2b * (2.0 / 7.0))));
  } else {
    const double lo = dmin(a.beta, b.beta), hi = dmax(a.beta, b.beta);
    beta_mean = (hi - lo) / log(hi / lo);
  }
  const double p_mean = div_rn(0.5 * sr, sb);
  const double vsq = a.kin + b.kin;
  const double f1 = rho_mean * 0.5 * (vna + vnb);
  double fe = f1 * 0.5 * (div_rn(inv_gm1, beta_mean) - vsq);
  f[0] = f1;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const double vavg = 0.5 * (a.v[k] + b.v[k]);
    f[k + 1] = f1 * vavg + p_mean * n[k];
    fe = fe + f[k + 1] * vavg;
  }
  f[DIM + 1] = fe;
}

// n . (F_1, ..., F_DIM)(u) (compressible_euler2D.hpp:25-57,
// compressible_euler3D.hpp:26-67) and the state's pressure.
template <int DIM>
__device__ __forceinline__ double physical_flux_exact(double g, const double* u, const double* n,
                                                      double* f) {
#pragma clang fp contract(off)
  constexpr int NV = DIM + 2;
  double v[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) v[k] = u[k + 1] / u[0];
  double v2 = v[0] * v[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) v2 = v2 + v[k] * v[k];
  const double p = (g - 1.0) * (u[NV - 1] - 0.5 * u[0] * v2);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double fd[NV];
    fd[0] = u[d + 1];
#pragma unroll
    for (int k = 0; k < DIM; ++k) fd[k + 1] = k == d ? u[d + 1] * v[k] + p : u[d + 1] * v[k];
    fd[NV - 1] = v[d] * (u[NV - 1] + p);
#pragma unroll
    for (int i = 0; i < NV; ++i) f[i] = d == 0 ? n[0] * fd[i] : f[i] + n[d] * fd[i];
  }
  return p;
}

// Local Lax-Friedrichs flux from conservative states (numerical_flux.hpp:54-68
// with compressible_euler{2D,3D}.hpp get_wave_speed).
template <int DIM>
__device__ __forceinline__ void flux_llf_exact(double g, const double* ul, const double* ur,
                                               const double* n, double* f) {
#pragma clang fp contract(off)
  constexpr int NV = DIM + 2;
  double fl[NV], fr[NV], vn[2], a[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double* u = s == 0 ? ul : ur;
    const double p = physical_flux_exact<DIM>(g, u, n, s == 0 ? fl : fr);
    a[s] = sqrt(g * p / u[0]);
    double vnn = (u[1] / u[0]) * n[0];
#pragma unroll
    for (int k = 1; k < DIM; ++k) vnn = vnn + (u[k + 1] / u[0]) * n[k];
    vn[s] = vnn;
  }
  double nn2 = n[0] * n[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) nn2 = nn2 + n[k] * n[k];
  const double lam = dmax(fabs(vn[0]), fabs(vn[1])) + dmax(a[0], a[1]) * sqrt(nn2);
#pragma unroll
  for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fl[v] + fr[v]) - 0.5 * lam * (ur[v] - ul[v]);
}

// HLLC flux with a normal vector (a build extension: the reference has
// HLLC for 1D Euler only, numerical_flux.hpp:283-378, whose structure this
// follows in the frame of the unit normal, scaled by |n|).
template <int DIM>
__device__ __forceinline__ void flux_hllc_exact(double g, const double* ul, const double* ur,
                                                const double* n, double* f) {
#pragma clang fp contract(off)
  constexpr int NV = DIM + 2;
  double nh[DIM], vl[DIM], vr[DIM], fl[NV], fr[NV];
  double nn2 = n[0] * n[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) nn2 = nn2 + n[k] * n[k];
  const double nrm = sqrt(nn2);
#pragma unroll
  for (int d = 0; d < DIM; ++d) nh[d] = n[d] / nrm;
  const double rho_ll = ul[0], rho_rr = ur[0], rhoE_ll = ul[NV - 1], rhoE_rr = ur[NV - 1];
  double v2l = 0.0, v2r = 0.0, vn_l = 0.0, vn_r = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    vl[d] = ul[d + 1] / rho_ll;
    vr[d] = ur[d + 1] / rho_rr;
    v2l += vl[d] * vl[d];
    v2r += vr[d] * vr[d];
    vn_l += vl[d] * nh[d];
    vn_r += vr[d] * nh[d];
  }
  const double E_ll = rhoE_ll / rho_ll, E_rr = rhoE_rr / rho_rr;
  const double p_ll = (g - 1.0) * (rhoE_ll - 0.5 * rho_ll * v2l);
  const double p_rr = (g - 1.0) * (rhoE_rr - 0.5 * rho_rr * v2r);
  const double c_ll = sqrt(g * p_ll / rho_ll), c_rr = sqrt(g * p_rr / rho_rr);
  physical_flux_exact<DIM>(g, ul, nh, fl);
  physical_flux_exact<DIM>(g, ur, nh, fr);
  const double srl = sqrt(rho_ll), srr = sqrt(rho_rr), ssum = srl + srr;
  double vroe2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double vroe_d = (srl * vl[d] + srr * vr[d]) / ssum;
    vroe2 += vroe_d * vroe_d;
  }
  const double vel_roe = (srl * vn_l + srr * vn_r) / ssum;
  const double ekin_roe = 0.5 * vroe2;
  const double H_ll = (rhoE_ll + p_ll) / rho_ll, H_rr = (rhoE_rr + p_rr) / rho_rr;
  const double H_roe = (srl * H_ll + srr * H_rr) / ssum;
  const double c_roe = sqrt((g - 1.0) * (H_roe - ekin_roe));
  const double Ssl = dmin(vn_l - c_ll, vel_roe - c_roe);
  const double Ssr = dmax(vn_r + c_rr, vel_roe + c_roe);
  const double sMu_L = Ssl - vn_l, sMu_R = Ssr - vn_r;
  if (Ssl >= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) f[i] = fl[i];
  } else if (Ssr <= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) f[i] = fr[i];
  } else {
    const double SStar = (p_rr - p_ll + rho_ll * vn_l * sMu_L - rho_rr * vn_r * sMu_R) /
                         (rho_ll * sMu_L - rho_rr * sMu_R);
    const bool left = SStar >= 0;
    const double* uk = left ? ul : ur;
    const double* fk = left ? fl : fr;
    const double* vk = left ? vl : vr;
    const double rk = left ? rho_ll : rho_rr, vnk = left ? vn_l : vn_r;
    const double Ek = left ? E_ll : E_rr, pk = left ? p_ll : p_rr;
    const double Sk = left ? Ssl : Ssr, sMu = left ? sMu_L : sMu_R;
    const double densStar = rk * sMu / (Sk - SStar);
    const double enerStar = Ek + (SStar - vnk) * (SStar + pk / (rk * sMu));
    f[0] = fk[0] + Sk * (densStar - uk[0]);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double mstar = densStar * (vk[d] + (SStar - vnk) * nh[d]);
      f[d + 1] = fk[d + 1] + Sk * (mstar - uk[d + 1]);
    }
    f[NV - 1] = fk[NV - 1] + Sk * (densStar * enerStar - uk[NV - 1]);
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) f[i] *= nrm;
}

template <int DIM, int N, bool NODAL, int SF>
__global__ void __launch_bounds__((Shape<DIM, N>::NN)) __attribute__((amdgpu_waves_per_eu(4)))
    rhs_node_kernel(const KParams p) {
  constexpr int n = N + 1, NV = DIM + 2, NN = Shape<DIM, N>::NN;
  constexpr int PN = NN + NN / n;  // padded plane (lds_pos)
  constexpr int DS = n + 1;        // padded row of the D_split table
  constexpr int QB = DIM + 1;      // plane of beta; kin follows
  __shared__ double s_q[(DIM + 3) * PN];             // rho, v[DIM], beta, kin
  __shared__ double s_ja[NODAL ? DIM * DIM * PN : 1];  // Ja^d_c / 2
  __shared__ double s_D[n * DS];
  __shared__ double s_Dh[n * DS];  // D_hat (Navier-Stokes viscous terms)

  const int a = threadIdx.x;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x);
  const int idx[3] = {a % n, (a / n) % n, DIM == 3 ? a / (n * n) : 0};
  constexpr int stride[3] = {1, n, n * n};
  const int pa = lds_pos<N>(a);
  const int nz = DIM == 3 ? p.cells[2] : 1;

  // 1. the node's state, metric rows and 1/J (the face data follows the
  // term loops, phase 4)
  double u[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) u[v] = p.u[(e * NV + v) * NN + a];
  double jm[DIM][DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int c = 0; c < DIM; ++c)
      jm[d][c] = NODAL ? p.Ja[((e * DIM + d) * DIM + c) * NN + a] : p.Ja[(e * DIM + d) * DIM + c];
  const double ij = NODAL ? p.invJ[e * NN + a] : p.invJ[e];

  // 2. the node's primitives and halved metric rows into LDS
  const NodePrims<DIM> q = node_prims_exact<DIM>(p.gamma, u);
  s_q[pa] = q.rho;
#pragma unroll
  for (int c = 0; c < DIM; ++c) s_q[(1 + c) * PN + pa] = q.v[c];
  s_q[QB * PN + pa] = q.beta;
  s_q[(QB + 1) * PN + pa] = q.kin;
  if constexpr (NODAL) {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
      for (int c = 0; c < DIM; ++c) s_ja[(d * DIM + c) * PN + pa] = 0.5 * jm[d][c];
  }
  if (a < n * n) {
    s_D[(a / n) * DS + a % n] = p.Ds[a / n][a % n];
    s_Dh[(a / n) * DS + a % n] = p.Dhat[a / n][a % n];
  }
  __syncthreads();

  // 3. flux differencing.  The reference (volume_flux.hpp:296-410) visits
  // the nodes in memory order and adds each pair's two terms as it goes, so
  // a node's sum takes its terms in the order: partners below it on its
  // z-line, y-line and x-line, then partners above it on its x-, y- and
  // z-lines, each line in increasing position.  Every node has 3N terms, so
  // each lane walks its own sequence in one uniform loop, and the face terms
  // follow in direction order (surface_integral.hpp:82-154).
  const int ix = idx[0], iy = idx[1], iz = idx[2];
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  // one term D_split(x_d, m) f(u_a, u_b) of line direction d, partner m
  auto term = [&](int d, int xd, int sd, int m, const double* jown) {
    const int pb = lds_pos<N>(a + (m - xd) * sd);
    NodePrims<DIM> qb;
    qb.rho = s_q[pb];
#pragma unroll
    for (int c = 0; c < DIM; ++c) qb.v[c] = s_q[(1 + c) * PN + pb];
    qb.beta = s_q[QB * PN + pb];
    qb.kin = s_q[(QB + 1) * PN + pb];
    qb.p = 0.0;
    qb.c = 0.0;
    double nrm[DIM], f[NV];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      // NODAL: (Ja_x + Ja_m) / 2 from the halved rows (halving is exact)
      nrm[c] = NODAL ? jown[c] + s_ja[(d * DIM + c) * PN + pb] : jown[c];
    }
    flux_chandrashekar_exact<DIM>(p.inv_gm1, q, qb, nrm, f);
    const double w = s_D[xd * DS + m];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
      acc[v] = acc[v] + w * f[v];
    }
  };
  // the node's own rows (NODAL: halved, from LDS; affine: the element's)
  double jx[DIM], jy[DIM], jz[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    jx[c] = NODAL ? s_ja[(0 * DIM + c) * PN + pa] : jm[0][c];
    jy[c] = NODAL ? s_ja[(1 * DIM + c) * PN + pa] : jm[1][c];
    jz[c] = DIM == 3 ? (NODAL ? s_ja[((DIM - 1) * DIM + c) * PN + pa] : jm[DIM - 1][c]) : 0.0;
  }
  // partners below on the z-line (3D)
  if constexpr (DIM == 3) {
#pragma unroll 1
    for (int m = 0; m < iz; ++m) term(2, iz, n * n, m, jz);
  }
  // below on the y-line, below and above on the x-line, above on the y-line:
  // 2N terms for every node
#pragma unroll 1
  for (int t = 0; t < 2 * N; ++t) {
    const bool yl = t < iy, xl = t < iy + ix, xu = t < iy + N;
    const bool dy = yl || !xu;
    const int m = yl ? t : (xl ? t - iy : (xu ? t - iy + 1 : t - N + 1));
    double jo[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) jo[c] = dy ? jy[c] : jx[c];
    term(dy ? 1 : 0, dy ? iy : ix, dy ? n : 1, m, jo);
  }
  // partners above on the z-line (3D)
  if constexpr (DIM == 3) {
#pragma unroll 1
    for (int m = iz + 1; m <= N; ++m) term(2, iz, n * n, m, jz);
  }
  // 3b. Navier-Stokes: the viscous volume integral adds
  // D_hat(x_d, m) Ja^d . (-F_v)(m) over every partner m of each line through
  // the node (the node itself included) into the same sums, after the
  // split-form terms (parabolic_viscous.hpp:103-163 3D, :412-447 2D).  It
  // visits the source nodes in memory order, so a node takes: sources below
  // it on its z-line and y-line, its x-line up to itself, its own y and z
  // terms, then the rest of its x-, y- and z-lines.  At a face node the sums
  // hold the uncancelled half of the face balance, about a thousand times
  // the RHS, so this order fixes the RHS's last bits whatever the viscosity.
  const double* vf = p.visc;  // [E][d][var][dof], or null (Euler)
  auto vterm = [&](int d, int xd, int sd, int m) {
    const int b = a + (m - xd) * sd;
    const double w = s_Dh[xd * DS + m];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
      acc[v] = acc[v] + w * vf[((e * DIM + d) * NV + v) * NN + b];
    }
  };
  if (vf) {
    if constexpr (DIM == 3) {
#pragma unroll 1
      for (int m = 0; m < iz; ++m) vterm(2, iz, n * n, m);
    }
#pragma unroll 1
    for (int m = 0; m < iy; ++m) vterm(1, iy, n, m);
#pragma unroll 1
    for (int m = 0; m <= ix; ++m) vterm(0, ix, 1, m);
    vterm(1, iy, n, iy);
    if constexpr (DIM == 3) vterm(2, iz, n * n, iz);
#pragma unroll 1
    for (int m = ix + 1; m <= N; ++m) vterm(0, ix, 1, m);
#pragma unroll 1
    for (int m = iy + 1; m <= N; ++m) vterm(1, iy, n, m);
    if constexpr (DIM == 3) {
#pragma unroll 1
      for (int m = iz + 1; m <= N; ++m) vterm(2, iz, n * n, m);
    }
  }
  const double own_sg = ij >= 0.0 ? 1.0 : -1.0;
  // 3c. the viscous interface terms in direction order
  // (structured_solver.hpp:359-360 with the one-sided flux of
  // parabolic_viscous.hpp:318-334 / :570-585): the lower face takes the
  // node's own (-F_v) . (sign(J) Ja^d), the upper face the upper
  // neighbour's at its node with index 0.
  if (vf) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const bool lower = idx[d] == 0, upper = idx[d] == N;
      if (lower || upper) {
        int64_t lo, hi;
        face_neighbours<DIM, false>(p, p.cells[0], p.cells[1], nz, e, d, lo, hi);
        const int64_t ve = lower ? e : hi;
        const int vn = lower ? a : a - N * stride[d];
        const double vij = lower ? ij : (NODAL ? p.invJ[ve * NN + vn] : p.invJ[ve]);
        const double sg = vij >= 0.0 ? 1.0 : -1.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
          const double fp = sg * vf[((ve * DIM + d) * NV + v) * NN + vn];
          acc[v] = lower ? acc[v] - fp * p.lhat0 : acc[v] + fp * p.lhatN;
        }
      }
    }
  }
  // 4. numerical fluxes on the faces this node lies on, lifted after the
  // volume terms in direction order.  Their inputs are loaded here, not at
  // the top, so that nothing but the sums lives through the term loops (a
  // value held across them would spill).  The lower face takes the lower
  // neighbour's node with index N, the upper face the upper neighbour's node
  // with index 0 and its metric row and 1/J, whose sign and row make the
  // interface normal; interior nodes load nothing.
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const bool lower = idx[d] == 0, upper = idx[d] == N;
    if (lower || upper) {
      int64_t lo, hi;
      face_neighbours<DIM, false>(p, p.cells[0], p.cells[1], nz, e, d, lo, hi);
      const int64_t ne = lower ? lo : hi;
      const int nn = lower ? a + N * stride[d] : a - N * stride[d];
      double un[NV], uo[NV], nrm[DIM], f[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        un[v] = p.u[(ne * NV + v) * NN + nn];
        uo[v] = p.u[(e * NV + v) * NN + a];
      }
      if (lower) {
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          nrm[c] = own_sg * (NODAL ? 2.0 * s_ja[(d * DIM + c) * PN + pa] : jm[d][c]);
      } else {
        const double nij = NODAL ? p.invJ[ne * NN + nn] : p.invJ[ne];
        const double nb_sg = nij >= 0.0 ? 1.0 : -1.0;
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          nrm[c] = nb_sg * (NODAL ? p.Ja[((ne * DIM + d) * DIM + c) * NN + nn]
                                  : p.Ja[(ne * DIM + d) * DIM + c]);
      }
      const double* ul = lower ? un : uo;
      const double* ur = lower ? uo : un;
      if constexpr (SF == kFluxLLF) {
        flux_llf_exact<DIM>(p.gamma, ul, ur, nrm, f);
      } else {
        flux_hllc_exact<DIM>(p.gamma, ul, ur, nrm, f);
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
        acc[v] = lower ? acc[v] - f[v] * p.lhat0 : acc[v] + f[v] * p.lhatN;
      }
    }
  }
  const double* tot = acc;

  // 5. -1/J and the stage combination.  1/J, the base
  // state and the result are touched once per launch: non-temporal accesses
  // keep them from displacing the neighbours' face traces in the L2.
  const int64_t g = e * NV * NN + a;
  const double fac = -ij;
  // the base state's loads go out together (a select between the stage
  // kinds inside the variable loop would serialise them)
  const bool with_base = p.mode == kModeStage3;
  double ub[NV];
  if (with_base) {
#pragma unroll
    for (int v = 0; v < NV; ++v) ub[v] = __builtin_nontemporal_load(&p.u_base[g + v * NN]);
  }
  double uk[NV];  // the stage input, reloaded (an L2 hit) rather than held
#pragma unroll
  for (int v = 0; v < NV; ++v) uk[v] = p.u[g + v * NN];
  double r[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) r[v] = tot[v] * fac;
  if (p.mode == kModeRhs) {
#pragma unroll
    for (int v = 0; v < NV; ++v) __builtin_nontemporal_store(r[v], &p.out[g + v * NN]);
  } else {
    // the stage combinations of SSPRK3.hpp:12-143, rounded term by term
    double s[NV];
    if (with_base) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
        s[v] = p.c1 * ub[v] + p.c2 * uk[v] + p.c3 * r[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
#pragma clang fp contract(off)
        s[v] = p.c1 * uk[v] + p.c2 * r[v];
      }
    }
    if (p.floor_value > 0.0) {
      if (s[0] < p.floor_value) s[0] = p.floor_value;
      double m2 = 0.0;
      for (int k = 0; k < DIM; ++k) m2 += s[k + 1] * s[k + 1];
      const double kin = 0.5 * m2 / s[0];
      if ((p.gamma - 1.0) * (s[NV - 1] - kin) < p.floor_value)
        s[NV - 1] = p.floor_value / (p.gamma - 1.0) + kin;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) __builtin_nontemporal_store(s[v], &p.out[g + v * NN]);
  }
}

template <int DIM, int N, bool NODAL, int SF>
hipError_t launch_node_kernel(const KParams& p, hipStream_t stream) {
  constexpr int NN = Shape<DIM, N>::NN;
  if (p.E <= 0) return hipSuccess;
  if (p.E > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((rhs_node_kernel<DIM, N, NODAL, SF>), dim3((unsigned)p.E), dim3(NN), 0, stream,
                     p);
  return hipGetLastError();
}

}  // namespace detail
}  // namespace dgsem
