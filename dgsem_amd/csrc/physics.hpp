// Pointwise physics for the compressible Euler / Navier-Stokes equations in
// 2D and 3D: physical flux, LLF and HLLC surface fluxes, the Chandrashekar
// entropy-conservative / kinetic-energy-preserving two-point flux and the
// logarithmic mean.  Expressions keep the reference's evaluation order so the
// results agree with it to rounding.
//
// Reference: src/base/math_utils.hpp:8-24 (ln_mean),
// src/equations/compressible_euler{2D,3D}.hpp (flux, wave speed),
// src/base/numerical_flux.hpp:39-135,198-228 (LLF), :436-594 (Chandrashekar),
// src/base/numerical_flux_detail.hpp:9-21.
#pragma once

#include <cmath>

// The functions are used by the HIP kernels and, compiled by a host C++
// compiler, by the CPU unit tests of their derivatives.
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DG_HD __host__ __device__ __forceinline__
#else
#define DG_HD inline
#endif

namespace dgsem {

using std::fabs;
using std::log;
using std::sqrt;

// Numerical flux kinds (include/dgsem_amd.h DGSEM_FLUX_*).  kFluxAny is a
// template value only: the kernel instantiation that reads the volume,
// surface and subcell flux kinds from its parameters at run time.
enum FluxKind : int { kFluxLLF = 0, kFluxHLLC = 1, kFluxCentral = 2, kFluxAny = 3 };
// volume flux kinds
enum VolumeFluxKind : int { kVolumeChandrashekar = 0, kVolumeCentral = 2 };

DG_HD double dmax(double a, double b) { return (a < b) ? b : a; }
DG_HD double dmin(double a, double b) { return (b < a) ? b : a; }
// the same for other scalar types (the derivative code's dual numbers)
template <class T>
DG_HD T dmax(T a, T b) {
  return (a < b) ? b : a;
}
template <class T>
DG_HD T dmin(T a, T b) {
  return (b < a) ? b : a;
}

// 1 / x: on the GPU the hardware reciprocal refined by two Newton steps,
// which is within an ulp of the IEEE quotient for half the instructions of
// the division sequence (the parity tests bound the effect on the RHS).
DG_HD double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;
#endif
}

// Logarithmic mean (y - x) / log(y / x) with the series branch for
// f2 = ((x - y) / (x + y))^2 < 1e-4, math_utils.hpp:8-24.
DG_HD double ln_mean(double x, double y) {
  const double num = x * (x - 2.0 * y) + y * y;
  const double den = x * (x + 2.0 * y) + y * y;
  const double f2 = num / den;
  if (f2 < 1.0e-4) {
    const double poly = 2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0)));
    return (x + y) / poly;
  }
  return (y - x) / log(y / x);
}

// Pressure of a conservative state.
template <int DIM>
DG_HD double pressure(double gamma, const double* u) {
  double v2 = 0.0;
  for (int k = 0; k < DIM; ++k) {
    const double v = u[k + 1] / u[0];
    v2 += v * v;
  }
  return (gamma - 1.0) * (u[DIM + 1] - 0.5 * u[0] * v2);
}

// Physical flux F(u) . n: compressible_euler2D.hpp:25-57,
// compressible_euler3D.hpp:26-67 (sum of the per-direction fluxes).  T is
// double, or the dual numbers of the derivative code (physics_ad.hpp).
template <int DIM, class T = double>
DG_HD void physical_flux_n(double gamma, const T* u, const double* n, T* f) {
  constexpr int NV = DIM + 2;
  const T rho = u[0], rhoE = u[NV - 1];
  T v[DIM];
  T v2 = 0.0;
  for (int k = 0; k < DIM; ++k) v[k] = u[k + 1] / rho;
  for (int k = 0; k < DIM; ++k) v2 += v[k] * v[k];
  const T p = (gamma - 1.0) * (rhoE - 0.5 * rho * v2);
  for (int i = 0; i < NV; ++i) f[i] = 0.0;
  for (int d = 0; d < DIM; ++d) {
    // flux in direction d: (m_d, m_d v + p e_d, v_d (E + p))
    T fd[NV];
    fd[0] = u[d + 1];
    for (int k = 0; k < DIM; ++k) fd[k + 1] = u[d + 1] * v[k] + (k == d ? p : 0.0);
    fd[NV - 1] = v[d] * (rhoE + p);
    for (int i = 0; i < NV; ++i) f[i] = f[i] + n[d] * fd[i];
  }
}

// max(|v_L.n|, |v_R.n|) + max(c_L, c_R) |n|, compressible_euler3D.hpp:121-166.
template <int DIM>
DG_HD double llf_wave_speed(double gamma, const double* ul, const double* ur, const double* n) {
  double vnl = 0.0, vnr = 0.0, nn = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vnl += (ul[k + 1] / ul[0]) * n[k];
    vnr += (ur[k + 1] / ur[0]) * n[k];
    nn += n[k] * n[k];
  }
  const double al = sqrt(gamma * pressure<DIM>(gamma, ul) / ul[0]);
  const double ar = sqrt(gamma * pressure<DIM>(gamma, ur) / ur[0]);
  return dmax(fabs(vnl), fabs(vnr)) + dmax(al, ar) * sqrt(nn);
}

// Local Lax-Friedrichs: 0.5 (F_L + F_R).n - 0.5 lambda (u_R - u_L).
template <int DIM>
DG_HD void flux_llf(double gamma, const double* ul, const double* ur, const double* n, double* f) {
  constexpr int NV = DIM + 2;
  double fl[NV], fr[NV];
  const double lam = llf_wave_speed<DIM>(gamma, ul, ur, n);
  physical_flux_n<DIM>(gamma, ul, n, fl);
  physical_flux_n<DIM>(gamma, ur, n, fr);
  for (int i = 0; i < NV; ++i) f[i] = 0.5 * (fl[i] + fr[i]) - 0.5 * lam * (ur[i] - ul[i]);
}

// HLLC with a normal vector.  The reference has HLLC for 1D Euler only
// (numerical_flux.hpp:283-378); this applies the same construction
// (Roe-averaged wave-speed bounds, contact speed, star states) in the frame of
// the unit normal with the tangential velocity carried through, and scales
// the result by |n|.
template <int DIM, class T = double>
DG_HD void flux_hllc(double gamma, const T* ul, const T* ur, const double* n, T* f) {
  constexpr int NV = DIM + 2;
  double nrm = 0.0, nh[DIM];
  T vl[DIM], vr[DIM];
  for (int d = 0; d < DIM; ++d) nrm += n[d] * n[d];
  nrm = sqrt(nrm);
  for (int d = 0; d < DIM; ++d) nh[d] = n[d] / nrm;
  const T rl = ul[0], rr = ur[0], El = ul[NV - 1], Er = ur[NV - 1];
  T v2l = 0.0, v2r = 0.0, vnl = 0.0, vnr = 0.0;
  for (int d = 0; d < DIM; ++d) {
    vl[d] = ul[d + 1] / rl;
    vr[d] = ur[d + 1] / rr;
    v2l += vl[d] * vl[d];
    v2r += vr[d] * vr[d];
    vnl += vl[d] * nh[d];
    vnr += vr[d] * nh[d];
  }
  const T el = El / rl, er = Er / rr;
  const T pl = (gamma - 1.0) * (El - 0.5 * rl * v2l);
  const T pr = (gamma - 1.0) * (Er - 0.5 * rr * v2r);
  const T cl = sqrt(gamma * pl / rl), cr = sqrt(gamma * pr / rr);
  const T sl = sqrt(rl), sr = sqrt(rr), ss = sl + sr;
  T vroe2 = 0.0;
  for (int d = 0; d < DIM; ++d) {
    const T vd = (sl * vl[d] + sr * vr[d]) / ss;
    vroe2 += vd * vd;
  }
  const T vroe = (sl * vnl + sr * vnr) / ss;
  const T Hl = (El + pl) / rl, Hr = (Er + pr) / rr;
  const T Hroe = (sl * Hl + sr * Hr) / ss;
  const T croe = sqrt((gamma - 1.0) * (Hroe - 0.5 * vroe2));
  const T SL = dmin(vnl - cl, vroe - croe);
  const T SR = dmax(vnr + cr, vroe + croe);
  const T mL = SL - vnl, mR = SR - vnr;
  // the side whose state (or star state) gives the flux; only that side's
  // physical flux is formed
  const bool star = !(SL >= 0) && !(SR <= 0);
  const T Ss = star ? (pr - pl + rl * vnl * mL - rr * vnr * mR) / (rl * mL - rr * mR) : T(0.0);
  const bool left = star ? Ss >= 0 : SL >= 0;
  const T* uk = left ? ul : ur;
  T fk[NV];
  physical_flux_n<DIM>(gamma, uk, nh, fk);
  if (!star) {
    for (int i = 0; i < NV; ++i) f[i] = fk[i];
  } else {
    const T* vk = left ? vl : vr;
    const T rk = left ? rl : rr, vnk = left ? vnl : vnr, ek = left ? el : er;
    const T pk = left ? pl : pr, Sk = left ? SL : SR, mk = left ? mL : mR;
    const T rs = rk * mk / (Sk - Ss);
    const T es = ek + (Ss - vnk) * (Ss + pk / (rk * mk));
    f[0] = fk[0] + Sk * (rs - uk[0]);
    for (int d = 0; d < DIM; ++d) f[d + 1] = fk[d + 1] + Sk * (rs * (vk[d] + (Ss - vnk) * nh[d]) - uk[d + 1]);
    f[NV - 1] = fk[NV - 1] + Sk * (rs * es - uk[NV - 1]);
  }
  for (int i = 0; i < NV; ++i) f[i] *= nrm;
}

// CentralFlux with a normal, numerical_flux.hpp:247-257 (2D in the
// reference): 0.5 (F(u_L).n + F(u_R).n), numerical_flux_detail.hpp:23-32.
template <int DIM>
DG_HD void flux_central(double gamma, const double* ul, const double* ur, const double* n, double* f) {
  constexpr int NV = DIM + 2;
  double fl[NV], fr[NV];
  physical_flux_n<DIM>(gamma, ul, n, fl);
  physical_flux_n<DIM>(gamma, ur, n, fr);
  for (int i = 0; i < NV; ++i) f[i] = 0.5 * (fl[i] + fr[i]);
}

template <int DIM>
DG_HD void surface_flux(int kind, double gamma, const double* ul, const double* ur, const double* n,
                        double* f) {
  if (kind == kFluxHLLC)
    flux_hllc<DIM>(gamma, ul, ur, n, f);
  else if (kind == kFluxCentral)
    flux_central<DIM>(gamma, ul, ur, n, f);
  else
    flux_llf<DIM>(gamma, ul, ur, n, f);
}

// Per-node quantities hoisted out of the two-point and surface fluxes:
// rho, v[DIM], beta = rho / (2 p), the specific kinetic energy |v|^2 / 2,
// the pressure and the speed of sound.
template <int DIM>
struct NodePrims {
  double rho, v[DIM], beta, kin, p, c;
};

template <int DIM>
DG_HD NodePrims<DIM> node_prims(double gamma, const double* u) {
  NodePrims<DIM> q;
  q.rho = u[0];
  double v2 = 0.0;
  const double ir = recip(u[0]);
  for (int k = 0; k < DIM; ++k) q.v[k] = u[k + 1] * ir;
  for (int k = 0; k < DIM; ++k) v2 += q.v[k] * q.v[k];
  q.p = (gamma - 1.0) * (u[DIM + 1] - 0.5 * u[0] * v2);
  q.beta = 0.5 * u[0] * recip(q.p);
  q.kin = 0.5 * v2;
  q.c = sqrt(gamma * q.p * ir);
  return q;
}

// Chandrashekar two-point flux from hoisted node quantities,
// numerical_flux.hpp:436-500 (2D) and :520-594 (3D).  inv_gm1 = 1/(gamma-1).
//
// The two logarithmic means (math_utils.hpp:8-24), the mean pressure
// 0.5 rho_avg / beta_avg and 1 / beta_mean need five quotients; in the usual
// case where both means take the series branch (f2 < 1e-4) they come from
// two reciprocals:  1/(s_rho^2 s_beta^2) gives f2 = (x - y)^2 / s^2 of both
// means, and 1/(poly_rho s_beta) gives rho_mean = s_rho/poly_rho,
// 1/beta_mean = poly_beta/s_beta and p_mean = s_rho/(2 s_beta), where
// s = x + y.  Results agree with the quotient form to a few ulp.  Otherwise
// the means are the reference's (y - x) / log(y / x) with x <= y.
// Every expression is symmetric in the two nodes, so f(a, b) and f(b, a)
// are the same bits: a kernel may evaluate a pair from either side.
template <int DIM>
DG_HD void flux_chandrashekar(double inv_gm1, const NodePrims<DIM>& a, const NodePrims<DIM>& b,
                              const double* n, double* f) {
  double vna = 0.0, vnb = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vna += a.v[k] * n[k];
    vnb += b.v[k] * n[k];
  }
  const double sr = a.rho + b.rho, sb = a.beta + b.beta;
  const double xr = a.rho - b.rho, xb = a.beta - b.beta;
  const double nr = xr * xr, dr = sr * sr;
  const double nb = xb * xb, db = sb * sb;
  const double r1 = recip(dr * db);
  const double f2r = nr * (db * r1), f2b = nb * (dr * r1);
  const double polyr = 2.0 + f2r * (2.0 / 3.0 + f2r * (2.0 / 5.0 + f2r * (2.0 / 7.0)));
  const double polyb = 2.0 + f2b * (2.0 / 3.0 + f2b * (2.0 / 5.0 + f2b * (2.0 / 7.0)));
  double rho_mean, inv_beta_mean, p_mean;
  if (f2r < 1.0e-4 && f2b < 1.0e-4) {
    const double r2 = recip(polyr * sb);
    rho_mean = sr * (sb * r2);
    inv_beta_mean = polyb * (polyr * r2);
    p_mean = 0.5 * sr * (polyr * r2);
  } else {
    const double rlo = dmin(a.rho, b.rho), rhi = dmax(a.rho, b.rho);
    const double blo = dmin(a.beta, b.beta), bhi = dmax(a.beta, b.beta);
    rho_mean = f2r < 1.0e-4 ? sr / polyr : (rhi - rlo) / log(rhi / rlo);
    inv_beta_mean = f2b < 1.0e-4 ? polyb / sb : log(bhi / blo) / (bhi - blo);
    p_mean = 0.5 * sr / sb;
  }
  const double vsq_avg = a.kin + b.kin;
  const double f1 = rho_mean * 0.5 * (vna + vnb);
  double fe = f1 * 0.5 * (inv_gm1 * inv_beta_mean - vsq_avg);
  f[0] = f1;
  for (int k = 0; k < DIM; ++k) {
    const double vavg = 0.5 * (a.v[k] + b.v[k]);
    f[k + 1] = f1 * vavg + p_mean * n[k];
    fe = fe + f[k + 1] * vavg;
  }
  f[DIM + 1] = fe;
}

// Primitive quantities of a state read from memory: one reciprocal of rho
// and one square root.
template <int DIM>
DG_HD NodePrims<DIM> face_prims(double gamma, const double* u) {
  NodePrims<DIM> q;
  const double ir = recip(u[0]);
  double v2 = 0.0;
  q.rho = u[0];
  for (int k = 0; k < DIM; ++k) q.v[k] = u[k + 1] * ir;
  for (int k = 0; k < DIM; ++k) v2 += q.v[k] * q.v[k];
  q.p = (gamma - 1.0) * (u[DIM + 1] - 0.5 * u[0] * v2);
  q.c = sqrt(gamma * q.p * ir);
  q.beta = 0.0;
  q.kin = 0.5 * v2;
  return q;
}

// Local Lax-Friedrichs flux from conservative states and their primitives:
// 0.5 (F_L + F_R).n - 0.5 lambda (u_R - u_L), lambda = max|v.n| + max(c)|n|
// (numerical_flux.hpp:54-68, compressible_euler3D.hpp:121-166).
template <int DIM>
DG_HD void flux_llf_prims(const double* ul, const NodePrims<DIM>& ql, const double* ur,
                          const NodePrims<DIM>& qr, const double* n, double* f) {
  constexpr int NV = DIM + 2;
  double vnl = 0.0, vnr = 0.0, nn = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vnl += ql.v[k] * n[k];
    vnr += qr.v[k] * n[k];
    nn += n[k] * n[k];
  }
  const double lam = dmax(fabs(vnl), fabs(vnr)) + dmax(ql.c, qr.c) * sqrt(nn);
  double fl[NV], fr[NV];
  fl[0] = ul[0] * vnl;
  fr[0] = ur[0] * vnr;
  for (int k = 0; k < DIM; ++k) {
    fl[k + 1] = ul[k + 1] * vnl + ql.p * n[k];
    fr[k + 1] = ur[k + 1] * vnr + qr.p * n[k];
  }
  fl[NV - 1] = (ul[NV - 1] + ql.p) * vnl;
  fr[NV - 1] = (ur[NV - 1] + qr.p) * vnr;
  for (int i = 0; i < NV; ++i) f[i] = 0.5 * (fl[i] + fr[i]) - 0.5 * lam * (ur[i] - ul[i]);
}

// gradient_variables (rho, v, T = p/rho), compressible_navier_stokes3D.hpp:55-70
template <int DIM>
DG_HD void grad_vars(double gamma, const double* u, double* q) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double inv_rho = 1.0 / u[0];
  double m2 = 0.0;
  for (int k = 0; k < DIM; ++k) m2 += u[k + 1] * u[k + 1];
  const double p = (gamma - 1.0) * (u[DIM + 1] - 0.5 * m2 * inv_rho);
  q[0] = u[0];
  for (int k = 0; k < DIM; ++k) q[k + 1] = u[k + 1] * inv_rho;
  q[DIM + 1] = p * inv_rho;
}

// Newtonian viscous flux in direction dim, compressible_navier_stokes{2D,3D}.hpp
// viscous_flux.  g[c][var] = d q_var / d x_c.
template <int DIM>
DG_HD void viscous_flux_dir(double mu, double kappa, const double* q,
                                                 const double g[DIM][DIM + 2], int dim, double* f) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  if constexpr (DIM == 2) {
    const double tau11 = (4.0 * g[0][1] - 2.0 * g[1][2]) / 3.0;
    const double tau22 = (4.0 * g[1][2] - 2.0 * g[0][1]) / 3.0;
    const double tau12 = g[1][1] + g[0][2];
    const double t1 = dim == 0 ? tau11 : tau12, t2 = dim == 0 ? tau12 : tau22;
    f[0] = 0.0;
    f[1] = mu * t1;
    f[2] = mu * t2;
    f[3] = mu * (q[1] * t1 + q[2] * t2 + kappa * g[dim][3]);
  } else {
    const double tau11 = (4.0 * g[0][1] - 2.0 * (g[1][2] + g[2][3])) / 3.0;
    const double tau22 = (4.0 * g[1][2] - 2.0 * (g[0][1] + g[2][3])) / 3.0;
    const double tau33 = (4.0 * g[2][3] - 2.0 * (g[0][1] + g[1][2])) / 3.0;
    const double tau12 = g[1][1] + g[0][2];
    const double tau13 = g[2][1] + g[0][3];
    const double tau23 = g[2][2] + g[1][3];
    const double t1 = dim == 0 ? tau11 : (dim == 1 ? tau12 : tau13);
    const double t2 = dim == 0 ? tau12 : (dim == 1 ? tau22 : tau23);
    const double t3 = dim == 0 ? tau13 : (dim == 1 ? tau23 : tau33);
    f[0] = 0.0;
    f[1] = mu * t1;
    f[2] = mu * t2;
    f[3] = mu * t3;
    f[4] = mu * (q[1] * t1 + q[2] * t2 + q[3] * t3 + kappa * g[dim][4]);
  }
}

// p * rho with gamma fixed at 1.4, the indicator variable of
// subcell_limiting/indicator.hpp:140-215.
template <int DIM>
DG_HD double indicator_variable(const double* u) {
  const double g = 1.4;
  double v2 = 0.0;
  for (int k = 0; k < DIM; ++k) {
    const double v = u[k + 1] / u[0];
    v2 += v * v;
  }
  const double p = (g - 1.0) * (u[DIM + 1] - 0.5 * u[0] * v2);
  return p * u[0];
}

}  // namespace dgsem
