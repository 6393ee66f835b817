// Tangent (JVP) and adjoint (VJP) of the pointwise physics in physics.hpp
// and parabolic_impl.hpp, derived by hand from the implemented formulas.
//
// Every nonlinear map y = F(x) here has a pair of functions:
//   *_jvp: dy = F'(x) dx
//   *_vjp: xbar += F'(x)^T ybar   (accumulates)
// built from the same partial derivatives, so <ybar, F' dx> = <F'^T ybar, dx>
// holds to rounding.  Branches (the series / logarithm forms of the
// logarithmic mean, the max in the LLF wave speed) follow the forward
// evaluation's choice at x: these are the derivatives of the function the
// forward kernels compute, i.e. a discrete tangent and adjoint.  Where the
// LLF wave speed's max() ties exactly, the flux has no derivative and the
// two sides share it (the mean of the one-sided derivatives).
//
// The blending factor of shock capturing is not differentiated (it is frozen
// at its forward value by the callers).
#pragma once

#include "physics.hpp"

namespace dgsem {

// Node primitives used by the two-point flux: q = (rho, v[DIM], beta, kin),
// beta = rho / (2 p), kin = |v|^2 / 2, p = (gamma - 1)(E - rho kin).
template <int DIM>
struct PrimTangent {
  double rho, v[DIM], beta, kin;
};

template <int DIM>
DG_HD void prims_jvp(double gamma, const double* u, const NodePrims<DIM>& q, const double* du,
                     PrimTangent<DIM>& dq) {
  const double ir = 1.0 / q.rho;
  dq.rho = du[0];
  double dkin = 0.0;
  for (int k = 0; k < DIM; ++k) {
    dq.v[k] = (du[k + 1] - q.v[k] * du[0]) * ir;
    dkin += q.v[k] * dq.v[k];
  }
  dq.kin = dkin;
  const double dp = (gamma - 1.0) * (du[DIM + 1] - q.kin * du[0] - q.rho * dkin);
  dq.beta = q.beta * (du[0] * ir - dp / q.p);
  (void)u;
}

// ubar += (dq/du)^T qbar; pbar is an extra cotangent of the pressure.
template <int DIM>
DG_HD void prims_vjp(double gamma, const NodePrims<DIM>& q, const PrimTangent<DIM>& qbar,
                     double pbar, double* ubar) {
  const double ir = 1.0 / q.rho;
  double rb = qbar.rho + qbar.beta * q.beta * ir;
  double pb = pbar - qbar.beta * q.beta / q.p;
  double vb[DIM];
  for (int k = 0; k < DIM; ++k) vb[k] = qbar.v[k] + qbar.kin * q.v[k];
  // p = (gamma - 1) (E - rho kin(v)), with v held as an independent variable
  ubar[DIM + 1] += (gamma - 1.0) * pb;
  rb -= (gamma - 1.0) * q.kin * pb;
  for (int k = 0; k < DIM; ++k) vb[k] -= (gamma - 1.0) * q.rho * q.v[k] * pb;
  // v = m / rho
  for (int k = 0; k < DIM; ++k) {
    ubar[k + 1] += vb[k] * ir;
    rb -= vb[k] * q.v[k] * ir;
  }
  ubar[0] += rb;
}

// Logarithmic mean and its partial derivatives on both branches, the one
// the forward evaluation took selected by `series`.  Both are computed and
// one is selected without a branch: divergent branches in each of the
// unrolled pair evaluations of a line would keep the line's state live across
// all of them and multiply the register count.  On the series branch the log
// form may be inf / nan (x == y); it is discarded by the select.
DG_HD void ln_mean_partials(double x, double y, bool series, double& L, double& Lx, double& Ly) {
  const double s = x + y, d = x - y;
  const double is = 1.0 / s;
  const double f = d * d * is * is;
  const double P = 2.0 + f * (2.0 / 3.0 + f * (2.0 / 5.0 + f * (2.0 / 7.0)));
  const double dP = 2.0 / 3.0 + f * (4.0 / 5.0 + f * (6.0 / 7.0));
  const double fx = 2.0 * d * is * is - 2.0 * d * d * is * is * is;
  const double fy = -2.0 * d * is * is - 2.0 * d * d * is * is * is;
  const double Ls = s / P;
  const double c = s * dP / (P * P);
  const double Lxs = 1.0 / P - c * fx, Lys = 1.0 / P - c * fy;
  const double D = log(y / x);
  const double Ll = (y - x) / D;
  const double Lxl = (Ll / x - 1.0) / D, Lyl = (1.0 - Ll / y) / D;
  L = series ? Ls : Ll;
  Lx = series ? Lxs : Lxl;
  Ly = series ? Lys : Lyl;
}

// Means of the Chandrashekar flux and their partial derivatives with
// respect to (rho_a, rho_b) and (beta_a, beta_b): rho_mean, its inverse
// beta mean B = 1 / ln_mean(beta_a, beta_b) and p_mean = (rho_a + rho_b) /
// (2 (beta_a + beta_b)).  The branch test is the forward's.
struct ChandrashekarMeans {
  double rho, rho_a, rho_b;  // rho_mean and partials
  double B, B_a, B_b;        // inverse beta mean and partials
  double p, p_rho, p_beta;   // p_mean, d/d(rho_a or rho_b), d/d(beta_a or beta_b)
};

template <int DIM>
DG_HD ChandrashekarMeans chandrashekar_means(const NodePrims<DIM>& a, const NodePrims<DIM>& b) {
  const double sr = a.rho + b.rho, sb = a.beta + b.beta;
  const double xr = a.rho - b.rho, xb = a.beta - b.beta;
  const double nr = xr * xr, dr = sr * sr;
  const double nb = xb * xb, db = sb * sb;
  const double r1 = 1.0 / (dr * db);
  const double f2r = nr * (db * r1), f2b = nb * (dr * r1);
  ChandrashekarMeans m;
  ln_mean_partials(a.rho, b.rho, f2r < 1.0e-4, m.rho, m.rho_a, m.rho_b);
  double Lb, Lba, Lbb;
  ln_mean_partials(a.beta, b.beta, f2b < 1.0e-4, Lb, Lba, Lbb);
  m.B = 1.0 / Lb;
  m.B_a = -m.B * m.B * Lba;
  m.B_b = -m.B * m.B * Lbb;
  m.p = 0.5 * sr / sb;
  m.p_rho = 0.5 / sb;
  m.p_beta = -m.p / sb;
  return m;
}

// Tangent of the Chandrashekar flux with respect to both nodes' primitives
// (the normal n is geometry, not differentiated).
template <int DIM>
DG_HD void chandrashekar_jvp(double inv_gm1, const NodePrims<DIM>& a, const NodePrims<DIM>& b,
                             const PrimTangent<DIM>& da, const PrimTangent<DIM>& db,
                             const double* n, double* df) {
  const ChandrashekarMeans m = chandrashekar_means<DIM>(a, b);
  double vna = 0.0, vnb = 0.0, dvna = 0.0, dvnb = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vna += a.v[k] * n[k];
    vnb += b.v[k] * n[k];
    dvna += da.v[k] * n[k];
    dvnb += db.v[k] * n[k];
  }
  const double drm = m.rho_a * da.rho + m.rho_b * db.rho;
  const double dB = m.B_a * da.beta + m.B_b * db.beta;
  const double dpm = m.p_rho * (da.rho + db.rho) + m.p_beta * (da.beta + db.beta);
  const double vsq = a.kin + b.kin, dvsq = da.kin + db.kin;
  const double f1 = m.rho * 0.5 * (vna + vnb);
  const double df1 = drm * 0.5 * (vna + vnb) + m.rho * 0.5 * (dvna + dvnb);
  const double A = 0.5 * (inv_gm1 * m.B - vsq);
  double dfe = df1 * A + f1 * 0.5 * (inv_gm1 * dB - dvsq);
  df[0] = df1;
  for (int k = 0; k < DIM; ++k) {
    const double vavg = 0.5 * (a.v[k] + b.v[k]);
    const double dvavg = 0.5 * (da.v[k] + db.v[k]);
    const double fk = f1 * vavg + m.p * n[k];
    const double dfk = df1 * vavg + f1 * dvavg + dpm * n[k];
    df[k + 1] = dfk;
    dfe += dfk * vavg + fk * dvavg;
  }
  df[DIM + 1] = dfe;
}

// abar += (df/dq_a)^T g, bbar += (df/dq_b)^T g.
template <int DIM>
DG_HD void chandrashekar_vjp(double inv_gm1, const NodePrims<DIM>& a, const NodePrims<DIM>& b,
                             const double* n, const double* g, PrimTangent<DIM>& abar,
                             PrimTangent<DIM>& bbar) {
  const ChandrashekarMeans m = chandrashekar_means<DIM>(a, b);
  double vna = 0.0, vnb = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vna += a.v[k] * n[k];
    vnb += b.v[k] * n[k];
  }
  const double f1 = m.rho * 0.5 * (vna + vnb);
  const double A = 0.5 * (inv_gm1 * m.B - (a.kin + b.kin));
  const double gE = g[DIM + 1];
  double f1b = g[0] + gE * A;
  double pmb = 0.0;
  for (int k = 0; k < DIM; ++k) {
    const double vavg = 0.5 * (a.v[k] + b.v[k]);
    const double fk = f1 * vavg + m.p * n[k];
    const double gk = g[k + 1] + gE * vavg;  // cotangent of f_k (own and in f_E)
    f1b += gk * vavg;
    pmb += gk * n[k];
    const double vavgb = gk * f1 + gE * fk;
    abar.v[k] += 0.5 * vavgb;
    bbar.v[k] += 0.5 * vavgb;
  }
  const double Bb = gE * f1 * 0.5 * inv_gm1;
  const double vsqb = -gE * f1 * 0.5;
  const double rmb = f1b * 0.5 * (vna + vnb);
  const double vnb_ = f1b * m.rho * 0.5;  // cotangent of vna and of vnb
  for (int k = 0; k < DIM; ++k) {
    abar.v[k] += vnb_ * n[k];
    bbar.v[k] += vnb_ * n[k];
  }
  abar.kin += vsqb;
  bbar.kin += vsqb;
  abar.rho += m.rho_a * rmb + m.p_rho * pmb;
  bbar.rho += m.rho_b * rmb + m.p_rho * pmb;
  abar.beta += m.B_a * Bb + m.p_beta * pmb;
  bbar.beta += m.B_b * Bb + m.p_beta * pmb;
}

// ---------------------------------------------------------------------
// Physical flux F(u).n and the LLF flux, in terms of u with the node
// primitives q (rho, v, p, c) of u.
// ---------------------------------------------------------------------

// Tangent of (v, p) of a state: dv_k = (dm_k - v_k drho) / rho,
// dp = (gamma - 1)(dE - v.dm + kin drho).
template <int DIM>
DG_HD void vp_jvp(double gamma, const NodePrims<DIM>& q, const double* du, double* dv, double& dp) {
  const double ir = 1.0 / q.rho;
  double vdm = 0.0;
  for (int k = 0; k < DIM; ++k) {
    dv[k] = (du[k + 1] - q.v[k] * du[0]) * ir;
    vdm += q.v[k] * du[k + 1];
  }
  dp = (gamma - 1.0) * (du[DIM + 1] - vdm + q.kin * du[0]);
}

// ubar += (d(v, p)/du)^T (vbar, pbar).
template <int DIM>
DG_HD void vp_vjp(double gamma, const NodePrims<DIM>& q, const double* vbar, double pbar,
                  double* ubar) {
  const double ir = 1.0 / q.rho;
  double rb = (gamma - 1.0) * q.kin * pbar;
  for (int k = 0; k < DIM; ++k) {
    ubar[k + 1] += vbar[k] * ir - (gamma - 1.0) * q.v[k] * pbar;
    rb -= vbar[k] * q.v[k] * ir;
  }
  ubar[0] += rb;
  ubar[DIM + 1] += (gamma - 1.0) * pbar;
}

// d(F(u).n) for F0 = rho vn, F_k = m_k vn + p n_k, F_E = (E + p) vn.
template <int DIM>
DG_HD void flux_n_jvp(double gamma, const double* u, const NodePrims<DIM>& q, const double* du,
                      const double* n, double* dF) {
  double dv[DIM], dp, vn = 0.0, dvn = 0.0;
  vp_jvp<DIM>(gamma, q, du, dv, dp);
  for (int k = 0; k < DIM; ++k) {
    vn += q.v[k] * n[k];
    dvn += dv[k] * n[k];
  }
  dF[0] = du[0] * vn + u[0] * dvn;
  for (int k = 0; k < DIM; ++k) dF[k + 1] = du[k + 1] * vn + u[k + 1] * dvn + dp * n[k];
  dF[DIM + 1] = (du[DIM + 1] + dp) * vn + (u[DIM + 1] + q.p) * dvn;
}

// ubar += (d(F(u).n)/du)^T g
template <int DIM>
DG_HD void flux_n_vjp(double gamma, const double* u, const NodePrims<DIM>& q, const double* n,
                      const double* g, double* ubar) {
  double vn = 0.0;
  for (int k = 0; k < DIM; ++k) vn += q.v[k] * n[k];
  const double gE = g[DIM + 1];
  ubar[0] += g[0] * vn;
  double vnb = g[0] * u[0] + gE * (u[DIM + 1] + q.p);
  double pb = gE * vn;
  for (int k = 0; k < DIM; ++k) {
    ubar[k + 1] += g[k + 1] * vn;
    vnb += g[k + 1] * u[k + 1];
    pb += g[k + 1] * n[k];
  }
  ubar[DIM + 1] += gE * vn;
  double vb[DIM];
  for (int k = 0; k < DIM; ++k) vb[k] = vnb * n[k];
  vp_vjp<DIM>(gamma, q, vb, pb, ubar);
}

// Gradient of the LLF wave speed max(|vn_L|, |vn_R|) + max(c_L, c_R)|n| with
// respect to one side's state: `left` selects the side, and only the terms
// whose max that side attains (the forward's dmax choice) contribute.
template <int DIM>
DG_HD void llf_lambda_grad(double gamma, const NodePrims<DIM>& ql, const NodePrims<DIM>& qr,
                           const double* n, bool left, double* grad) {
  double vnl = 0.0, vnr = 0.0, nn = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vnl += ql.v[k] * n[k];
    vnr += qr.v[k] * n[k];
    nn += n[k] * n[k];
  }
  const NodePrims<DIM>& q = left ? ql : qr;
  for (int i = 0; i < DIM + 2; ++i) grad[i] = 0.0;
  // The side a max() takes gets the whole gradient.  On an exact tie the max
  // has no derivative; each side then gets half of its own, the mean of the
  // two one-sided derivatives, which is what a central difference sees.
  // Symmetric states put ties on many faces (equal traces, or a velocity or
  // sound speed that does not vary along the face normal).
  const double avl = fabs(vnl), avr = fabs(vnr);
  const double wvn = avl == avr ? 0.5 : ((avl > avr) == left ? 1.0 : 0.0);
  const double wc = ql.c == qr.c ? 0.5 : ((ql.c > qr.c) == left ? 1.0 : 0.0);
  double vb[DIM] = {};
  double pb = 0.0;
  if (wvn != 0.0) {
    const double vn = left ? vnl : vnr;
    const double s = vn > 0.0 ? 1.0 : (vn < 0.0 ? -1.0 : 0.0);
    for (int k = 0; k < DIM; ++k) vb[k] += wvn * s * n[k];
  }
  if (wc != 0.0) {
    // c = sqrt(gamma p / rho): dc = c/2 (dp/p - drho/rho)
    const double w = wc * sqrt(nn) * 0.5 * q.c;
    pb += w / q.p;
    grad[0] -= w / q.rho;
  }
  vp_vjp<DIM>(gamma, q, vb, pb, grad);
}

// Tangent of the LLF flux 0.5 (F_L + F_R).n - 0.5 lambda (u_R - u_L).
template <int DIM>
DG_HD void llf_jvp(double gamma, const double* ul, const NodePrims<DIM>& ql, const double* ur,
                   const NodePrims<DIM>& qr, const double* dul, const double* dur, const double* n,
                   double* df) {
  constexpr int NV = DIM + 2;
  double dfl[NV], dfr[NV], gl[NV], gr[NV];
  flux_n_jvp<DIM>(gamma, ul, ql, dul, n, dfl);
  flux_n_jvp<DIM>(gamma, ur, qr, dur, n, dfr);
  llf_lambda_grad<DIM>(gamma, ql, qr, n, true, gl);
  llf_lambda_grad<DIM>(gamma, ql, qr, n, false, gr);
  double vnl = 0.0, vnr = 0.0, nn = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vnl += ql.v[k] * n[k];
    vnr += qr.v[k] * n[k];
    nn += n[k] * n[k];
  }
  const double lam = dmax(fabs(vnl), fabs(vnr)) + dmax(ql.c, qr.c) * sqrt(nn);
  double dlam = 0.0;
  for (int i = 0; i < NV; ++i) dlam += gl[i] * dul[i] + gr[i] * dur[i];
  for (int i = 0; i < NV; ++i)
    df[i] = 0.5 * (dfl[i] + dfr[i]) - 0.5 * dlam * (ur[i] - ul[i]) - 0.5 * lam * (dur[i] - dul[i]);
}

// ubar_side += (df/du_side)^T g for one side of the LLF flux.
template <int DIM>
DG_HD void llf_vjp_side(double gamma, const double* ul, const NodePrims<DIM>& ql, const double* ur,
                        const NodePrims<DIM>& qr, const double* n, const double* g, bool left,
                        double* ubar) {
  constexpr int NV = DIM + 2;
  double vnl = 0.0, vnr = 0.0, nn = 0.0;
  for (int k = 0; k < DIM; ++k) {
    vnl += ql.v[k] * n[k];
    vnr += qr.v[k] * n[k];
    nn += n[k] * n[k];
  }
  const double lam = dmax(fabs(vnl), fabs(vnr)) + dmax(ql.c, qr.c) * sqrt(nn);
  double gh[NV], lg[NV], gdu = 0.0;
  for (int i = 0; i < NV; ++i) {
    gh[i] = 0.5 * g[i];
    gdu += g[i] * (ur[i] - ul[i]);
  }
  flux_n_vjp<DIM>(gamma, left ? ul : ur, left ? ql : qr, n, gh, ubar);
  llf_lambda_grad<DIM>(gamma, ql, qr, n, left, lg);
  const double sgn = left ? 0.5 : -0.5;
  for (int i = 0; i < NV; ++i) ubar[i] += sgn * lam * g[i] - 0.5 * gdu * lg[i];
}

// ---------------------------------------------------------------------
// HLLC: the tangent by forward-mode dual numbers through the flux as
// physics.hpp writes it (flux_hllc is templated on its scalar type), so it
// follows the branches (wave-speed bounds, star region, side) the forward
// evaluation takes; the adjoint by hand in reverse mode, through the same
// branches.
// ---------------------------------------------------------------------

// value and derivative; arithmetic by the usual rules, comparisons on values
struct Dual {
  double v, d;
  DG_HD Dual(double x = 0.0, double dx = 0.0) : v(x), d(dx) {}
  DG_HD Dual& operator+=(Dual b) {
    v += b.v;
    d += b.d;
    return *this;
  }
  DG_HD Dual& operator*=(Dual b) {
    d = d * b.v + v * b.d;
    v *= b.v;
    return *this;
  }
};
DG_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
DG_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
DG_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
DG_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
DG_HD Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
DG_HD Dual sqrt(Dual a) {
  const double r = sqrt(a.v);
  return Dual(r, 0.5 * a.d / r);
}
DG_HD bool operator<(Dual a, Dual b) { return a.v < b.v; }
DG_HD bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
DG_HD bool operator>=(Dual a, Dual b) { return a.v >= b.v; }

// df = (dF/du_l) du_l + (dF/du_r) du_r for the HLLC flux F(u_l, u_r; n)
template <int DIM>
DG_HD void hllc_jvp(double gamma, const double* ul, const double* ur, const double* dul,
                    const double* dur, const double* n, double* df) {
  constexpr int NV = DIM + 2;
  Dual a[NV], b[NV], f[NV];
  for (int i = 0; i < NV; ++i) {
    a[i] = Dual(ul[i], dul[i]);
    b[i] = Dual(ur[i], dur[i]);
  }
  flux_hllc<DIM, Dual>(gamma, a, b, n, f);
  for (int i = 0; i < NV; ++i) df[i] = f[i].d;
}

// ubar_side += (dF/du_side)^T g for the sides selected, by forward-mode
// dual numbers: one column of the Jacobian per input component.  The
// kernels use the reverse-mode hllc_vjp below; this form serves as its check.
template <int DIM>
DG_HD void hllc_vjp_columns(double gamma, const double* ul, const double* ur, const double* n,
                            const double* g, bool with_left, bool with_right, double* ubar_l,
                            double* ubar_r) {
  constexpr int NV = DIM + 2;
  for (int side = 0; side < 2; ++side) {
    if (side == 0 ? !with_left : !with_right) continue;
    double* ub = side == 0 ? ubar_l : ubar_r;
    for (int j = 0; j < NV; ++j) {
      Dual a[NV], b[NV], f[NV];
      for (int i = 0; i < NV; ++i) {
        a[i] = Dual(ul[i], side == 0 && i == j ? 1.0 : 0.0);
        b[i] = Dual(ur[i], side == 1 && i == j ? 1.0 : 0.0);
      }
      flux_hllc<DIM, Dual>(gamma, a, b, n, f);
      double s = 0.0;
      for (int i = 0; i < NV; ++i) s += g[i] * f[i].d;
      ub[j] += s;
    }
  }
}

// ubar_side += (dF/du_side)^T g for the sides selected, in reverse mode: the
// forward pass of flux_hllc (the same expressions, so the same wave-speed,
// star-region and side choices), then the cotangents back through it, about
// the cost of two flux evaluations where the dual-number columns cost ten.
template <int DIM>
DG_HD void hllc_vjp(double gamma, const double* ul, const double* ur, const double* n,
                    const double* g, bool with_left, bool with_right, double* ubar_l,
                    double* ubar_r) {
  constexpr int NV = DIM + 2;
  const double gm1 = gamma - 1.0;
  // forward
  double nrm = 0.0, nh[DIM], vl[DIM], vr[DIM];
  for (int d = 0; d < DIM; ++d) nrm += n[d] * n[d];
  nrm = sqrt(nrm);
  for (int d = 0; d < DIM; ++d) nh[d] = n[d] / nrm;
  const double rl = ul[0], rr = ur[0], El = ul[NV - 1], Er = ur[NV - 1];
  double v2l = 0.0, v2r = 0.0, vnl = 0.0, vnr = 0.0;
  for (int d = 0; d < DIM; ++d) {
    vl[d] = ul[d + 1] / rl;
    vr[d] = ur[d + 1] / rr;
    v2l += vl[d] * vl[d];
    v2r += vr[d] * vr[d];
    vnl += vl[d] * nh[d];
    vnr += vr[d] * nh[d];
  }
  const double el = El / rl, er = Er / rr;
  const double pl = (gamma - 1.0) * (El - 0.5 * rl * v2l);
  const double pr = (gamma - 1.0) * (Er - 0.5 * rr * v2r);
  const double cl = sqrt(gamma * pl / rl), cr = sqrt(gamma * pr / rr);
  const double sl = sqrt(rl), sr = sqrt(rr), ss = sl + sr;
  double vd[DIM], vroe2 = 0.0;
  for (int d = 0; d < DIM; ++d) {
    vd[d] = (sl * vl[d] + sr * vr[d]) / ss;
    vroe2 += vd[d] * vd[d];
  }
  const double vroe = (sl * vnl + sr * vnr) / ss;
  const double Hl = (El + pl) / rl, Hr = (Er + pr) / rr;
  const double Hroe = (sl * Hl + sr * Hr) / ss;
  const double croe = sqrt((gamma - 1.0) * (Hroe - 0.5 * vroe2));
  // SL = dmin(vnl - cl, vroe - croe), SR = dmax(vnr + cr, vroe + croe)
  const double aL = vnl - cl, bL = vroe - croe, aR = vnr + cr, bR = vroe + croe;
  const bool SL_roe = bL < aL, SR_roe = aR < bR;
  const double SL = SL_roe ? bL : aL, SR = SR_roe ? bR : aR;
  const double mL = SL - vnl, mR = SR - vnr;
  const bool star = !(SL >= 0) && !(SR <= 0);
  double Ss = 0.0, Den = 1.0;
  if (star) {
    Den = rl * mL - rr * mR;
    Ss = (pr - pl + rl * vnl * mL - rr * vnr * mR) / (rl * mL - rr * mR);
  }
  const bool left = star ? Ss >= 0 : SL >= 0;
  const double* uk = left ? ul : ur;

  // reverse: f = nrm * (fk(uk) [+ Sk (star terms)])
  double G[NV];
  for (int i = 0; i < NV; ++i) G[i] = g[i] * nrm;
  double ulb[NV], urb[NV], ukb[NV];
  for (int i = 0; i < NV; ++i) ulb[i] = urb[i] = ukb[i] = 0.0;
  flux_n_vjp<DIM>(gamma, uk, node_prims<DIM>(gamma, uk), nh, G, ukb);
  double rl_b = 0.0, rr_b = 0.0, El_b = 0.0, Er_b = 0.0, vnl_b = 0.0, vnr_b = 0.0;
  double pl_b = 0.0, pr_b = 0.0, el_b = 0.0, er_b = 0.0, vl_b[DIM], vr_b[DIM];
  for (int d = 0; d < DIM; ++d) vl_b[d] = vr_b[d] = 0.0;
  if (star) {
    const double* vk = left ? vl : vr;
    const double rk = left ? rl : rr, vnk = left ? vnl : vnr, ek = left ? el : er;
    const double pk = left ? pl : pr, Sk = left ? SL : SR, mk = left ? mL : mR;
    const double D = Sk - Ss, rs = rk * mk / D;
    const double X = Ss - vnk, rkmk = rk * mk, Z = pk / rkmk, Y = Ss + Z, es = ek + X * Y;
    // f0 = fk0 + Sk (rs - uk0), f_d = fk_d + Sk (rs (vk_d + X nh_d) - uk_d),
    // fE = fkE + Sk (rs es - ukE)
    double Sk_b = G[0] * (rs - uk[0]) + G[NV - 1] * (rs * es - uk[NV - 1]);
    double rs_b = Sk * (G[0] + G[NV - 1] * es), gn = 0.0, vk_b[DIM];
    for (int d = 0; d < DIM; ++d) {
      const double w = vk[d] + X * nh[d];
      Sk_b += G[d + 1] * (rs * w - uk[d + 1]);
      rs_b += Sk * G[d + 1] * w;
      vk_b[d] = Sk * rs * G[d + 1];
      gn += G[d + 1] * nh[d];
    }
    for (int i = 0; i < NV; ++i) ukb[i] -= Sk * G[i];
    const double es_b = Sk * rs * G[NV - 1];
    double X_b = Sk * rs * gn + es_b * Y, Ss_b = 0.0;
    const double Y_b = es_b * X, ek_b = es_b;
    // Y = Ss + pk / (rk mk)
    Ss_b += Y_b;
    const double pk_b = Y_b / rkmk, rkmk_b = -Y_b * Z / rkmk;
    // X = Ss - vnk
    Ss_b += X_b;
    double vnk_b = -X_b;
    // rs = rk mk / (Sk - Ss)
    const double rk_b = rs_b * mk / D + rkmk_b * mk;
    double mk_b = rs_b * rk / D + rkmk_b * rk;
    const double D_b = -rs_b * rs / D;
    Sk_b += D_b;
    Ss_b -= D_b;
    double SL_b = 0.0, SR_b = 0.0, mL_b = 0.0, mR_b = 0.0;
    if (left) {
      rl_b += rk_b; vnl_b += vnk_b; el_b += ek_b; pl_b += pk_b; SL_b += Sk_b; mL_b += mk_b;
      for (int d = 0; d < DIM; ++d) vl_b[d] += vk_b[d];
    } else {
      rr_b += rk_b; vnr_b += vnk_b; er_b += ek_b; pr_b += pk_b; SR_b += Sk_b; mR_b += mk_b;
      for (int d = 0; d < DIM; ++d) vr_b[d] += vk_b[d];
    }
    // Ss = (pr - pl + rl vnl mL - rr vnr mR) / (rl mL - rr mR)
    const double Num_b = Ss_b / Den, Den_b = -Ss_b * Ss / Den;
    pr_b += Num_b;
    pl_b -= Num_b;
    rl_b += Num_b * vnl * mL + Den_b * mL;
    vnl_b += Num_b * rl * mL;
    mL_b += Num_b * rl * vnl + Den_b * rl;
    rr_b -= Num_b * vnr * mR + Den_b * mR;
    vnr_b -= Num_b * rr * mR;
    mR_b -= Num_b * rr * vnr + Den_b * rr;
    // mL = SL - vnl, mR = SR - vnr
    SL_b += mL_b;
    vnl_b -= mL_b;
    SR_b += mR_b;
    vnr_b -= mR_b;
    // the wave-speed bounds, through the branch the forward took
    double vroe_b = 0.0, croe_b = 0.0, cl_b = 0.0, cr_b = 0.0;
    if (SL_roe) {
      vroe_b += SL_b;
      croe_b -= SL_b;
    } else {
      vnl_b += SL_b;
      cl_b -= SL_b;
    }
    if (SR_roe) {
      vroe_b += SR_b;
      croe_b += SR_b;
    } else {
      vnr_b += SR_b;
      cr_b += SR_b;
    }
    // croe = sqrt((gamma - 1)(Hroe - vroe2 / 2))
    const double W_b = croe_b / (2.0 * croe);
    const double Hroe_b = gm1 * W_b, vroe2_b = -0.5 * gm1 * W_b;
    double sl_b = Hroe_b * (Hl - Hroe) / ss + vroe_b * (vnl - vroe) / ss;
    double sr_b = Hroe_b * (Hr - Hroe) / ss + vroe_b * (vnr - vroe) / ss;
    const double Hl_b = Hroe_b * sl / ss, Hr_b = Hroe_b * sr / ss;
    vnl_b += vroe_b * sl / ss;
    vnr_b += vroe_b * sr / ss;
    for (int d = 0; d < DIM; ++d) {
      const double vd_b = 2.0 * vd[d] * vroe2_b;
      vl_b[d] += vd_b * sl / ss;
      vr_b[d] += vd_b * sr / ss;
      sl_b += vd_b * (vl[d] - vd[d]) / ss;
      sr_b += vd_b * (vr[d] - vd[d]) / ss;
    }
    // H = (E + p) / rho, s = sqrt(rho), c = sqrt(gamma p / rho)
    El_b += Hl_b / rl;
    pl_b += Hl_b / rl;
    rl_b -= Hl_b * Hl / rl;
    Er_b += Hr_b / rr;
    pr_b += Hr_b / rr;
    rr_b -= Hr_b * Hr / rr;
    rl_b += sl_b / (2.0 * sl);
    rr_b += sr_b / (2.0 * sr);
    if (cl_b != 0.0) {
      const double Q_b = cl_b / (2.0 * cl);
      pl_b += Q_b * gamma / rl;
      rl_b -= Q_b * cl * cl / rl;
    }
    if (cr_b != 0.0) {
      const double Q_b = cr_b / (2.0 * cr);
      pr_b += Q_b * gamma / rr;
      rr_b -= Q_b * cr * cr / rr;
    }
  }
  // e = E / rho, p = (gamma - 1)(E - rho |v|^2 / 2), v = m / rho
  El_b += el_b / rl + gm1 * pl_b;
  rl_b -= el_b * el / rl + 0.5 * gm1 * v2l * pl_b;
  Er_b += er_b / rr + gm1 * pr_b;
  rr_b -= er_b * er / rr + 0.5 * gm1 * v2r * pr_b;
  const double v2l_b = -0.5 * gm1 * rl * pl_b, v2r_b = -0.5 * gm1 * rr * pr_b;
  for (int d = 0; d < DIM; ++d) {
    vl_b[d] += 2.0 * vl[d] * v2l_b + vnl_b * nh[d];
    vr_b[d] += 2.0 * vr[d] * v2r_b + vnr_b * nh[d];
    ulb[d + 1] += vl_b[d] / rl;
    rl_b -= vl_b[d] * vl[d] / rl;
    urb[d + 1] += vr_b[d] / rr;
    rr_b -= vr_b[d] * vr[d] / rr;
  }
  ulb[0] += rl_b;
  ulb[NV - 1] += El_b;
  urb[0] += rr_b;
  urb[NV - 1] += Er_b;
  double* ukside = left ? ulb : urb;
  for (int i = 0; i < NV; ++i) ukside[i] += ukb[i];
  if (with_left)
    for (int i = 0; i < NV; ++i) ubar_l[i] += ulb[i];
  if (with_right)
    for (int i = 0; i < NV; ++i) ubar_r[i] += urb[i];
}

// ---------------------------------------------------------------------
// Navier-Stokes: gradient variables q = (rho, v, T = p / rho) and the
// Newtonian viscous flux (parabolic_impl.hpp grad_vars, viscous_flux_dir).
// ---------------------------------------------------------------------

template <int DIM>
DG_HD void grad_vars_jvp(double gamma, const double* u, const double* du, double* dq) {
  const double ir = 1.0 / u[0];
  double m2 = 0.0, mdm = 0.0;
  for (int k = 0; k < DIM; ++k) {
    m2 += u[k + 1] * u[k + 1];
    mdm += u[k + 1] * du[k + 1];
  }
  const double p = (gamma - 1.0) * (u[DIM + 1] - 0.5 * m2 * ir);
  const double dp = (gamma - 1.0) * (du[DIM + 1] - mdm * ir + 0.5 * m2 * ir * ir * du[0]);
  dq[0] = du[0];
  for (int k = 0; k < DIM; ++k) dq[k + 1] = (du[k + 1] - u[k + 1] * ir * du[0]) * ir;
  dq[DIM + 1] = (dp - p * ir * du[0]) * ir;
}

template <int DIM>
DG_HD void grad_vars_vjp(double gamma, const double* u, const double* qbar, double* ubar) {
  const double ir = 1.0 / u[0];
  double m2 = 0.0;
  for (int k = 0; k < DIM; ++k) m2 += u[k + 1] * u[k + 1];
  const double p = (gamma - 1.0) * (u[DIM + 1] - 0.5 * m2 * ir);
  const double Tb = qbar[DIM + 1];
  // T = p / rho
  const double pb = Tb * ir;
  double rb = qbar[0] - Tb * p * ir * ir;
  // p = (gamma - 1)(E - |m|^2 / (2 rho))
  ubar[DIM + 1] += (gamma - 1.0) * pb;
  rb += (gamma - 1.0) * 0.5 * m2 * ir * ir * pb;
  for (int k = 0; k < DIM; ++k) {
    ubar[k + 1] += -(gamma - 1.0) * u[k + 1] * ir * pb + qbar[k + 1] * ir;
    rb -= qbar[k + 1] * u[k + 1] * ir * ir;
  }
  ubar[0] += rb;
}

// Symmetric viscous stress tau (without mu) from the velocity gradients
// g[c][1 + k] = d v_k / d x_c, as viscous_flux_dir forms it.
template <int DIM>
DG_HD void stress(const double g[DIM][DIM + 2], double tau[DIM][DIM]) {
  if constexpr (DIM == 2) {
    tau[0][0] = (4.0 * g[0][1] - 2.0 * g[1][2]) / 3.0;
    tau[1][1] = (4.0 * g[1][2] - 2.0 * g[0][1]) / 3.0;
    tau[0][1] = tau[1][0] = g[1][1] + g[0][2];
  } else {
    tau[0][0] = (4.0 * g[0][1] - 2.0 * (g[1][2] + g[2][3])) / 3.0;
    tau[1][1] = (4.0 * g[1][2] - 2.0 * (g[0][1] + g[2][3])) / 3.0;
    tau[2][2] = (4.0 * g[2][3] - 2.0 * (g[0][1] + g[1][2])) / 3.0;
    tau[0][1] = tau[1][0] = g[1][1] + g[0][2];
    tau[0][2] = tau[2][0] = g[2][1] + g[0][3];
    tau[1][2] = tau[2][1] = g[2][2] + g[1][3];
  }
}

// gbar += (dtau/dg)^T taubar (taubar as a full, not necessarily symmetric,
// matrix of cotangents of the entries tau[c][k]).
template <int DIM>
DG_HD void stress_vjp(const double taubar[DIM][DIM], double gbar[DIM][DIM + 2]) {
  if constexpr (DIM == 2) {
    const double t00 = taubar[0][0], t11 = taubar[1][1], t01 = taubar[0][1] + taubar[1][0];
    gbar[0][1] += (4.0 * t00 - 2.0 * t11) / 3.0;
    gbar[1][2] += (4.0 * t11 - 2.0 * t00) / 3.0;
    gbar[1][1] += t01;
    gbar[0][2] += t01;
  } else {
    const double t00 = taubar[0][0], t11 = taubar[1][1], t22 = taubar[2][2];
    const double t01 = taubar[0][1] + taubar[1][0];
    const double t02 = taubar[0][2] + taubar[2][0];
    const double t12 = taubar[1][2] + taubar[2][1];
    gbar[0][1] += (4.0 * t00 - 2.0 * (t11 + t22)) / 3.0;
    gbar[1][2] += (4.0 * t11 - 2.0 * (t00 + t22)) / 3.0;
    gbar[2][3] += (4.0 * t22 - 2.0 * (t00 + t11)) / 3.0;
    gbar[1][1] += t01;
    gbar[0][2] += t01;
    gbar[2][1] += t02;
    gbar[0][3] += t02;
    gbar[2][2] += t12;
    gbar[1][3] += t12;
  }
}

// Tangent of the viscous flux in every direction c: f[c] = (0, mu tau[c][.],
// mu (v . tau[c][.] + kappa dT/dx_c)).
template <int DIM>
DG_HD void viscous_flux_jvp(double mu, double kappa, const double* q, const double g[DIM][DIM + 2],
                            const double* dq, const double dg[DIM][DIM + 2],
                            double df[DIM][DIM + 2]) {
  double tau[DIM][DIM], dtau[DIM][DIM];
  stress<DIM>(g, tau);
  stress<DIM>(dg, dtau);
  for (int c = 0; c < DIM; ++c) {
    df[c][0] = 0.0;
    double e = kappa * dg[c][DIM + 1];
    for (int k = 0; k < DIM; ++k) {
      df[c][k + 1] = mu * dtau[c][k];
      e += dq[k + 1] * tau[c][k] + q[k + 1] * dtau[c][k];
    }
    df[c][DIM + 1] = mu * e;
  }
}

// qbar += (df/dq)^T fbar, gbar += (df/dg)^T fbar over all directions c.
template <int DIM>
DG_HD void viscous_flux_vjp(double mu, double kappa, const double* q, const double g[DIM][DIM + 2],
                            const double fbar[DIM][DIM + 2], double* qbar,
                            double gbar[DIM][DIM + 2]) {
  double tau[DIM][DIM], taubar[DIM][DIM];
  stress<DIM>(g, tau);
  for (int c = 0; c < DIM; ++c) {
    const double fe = mu * fbar[c][DIM + 1];
    gbar[c][DIM + 1] += kappa * fe;
    for (int k = 0; k < DIM; ++k) {
      taubar[c][k] = mu * fbar[c][k + 1] + fe * q[k + 1];
      qbar[k + 1] += fe * tau[c][k];
    }
  }
  stress_vjp<DIM>(taubar, gbar);
}

}  // namespace dgsem
