// CPU checks of the hand-written tangents and adjoints of the pointwise
// physics (dgsem_amd/csrc/physics_ad.hpp):
//   * tangent against central finite differences of the forward function,
//   * adjoint against tangent by the dot-product identity
//     <g, F'(x) dx> = <F'(x)^T g, dx>,
// at random states, with both branches of the logarithmic mean.
// Prints one line per check, "<name> <fd_rel_err> <dot_rel_err>", and exits
// non-zero if a bound is exceeded.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../dgsem_amd/csrc/physics_ad.hpp"

using namespace dgsem;

namespace {

std::mt19937_64 rng(12345);
double uni(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

template <int DIM>
void random_state(double* u, double rho_lo = 0.5, double rho_hi = 2.0) {
  const double rho = uni(rho_lo, rho_hi), p = uni(0.5, 2.0);
  double v2 = 0.0;
  u[0] = rho;
  for (int k = 0; k < DIM; ++k) {
    const double v = uni(-1.0, 1.0);
    u[k + 1] = rho * v;
    v2 += v * v;
  }
  u[DIM + 1] = p / 0.4 + 0.5 * rho * v2;
}

// a state close to u (relative perturbation eps) so the means take the series branch
template <int DIM>
void near_state(const double* u, double* w, double eps) {
  for (int i = 0; i < DIM + 2; ++i) w[i] = u[i] * (1.0 + eps * uni(-1.0, 1.0));
}

int failures = 0;

void report(const char* name, double fd, double dot, double fd_tol, double dot_tol) {
  const bool ok = fd <= fd_tol && dot <= dot_tol;
  std::printf("%-28s fd %.3e  dot %.3e  %s\n", name, fd, dot, ok ? "ok" : "FAIL");
  if (!ok) ++failures;
}

double rel(double a, double b, double scale) { return std::fabs(a - b) / scale; }

// Chandrashekar flux of two conservative states (forward, as the kernels
// evaluate it from node primitives).
template <int DIM>
void cflux(const double* ua, const double* ub, const double* n, double* f) {
  flux_chandrashekar<DIM>(1.0 / 0.4, node_prims<DIM>(1.4, ua), node_prims<DIM>(1.4, ub), n, f);
}

template <int DIM>
void check_chandrashekar(bool near) {
  constexpr int NV = DIM + 2;
  double worst_fd = 0.0, worst_dot = 0.0;
  for (int trial = 0; trial < 200; ++trial) {
    double ua[NV], ub[NV], dua[NV], dub[NV], n[DIM], g[NV];
    random_state<DIM>(ua);
    if (near)
      near_state<DIM>(ua, ub, 1e-3);
    else
      random_state<DIM>(ub);
    for (int i = 0; i < NV; ++i) {
      dua[i] = uni(-1, 1);
      dub[i] = uni(-1, 1);
      g[i] = uni(-1, 1);
    }
    for (int k = 0; k < DIM; ++k) n[k] = uni(-1, 1);
    const NodePrims<DIM> qa = node_prims<DIM>(1.4, ua), qb = node_prims<DIM>(1.4, ub);
    PrimTangent<DIM> da, db;
    prims_jvp<DIM>(1.4, ua, qa, dua, da);
    prims_jvp<DIM>(1.4, ub, qb, dub, db);
    double df[NV];
    chandrashekar_jvp<DIM>(1.0 / 0.4, qa, qb, da, db, n, df);
    // finite differences of the full map (ua, ub) -> f
    const double h = 1e-6;
    double up[NV], um[NV], vp[NV], vm[NV], fp[NV], fm[NV];
    for (int i = 0; i < NV; ++i) {
      up[i] = ua[i] + h * dua[i];
      um[i] = ua[i] - h * dua[i];
      vp[i] = ub[i] + h * dub[i];
      vm[i] = ub[i] - h * dub[i];
    }
    cflux<DIM>(up, vp, n, fp);
    cflux<DIM>(um, vm, n, fm);
    double scale = 1e-300;
    for (int i = 0; i < NV; ++i) scale = std::fmax(scale, std::fabs(df[i]));
    for (int i = 0; i < NV; ++i)
      worst_fd = std::fmax(worst_fd, rel((fp[i] - fm[i]) / (2 * h), df[i], scale));
    // adjoint: q-space cotangents, then back to u
    PrimTangent<DIM> ab{}, bb{};
    chandrashekar_vjp<DIM>(1.0 / 0.4, qa, qb, n, g, ab, bb);
    double uab[NV] = {}, ubb[NV] = {};
    prims_vjp<DIM>(1.4, qa, ab, 0.0, uab);
    prims_vjp<DIM>(1.4, qb, bb, 0.0, ubb);
    double lhs = 0.0, rhs = 0.0, mag = 0.0;
    for (int i = 0; i < NV; ++i) {
      lhs += g[i] * df[i];
      rhs += uab[i] * dua[i] + ubb[i] * dub[i];
      mag += std::fabs(uab[i] * dua[i]) + std::fabs(ubb[i] * dub[i]);
    }
    worst_dot = std::fmax(worst_dot, std::fabs(lhs - rhs) / mag);
  }
  char name[64];
  std::snprintf(name, sizeof(name), "chandrashekar%dD%s", DIM, near ? " series" : " log");
  report(name, worst_fd, worst_dot, 1e-6, 1e-13);
}

template <int DIM>
void llf(const double* ul, const double* ur, const double* n, double* f) {
  flux_llf_prims<DIM>(ul, face_prims<DIM>(1.4, ul), ur, face_prims<DIM>(1.4, ur), n, f);
}

template <int DIM>
void check_llf() {
  constexpr int NV = DIM + 2;
  double worst_fd = 0.0, worst_dot = 0.0;
  for (int trial = 0; trial < 200; ++trial) {
    double ul[NV], ur[NV], dl[NV], dr[NV], n[DIM], g[NV];
    random_state<DIM>(ul);
    random_state<DIM>(ur);
    for (int i = 0; i < NV; ++i) {
      dl[i] = uni(-1, 1);
      dr[i] = uni(-1, 1);
      g[i] = uni(-1, 1);
    }
    for (int k = 0; k < DIM; ++k) n[k] = uni(-1, 1);
    const NodePrims<DIM> ql = node_prims<DIM>(1.4, ul), qr = node_prims<DIM>(1.4, ur);
    double df[NV];
    llf_jvp<DIM>(1.4, ul, ql, ur, qr, dl, dr, n, df);
    const double h = 1e-7;
    double lp[NV], lm[NV], rp[NV], rm[NV], fp[NV], fm[NV];
    for (int i = 0; i < NV; ++i) {
      lp[i] = ul[i] + h * dl[i];
      lm[i] = ul[i] - h * dl[i];
      rp[i] = ur[i] + h * dr[i];
      rm[i] = ur[i] - h * dr[i];
    }
    llf<DIM>(lp, rp, n, fp);
    llf<DIM>(lm, rm, n, fm);
    double scale = 1e-300;
    for (int i = 0; i < NV; ++i) scale = std::fmax(scale, std::fabs(df[i]));
    for (int i = 0; i < NV; ++i)
      worst_fd = std::fmax(worst_fd, rel((fp[i] - fm[i]) / (2 * h), df[i], scale));
    double bl[NV] = {}, br[NV] = {};
    llf_vjp_side<DIM>(1.4, ul, ql, ur, qr, n, g, true, bl);
    llf_vjp_side<DIM>(1.4, ul, ql, ur, qr, n, g, false, br);
    double lhs = 0.0, rhs = 0.0, mag = 0.0;
    for (int i = 0; i < NV; ++i) {
      lhs += g[i] * df[i];
      rhs += bl[i] * dl[i] + br[i] * dr[i];
      mag += std::fabs(bl[i] * dl[i]) + std::fabs(br[i] * dr[i]);
    }
    worst_dot = std::fmax(worst_dot, std::fabs(lhs - rhs) / mag);
  }
  char name[64];
  std::snprintf(name, sizeof(name), "llf%dD", DIM);
  report(name, worst_fd, worst_dot, 1e-6, 1e-13);
}

// HLLC through dual numbers: states drawn so that every branch occurs (the
// supersonic sides and the star region on either side of the contact).
template <int DIM>
void check_hllc() {
  constexpr int NV = DIM + 2;
  double worst_fd = 0.0, worst_dot = 0.0, worst_cols = 0.0;
  for (int trial = 0; trial < 400; ++trial) {
    double ul[NV], ur[NV], dl[NV], dr[NV], n[DIM], g[NV];
    random_state<DIM>(ul);
    random_state<DIM>(ur);
    // a bulk velocity along the normal, up to supersonic either way
    for (int k = 0; k < DIM; ++k) n[k] = uni(-1, 1);
    const double shift = uni(-4.0, 4.0);
    for (int s = 0; s < 2; ++s) {
      double* u = s == 0 ? ul : ur;
      double m2 = 0.0;
      for (int k = 0; k < DIM; ++k) m2 += u[k + 1] * u[k + 1];
      const double p = 0.4 * (u[NV - 1] - 0.5 * m2 / u[0]);
      m2 = 0.0;
      for (int k = 0; k < DIM; ++k) {
        u[k + 1] += u[0] * shift * n[k];
        m2 += u[k + 1] * u[k + 1];
      }
      u[NV - 1] = p / 0.4 + 0.5 * m2 / u[0];
    }
    for (int i = 0; i < NV; ++i) {
      dl[i] = uni(-1, 1);
      dr[i] = uni(-1, 1);
      g[i] = uni(-1, 1);
    }
    double df[NV];
    hllc_jvp<DIM>(1.4, ul, ur, dl, dr, n, df);
    const double h = 1e-7;
    double lp[NV], lm[NV], rp[NV], rm[NV], fp[NV], fm[NV];
    for (int i = 0; i < NV; ++i) {
      lp[i] = ul[i] + h * dl[i];
      lm[i] = ul[i] - h * dl[i];
      rp[i] = ur[i] + h * dr[i];
      rm[i] = ur[i] - h * dr[i];
    }
    flux_hllc<DIM>(1.4, lp, rp, n, fp);
    flux_hllc<DIM>(1.4, lm, rm, n, fm);
    double scale = 1e-300;
    for (int i = 0; i < NV; ++i) scale = std::fmax(scale, std::fabs(df[i]));
    for (int i = 0; i < NV; ++i)
      worst_fd = std::fmax(worst_fd, rel((fp[i] - fm[i]) / (2 * h), df[i], scale));
    double bl[NV] = {}, br[NV] = {}, cl[NV] = {}, cr[NV] = {};
    hllc_vjp<DIM>(1.4, ul, ur, n, g, true, true, bl, br);
    // the reverse-mode adjoint against the dual-number columns
    hllc_vjp_columns<DIM>(1.4, ul, ur, n, g, true, true, cl, cr);
    double cmag = 1e-300;
    for (int i = 0; i < NV; ++i) cmag = std::fmax(cmag, std::fmax(std::fabs(cl[i]), std::fabs(cr[i])));
    for (int i = 0; i < NV; ++i)
      worst_cols = std::fmax(worst_cols, std::fmax(std::fabs(bl[i] - cl[i]), std::fabs(br[i] - cr[i])) / cmag);
    double lhs = 0.0, rhs = 0.0, mag = 0.0;
    for (int i = 0; i < NV; ++i) {
      lhs += g[i] * df[i];
      rhs += bl[i] * dl[i] + br[i] * dr[i];
      mag += std::fabs(bl[i] * dl[i]) + std::fabs(br[i] * dr[i]);
    }
    worst_dot = std::fmax(worst_dot, std::fabs(lhs - rhs) / mag);
  }
  char name[64];
  std::snprintf(name, sizeof(name), "hllc%dD", DIM);
  report(name, worst_fd, worst_dot, 1e-6, 1e-13);
  std::snprintf(name, sizeof(name), "hllc%dD reverse vs columns", DIM);
  report(name, 0.0, worst_cols, 1e-6, 1e-13);
}

template <int DIM>
void visc(const double* u, const double g[DIM][DIM + 2], double f[DIM][DIM + 2]) {
  double q[DIM + 2];
  // gradient variables as the kernels form them
  const double ir = 1.0 / u[0];
  double m2 = 0.0;
  for (int k = 0; k < DIM; ++k) m2 += u[k + 1] * u[k + 1];
  const double p = 0.4 * (u[DIM + 1] - 0.5 * m2 * ir);
  q[0] = u[0];
  for (int k = 0; k < DIM; ++k) q[k + 1] = u[k + 1] * ir;
  q[DIM + 1] = p * ir;
  double tau[DIM][DIM];
  stress<DIM>(g, tau);
  for (int c = 0; c < DIM; ++c) {
    f[c][0] = 0.0;
    double e = 0.7 * g[c][DIM + 1];
    for (int k = 0; k < DIM; ++k) {
      f[c][k + 1] = 0.01 * tau[c][k];
      e += q[k + 1] * tau[c][k];
    }
    f[c][DIM + 1] = 0.01 * e;
  }
}

// The viscous flux as a function of (u, g) through the gradient variables.
template <int DIM>
void check_viscous() {
  constexpr int NV = DIM + 2;
  const double mu = 0.01, kappa = 0.7;
  double worst_fd = 0.0, worst_dot = 0.0;
  for (int trial = 0; trial < 200; ++trial) {
    double u[NV], du[NV], g[DIM][NV], dg[DIM][NV], fb[DIM][NV];
    random_state<DIM>(u);
    for (int i = 0; i < NV; ++i) du[i] = uni(-1, 1);
    for (int c = 0; c < DIM; ++c)
      for (int i = 0; i < NV; ++i) {
        g[c][i] = uni(-1, 1);
        dg[c][i] = uni(-1, 1);
        fb[c][i] = uni(-1, 1);
      }
    double q[NV], dq[NV], df[DIM][NV];
    const double ir = 1.0 / u[0];
    double m2 = 0.0;
    for (int k = 0; k < DIM; ++k) m2 += u[k + 1] * u[k + 1];
    q[0] = u[0];
    for (int k = 0; k < DIM; ++k) q[k + 1] = u[k + 1] * ir;
    q[DIM + 1] = 0.4 * (u[DIM + 1] - 0.5 * m2 * ir) * ir;
    grad_vars_jvp<DIM>(1.4, u, du, dq);
    viscous_flux_jvp<DIM>(mu, kappa, q, g, dq, dg, df);
    const double h = 1e-6;
    double up[NV], um[NV], gp[DIM][NV], gm[DIM][NV], fp[DIM][NV], fm[DIM][NV];
    for (int i = 0; i < NV; ++i) {
      up[i] = u[i] + h * du[i];
      um[i] = u[i] - h * du[i];
    }
    for (int c = 0; c < DIM; ++c)
      for (int i = 0; i < NV; ++i) {
        gp[c][i] = g[c][i] + h * dg[c][i];
        gm[c][i] = g[c][i] - h * dg[c][i];
      }
    visc<DIM>(up, gp, fp);
    visc<DIM>(um, gm, fm);
    double scale = 1e-300;
    for (int c = 0; c < DIM; ++c)
      for (int i = 0; i < NV; ++i) scale = std::fmax(scale, std::fabs(df[c][i]));
    for (int c = 0; c < DIM; ++c)
      for (int i = 0; i < NV; ++i)
        worst_fd = std::fmax(worst_fd, rel((fp[c][i] - fm[c][i]) / (2 * h), df[c][i], scale));
    double qb[NV] = {}, gb[DIM][NV] = {}, ub[NV] = {};
    viscous_flux_vjp<DIM>(mu, kappa, q, g, fb, qb, gb);
    grad_vars_vjp<DIM>(1.4, u, qb, ub);
    double lhs = 0.0, rhs = 0.0, mag = 0.0;
    for (int c = 0; c < DIM; ++c)
      for (int i = 0; i < NV; ++i) {
        lhs += fb[c][i] * df[c][i];
        rhs += gb[c][i] * dg[c][i];
        mag += std::fabs(gb[c][i] * dg[c][i]);
      }
    for (int i = 0; i < NV; ++i) {
      rhs += ub[i] * du[i];
      mag += std::fabs(ub[i] * du[i]);
    }
    worst_dot = std::fmax(worst_dot, std::fabs(lhs - rhs) / mag);
  }
  char name[64];
  std::snprintf(name, sizeof(name), "viscous%dD", DIM);
  report(name, worst_fd, worst_dot, 1e-6, 1e-13);
}

}  // namespace

int main() {
  check_chandrashekar<2>(false);
  check_chandrashekar<2>(true);
  check_chandrashekar<3>(false);
  check_chandrashekar<3>(true);
  check_llf<2>();
  check_llf<3>();
  check_hllc<2>();
  check_hllc<3>();
  check_viscous<2>();
  check_viscous<3>();
  return failures ? 1 : 0;
}
