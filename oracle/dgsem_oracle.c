/*
 * dgsem_oracle.c -- CPU restatement of the reference DGSEM right-hand side
 * and SSPRK3 step.  TEST INFRASTRUCTURE ONLY (see dgsem_oracle.h).
 *
 * Every routine names the reference file:line it restates.  Operation order
 * and rounding follow the reference source expression by expression; build
 * with -O2 -ffp-contract=off.  The per-element loops carry an OpenMP pragma
 * where the reference issues a Kokkos parallel_for over elements, so the
 * timed CPU baseline has the reference's parallel structure.
 *
 * Pieces with no counterpart in the reference (HLLC with a normal vector,
 * the positivity floor, the Gaussian Sedov initial state) are marked
 * "build extension" and follow the structure of the nearest reference code.
 */
#include "dgsem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 16
#define PI_D 3.14159265358979323846

struct orc_case {
  int dim, N, n, nn, nvars, nx, ny, nz;
  int64_t E;
  int eq, volume, fv_flux, surface_flux, volume_flux, nthreads;
  double gamma, mu, prandtl, kappa;
  double alpha_max, alpha_min, floor_value;
  double nodes[MAXN], weights[MAXN], inv_weights[MAXN];
  double D[MAXN * MAXN], Dsplit[MAXN * MAXN], Dhat[MAXN * MAXN];
  double Vinv[MAXN * MAXN], lhat_l[MAXN], lhat_r[MAXN];
  double *coords, *jac, *Ja, *invJ, *sn1, *sn2, *sn3;
  int64_t* nbr;
  double *sflux, *grad, *gflux, *vflux, *vsflux; /* scratch */
  double *tmp1, *tmp2, *du;                    /* SSPRK3 scratch */
  int bc[6];                                   /* per face: ORC_BC_*, -x,+x,-y,+y,-z,+z */
  double bc_state[6][5];                       /* constant Dirichlet states */
  const double* bc_values[6];                  /* per-node Dirichlet states, or NULL */
  const double* alpha_override;                /* test hook: fixed blending factor, or NULL */
};

/* ------------------------------------------------------------------------ */
/* Basis: src/polynomial_basis/basis_detail.hpp                              */
/* ------------------------------------------------------------------------ */

/* evaluate_q_and_l, basis_detail.hpp:27-50 */
static void q_and_l(size_t polydeg, double x, double* q, double* qder,
                    double* LN) {
  double L_Nm2 = 1.0, L_Nm1 = x, Lder_Nm2 = 0.0, Lder_Nm1 = 1.0;
  double L_N = 0.0, Lder_N = 0.0;
  for (size_t i = 2; i <= polydeg; ++i) {
    L_N = ((double)(2.0 * i - 1.0) * x * L_Nm1 - (double)(i - 1.0) * L_Nm2) /
          (double)i;
    Lder_N = Lder_Nm2 + (double)(2.0 * i - 1.0) * L_Nm1;
    L_Nm2 = L_Nm1;
    L_Nm1 = L_N;
    Lder_Nm2 = Lder_Nm1;
    Lder_Nm1 = Lder_N;
  }
  *q = (double)(2.0 * polydeg + 1.0) / ((double)polydeg + 1.0) *
       (x * L_N - L_Nm2);
  *qder = (2.0 * polydeg + 1.0) * L_N;
  *LN = L_N;
}

/* gauss_lobatto_nodes_weights, basis_detail.hpp:52-109 */
static void gll(int nnodes, double* nodes, double* w, double* iw) {
  const size_t n_iterations = 20;
  const size_t Polydeg = (size_t)nnodes - 1;
  const double tolerance = 2.0 * 2.220446049250313080847e-16;
  if (nnodes == 1) {
    nodes[0] = 0;
    w[0] = 2;
    iw[0] = 1 / w[0];
    return;
  }
  nodes[0] = -1.0;
  nodes[nnodes - 1] = 1.0;
  w[0] = 2.0 / ((double)Polydeg * ((double)Polydeg + 1.0));
  w[nnodes - 1] = w[0];
  {
    double cont1 = PI_D / (double)Polydeg;
    double cont2 = 3.0 / (8.0 * (double)Polydeg * PI_D);
    for (size_t i = 0; i < (size_t)(nnodes / 2 - 1); ++i) {
      nodes[i + 1] = -cos(cont1 * ((double)(i + 1) + 0.25) -
                          cont2 / ((double)(i + 1) + 0.25));
      for (size_t it = 0; it < n_iterations; ++it) {
        double q, qder, LN;
        q_and_l(Polydeg, nodes[i + 1], &q, &qder, &LN);
        double dx = -q / qder;
        nodes[i + 1] += dx;
        if (fabs(dx) < tolerance * fabs(nodes[i + 1])) break;
      }
      double q, qder, LN;
      q_and_l(Polydeg, nodes[i + 1], &q, &qder, &LN);
      w[i + 1] = w[0] / (LN * LN);
      nodes[nnodes - 2 - i] = -nodes[i + 1];
      w[nnodes - 2 - i] = w[i + 1];
    }
  }
  if (nnodes % 2 == 1) {
    double q, qder, LN;
    nodes[nnodes / 2] = 0.0;
    q_and_l(Polydeg, 0.0, &q, &qder, &LN);
    w[nnodes / 2] = w[0] / (LN * LN);
  }
  for (int i = 0; i < nnodes; ++i) iw[i] = 1.0 / w[i];
}

/* barycenteric_weights, basis_detail.hpp:111-128 */
static void bary(int n, const double* x, double* wb) {
  for (int i = 0; i < n; ++i) wb[i] = 1.0;
  for (int j = 1; j < n; ++j)
    for (int k = 0; k < j; ++k) {
      wb[k] *= (x[k] - x[j]);
      wb[j] *= (x[j] - x[k]);
    }
  for (int i = 0; i < n; ++i) wb[i] = 1.0 / wb[i];
}

/* lagrange_interpolating_polynomials + calc_lhat, basis_detail.hpp:130-168 */
static void lhat_at(int n, double x, const double* nodes, const double* w,
                    double* out) {
  double wb[MAXN], tot = 0.0;
  int hit = -1;
  bary(n, nodes, wb);
  for (int i = 0; i < n; ++i) out[i] = 0.0;
  for (int i = 0; i < n; ++i)
    if (fabs(x - nodes[i]) < 2.220446049250313080847e-16) {
      hit = i;
      break;
    }
  if (hit >= 0) {
    out[hit] = 1.0;
  } else {
    for (int i = 0; i < n; ++i) out[i] = wb[i] / (x - nodes[i]);
    for (int i = 0; i < n; ++i) tot += out[i];
    for (int i = 0; i < n; ++i) out[i] /= tot;
  }
  for (int i = 0; i < n; ++i) out[i] /= w[i];
}

/* polynomial_derivative_matrix, basis_detail.hpp:170-188 */
static void dmat(int n, const double* x, double* D) {
  double wb[MAXN];
  bary(n, x, wb);
  for (int i = 0; i < n * n; ++i) D[i] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j) {
        D[i * n + j] = wb[j] / wb[i] * 1.0 / (x[i] - x[j]);
        D[i * n + i] -= D[i * n + j];
      }
}

/* legendre_polynomial_and_derivative, basis_detail.hpp:217-248 */
static double legendre_norm(size_t n, double x) {
  double poly = 0.0;
  if (n == 0) {
    poly = 1.0;
  } else if (n == 1) {
    poly = x;
  } else {
    double Pm2 = 1.0, Pm1 = x;
    for (size_t i = 2; i <= n; ++i) {
      poly = ((double)(2.0 * i - 1.0) * x * Pm1 - (double)(i - 1.0) * Pm2) /
             (double)i;
      Pm2 = Pm1;
      Pm1 = poly;
    }
  }
  return poly * sqrt((double)n + 0.5);
}

/* Mat::inverse, Gauss-Jordan with partial pivoting,
 * src/base/container_fixed.hpp:117-162 */
static void mat_inverse(int n, const double* Ain, double* inv) {
  double A[MAXN * MAXN];
  memcpy(A, Ain, sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i) inv[i] = 0.0;
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int col = 0; col < n; ++col) {
    int pivot = col;
    double maxv = (A[col * n + col] < 0 ? -A[col * n + col] : A[col * n + col]);
    for (int row = col + 1; row < n; ++row) {
      double v = (A[row * n + col] < 0 ? -A[row * n + col] : A[row * n + col]);
      if (v > maxv) {
        maxv = v;
        pivot = row;
      }
    }
    if (pivot != col)
      for (int k = 0; k < n; ++k) {
        double t = A[col * n + k];
        A[col * n + k] = A[pivot * n + k];
        A[pivot * n + k] = t;
        t = inv[col * n + k];
        inv[col * n + k] = inv[pivot * n + k];
        inv[pivot * n + k] = t;
      }
    double diag = A[col * n + col];
    for (int k = 0; k < n; ++k) {
      A[col * n + k] /= diag;
      inv[col * n + k] /= diag;
    }
    for (int row = 0; row < n; ++row) {
      if (row == col) continue;
      double factor = A[row * n + col];
      for (int k = 0; k < n; ++k) {
        A[row * n + k] -= factor * A[col * n + k];
        inv[row * n + k] -= factor * inv[col * n + k];
      }
    }
  }
}

/* LobattoLegendreBasis::initialize, basis.hpp:69-153 (calc_dsplit :190-200,
 * calc_dhat :202-215, inverse_vandermonde_legendre :250-262) */
void orc_basis(int N, double* nodes, double* weights, double* inv_weights,
               double* D, double* Dsplit, double* Dhat, double* Vinv,
               double* lhat_left, double* lhat_right) {
  const int n = N + 1;
  double V[MAXN * MAXN];
  gll(n, nodes, weights, inv_weights);
  dmat(n, nodes, D);
  for (int i = 0; i < n * n; ++i) Dsplit[i] = 2.0 * D[i];
  Dsplit[0] += 1.0 / weights[0];
  Dsplit[(n - 1) * n + (n - 1)] -= 1.0 / weights[n - 1];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Dhat[j * n + i] = D[i * n + j];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Dhat[j * n + i] *= -weights[i] / weights[j];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i * n + j] = legendre_norm((size_t)j, nodes[i]);
  mat_inverse(n, V, Vinv);
  lhat_at(n, -1.0, nodes, weights, lhat_left);
  lhat_at(n, 1.0, nodes, weights, lhat_right);
}

/* ------------------------------------------------------------------------ */
/* Pointwise physics: src/base/math_utils.hpp, src/equations, numerical_flux */
/* ------------------------------------------------------------------------ */

/* ln_mean, math_utils.hpp:8-24 */
double orc_ln_mean(double x, double y) {
  const double epsilon_f2 = 1.0e-4;
  double numerator = x * (x - 2.0 * y) + y * y;
  double denominator = x * (x + 2.0 * y) + y * y;
  double f2 = numerator / denominator;
  if (f2 < epsilon_f2) {
    double poly = 2.0 + f2 * (2.0 / 3.0 + f2 * (2.0 / 5.0 + f2 * (2.0 / 7.0)));
    return (x + y) / poly;
  }
  return (y - x) / log(y / x);
}

static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

/* CompressibleEuler2D::flux(u, dim), compressible_euler2D.hpp:25-45 */
static void euler2d_flux_dim(double g, const double* u, int dim, double* f) {
  const double rho = u[0], rhou = u[1], rhov = u[2], rhoE = u[3];
  const double uv = rhou / rho, vv = rhov / rho;
  const double kinetic = 0.5 * rho * (uv * uv + vv * vv);
  const double p = (g - 1.0) * (rhoE - kinetic);
  if (dim == 0) {
    f[0] = rhou; f[1] = rhou * uv + p; f[2] = rhou * vv; f[3] = uv * (rhoE + p);
  } else {
    f[0] = rhov; f[1] = rhov * uv; f[2] = rhov * vv + p; f[3] = vv * (rhoE + p);
  }
}

/* CompressibleEuler3D::flux(u, dim), compressible_euler3D.hpp:26-52 */
static void euler3d_flux_dim(double g, const double* u, int dim, double* f) {
  const double rho = u[0], rhou = u[1], rhov = u[2], rhow = u[3], rhoE = u[4];
  const double uv = rhou / rho, vv = rhov / rho, wv = rhow / rho;
  const double kinetic = 0.5 * rho * (uv * uv + vv * vv + wv * wv);
  const double p = (g - 1.0) * (rhoE - kinetic);
  if (dim == 0) {
    f[0] = rhou; f[1] = rhou * uv + p; f[2] = rhou * vv; f[3] = rhou * wv;
    f[4] = uv * (rhoE + p);
  } else if (dim == 1) {
    f[0] = rhov; f[1] = rhov * uv; f[2] = rhov * vv + p; f[3] = rhov * wv;
    f[4] = vv * (rhoE + p);
  } else {
    f[0] = rhow; f[1] = rhow * uv; f[2] = rhow * vv; f[3] = rhow * wv + p;
    f[4] = wv * (rhoE + p);
  }
}

/* flux(u, normal): compressible_euler2D.hpp:47-57, compressible_euler3D.hpp:54-67 */
void orc_physical_flux(int dim, double g, const double* u, const double* n,
                       double* f) {
  double fx[5], fy[5], fz[5];
  if (dim == 2) {
    euler2d_flux_dim(g, u, 0, fx);
    euler2d_flux_dim(g, u, 1, fy);
    for (int v = 0; v < 4; ++v) f[v] = n[0] * fx[v] + n[1] * fy[v];
  } else {
    euler3d_flux_dim(g, u, 0, fx);
    euler3d_flux_dim(g, u, 1, fy);
    euler3d_flux_dim(g, u, 2, fz);
    for (int v = 0; v < 5; ++v)
      f[v] = n[0] * fx[v] + n[1] * fy[v] + n[2] * fz[v];
  }
}

/* get_wave_speed(u_ll, u_rr, normal): compressible_euler2D.hpp:68-106,
 * compressible_euler3D.hpp:121-166 */
static double wave_speed(int dim, double g, const double* ul, const double* ur,
                         const double* n) {
  if (dim == 2) {
    const double ul0 = ul[1] / ul[0], vl0 = ul[2] / ul[0];
    const double ur0 = ur[1] / ur[0], vr0 = ur[2] / ur[0];
    const double p_ll = (g - 1.0) * (ul[3] - 0.5 * ul[0] * (ul0 * ul0 + vl0 * vl0));
    const double p_rr = (g - 1.0) * (ur[3] - 0.5 * ur[0] * (ur0 * ur0 + vr0 * vr0));
    const double a_ll = sqrt(g * p_ll / ul[0]);
    const double a_rr = sqrt(g * p_rr / ur[0]);
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1]);
    const double vn_ll = ul0 * n[0] + vl0 * n[1];
    const double vn_rr = ur0 * n[0] + vr0 * n[1];
    return dmax(fabs(vn_ll), fabs(vn_rr)) + dmax(a_ll, a_rr) * nn;
  } else {
    const double ul0 = ul[1] / ul[0], vl0 = ul[2] / ul[0], wl0 = ul[3] / ul[0];
    const double ur0 = ur[1] / ur[0], vr0 = ur[2] / ur[0], wr0 = ur[3] / ur[0];
    const double p_ll =
        (g - 1.0) * (ul[4] - 0.5 * ul[0] * (ul0 * ul0 + vl0 * vl0 + wl0 * wl0));
    const double p_rr =
        (g - 1.0) * (ur[4] - 0.5 * ur[0] * (ur0 * ur0 + vr0 * vr0 + wr0 * wr0));
    const double a_ll = sqrt(g * p_ll / ul[0]);
    const double a_rr = sqrt(g * p_rr / ur[0]);
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double vn_ll = ul0 * n[0] + vl0 * n[1] + wl0 * n[2];
    const double vn_rr = ur0 * n[0] + vr0 * n[1] + wr0 * n[2];
    return dmax(fabs(vn_ll), fabs(vn_rr)) + dmax(a_ll, a_rr) * nn;
  }
}

/* LaxFriedrichsFlux<Euler2D/3D/NS>::numerical_flux(eq,uL,uR,normal),
 * numerical_flux.hpp:54-68,216-227 + numerical_flux_detail.hpp:9-21 */
void orc_flux_llf(int dim, double g, const double* ul, const double* ur,
                  const double* n, double* f) {
  const int nv = dim + 2;
  double fl[5], fr[5];
  const double max_speed = wave_speed(dim, g, ul, ur, n);
  orc_physical_flux(dim, g, ul, n, fl);
  orc_physical_flux(dim, g, ur, n, fr);
  for (int i = 0; i < nv; ++i)
    f[i] = 0.5 * (fl[i] + fr[i]) - 0.5 * max_speed * (ur[i] - ul[i]);
}

/* ChandrashekarFlux<Euler2D>, numerical_flux.hpp:436-500 and
 * ChandrashekarFlux<Euler3D>, numerical_flux.hpp:520-594 */
void orc_flux_chandrashekar(int dim, double g, const double* u_ll,
                            const double* u_rr, const double* normal,
                            double* f) {
  if (dim == 2) {
    const double rho_ll = u_ll[0], rhou_ll = u_ll[1], rhov_ll = u_ll[2],
                 rhoE_ll = u_ll[3];
    const double rho_rr = u_rr[0], rhou_rr = u_rr[1], rhov_rr = u_rr[2],
                 rhoE_rr = u_rr[3];
    const double v1_ll = rhou_ll / rho_ll, v2_ll = rhov_ll / rho_ll;
    const double v1_rr = rhou_rr / rho_rr, v2_rr = rhov_rr / rho_rr;
    const double v_dot_n_ll = v1_ll * normal[0] + v2_ll * normal[1];
    const double v_dot_n_rr = v1_rr * normal[0] + v2_rr * normal[1];
    const double specific_kin_ll = 0.5 * (v1_ll * v1_ll + v2_ll * v2_ll);
    const double specific_kin_rr = 0.5 * (v1_rr * v1_rr + v2_rr * v2_rr);
    const double p_ll =
        (g - 1.0) * (rhoE_ll - 0.5 * rho_ll * (v1_ll * v1_ll + v2_ll * v2_ll));
    const double p_rr =
        (g - 1.0) * (rhoE_rr - 0.5 * rho_rr * (v1_rr * v1_rr + v2_rr * v2_rr));
    const double beta_ll = 0.5 * rho_ll / p_ll;
    const double beta_rr = 0.5 * rho_rr / p_rr;
    const double rho_avg = 0.5 * (rho_ll + rho_rr);
    const double rho_mean = orc_ln_mean(rho_ll, rho_rr);
    const double beta_mean = orc_ln_mean(beta_ll, beta_rr);
    const double beta_avg = 0.5 * (beta_ll + beta_rr);
    const double v1_avg = 0.5 * (v1_ll + v1_rr);
    const double v2_avg = 0.5 * (v2_ll + v2_rr);
    const double p_mean = 0.5 * rho_avg / beta_avg;
    const double velocity_square_avg = specific_kin_ll + specific_kin_rr;
    const double f1 = rho_mean * 0.5 * (v_dot_n_ll + v_dot_n_rr);
    const double f2 = f1 * v1_avg + p_mean * normal[0];
    const double f3 = f1 * v2_avg + p_mean * normal[1];
    const double f4 =
        f1 * 0.5 * (1.0 / (g - 1.0) / beta_mean - velocity_square_avg) +
        f2 * v1_avg + f3 * v2_avg;
    f[0] = f1; f[1] = f2; f[2] = f3; f[3] = f4;
  } else {
    const double rho_ll = u_ll[0], rhou_ll = u_ll[1], rhov_ll = u_ll[2],
                 rhow_ll = u_ll[3], rhoE_ll = u_ll[4];
    const double rho_rr = u_rr[0], rhou_rr = u_rr[1], rhov_rr = u_rr[2],
                 rhow_rr = u_rr[3], rhoE_rr = u_rr[4];
    const double v1_ll = rhou_ll / rho_ll, v2_ll = rhov_ll / rho_ll,
                 v3_ll = rhow_ll / rho_ll;
    const double v1_rr = rhou_rr / rho_rr, v2_rr = rhov_rr / rho_rr,
                 v3_rr = rhow_rr / rho_rr;
    const double v_dot_n_ll = v1_ll * normal[0] + v2_ll * normal[1] + v3_ll * normal[2];
    const double v_dot_n_rr = v1_rr * normal[0] + v2_rr * normal[1] + v3_rr * normal[2];
    const double specific_kin_ll = 0.5 * (v1_ll * v1_ll + v2_ll * v2_ll + v3_ll * v3_ll);
    const double specific_kin_rr = 0.5 * (v1_rr * v1_rr + v2_rr * v2_rr + v3_rr * v3_rr);
    const double p_ll =
        (g - 1.0) *
        (rhoE_ll - 0.5 * rho_ll * (v1_ll * v1_ll + v2_ll * v2_ll + v3_ll * v3_ll));
    const double p_rr =
        (g - 1.0) *
        (rhoE_rr - 0.5 * rho_rr * (v1_rr * v1_rr + v2_rr * v2_rr + v3_rr * v3_rr));
    const double beta_ll = 0.5 * rho_ll / p_ll;
    const double beta_rr = 0.5 * rho_rr / p_rr;
    const double rho_avg = 0.5 * (rho_ll + rho_rr);
    const double rho_mean = orc_ln_mean(rho_ll, rho_rr);
    const double beta_mean = orc_ln_mean(beta_ll, beta_rr);
    const double beta_avg = 0.5 * (beta_ll + beta_rr);
    const double v1_avg = 0.5 * (v1_ll + v1_rr);
    const double v2_avg = 0.5 * (v2_ll + v2_rr);
    const double v3_avg = 0.5 * (v3_ll + v3_rr);
    const double p_mean = 0.5 * rho_avg / beta_avg;
    const double velocity_square_avg = specific_kin_ll + specific_kin_rr;
    const double f1 = rho_mean * 0.5 * (v_dot_n_ll + v_dot_n_rr);
    const double f2 = f1 * v1_avg + p_mean * normal[0];
    const double f3 = f1 * v2_avg + p_mean * normal[1];
    const double f4 = f1 * v3_avg + p_mean * normal[2];
    const double f5 =
        f1 * 0.5 * (1.0 / (g - 1.0) / beta_mean - velocity_square_avg) +
        f2 * v1_avg + f3 * v2_avg + f4 * v3_avg;
    f[0] = f1; f[1] = f2; f[2] = f3; f[3] = f4; f[4] = f5;
  }
}

/* HLLC with a normal vector -- build extension.  The reference has HLLC only
 * for 1D Euler (numerical_flux.hpp:283-378); this follows that structure
 * (Roe-averaged bounds, star states) in the frame of the unit normal, with
 * the tangential velocity carried through the star state and the result
 * scaled by |n|.  The Roe kinetic energy uses the full averaged velocity. */
void orc_flux_hllc(int dim, double g, const double* ul, const double* ur,
                   const double* n, double* f) {
  const int nv = dim + 2;
  double nrm = 0.0, nh[3], vl[3], vr[3], fl[5], fr[5];
  for (int d = 0; d < dim; ++d) nrm += n[d] * n[d];
  nrm = sqrt(nrm);
  for (int d = 0; d < dim; ++d) nh[d] = n[d] / nrm;
  const double rho_ll = ul[0], rho_rr = ur[0];
  const double rhoE_ll = ul[nv - 1], rhoE_rr = ur[nv - 1];
  double v2l = 0.0, v2r = 0.0, vn_l = 0.0, vn_r = 0.0;
  for (int d = 0; d < dim; ++d) {
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
  const double c_ll = sqrt(g * p_ll / rho_ll);
  const double c_rr = sqrt(g * p_rr / rho_rr);
  orc_physical_flux(dim, g, ul, nh, fl);
  orc_physical_flux(dim, g, ur, nh, fr);
  const double srl = sqrt(rho_ll), srr = sqrt(rho_rr), ssum = srl + srr;
  double vroe2 = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double vroe_d = (srl * vl[d] + srr * vr[d]) / ssum;
    vroe2 += vroe_d * vroe_d;
  }
  const double vel_roe = (srl * vn_l + srr * vn_r) / ssum;
  const double ekin_roe = 0.5 * vroe2;
  const double H_ll = (rhoE_ll + p_ll) / rho_ll;
  const double H_rr = (rhoE_rr + p_rr) / rho_rr;
  const double H_roe = (srl * H_ll + srr * H_rr) / ssum;
  const double c_roe = sqrt((g - 1.0) * (H_roe - ekin_roe));
  const double Ssl = dmin(vn_l - c_ll, vel_roe - c_roe);
  const double Ssr = dmax(vn_r + c_rr, vel_roe + c_roe);
  const double sMu_L = Ssl - vn_l, sMu_R = Ssr - vn_r;
  if (Ssl >= 0) {
    for (int i = 0; i < nv; ++i) f[i] = fl[i];
  } else if (Ssr <= 0) {
    for (int i = 0; i < nv; ++i) f[i] = fr[i];
  } else {
    const double SStar =
        (p_rr - p_ll + rho_ll * vn_l * sMu_L - rho_rr * vn_r * sMu_R) /
        (rho_ll * sMu_L - rho_rr * sMu_R);
    const double* uk = (SStar >= 0) ? ul : ur;
    const double* fk = (SStar >= 0) ? fl : fr;
    const double* vk = (SStar >= 0) ? vl : vr;
    const double rk = (SStar >= 0) ? rho_ll : rho_rr;
    const double vnk = (SStar >= 0) ? vn_l : vn_r;
    const double Ek = (SStar >= 0) ? E_ll : E_rr;
    const double pk = (SStar >= 0) ? p_ll : p_rr;
    const double Sk = (SStar >= 0) ? Ssl : Ssr;
    const double sMu = (SStar >= 0) ? sMu_L : sMu_R;
    const double densStar = rk * sMu / (Sk - SStar);
    const double enerStar = Ek + (SStar - vnk) * (SStar + pk / (rk * sMu));
    f[0] = fk[0] + Sk * (densStar - uk[0]);
    for (int d = 0; d < dim; ++d) {
      const double mstar = densStar * (vk[d] + (SStar - vnk) * nh[d]);
      f[d + 1] = fk[d + 1] + Sk * (mstar - uk[d + 1]);
    }
    f[nv - 1] = fk[nv - 1] + Sk * (densStar * enerStar - uk[nv - 1]);
  }
  for (int i = 0; i < nv; ++i) f[i] *= nrm;
}

/* CentralFlux::numerical_flux with a normal (numerical_flux.hpp:247-257,
 * 2D in the reference): eq.flux(u, normal) of both states, then
 * central_flux_from_physical_fluxes, 0.5 (f_L + f_R)
 * (numerical_flux_detail.hpp:23-32). */
void orc_flux_central(int dim, double g, const double* ul, const double* ur,
                      const double* n, double* f) {
  double fl[5], fr[5];
  orc_physical_flux(dim, g, ul, n, fl);
  orc_physical_flux(dim, g, ur, n, fr);
  for (int i = 0; i < dim + 2; ++i) f[i] = 0.5 * (fl[i] + fr[i]);
}

static void surface_flux_fn(const orc_case* c, int kind, const double* ul,
                            const double* ur, const double* n, double* f) {
  if (kind == ORC_FLUX_HLLC)
    orc_flux_hllc(c->dim, c->gamma, ul, ur, n, f);
  else if (kind == ORC_FLUX_CENTRAL)
    orc_flux_central(c->dim, c->gamma, ul, ur, n, f);
  else
    orc_flux_llf(c->dim, c->gamma, ul, ur, n, f);
}

/* ------------------------------------------------------------------------ */
/* Indexing helpers (reference LayoutRight, local_dof.hpp:9-19)              */
/* ------------------------------------------------------------------------ */

static inline int ldof2(int n, int i, int j) { return j * n + i; }
static inline int ldof3(int n, int i, int j, int k) { return (k * n + j) * n + i; }
static inline int64_t eidx(const orc_case* c, int ie, int je, int ke) {
  return c->dim == 2 ? (int64_t)ie * c->ny + je
                     : ((int64_t)ie * c->ny + je) * c->nz + ke;
}
#define U(arr, e, dof, var) (arr)[((e) * (int64_t)c->nn + (dof)) * c->nvars + (var)]
#define JA(e, dof, dir, comp) \
  c->Ja[(((e) * (int64_t)c->nn + (dof)) * c->dim + (dir)) * c->dim + (comp)]
#define INVJ(e, dof) c->invJ[(e) * (int64_t)c->nn + (dof)]

/* ------------------------------------------------------------------------ */
/* Geometry: src/containers/data_container.hpp, src/base/global_def.hpp      */
/* ------------------------------------------------------------------------ */

static void build_metrics(orc_case* c);

static void build_geometry(orc_case* c, const double* L, const double* R,
                           const int* periodic) {
  const int n = c->n, nn = c->nn, d = c->dim;
  const int64_t E = c->E;
  const int ncell[3] = {c->nx, c->ny, c->nz};
  c->coords = (double*)calloc((size_t)(E * nn * d), sizeof(double));
  c->jac = (double*)calloc((size_t)(E * nn * d * d), sizeof(double));
  c->Ja = (double*)calloc((size_t)(E * nn * d * d), sizeof(double));
  c->invJ = (double*)calloc((size_t)(E * nn), sizeof(double));
  c->nbr = (int64_t*)calloc((size_t)(E * d * d), sizeof(int64_t));
  /* calc_node_coordinates: data_container.hpp:193-221 (2D), :371-414 (3D);
   * LinearMapping::eval -> linear_interpolate, mapping.hpp:7-20 */
  double delta[3];
  for (int k = 0; k < d; ++k) delta[k] = 2.0 / (double)ncell[k];
  for (int ie = 0; ie < c->nx; ++ie)
    for (int je = 0; je < c->ny; ++je)
      for (int ke = 0; ke < (d == 3 ? c->nz : 1); ++ke) {
        const int64_t e = eidx(c, ie, je, ke);
        const int idx[3] = {ie, je, ke};
        double off[3];
        for (int k = 0; k < d; ++k)
          off[k] = -1.0 + (double)idx[k] * delta[k] + 0.5 * delta[k];
        for (int dof = 0; dof < nn; ++dof) {
          const int nd[3] = {dof % n, (dof / n) % n, dof / (n * n)};
          for (int k = 0; k < d; ++k) {
            const double ref = off[k] + 0.5 * delta[k] * c->nodes[nd[k]];
            c->coords[(e * nn + dof) * d + k] =
                0.5 * ((1.0 - ref) * L[k] + (1.0 + ref) * R[k]);
          }
        }
      }
  build_metrics(c);
  /* initialize_left_neighbor: :314-331 (2D), :695-724 (3D); -1 = none */
  for (int ie = 0; ie < c->nx; ++ie)
    for (int je = 0; je < c->ny; ++je)
      for (int ke = 0; ke < (d == 3 ? c->nz : 1); ++ke) {
        const int64_t e = eidx(c, ie, je, ke);
        const int idx[3] = {ie, je, ke};
        for (int dir = 0; dir < d; ++dir)
          for (int comp = 0; comp < d; ++comp) {
            int64_t v = idx[comp];
            if (comp == dir)
              v = idx[dir] > 0 ? idx[dir] - 1 : (periodic[dir] ? ncell[dir] - 1 : -1);
            c->nbr[(e * d + dir) * d + comp] = v;
          }
      }
}

/* Jacobian matrix, contravariant vectors and inverse Jacobian from the node
 * coordinates c->coords (whatever mapping produced them). */
static void build_metrics(orc_case* c) {
  const int n = c->n, nn = c->nn, d = c->dim;
  const int64_t E = c->E;
#define X(e, dof, comp) c->coords[((e) * (int64_t)nn + (dof)) * d + (comp)]
#define JM(e, dof, dir, comp) c->jac[(((e) * (int64_t)nn + (dof)) * d + (dir)) * d + (comp)]
  /* calc_jacobian_matrix: :223-258 (2D), :416-482 (3D) */
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e)
    for (int dof = 0; dof < nn; ++dof) {
      const int nd[3] = {dof % n, (dof / n) % n, dof / (n * n)};
      for (int dir = 0; dir < d; ++dir) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < n; ++m) {
          int md[3] = {nd[0], nd[1], nd[2]};
          md[dir] = m;
          const int dm = (d == 2) ? ldof2(n, md[0], md[1]) : ldof3(n, md[0], md[1], md[2]);
          for (int comp = 0; comp < d; ++comp)
            acc[comp] += c->D[nd[dir] * n + m] * X(e, dm, comp);
        }
        for (int comp = 0; comp < d; ++comp) JM(e, dof, dir, comp) = acc[comp];
      }
    }
  /* calc_contravariant_vectors: :260-291 (2D adjugate), :484-659 (3D curl) */
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    if (d == 2) {
      for (int dof = 0; dof < nn; ++dof) {
        JA(e, dof, 0, 0) = JM(e, dof, 1, 1);
        JA(e, dof, 0, 1) = -JM(e, dof, 1, 0);
        JA(e, dof, 1, 0) = -JM(e, dof, 0, 1);
        JA(e, dof, 1, 1) = JM(e, dof, 0, 0);
      }
    } else {
      for (int nc = 0; nc < 3; ++nc) {
        const int m = (nc + 1) % 3, l = (nc + 2) % 3;
        for (int dof = 0; dof < nn; ++dof) {
          const int i = dof % n, j = (dof / n) % n, k = dof / (n * n);
          double r1 = 0.0, r2 = 0.0, r3 = 0.0;
          /* Ja^1_n: (A)_eta - (B)_zeta */
          for (int ii = 0; ii < n; ++ii) {
            const int de = ldof3(n, i, ii, k);
            r1 += 0.5 * c->D[j * n + ii] *
                  (X(e, de, m) * JM(e, de, 2, l) - X(e, de, l) * JM(e, de, 2, m));
          }
          for (int ii = 0; ii < n; ++ii) {
            const int dz = ldof3(n, i, j, ii);
            r1 -= 0.5 * c->D[k * n + ii] *
                  (X(e, dz, m) * JM(e, dz, 1, l) - X(e, dz, l) * JM(e, dz, 1, m));
          }
          /* Ja^2_n: (A)_zeta - (B)_xi */
          for (int ii = 0; ii < n; ++ii) {
            const int dz = ldof3(n, i, j, ii);
            r2 += 0.5 * c->D[k * n + ii] *
                  (X(e, dz, m) * JM(e, dz, 0, l) - X(e, dz, l) * JM(e, dz, 0, m));
          }
          for (int ii = 0; ii < n; ++ii) {
            const int dx = ldof3(n, ii, j, k);
            r2 -= 0.5 * c->D[i * n + ii] *
                  (X(e, dx, m) * JM(e, dx, 2, l) - X(e, dx, l) * JM(e, dx, 2, m));
          }
          /* Ja^3_n: (A)_xi - (B)_eta */
          for (int ii = 0; ii < n; ++ii) {
            const int dx = ldof3(n, ii, j, k);
            r3 += 0.5 * c->D[i * n + ii] *
                  (X(e, dx, m) * JM(e, dx, 1, l) - X(e, dx, l) * JM(e, dx, 1, m));
          }
          for (int ii = 0; ii < n; ++ii) {
            const int de = ldof3(n, i, ii, k);
            r3 -= 0.5 * c->D[j * n + ii] *
                  (X(e, de, m) * JM(e, de, 0, l) - X(e, de, l) * JM(e, de, 0, m));
          }
          JA(e, dof, 0, nc) = r1;
          JA(e, dof, 1, nc) = r2;
          JA(e, dof, 2, nc) = r3;
        }
      }
    }
  }
  /* calc_inverse_jacobian: :293-312 (2D), :661-693 (3D) */
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e)
    for (int dof = 0; dof < nn; ++dof) {
      if (d == 2) {
        INVJ(e, dof) = 1.0 / (JM(e, dof, 0, 0) * JM(e, dof, 1, 1) -
                              JM(e, dof, 1, 0) * JM(e, dof, 0, 1));
      } else {
        const double j11 = JM(e, dof, 0, 0), j12 = JM(e, dof, 0, 1), j13 = JM(e, dof, 0, 2);
        const double j21 = JM(e, dof, 1, 0), j22 = JM(e, dof, 1, 1), j23 = JM(e, dof, 1, 2);
        const double j31 = JM(e, dof, 2, 0), j32 = JM(e, dof, 2, 1), j33 = JM(e, dof, 2, 2);
        const double det = j11 * (j22 * j33 - j23 * j32) -
                           j12 * (j21 * j33 - j23 * j31) +
                           j13 * (j21 * j32 - j22 * j31);
        INVJ(e, dof) = 1.0 / det;
      }
    }
#undef X
#undef JM
}

/* SubcellNormalVectors<T,2/3>::initialize_host, global_def.hpp:175-251 (2D),
 * :299-439 (3D).  Stored as normal_dir[e][dir-line node ...][comp] with the
 * face index in the position of the line direction. */
static void build_subcell_normals(orc_case* c) {
  const int n = c->n, d = c->dim;
  const int64_t E = c->E;
  const int64_t per = (int64_t)(n - 1) * (d == 3 ? n * n : n) * d;
  free(c->sn1);
  free(c->sn2);
  free(c->sn3);
  c->sn1 = (double*)calloc((size_t)(E * per), sizeof(double));
  c->sn2 = (double*)calloc((size_t)(E * per), sizeof(double));
  c->sn3 = d == 3 ? (double*)calloc((size_t)(E * per), sizeof(double)) : NULL;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    for (int dir = 0; dir < d; ++dir) {
      double* sn = dir == 0 ? c->sn1 : (dir == 1 ? c->sn2 : c->sn3);
      const int nl = (d == 3) ? n * n : n; /* lines per direction */
      for (int line = 0; line < nl; ++line) {
        /* line coordinates: the two indices other than dir, in order */
        int a = line % n, b = line / n;
        for (int f = 0; f < n - 1; ++f) {
          int idx[3];
          double acc[3];
          /* sn index: position of (f, a, b) in the array shape where the
           * face index replaces the dir-th node index */
          int pos[3];
          if (d == 2) {
            pos[dir] = f;
            pos[1 - dir] = a;
          } else {
            int o = 0;
            for (int k = 0; k < 3; ++k) {
              if (k == dir) pos[k] = f;
              else pos[k] = (o++ == 0) ? a : b;
            }
          }
          int64_t slot;
          if (d == 2) {
            slot = dir == 0 ? ((int64_t)pos[0] * n + pos[1]) : ((int64_t)pos[0] * (n - 1) + pos[1]);
          } else {
            const int s0 = dir == 0 ? n - 1 : n, s1 = dir == 1 ? n - 1 : n,
                      s2 = dir == 2 ? n - 1 : n;
            (void)s0;
            slot = ((int64_t)pos[0] * s1 + pos[1]) * s2 + pos[2];
          }
          double* out = sn + (e * (per / d) + slot) * d;
          if (f == 0) {
            for (int k = 0; k < d; ++k) idx[k] = pos[k];
            idx[dir] = 0;
            const int dof0 = d == 2 ? ldof2(n, idx[0], idx[1]) : ldof3(n, idx[0], idx[1], idx[2]);
            for (int comp = 0; comp < d; ++comp) acc[comp] = JA(e, dof0, dir, comp);
          } else {
            const int64_t prev = slot - (d == 2 ? (dir == 0 ? n : 1)
                                                 : (dir == 0 ? n * n : (dir == 1 ? n : 1)));
            double* po = sn + (e * (per / d) + prev) * d;
            for (int comp = 0; comp < d; ++comp) acc[comp] = po[comp];
          }
          for (int m = 0; m < n; ++m) {
            for (int k = 0; k < d; ++k) idx[k] = pos[k];
            idx[dir] = m;
            const int dm = d == 2 ? ldof2(n, idx[0], idx[1]) : ldof3(n, idx[0], idx[1], idx[2]);
            const double wD = c->weights[f] * c->D[f * n + m];
            for (int comp = 0; comp < d; ++comp) acc[comp] += wD * JA(e, dm, dir, comp);
          }
          for (int comp = 0; comp < d; ++comp) out[comp] = acc[comp];
        }
      }
    }
  }
}

/* Subcell normal of interface f (between node f and f+1) along direction dir
 * on the line through (a,b) (the other two node indices in increasing
 * dimension order). */
static const double* subcell_normal(const orc_case* c, int64_t e, int dir, int f,
                                    int a, int b) {
  const int n = c->n, d = c->dim;
  const int64_t per = (int64_t)(n - 1) * (d == 3 ? n * n : n);
  int64_t slot;
  if (d == 2) {
    slot = dir == 0 ? ((int64_t)f * n + a) : ((int64_t)a * (n - 1) + f);
  } else {
    int pos[3], o = 0;
    for (int k = 0; k < 3; ++k) pos[k] = (k == dir) ? f : ((o++ == 0) ? a : b);
    const int s1 = dir == 1 ? n - 1 : n, s2 = dir == 2 ? n - 1 : n;
    slot = ((int64_t)pos[0] * s1 + pos[1]) * s2 + pos[2];
  }
  const double* sn = dir == 0 ? c->sn1 : (dir == 1 ? c->sn2 : c->sn3);
  return sn + (e * per + slot) * d;
}

/* ------------------------------------------------------------------------ */
/* Case lifecycle                                                            */
/* ------------------------------------------------------------------------ */

orc_case* orc_create(int dim, int N, int nx, int ny, int nz, const double* L,
                     const double* R, const int* periodic) {
  if ((dim != 2 && dim != 3) || N < 1 || N + 1 > MAXN) return NULL;
  orc_case* c = (orc_case*)calloc(1, sizeof(orc_case));
  c->dim = dim;
  c->N = N;
  c->n = N + 1;
  c->nn = dim == 2 ? c->n * c->n : c->n * c->n * c->n;
  c->nvars = dim + 2;
  c->nx = nx;
  c->ny = ny;
  c->nz = dim == 3 ? nz : 1;
  c->E = (int64_t)nx * ny * c->nz;
  c->gamma = 1.4;
  c->mu = dim == 3 ? 6.25e-4 : 5.0e-3;
  c->prandtl = dim == 3 ? 0.72 : 0.73;
  c->kappa = c->gamma / ((c->gamma - 1.0) * c->prandtl);
  c->alpha_max = 0.5;
  c->alpha_min = 0.001;
  c->floor_value = 0.0;
  c->nthreads = 0;
  orc_basis(N, c->nodes, c->weights, c->inv_weights, c->D, c->Dsplit, c->Dhat,
            c->Vinv, c->lhat_l, c->lhat_r);
  build_geometry(c, L, R, periodic);
  build_subcell_normals(c);
  return c;
}

void orc_destroy(orc_case* c) {
  if (!c) return;
  free(c->coords); free(c->jac); free(c->Ja); free(c->invJ); free(c->nbr);
  free(c->sn1); free(c->sn2); free(c->sn3);
  free(c->sflux); free(c->grad); free(c->gflux); free(c->vflux); free(c->vsflux);
  free(c->tmp1); free(c->tmp2); free(c->du);
  free(c);
}

void orc_set_equation(orc_case* c, int eq, double gamma, double mu,
                      double prandtl) {
  c->eq = eq;
  c->gamma = gamma;
  c->mu = mu;
  c->prandtl = prandtl;
  /* kappa = gamma / ((gamma - 1) Pr), compressible_navier_stokes3D.hpp:19-22 */
  c->kappa = gamma / ((gamma - 1.0) * prandtl);
}
void orc_set_volume(orc_case* c, int volume, int fv_flux, double amax,
                    double amin) {
  c->volume = volume;
  c->fv_flux = fv_flux;
  c->alpha_max = amax;
  c->alpha_min = amin;
}
void orc_set_surface_flux(orc_case* c, int sf) { c->surface_flux = sf; }

void orc_set_volume_flux(orc_case* c, int vf) { c->volume_flux = vf; }
void orc_set_positivity_floor(orc_case* c, double v) { c->floor_value = v; }
void orc_set_threads(orc_case* c, int t) { c->nthreads = t; }
int64_t orc_num_elements(const orc_case* c) { return c->E; }
int orc_nvars(const orc_case* c) { return c->nvars; }
int orc_ndofs(const orc_case* c) { return c->nn; }

void orc_get_coordinates(const orc_case* c, double* o) {
  memcpy(o, c->coords, sizeof(double) * (size_t)(c->E * c->nn * c->dim));
}
void orc_get_jacobian(const orc_case* c, double* o) {
  memcpy(o, c->jac, sizeof(double) * (size_t)(c->E * c->nn * c->dim * c->dim));
}
void orc_get_contravariant(const orc_case* c, double* o) {
  memcpy(o, c->Ja, sizeof(double) * (size_t)(c->E * c->nn * c->dim * c->dim));
}
void orc_get_inverse_jacobian(const orc_case* c, double* o) {
  memcpy(o, c->invJ, sizeof(double) * (size_t)(c->E * c->nn));
}
void orc_get_left_neighbors(const orc_case* c, int64_t* o) {
  memcpy(o, c->nbr, sizeof(int64_t) * (size_t)(c->E * c->dim * c->dim));
}
/* Node coordinates from any mapping (the reference's StructuredElementInitializer
 * with a user Mapping, data_container.hpp:729-768): the metrics and subcell
 * normals are recomputed from them as for the linear mapping. */
void orc_set_coordinates(orc_case* c, const double* coords) {
  memcpy(c->coords, coords, sizeof(double) * (size_t)(c->E * c->nn * c->dim));
  build_metrics(c);
  build_subcell_normals(c);
}

void orc_set_metrics(orc_case* c, const double* Ja, const double* invJ) {
  memcpy(c->Ja, Ja, sizeof(double) * (size_t)(c->E * c->nn * c->dim * c->dim));
  memcpy(c->invJ, invJ, sizeof(double) * (size_t)(c->E * c->nn));
  build_subcell_normals(c);
}

static int nthreads_of(const orc_case* c) {
#ifdef _OPENMP
  return c->nthreads > 0 ? c->nthreads : omp_get_max_threads();
#else
  (void)c;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Initial conditions                                                        */
/* ------------------------------------------------------------------------ */

/* prim_to_cons, src/utils/state_conversion.hpp:12-28 */
static void prim_to_cons(int dim, const double* prim, double g, double* cons) {
  double vsq = 0.0;
  cons[0] = prim[0];
  for (int k = 0; k < dim; ++k) {
    const double v = prim[k + 1];
    cons[k + 1] = prim[0] * v;
    vsq += v * v;
  }
  cons[dim + 1] = prim[dim + 1] / (g - 1.0) + 0.5 * prim[0] * vsq;
}

/* KelvinHelmholtzInitial, examples/Euler2D/kelvin_helmholtz_hg/
 * kelvin_helmholtz_hg.cpp:16-31 */
void orc_init_khi(const orc_case* c, double* u) {
  for (int64_t e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      const double x = c->coords[(e * c->nn + dof) * 2 + 0];
      const double y = c->coords[(e * c->nn + dof) * 2 + 1];
      const double slope = 15.0;
      const double B = tanh(slope * y + 7.5) - tanh(slope * y - 7.5);
      double prim[4], cons[4];
      prim[0] = 0.5 + 0.75 * B;
      prim[1] = 0.5 * (B - 1.0);
      prim[2] = 0.1 * sin(2.0 * PI_D * x);
      prim[3] = 1.0;
      prim_to_cons(2, prim, 1.4, cons);
      for (int v = 0; v < 4; ++v) U(u, e, dof, v) = cons[v];
    }
}

/* InviscidTaylorGreenVortex, examples/Euler3D/inviscid_tgv/
 * inviscid_tgv.cpp:17-35 */
void orc_init_tgv(const orc_case* c, double* u) {
  for (int64_t e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      const double x = c->coords[(e * c->nn + dof) * 3 + 0];
      const double y = c->coords[(e * c->nn + dof) * 3 + 1];
      const double z = c->coords[(e * c->nn + dof) * 3 + 2];
      double prim[5], cons[5];
      prim[0] = 1.0;
      prim[1] = sin(x) * cos(y) * cos(z);
      prim[2] = -cos(x) * sin(y) * cos(z);
      prim[3] = 0.0;
      prim[4] = 100.0 / 1.4 +
                1.0 / 16.0 *
                    (cos(2.0 * x) * cos(2.0 * z) + 2.0 * cos(2.0 * y) +
                     2.0 * cos(2.0 * x) + cos(2.0 * y) * cos(2.0 * z));
      prim_to_cons(3, prim, 1.4, cons);
      for (int v = 0; v < 5; ++v) U(u, e, dof, v) = cons[v];
    }
}

/* Gaussian Sedov state -- build extension (the reference's sedov example,
 * examples/Euler3D/sedov_blast_wave/sedov_blast_wave.cpp:8-39, uses a
 * top-hat).  p = p_bg + (gamma-1) E0 exp(-r^2/(2 s^2)) / ((2 pi)^1.5 s^3). */
void orc_init_sedov(const orc_case* c, double* u, double E0, double sigma,
                    double p_bg) {
  const double norm = pow(2.0 * PI_D, 1.5) * sigma * sigma * sigma;
  for (int64_t e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      double r2 = 0.0;
      for (int k = 0; k < c->dim; ++k) {
        const double x = c->coords[(e * c->nn + dof) * c->dim + k];
        r2 += x * x;
      }
      double prim[5] = {0, 0, 0, 0, 0}, cons[5];
      prim[0] = 1.0;
      prim[c->dim + 1] =
          p_bg + (c->gamma - 1.0) * E0 * exp(-r2 / (2.0 * sigma * sigma)) / norm;
      prim_to_cons(c->dim, prim, c->gamma, cons);
      for (int v = 0; v < c->nvars; ++v) U(u, e, dof, v) = cons[v];
    }
}

/* ------------------------------------------------------------------------ */
/* Hennemann-Gassner indicator: src/subcell_limiting/indicator.hpp           */
/* ------------------------------------------------------------------------ */

/* IndicatorValueAccessor: p*rho with gamma hard-coded to 1.4,
 * indicator.hpp:140-215 */
static double indicator_value(int dim, const double* s) {
  const double gamma = 1.4;
  if (dim == 2) {
    const double uv = s[1] / s[0], vv = s[2] / s[0];
    const double p = (gamma - 1.0) * (s[3] - 0.5 * s[0] * (uv * uv + vv * vv));
    return p * s[0];
  } else {
    const double uv = s[1] / s[0], vv = s[2] / s[0], wv = s[3] / s[0];
    const double p =
        (gamma - 1.0) * (s[4] - 0.5 * s[0] * (uv * uv + vv * vv + wv * wv));
    return p * s[0];
  }
}

/* HGIndicator<.,2>::operator() (indicator.hpp:337-417) and
 * HGIndicator<.,3>::operator() (indicator.hpp:454-574) for one element */
static double hg_alpha(const orc_case* c, const double* u, int64_t e) {
  const int n = c->n, dim = c->dim;
  double ind[MAXN * MAXN * MAXN], tmp1[MAXN * MAXN * MAXN],
      tmp2[MAXN * MAXN * MAXN], modal[MAXN * MAXN * MAXN];
  double s[5];
  /* threshold = 0.5 * 10^(-1.8 * NNodes^0.25), s = log((1-1e-4)/1e-4):
   * indicator.hpp:323-335 / :443-452 */
  const double threshold = 0.5 * pow(10.0, -1.8 * pow((double)n, 0.25));
  const double sfac = log((1.0 - 0.0001) / 0.0001);
  for (int dof = 0; dof < c->nn; ++dof) {
    for (int v = 0; v < c->nvars; ++v) s[v] = U(u, e, dof, v);
    ind[dof] = indicator_value(dim, s);
  }
  double total = 0.0, clip1 = 0.0, clip2 = 0.0;
  if (dim == 2) {
    /* multiply_scalar_dimensionwise<.,2>, indicator.hpp:33-59 */
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double r = 0;
        for (int ii = 0; ii < n; ++ii) r += c->Vinv[i * n + ii] * ind[j * n + ii];
        tmp1[j * n + i] = r;
      }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double r = 0;
        for (int jj = 0; jj < n; ++jj) r += c->Vinv[j * n + jj] * tmp1[jj * n + i];
        modal[j * n + i] = r;
      }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) total += modal[j * n + i] * modal[j * n + i];
    for (int j = 0; j + 1 < n; ++j)
      for (int i = 0; i + 1 < n; ++i) clip1 += modal[j * n + i] * modal[j * n + i];
    for (int j = 0; j + 2 < n; ++j)
      for (int i = 0; i + 2 < n; ++i) clip2 += modal[j * n + i] * modal[j * n + i];
  } else {
    /* multiply_scalar_dimensionwise<.,3>, indicator.hpp:62-111 */
#define I3(k, j, i) (((k) * n + (j)) * n + (i))
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double r = 0.0;
          for (int ii = 0; ii < n; ++ii) r += c->Vinv[i * n + ii] * ind[I3(k, j, ii)];
          tmp1[I3(k, j, i)] = r;
        }
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double r = 0.0;
          for (int jj = 0; jj < n; ++jj) r += c->Vinv[j * n + jj] * tmp1[I3(k, jj, i)];
          tmp2[I3(k, j, i)] = r;
        }
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double r = 0.0;
          for (int kk = 0; kk < n; ++kk) r += c->Vinv[k * n + kk] * tmp2[I3(kk, j, i)];
          modal[I3(k, j, i)] = r;
        }
    /* energy sums, indicator.hpp:505-555, including the omission of the
     * modes (i=n-1, j<n-1, k<n-1) from the total */
    for (int k = 0; k < n - 2; ++k)
      for (int j = 0; j < n - 2; ++j)
        for (int i = 0; i < n - 2; ++i) clip2 += modal[I3(k, j, i)] * modal[I3(k, j, i)];
    clip1 = clip2;
    for (int j = 0; j < n - 1; ++j)
      for (int i = 0; i < n - 1; ++i)
        clip1 += modal[I3(n - 2, j, i)] * modal[I3(n - 2, j, i)];
    for (int k = 0; k < n - 2; ++k)
      for (int i = 0; i < n - 1; ++i)
        clip1 += modal[I3(k, n - 2, i)] * modal[I3(k, n - 2, i)];
    for (int k = 0; k < n - 2; ++k)
      for (int j = 0; j < n - 2; ++j)
        clip1 += modal[I3(k, j, n - 2)] * modal[I3(k, j, n - 2)];
    total = clip1;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        total += modal[I3(n - 1, j, i)] * modal[I3(n - 1, j, i)];
    for (int k = 0; k < n - 1; ++k)
      for (int i = 0; i < n; ++i)
        total += modal[I3(k, n - 1, i)] * modal[I3(k, n - 1, i)];
#undef I3
  }
  double ef1, ef2;
  if (total != 0) ef1 = (total - clip1) / total;
  else ef1 = 0.0;
  if (clip1 != 0) ef2 = (clip1 - clip2) / clip1;
  else ef2 = 0.0;
  const double energy = dmax(ef1, ef2);
  double alpha = 1.0 / (1.0 + exp(-sfac / threshold * (energy - threshold)));
  if (alpha < c->alpha_min) alpha = 0.0;
  if (alpha > 1.0 - c->alpha_min) alpha = 1.0;
  return dmin(c->alpha_max, alpha);
}

void orc_calc_alpha(orc_case* c, const double* u, double* alpha) {
  int64_t e;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e) alpha[e] = hg_alpha(c, u, e);
}

/* ------------------------------------------------------------------------ */
/* Volume integral: src/space_integral/volume_flux.hpp                       */
/* ------------------------------------------------------------------------ */

/* VolumeIntegralSplit<.,2>::split_form_kernel (volume_flux.hpp:221-294) and
 * VolumeIntegralSplit<.,3>::split_form_kernel (volume_flux.hpp:296-410) */
static void split_form(const orc_case* c, int64_t e, const double* u,
                       double* du, double alpha) {
  const int n = c->n, d = c->dim, nv = c->nvars;
  double ui[5], ul[5], f[5], nrm[3];
  for (int k = 0; k < (d == 3 ? n : 1); ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const int dof_i = d == 2 ? ldof2(n, i, j) : ldof3(n, i, j, k);
        for (int v = 0; v < nv; ++v) ui[v] = U(u, e, dof_i, v);
        const int idx[3] = {i, j, k};
        for (int dir = 0; dir < d; ++dir) {
          for (int l = idx[dir] + 1; l < n; ++l) {
            int li[3] = {i, j, k};
            li[dir] = l;
            const int dof_l = d == 2 ? ldof2(n, li[0], li[1]) : ldof3(n, li[0], li[1], li[2]);
            for (int v = 0; v < nv; ++v) ul[v] = U(u, e, dof_l, v);
            for (int comp = 0; comp < d; ++comp)
              nrm[comp] = 0.5 * (JA(e, dof_i, dir, comp) + JA(e, dof_l, dir, comp));
            if (c->volume_flux == ORC_VOLUME_CENTRAL)
              orc_flux_central(d, c->gamma, ui, ul, nrm, f);
            else
              orc_flux_chandrashekar(d, c->gamma, ui, ul, nrm, f);
            const int a = idx[dir];
            for (int v = 0; v < nv; ++v) {
              U(du, e, dof_i, v) += alpha * c->Dsplit[a * n + l] * f[v];
              U(du, e, dof_l, v) += alpha * c->Dsplit[l * n + a] * f[v];
            }
          }
        }
      }
}

/* FinitVolumeIntegral<.,2>::fv_kernel (volume_flux.hpp:449-511) and
 * FinitVolumeIntegral<.,3>::fv_kernel (volume_flux.hpp:513-630) */
static void fv_subcell(const orc_case* c, int64_t e, const double* u, double* du,
                       double alpha) {
  const int n = c->n, d = c->dim, nv = c->nvars;
  double ull[5], urr[5], f[5];
  for (int dir = 0; dir < d; ++dir) {
    /* Loop nest order of the reference: 3D dir0: k, j, iface; dir1: k, jface,
     * i; dir2: kface, j, i.  2D dir0: j, iface; dir1: i, jface. */
    const int outer_n = d == 3 ? n : 1;
    for (int o = 0; o < outer_n; ++o)
      for (int mid = 0; mid < n; ++mid)
        for (int in = 0; in < n; ++in) {
          int li[3], ri[3], f_idx, a, b;
          if (d == 2) {
            if (dir == 0) { /* jnode = mid? reference: for j, for iface */
              if (in == 0) continue;
              f_idx = in; a = mid; b = 0;
              li[0] = f_idx - 1; li[1] = mid; ri[0] = f_idx; ri[1] = mid;
            } else { /* for inode, for iface */
              if (in == 0) continue;
              f_idx = in; a = mid; b = 0;
              li[0] = mid; li[1] = f_idx - 1; ri[0] = mid; ri[1] = f_idx;
            }
            li[2] = ri[2] = 0;
          } else {
            if (dir == 0) { /* k=o, j=mid, iface=in */
              if (in == 0) continue;
              f_idx = in; a = mid; b = o;
              li[0] = f_idx - 1; li[1] = mid; li[2] = o;
              ri[0] = f_idx; ri[1] = mid; ri[2] = o;
            } else if (dir == 1) { /* k=o, jface=mid, i=in */
              if (mid == 0) continue;
              f_idx = mid; a = in; b = o;
              li[0] = in; li[1] = f_idx - 1; li[2] = o;
              ri[0] = in; ri[1] = f_idx; ri[2] = o;
            } else { /* kface=o, j=mid, i=in */
              if (o == 0) continue;
              f_idx = o; a = in; b = mid;
              li[0] = in; li[1] = mid; li[2] = f_idx - 1;
              ri[0] = in; ri[1] = mid; ri[2] = f_idx;
            }
          }
          const int ld = d == 2 ? ldof2(n, li[0], li[1]) : ldof3(n, li[0], li[1], li[2]);
          const int rd = d == 2 ? ldof2(n, ri[0], ri[1]) : ldof3(n, ri[0], ri[1], ri[2]);
          for (int v = 0; v < nv; ++v) {
            ull[v] = U(u, e, ld, v);
            urr[v] = U(u, e, rd, v);
          }
          const double* nrm = subcell_normal(c, e, dir, f_idx - 1, a, b);
          surface_flux_fn(c, c->fv_flux, ull, urr, nrm, f);
          for (int v = 0; v < nv; ++v) {
            U(du, e, ld, v) += alpha * c->inv_weights[f_idx - 1] * f[v];
            U(du, e, rd, v) -= alpha * c->inv_weights[f_idx] * f[v];
          }
        }
  }
}

/* VolumeIntegralShockCapturingHG::operator() (volume_flux.hpp:843-869) and
 * VolumeIntegralSplitForm::operator() (volume_flux.hpp:725-735) */
void orc_volume_integral(orc_case* c, const double* u, double* du,
                         double* alpha_out) {
  int64_t e;
  if (c->volume == ORC_SHOCK_CAPTURING) {
    /* atol = max(100 eps, eps^0.75), volume_flux.hpp:772-773 */
    const double eps = 2.220446049250313080847e-16;
    const double atol = dmax(100.0 * eps, pow(eps, 0.75));
    double* alpha = alpha_out ? alpha_out : (double*)malloc(sizeof(double) * (size_t)c->E);
    if (c->alpha_override)
      memcpy(alpha, c->alpha_override, sizeof(double) * (size_t)c->E);
    else
      orc_calc_alpha(c, u, alpha);
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
    for (e = 0; e < c->E; ++e) {
      const double a = alpha[e];
      const int dg_only = fabs(a - 0.0) <= atol;
      if (dg_only) {
        split_form(c, e, u, du, 1.0);
      } else {
        split_form(c, e, u, du, 1.0 - a);
        fv_subcell(c, e, u, du, a);
      }
    }
    if (!alpha_out) free(alpha);
  } else {
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
    for (e = 0; e < c->E; ++e) split_form(c, e, u, du, 1.0);
  }
}

/* ------------------------------------------------------------------------ */
/* Interface flux, surface integral, Jacobian                                */
/* ------------------------------------------------------------------------ */

static int face_nodes(const orc_case* c) { return c->dim == 2 ? c->n : c->n * c->n; }
#define SF(arr, e, slot, var) \
  (arr)[((e) * (int64_t)(2 * c->dim * face_nodes(c)) + (slot)) * nv + (var)]

/* dof of face node (a, b) on the 'low' (side 0) or 'high' (side 1) face of
 * direction dir; (a, b) are the other indices in increasing dimension order */
static int face_dof_node(const orc_case* c, int dir, int side, int a, int b) {
  const int n = c->n, fixed = side ? n - 1 : 0;
  if (c->dim == 2) return dir == 0 ? ldof2(n, fixed, a) : ldof2(n, a, fixed);
  if (dir == 0) return ldof3(n, fixed, a, b);
  if (dir == 1) return ldof3(n, a, fixed, b);
  return ldof3(n, a, b, fixed);
}

/* InterfaceHelper<.,StructuredElementContainer<T,2/3>>::interface_flux,
 * surface_flux.hpp:45-127 (2D), :129-271 (3D); launched per element by
 * InterfaceFluxFunctor (surface_flux_functor.hpp:40-111) */
static void interface_flux_generic(orc_case* c, const double* u, double* sflux,
                                   int kind) {
  const int d = c->dim, n = c->n, nv = c->nvars, fn = face_nodes(c);
  int64_t e;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e) {
    double ul[5], ur[5], f[5], nrm[3];
    for (int dir = 0; dir < d; ++dir) {
      int64_t nb[3], valid = 1;
      for (int comp = 0; comp < d; ++comp) {
        nb[comp] = c->nbr[(e * d + dir) * d + comp];
        if (nb[comp] < 0) valid = 0;
      }
      if (!valid) continue;
      const int64_t le = d == 2 ? nb[0] * c->ny + nb[1]
                                : (nb[0] * c->ny + nb[1]) * c->nz + nb[2];
      for (int q = 0; q < fn; ++q) {
        const int a = q % n, b = q / n;
        const int ldof = face_dof_node(c, dir, 1, a, b);
        const int rdof = face_dof_node(c, dir, 0, a, b);
        const double sgn = INVJ(e, rdof) >= 0.0 ? 1.0 : -1.0;
        for (int comp = 0; comp < d; ++comp) nrm[comp] = sgn * JA(e, rdof, dir, comp);
        for (int v = 0; v < nv; ++v) {
          ul[v] = U(u, le, ldof, v);
          ur[v] = U(u, e, rdof, v);
        }
        surface_flux_fn(c, kind, ul, ur, nrm, f);
        for (int v = 0; v < nv; ++v) {
          SF(sflux, le, (2 * dir + 1) * fn + q, v) = f[v];
          SF(sflux, e, (2 * dir) * fn + q, v) = f[v];
        }
      }
    }
  }
}

void orc_interface_flux(orc_case* c, const double* u, double* sflux) {
  interface_flux_generic(c, u, sflux, c->surface_flux);
}

/* Boundary-face fluxes of a 2D mesh (build of the reference's BoundarySet::
 * apply, boundary_set.hpp:88-111, for 2D compressible flow):
 *   slip wall   slip_wall_boundary.hpp:12-62  (outer state = inner state with
 *               the normal velocity reflected),
 *   outflow     outflow_boundary.hpp:61-105   (outer state = mean of the
 *               inner element's nodes on the line normal to the face),
 *   Dirichlet   dirichlet_boundary.hpp:70-118 (outer state given).
 * Node order and normals follow utils/boundary_helper.hpp:69-150: the inner
 * element's sign(J) Ja^dim at its boundary node, the inner state on the left
 * of the flux on the upper faces and on the right on the lower faces.
 * Per-node Dirichlet states are indexed by (element position along the face)
 * * (N+1) + (node position along the face). */
static void boundary_fluxes(orc_case* c, const double* u, double* sflux) {
  if (c->dim != 2) return;
  const int n = c->n, nv = c->nvars, fn = face_nodes(c);
  for (int face = 0; face < 4; ++face) {
    const int kind = c->bc[face];
    if (kind == ORC_BC_PERIODIC) continue;
    const int dim = face / 2, upper = face % 2;
    const int count = dim == 0 ? c->ny : c->nx;
    for (int idx = 0; idx < count; ++idx) {
      const int ie = dim == 0 ? (upper ? c->nx - 1 : 0) : idx;
      const int je = dim == 0 ? idx : (upper ? c->ny - 1 : 0);
      const int64_t e = (int64_t)ie * c->ny + je;
      for (int q = 0; q < n; ++q) {
        /* boundary node and the line through it, normal to the face */
        const int bi = dim == 0 ? (upper ? n - 1 : 0) : q;
        const int bj = dim == 0 ? q : (upper ? n - 1 : 0);
        const int bdof = bj * n + bi;
        double uin[4], uout[4], nrm[2], f[4];
        for (int v = 0; v < nv; ++v) uin[v] = U(u, e, bdof, v);
        const double sgn = INVJ(e, bdof) >= 0.0 ? 1.0 : -1.0;
        for (int comp = 0; comp < 2; ++comp) nrm[comp] = sgn * JA(e, bdof, dim, comp);
        const int bnode = idx * n + q;
        const int mixed_dirichlet = kind == ORC_BC_MIXED_DIRICHLET_SLIP_WALL &&
                                    c->bc_values[face][bnode * (nv + 1) + nv] != 0.0;
        if (kind == ORC_BC_SLIP_WALL || kind == ORC_BC_NS_NOSLIP_WALL || kind == ORC_BC_NS_SLIP_WALL ||
            (kind == ORC_BC_MIXED_DIRICHLET_SLIP_WALL && !mixed_dirichlet)) {
          /* reflect_state, slip_wall_boundary.hpp:12-24; the NS walls use
           * reflect_normal_velocity_2d (navier_stokes_wall_boundary.hpp:12-27) */
          const double rho = uin[0], vx = uin[1] / rho, vy = uin[2] / rho;
          const double nn = nrm[0] * nrm[0] + nrm[1] * nrm[1];
          const double vn = (vx * nrm[0] + vy * nrm[1]) / nn;
          uout[0] = rho;
          uout[1] = rho * (vx - 2.0 * vn * nrm[0]);
          uout[2] = rho * (vy - 2.0 * vn * nrm[1]);
          uout[3] = uin[3];
        } else if (kind == ORC_BC_OUTFLOW) {
          for (int v = 0; v < nv; ++v) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) {
              const int dof = dim == 0 ? bj * n + k : k * n + bi;
              acc += U(u, e, dof, v);
            }
            uout[v] = acc / (double)n;
          }
        } else if (kind == ORC_BC_MIXED_DIRICHLET_SLIP_WALL) {
          for (int v = 0; v < nv; ++v) uout[v] = c->bc_values[face][bnode * (nv + 1) + v];
        } else {
          for (int v = 0; v < nv; ++v)
            uout[v] = c->bc_values[face] ? c->bc_values[face][bnode * nv + v] : c->bc_state[face][v];
        }
        if (upper)
          surface_flux_fn(c, c->surface_flux, uin, uout, nrm, f);
        else
          surface_flux_fn(c, c->surface_flux, uout, uin, nrm, f);
        for (int v = 0; v < nv; ++v) SF(sflux, e, face * fn + q, v) = f[v];
      }
    }
  }
}

void orc_set_boundary(orc_case* c, int face, int kind, const double* state,
                      const double* values) {
  c->bc[face] = kind;
  for (int v = 0; v < 5; ++v) c->bc_state[face][v] = state ? state[v] : 0.0;
  c->bc_values[face] = values;
}

/* SurfaceIntegral<.,2/3>::integral, surface_integral.hpp:35-80,82-154 */
static void surface_integral_generic(orc_case* c, const double* sflux,
                                     double* du, int nv) {
  const int d = c->dim, fn = face_nodes(c);
  const double f1 = c->lhat_l[0], f2 = c->lhat_r[c->n - 1];
  int64_t e;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int dir = 0; dir < d; ++dir)
      for (int q = 0; q < fn; ++q) {
        const int a = q % c->n, b = q / c->n;
        const int lo = face_dof_node(c, dir, 0, a, b);
        const int hi = face_dof_node(c, dir, 1, a, b);
        for (int v = 0; v < nv; ++v) {
          du[(e * c->nn + lo) * nv + v] -= SF(sflux, e, (2 * dir) * fn + q, v) * f1;
          du[(e * c->nn + hi) * nv + v] += SF(sflux, e, (2 * dir + 1) * fn + q, v) * f2;
        }
      }
}

void orc_surface_integral(orc_case* c, const double* sflux, double* du) {
  surface_integral_generic(c, sflux, du, c->nvars);
}

/* JacobianProjFunctor, jacobian.hpp:65-101 */
void orc_apply_jacobian(orc_case* c, double* du) {
  int64_t e;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      const double factor = -INVJ(e, dof);
      for (int v = 0; v < c->nvars; ++v) U(du, e, dof, v) *= factor;
    }
}

/* ------------------------------------------------------------------------ */
/* BR1 parabolic terms: parabolic_gradient.hpp, parabolic_viscous.hpp        */
/* ------------------------------------------------------------------------ */

/* gradient_variables: compressible_navier_stokes2D.hpp:56-69,
 * compressible_navier_stokes3D.hpp:55-70 */
static void gradient_variables(const orc_case* c, const double* u, double* q) {
  const double g = c->gamma;
  const double rho = u[0];
  const double inv_rho = 1.0 / rho;
  if (c->dim == 2) {
    const double v1 = u[1] * inv_rho, v2 = u[2] * inv_rho;
    const double p = (g - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) * inv_rho);
    q[0] = rho; q[1] = v1; q[2] = v2; q[3] = p * inv_rho;
  } else {
    const double v1 = u[1] * inv_rho, v2 = u[2] * inv_rho, v3 = u[3] * inv_rho;
    const double p = (g - 1.0) *
                     (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * inv_rho);
    q[0] = rho; q[1] = v1; q[2] = v2; q[3] = v3; q[4] = p * inv_rho;
  }
}

/* viscous_flux: compressible_navier_stokes2D.hpp:71-101,
 * compressible_navier_stokes3D.hpp:72-123.  grad[dim][var]. */
static void viscous_flux_fn(const orc_case* c, const double* q,
                            const double grad[3][5], int dim, double* f) {
  const double mu = c->mu, kappa = c->kappa;
  if (c->dim == 2) {
    const double v1 = q[1], v2 = q[2];
    const double dv1dx = grad[0][1], dv2dx = grad[0][2], dTdx = grad[0][3];
    const double dv1dy = grad[1][1], dv2dy = grad[1][2], dTdy = grad[1][3];
    const double tau11 = (4.0 * dv1dx - 2.0 * dv2dy) / 3.0;
    const double tau22 = (4.0 * dv2dy - 2.0 * dv1dx) / 3.0;
    const double tau12 = dv1dy + dv2dx;
    const double q1 = kappa * dTdx, q2 = kappa * dTdy;
    if (dim == 0) {
      f[0] = 0.0; f[1] = mu * tau11; f[2] = mu * tau12;
      f[3] = mu * (v1 * tau11 + v2 * tau12 + q1);
    } else {
      f[0] = 0.0; f[1] = mu * tau12; f[2] = mu * tau22;
      f[3] = mu * (v1 * tau12 + v2 * tau22 + q2);
    }
  } else {
    const double v1 = q[1], v2 = q[2], v3 = q[3];
    const double dv1dx = grad[0][1], dv2dx = grad[0][2], dv3dx = grad[0][3], dTdx = grad[0][4];
    const double dv1dy = grad[1][1], dv2dy = grad[1][2], dv3dy = grad[1][3], dTdy = grad[1][4];
    const double dv1dz = grad[2][1], dv2dz = grad[2][2], dv3dz = grad[2][3], dTdz = grad[2][4];
    const double tau11 = (4.0 * dv1dx - 2.0 * (dv2dy + dv3dz)) / 3.0;
    const double tau22 = (4.0 * dv2dy - 2.0 * (dv1dx + dv3dz)) / 3.0;
    const double tau33 = (4.0 * dv3dz - 2.0 * (dv1dx + dv2dy)) / 3.0;
    const double tau12 = dv1dy + dv2dx;
    const double tau13 = dv1dz + dv3dx;
    const double tau23 = dv2dz + dv3dy;
    const double q1 = kappa * dTdx, q2 = kappa * dTdy, q3 = kappa * dTdz;
    if (dim == 0) {
      f[0] = 0.0; f[1] = mu * tau11; f[2] = mu * tau12; f[3] = mu * tau13;
      f[4] = mu * (v1 * tau11 + v2 * tau12 + v3 * tau13 + q1);
    } else if (dim == 1) {
      f[0] = 0.0; f[1] = mu * tau12; f[2] = mu * tau22; f[3] = mu * tau23;
      f[4] = mu * (v1 * tau12 + v2 * tau22 + v3 * tau23 + q2);
    } else {
      f[0] = 0.0; f[1] = mu * tau13; f[2] = mu * tau23; f[3] = mu * tau33;
      f[4] = mu * (v1 * tau13 + v2 * tau23 + v3 * tau33 + q3);
    }
  }
}

/* The two pointwise functions above, exported for unit tests; grad is
 * [dim][nvars] row-major. */
void orc_gradient_variables(const orc_case* c, const double* u, double* q) {
  gradient_variables(c, u, q);
}
void orc_viscous_flux(const orc_case* c, const double* q, const double* grad, int dim, double* f) {
  double g[3][5] = {{0}};
  for (int d = 0; d < c->dim; ++d)
    for (int v = 0; v < c->nvars; ++v) g[d][v] = grad[d * c->nvars + v];
  viscous_flux_fn(c, q, (const double(*)[5])g, dim, f);
}

static void ensure_parabolic_scratch(orc_case* c) {
  const int64_t E = c->E;
  const int nv = c->nvars, d = c->dim, fn = face_nodes(c);
  if (!c->grad) c->grad = (double*)calloc((size_t)(E * c->nn * nv * d), sizeof(double));
  if (!c->vflux) c->vflux = (double*)calloc((size_t)(E * c->nn * nv * d), sizeof(double));
  if (!c->gflux) c->gflux = (double*)calloc((size_t)(E * 2 * d * fn * nv), sizeof(double));
  if (!c->vsflux) c->vsflux = (double*)calloc((size_t)(E * 2 * d * fn * nv), sizeof(double));
}

#define G(arr, e, dof, var, dm) (arr)[(((e) * (int64_t)c->nn + (dof)) * nv + (var)) * d + (dm)]

/* Boundary node q of element position idx along a face of a 2D mesh:
 * element, boundary dof and the normal sign(J) Ja^dim there
 * (utils/boundary_helper.hpp:69-150). */
static void wall_node(const orc_case* c, int face, int idx, int q, int64_t* e, int* bdof,
                      double nrm[2]) {
  const int n = c->n, dim = face / 2, upper = face % 2;
  const int ie = dim == 0 ? (upper ? c->nx - 1 : 0) : idx;
  const int je = dim == 0 ? idx : (upper ? c->ny - 1 : 0);
  *e = (int64_t)ie * c->ny + je;
  const int bi = dim == 0 ? (upper ? n - 1 : 0) : q;
  const int bj = dim == 0 ? q : (upper ? n - 1 : 0);
  *bdof = bj * n + bi;
  const double sgn = INVJ(*e, *bdof) >= 0.0 ? 1.0 : -1.0;
  for (int comp = 0; comp < 2; ++comp) nrm[comp] = sgn * JA(*e, *bdof, dim, comp);
}

/* Gradient face values (apply_gradient_device hooks, dispatched by
 * boundary_set.hpp:16-33):
 *   adiabatic walls: the inner node's gradient variables with the velocity
 *     set to zero (NSNoSlipAdiabaticWallBC, navier_stokes_wall_boundary.hpp:
 *     81-114) or with its normal component removed (NSSlipAdiabaticWallBC,
 *     :163-197 with remove_normal_gradient_velocity_2d :28-36);
 *   Dirichlet: the gradient variables of the outer state
 *     (DirichletBC::apply_gradient_device, dirichlet_boundary.hpp:122-162).
 * Faces of the other kinds (slip wall, outflow, mixed) have no gradient hook
 * and keep the zero the face array was cleared to (structured_solver.hpp:250). */
static void wall_gradient_values(orc_case* c, const double* u) {
  if (c->dim != 2) return;
  const int n = c->n, nv = c->nvars, fn = face_nodes(c);
  for (int face = 0; face < 4; ++face) {
    const int kind = c->bc[face];
    if (kind != ORC_BC_NS_NOSLIP_WALL && kind != ORC_BC_NS_SLIP_WALL && kind != ORC_BC_DIRICHLET)
      continue;
    const int count = face < 2 ? c->ny : c->nx;
    for (int idx = 0; idx < count; ++idx)
      for (int q = 0; q < n; ++q) {
        int64_t e;
        int bdof;
        double nrm[2], s[4], qw[4];
        wall_node(c, face, idx, q, &e, &bdof, nrm);
        const int bnode = idx * n + q;
        for (int v = 0; v < nv; ++v)
          s[v] = kind != ORC_BC_DIRICHLET ? U(u, e, bdof, v)
                 : c->bc_values[face]    ? c->bc_values[face][bnode * nv + v]
                                         : c->bc_state[face][v];
        gradient_variables(c, s, qw);
        if (kind == ORC_BC_DIRICHLET) {
          /* the outer state's variables as they are */
        } else if (kind == ORC_BC_NS_NOSLIP_WALL) {
          qw[1] = 0.0;
          qw[2] = 0.0;
        } else {
          const double nn = nrm[0] * nrm[0] + nrm[1] * nrm[1];
          const double vn = (qw[1] * nrm[0] + qw[2] * nrm[1]) / nn;
          qw[1] -= vn * nrm[0];
          qw[2] -= vn * nrm[1];
        }
        for (int v = 0; v < nv; ++v) SF(c->gflux, e, face * fn + q, v) = qw[v];
      }
  }
}

/* Viscous face values (apply_viscous_device hooks, boundary_set.hpp:35-51):
 * the no-slip adiabatic wall takes the inner node's (-F_v).n for the
 * momentum components and zero mass and energy flux
 * (NSNoSlipAdiabaticWallBC, navier_stokes_wall_boundary.hpp:116-145);
 * Dirichlet takes the inner node's (-F_v).n for every variable
 * (dirichlet_boundary.hpp:164-194); the slip adiabatic wall's values are all
 * zero (:199-216), as are those of kinds without a viscous hook. */
static void wall_viscous_values(orc_case* c) {
  if (c->dim != 2) return;
  const int n = c->n, nv = c->nvars, d = c->dim, fn = face_nodes(c);
  for (int face = 0; face < 4; ++face) {
    const int kind = c->bc[face];
    if (kind != ORC_BC_NS_NOSLIP_WALL && kind != ORC_BC_DIRICHLET) continue;
    const int lo = kind == ORC_BC_DIRICHLET ? 0 : 1, hi = kind == ORC_BC_DIRICHLET ? nv : nv - 1;
    const int count = face < 2 ? c->ny : c->nx;
    for (int idx = 0; idx < count; ++idx)
      for (int q = 0; q < n; ++q) {
        int64_t e;
        int bdof;
        double nrm[2];
        wall_node(c, face, idx, q, &e, &bdof, nrm);
        for (int v = 0; v < nv; ++v)
          SF(c->vsflux, e, face * fn + q, v) =
              v >= lo && v < hi
                  ? G(c->vflux, e, bdof, v, 0) * nrm[0] + G(c->vflux, e, bdof, v, 1) * nrm[1]
                  : 0.0;
      }
  }
}

/* calc_gradients sequence, structured_solver.hpp:437-449 with
 * GradientVolumeIntegralFunctor (parabolic_gradient.hpp:40-79 / :502-527),
 * GradientPhysicalTransformFunctor (:121-150 / :567-584),
 * GradientInterfaceFluxFunctor (:186-291 / :628-685),
 * GradientSurfaceIntegralFunctor (:348-430 / :737-788),
 * GradientJacobianProjFunctor (:462-474 / :806-815) */
void orc_gradients(orc_case* c, const double* u, double* grad) {
  const int d = c->dim, n = c->n, nv = c->nvars, fn = face_nodes(c);
  int64_t e;
  ensure_parabolic_scratch(c);
  memset(grad, 0, sizeof(double) * (size_t)(c->E * c->nn * nv * d));
  /* volume: Dhat * q along each direction */
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e) {
    double s[5], q[5];
    for (int k = 0; k < (d == 3 ? n : 1); ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int dof = d == 2 ? ldof2(n, i, j) : ldof3(n, i, j, k);
          for (int v = 0; v < nv; ++v) s[v] = U(u, e, dof, v);
          gradient_variables(c, s, q);
          const int idx[3] = {i, j, k};
          for (int dir = 0; dir < d; ++dir)
            for (int o = 0; o < n; ++o) {
              int oi[3] = {i, j, k};
              oi[dir] = o;
              const int od = d == 2 ? ldof2(n, oi[0], oi[1]) : ldof3(n, oi[0], oi[1], oi[2]);
              for (int v = 0; v < nv; ++v)
                G(grad, e, od, v, dir) += c->Dhat[o * n + idx[dir]] * q[v];
            }
        }
  }
  /* transform to physical: Ja^T */
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof)
      for (int v = 0; v < nv; ++v) {
        if (d == 2) {
          const double rx = G(grad, e, dof, v, 0), ry = G(grad, e, dof, v, 1);
          G(grad, e, dof, v, 0) = JA(e, dof, 0, 0) * rx + JA(e, dof, 1, 0) * ry;
          G(grad, e, dof, v, 1) = JA(e, dof, 0, 1) * rx + JA(e, dof, 1, 1) * ry;
        } else {
          const double rx = G(grad, e, dof, v, 0), ry = G(grad, e, dof, v, 1),
                       rz = G(grad, e, dof, v, 2);
          G(grad, e, dof, v, 0) =
              JA(e, dof, 0, 0) * rx + JA(e, dof, 1, 0) * ry + JA(e, dof, 2, 0) * rz;
          G(grad, e, dof, v, 1) =
              JA(e, dof, 0, 1) * rx + JA(e, dof, 1, 1) * ry + JA(e, dof, 2, 1) * rz;
          G(grad, e, dof, v, 2) =
              JA(e, dof, 0, 2) * rx + JA(e, dof, 1, 2) * ry + JA(e, dof, 2, 2) * rz;
        }
      }
  /* interface: qhat = 0.5 (q_L + q_R), stored on both sides */
  memset(c->gflux, 0, sizeof(double) * (size_t)(c->E * 2 * d * fn * nv));
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e) {
    double s[5], ql[5], qr[5];
    for (int dir = 0; dir < d; ++dir) {
      int64_t nb[3], valid = 1;
      for (int comp = 0; comp < d; ++comp) {
        nb[comp] = c->nbr[(e * d + dir) * d + comp];
        if (nb[comp] < 0) valid = 0;
      }
      if (!valid) continue;
      const int64_t le = d == 2 ? nb[0] * c->ny + nb[1]
                                : (nb[0] * c->ny + nb[1]) * c->nz + nb[2];
      for (int q = 0; q < fn; ++q) {
        const int a = q % n, b = q / n;
        const int ldof = face_dof_node(c, dir, 1, a, b);
        const int rdof = face_dof_node(c, dir, 0, a, b);
        for (int v = 0; v < nv; ++v) s[v] = U(u, le, ldof, v);
        gradient_variables(c, s, ql);
        for (int v = 0; v < nv; ++v) s[v] = U(u, e, rdof, v);
        gradient_variables(c, s, qr);
        for (int v = 0; v < nv; ++v) {
          const double qhat = 0.5 * (ql[v] + qr[v]);
          SF(c->gflux, le, (2 * dir + 1) * fn + q, v) = qhat;
          SF(c->gflux, e, (2 * dir) * fn + q, v) = qhat;
        }
      }
    }
  }
  wall_gradient_values(c, u);
  /* surface: grad(face node) += (-+ lhat) * qhat * sign(J) Ja_dir */
  {
    const double f1 = c->lhat_l[0], f2 = c->lhat_r[n - 1];
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
    for (e = 0; e < c->E; ++e)
      for (int dir = 0; dir < d; ++dir)
        for (int q = 0; q < fn; ++q) {
          const int a = q % n, b = q / n;
          for (int side = 0; side < 2; ++side) {
            const int dof = face_dof_node(c, dir, side, a, b);
            const double factor = side ? f2 : -f1;
            const double sgn = INVJ(e, dof) >= 0.0 ? 1.0 : -1.0;
            double nrm[3];
            for (int comp = 0; comp < d; ++comp) nrm[comp] = sgn * JA(e, dof, dir, comp);
            for (int v = 0; v < nv; ++v) {
              const double qhat = SF(c->gflux, e, (2 * dir + side) * fn + q, v);
              for (int dm = 0; dm < d; ++dm) G(grad, e, dof, v, dm) += factor * qhat * nrm[dm];
            }
          }
        }
  }
  /* x 1/J */
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      const double factor = INVJ(e, dof);
      for (int v = 0; v < nv; ++v)
        for (int dm = 0; dm < d; ++dm) G(grad, e, dof, v, dm) *= factor;
    }
}

/* Viscous part of calc_parabolic_rhs_terms, structured_solver.hpp:451-460:
 * ViscousFluxFunctor (parabolic_viscous.hpp:39-66 / :367-386) stores -F_v,
 * ViscousVolumeIntegralFunctor (:103-163 / :412-447),
 * ViscousInterfaceFluxFunctor (:213-334 / :480-585): one-sided, primary
 * (right/top/front) element's (-F_v).n only, and the viscous surface integral
 * (structured_solver.hpp:359-360). */
static void viscous_terms(orc_case* c, const double* u, double* du) {
  const int d = c->dim, n = c->n, nv = c->nvars, fn = face_nodes(c);
  int64_t e;
  ensure_parabolic_scratch(c);
  orc_gradients(c, u, c->grad);
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      double s[5], q[5], g[3][5], f[5];
      for (int v = 0; v < nv; ++v) s[v] = U(u, e, dof, v);
      gradient_variables(c, s, q);
      for (int dm = 0; dm < d; ++dm)
        for (int v = 0; v < nv; ++v) g[dm][v] = G(c->grad, e, dof, v, dm);
      for (int dm = 0; dm < d; ++dm) {
        viscous_flux_fn(c, q, g, dm, f);
        for (int v = 0; v < nv; ++v) G(c->vflux, e, dof, v, dm) = -f[v];
      }
    }
  /* viscous volume integral */
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int k = 0; k < (d == 3 ? n : 1); ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int dof = d == 2 ? ldof2(n, i, j) : ldof3(n, i, j, k);
          double fref[3][5];
          for (int v = 0; v < nv; ++v) {
            if (d == 2) {
              const double f1 = G(c->vflux, e, dof, v, 0), f2 = G(c->vflux, e, dof, v, 1);
              fref[0][v] = JA(e, dof, 0, 0) * f1 + JA(e, dof, 0, 1) * f2;
              fref[1][v] = JA(e, dof, 1, 0) * f1 + JA(e, dof, 1, 1) * f2;
            } else {
              const double f1 = G(c->vflux, e, dof, v, 0), f2 = G(c->vflux, e, dof, v, 1),
                           f3 = G(c->vflux, e, dof, v, 2);
              fref[0][v] = JA(e, dof, 0, 0) * f1 + JA(e, dof, 0, 1) * f2 + JA(e, dof, 0, 2) * f3;
              fref[1][v] = JA(e, dof, 1, 0) * f1 + JA(e, dof, 1, 1) * f2 + JA(e, dof, 1, 2) * f3;
              fref[2][v] = JA(e, dof, 2, 0) * f1 + JA(e, dof, 2, 1) * f2 + JA(e, dof, 2, 2) * f3;
            }
          }
          const int idx[3] = {i, j, k};
          for (int dir = 0; dir < d; ++dir)
            for (int o = 0; o < n; ++o) {
              int oi[3] = {i, j, k};
              oi[dir] = o;
              const int od = d == 2 ? ldof2(n, oi[0], oi[1]) : ldof3(n, oi[0], oi[1], oi[2]);
              for (int v = 0; v < nv; ++v)
                U(du, e, od, v) += c->Dhat[o * n + idx[dir]] * fref[dir][v];
            }
        }
  /* viscous interface flux (one-sided) */
  memset(c->vsflux, 0, sizeof(double) * (size_t)(c->E * 2 * d * fn * nv));
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int dir = 0; dir < d; ++dir) {
      int64_t nb[3], valid = 1;
      for (int comp = 0; comp < d; ++comp) {
        nb[comp] = c->nbr[(e * d + dir) * d + comp];
        if (nb[comp] < 0) valid = 0;
      }
      if (!valid) continue;
      const int64_t le = d == 2 ? nb[0] * c->ny + nb[1]
                                : (nb[0] * c->ny + nb[1]) * c->nz + nb[2];
      for (int q = 0; q < fn; ++q) {
        const int a = q % n, b = q / n;
        const int rdof = face_dof_node(c, dir, 0, a, b);
        const double sgn = INVJ(e, rdof) >= 0.0 ? 1.0 : -1.0;
        double nrm[3];
        for (int comp = 0; comp < d; ++comp) nrm[comp] = sgn * JA(e, rdof, dir, comp);
        for (int v = 0; v < nv; ++v) {
          double fp;
          if (d == 2)
            fp = G(c->vflux, e, rdof, v, 0) * nrm[0] + G(c->vflux, e, rdof, v, 1) * nrm[1];
          else
            fp = G(c->vflux, e, rdof, v, 0) * nrm[0] + G(c->vflux, e, rdof, v, 1) * nrm[1] +
                 G(c->vflux, e, rdof, v, 2) * nrm[2];
          SF(c->vsflux, e, (2 * dir) * fn + q, v) = fp;
          SF(c->vsflux, le, (2 * dir + 1) * fn + q, v) = fp;
        }
      }
    }
  wall_viscous_values(c);
  surface_integral_generic(c, c->vsflux, du, nv);
}

/* ------------------------------------------------------------------------ */
/* calc_rhs and SSPRK3                                                       */
/* ------------------------------------------------------------------------ */

/* StructuredSolver::calc_rhs, structured_solver.hpp:463-480 */
void orc_rhs(orc_case* c, const double* u, double* du, double* alpha) {
  const int64_t total = c->E * c->nn * c->nvars;
  const int nv = c->nvars, d = c->dim;
  (void)d;
  if (!c->sflux)
    c->sflux = (double*)calloc((size_t)(c->E * 2 * c->dim * face_nodes(c) * nv),
                               sizeof(double));
  memset(du, 0, sizeof(double) * (size_t)total);
  orc_volume_integral(c, u, du, alpha);
  orc_interface_flux(c, u, c->sflux);
  /* periodic boundary conditions are no-ops (periodic_boundary.hpp:11-21) */
  boundary_fluxes(c, u, c->sflux);
  if (c->eq == ORC_NAVIER_STOKES) viscous_terms(c, u, du);
  orc_surface_integral(c, c->sflux, du);
  orc_apply_jacobian(c, du);
}

/* Positivity floor after a stage update -- build extension: density is
 * clamped to the floor and the energy is raised so that the pressure is at
 * least the floor. */
static void apply_floor(const orc_case* c, double* u) {
  if (!(c->floor_value > 0.0)) return;
  const int nv = c->nvars, d = c->dim;
  const double fl = c->floor_value;
  int64_t e;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
  for (e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      double* s = &U(u, e, dof, 0);
      if (s[0] < fl) s[0] = fl;
      double m2 = 0.0;
      for (int k = 0; k < d; ++k) m2 += s[k + 1] * s[k + 1];
      const double kin = 0.5 * m2 / s[0];
      const double p = (c->gamma - 1.0) * (s[nv - 1] - kin);
      if (p < fl) s[nv - 1] = fl / (c->gamma - 1.0) + kin;
    }
}

/* Test hook, no counterpart in the reference: every later shock-capturing
 * volume integral blends with the given alpha ([E], kept by reference)
 * instead of the indicator's, so that finite differences of the RHS hold
 * alpha fixed.  NULL restores the indicator. */
void orc_set_alpha_override(orc_case* c, const double* alpha) { c->alpha_override = alpha; }

/* The floor alone on a state, as every stage applies it. */
void orc_apply_floor(const orc_case* c, double* u) { apply_floor(c, u); }

/* parallel_ma2 / parallel_ma3 (SSPRK3.hpp:12-143) and SSPRK3::step
 * (SSPRK3.hpp:159-181) */
void orc_ssprk3(orc_case* c, double* u, double dt, int nsteps) {
  const int64_t total = c->E * c->nn * c->nvars;
  int64_t i;
  if (!c->tmp1) {
    c->tmp1 = (double*)calloc((size_t)total, sizeof(double));
    c->tmp2 = (double*)calloc((size_t)total, sizeof(double));
    c->du = (double*)calloc((size_t)total, sizeof(double));
  }
  double *tmp1 = c->tmp1, *tmp2 = c->tmp2, *du = c->du;
  for (int s = 0; s < nsteps; ++s) {
    orc_rhs(c, u, du, NULL);
    {
      const double a1 = 1.0, a2 = dt;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
      for (i = 0; i < total; ++i) tmp1[i] = a1 * u[i] + a2 * du[i];
    }
    apply_floor(c, tmp1);
    orc_rhs(c, tmp1, du, NULL);
    {
      const double a1 = 3.0 / 4.0, a2 = 1.0 / 4.0, a3 = 1.0 / 4.0 * dt;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
      for (i = 0; i < total; ++i) tmp2[i] = a1 * u[i] + a2 * tmp1[i] + a3 * du[i];
    }
    apply_floor(c, tmp2);
    orc_rhs(c, tmp2, du, NULL);
    {
      const double a1 = 1.0 / 3.0, a2 = 2.0 / 3.0, a3 = 2.0 / 3.0 * dt;
#pragma omp parallel for num_threads(nthreads_of(c)) schedule(static)
      for (i = 0; i < total; ++i) u[i] = a1 * u[i] + a2 * tmp2[i] + a3 * du[i];
    }
    apply_floor(c, u);
  }
}

/* VolumeAverageEuler kinetic-energy integral with the volume weights of
 * VolumeWeightedAnalyzerFunctor (volume_average_euler.hpp:30-53,
 * volume_weighted_analyzer_functor.hpp:108-134), summed in element order. */
double orc_kinetic_energy(const orc_case* c, const double* u) {
  const int n = c->n, d = c->dim;
  double total = 0.0;
  for (int64_t e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      const int i = dof % n, j = (dof / n) % n, k = dof / (n * n);
      double w = c->weights[i] * c->weights[j];
      if (d == 3) w = w * c->weights[k];
      const double vw = w / fabs(INVJ(e, dof));
      double m2 = 0.0;
      for (int k2 = 0; k2 < d; ++k2) m2 += U(u, e, dof, k2 + 1) * U(u, e, dof, k2 + 1);
      total += vw * (0.5 * m2 / U(u, e, dof, 0));
    }
  return total;
}


/* VolumeAverageEuler (volume_average_euler.hpp:30-53) over the volume
 * weights of VolumeWeightedAnalyzerFunctor (volume_weighted_analyzer_
 * functor.hpp:108-134), DivergenceChecker's non-finite test
 * (divergence_checker.hpp:6-27) and the largest max_d |v_d| + c, in element
 * order: out = [volume, kinetic energy, entropy, non-finite nodes, max speed].
 * Non-finite nodes are left out of the sums. */
void orc_analyze(const orc_case* c, const double* u, double* out) {
  const int n = c->n, d = c->dim;
  double vol = 0.0, ke = 0.0, ent = 0.0, bad = 0.0, vmax = 0.0;
  for (int64_t e = 0; e < c->E; ++e)
    for (int dof = 0; dof < c->nn; ++dof) {
      int finite = 1;
      for (int v = 0; v < c->nvars; ++v)
        if (!isfinite(U(u, e, dof, v))) finite = 0;
      if (!finite) {
        bad += 1.0;
        continue;
      }
      const int i = dof % n, j = (dof / n) % n, k = dof / (n * n);
      double w = c->weights[i] * c->weights[j];
      if (d == 3) w = w * c->weights[k];
      const double vw = w / fabs(INVJ(e, dof));
      const double rho = U(u, e, dof, 0);
      double m2 = 0.0, v2 = 0.0, vabs = 0.0;
      for (int k2 = 0; k2 < d; ++k2) {
        const double m = U(u, e, dof, k2 + 1), vk = m / rho;
        m2 += m * m;
        v2 += vk * vk;
        if (fabs(vk) > vabs) vabs = fabs(vk);
      }
      const double p = (c->gamma - 1.0) * (U(u, e, dof, d + 1) - 0.5 * rho * v2);
      const double s = log(p) - c->gamma * log(rho);
      vol += vw;
      ke += vw * (0.5 * m2 / rho);
      ent += vw * (-rho * s / (c->gamma - 1.0));
      const double sp = vabs + sqrt(c->gamma * p / rho);
      if (sp > vmax) vmax = sp;
    }
  out[0] = vol;
  out[1] = ke;
  out[2] = ent;
  out[3] = bad;
  out[4] = vmax;
}
