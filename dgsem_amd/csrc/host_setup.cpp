// Host construction of basis tables and geometry.  See host_setup.hpp.
// Built with g++ -O2 -ffp-contract=off (no FMA contraction) so the values are
// bitwise the ones the reference computes on its host.
#include "host_setup.hpp"

#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <numbers>
#include <stdexcept>
#include <utility>

namespace dgsem {
namespace {

struct QL {
  double q, qder, LN;
};

// Legendre recursion for q = (2N+1)/(N+1) (x L_N - L_{N-1}) and its
// derivative, polynomial_basis/basis_detail.hpp:27-50.
QL eval_q_and_l(std::size_t p, double x) {
  double l2 = 1.0, l1 = x, d2 = 0.0, d1 = 1.0, lN = 0.0, dN = 0.0;
  for (std::size_t i = 2; i <= p; ++i) {
    lN = ((double)(2.0 * i - 1.0) * x * l1 - (double)(i - 1.0) * l2) / (double)i;
    dN = d2 + (double)(2.0 * i - 1.0) * l1;
    l2 = l1;
    l1 = lN;
    d2 = d1;
    d1 = dN;
  }
  QL r;
  r.q = (double)(2.0 * p + 1.0) / ((double)p + 1.0) * (x * lN - l2);
  r.qder = (2.0 * p + 1.0) * lN;
  r.LN = lN;
  return r;
}

// Newton iteration from a Chebyshev-like guess, basis_detail.hpp:52-109.
void gauss_lobatto(int nn, double* x, double* w, double* iw) {
  const std::size_t P = (std::size_t)nn - 1;
  const double tol = 2.0 * std::numeric_limits<double>::epsilon();
  x[0] = -1.0;
  x[nn - 1] = 1.0;
  w[0] = 2.0 / ((double)P * ((double)P + 1.0));
  w[nn - 1] = w[0];
  const double c1 = std::numbers::pi / (double)P;
  const double c2 = 3.0 / (8.0 * (double)P * std::numbers::pi);
  for (std::size_t i = 0; i < (std::size_t)(nn / 2 - 1); ++i) {
    const double s = (double)(i + 1) + 0.25;
    double xi = -std::cos(c1 * s - c2 / s);
    for (int it = 0; it < 20; ++it) {
      const QL r = eval_q_and_l(P, xi);
      const double dx = -r.q / r.qder;
      xi += dx;
      if (std::abs(dx) < tol * std::abs(xi)) break;
    }
    const QL r = eval_q_and_l(P, xi);
    x[i + 1] = xi;
    w[i + 1] = w[0] / (r.LN * r.LN);
    x[nn - 2 - i] = -xi;
    w[nn - 2 - i] = w[i + 1];
  }
  if (nn % 2 == 1) {
    x[nn / 2] = 0.0;
    const QL r = eval_q_and_l(P, 0.0);
    w[nn / 2] = w[0] / (r.LN * r.LN);
  }
  for (int i = 0; i < nn; ++i) iw[i] = 1.0 / w[i];
}

std::vector<double> barycentric(const std::vector<double>& x) {
  const std::size_t n = x.size();
  std::vector<double> wb(n, 1.0);
  for (std::size_t j = 1; j < n; ++j)
    for (std::size_t k = 0; k < j; ++k) {
      wb[k] *= (x[k] - x[j]);
      wb[j] *= (x[j] - x[k]);
    }
  for (auto& v : wb) v = 1.0 / v;
  return wb;
}

// l_i(x)/w_i, basis_detail.hpp:130-168.
std::vector<double> boundary_lift(double xe, const std::vector<double>& x,
                                  const std::vector<double>& w) {
  const std::size_t n = x.size();
  std::vector<double> l(n, 0.0);
  bool hit = false;
  for (std::size_t i = 0; i < n && !hit; ++i)
    if (std::abs(xe - x[i]) < std::numeric_limits<double>::epsilon()) {
      l[i] = 1.0;
      hit = true;
    }
  if (!hit) {
    const auto wb = barycentric(x);
    double total = 0.0;
    for (std::size_t i = 0; i < n; ++i) l[i] = wb[i] / (xe - x[i]);
    for (std::size_t i = 0; i < n; ++i) total += l[i];
    for (auto& v : l) v /= total;
  }
  for (std::size_t i = 0; i < n; ++i) l[i] /= w[i];
  return l;
}

// Normalised Legendre P_k(x) sqrt(k+1/2), basis_detail.hpp:217-248.
double legendre_normalised(std::size_t k, double x) {
  double p = 0.0;
  if (k == 0) {
    p = 1.0;
  } else if (k == 1) {
    p = x;
  } else {
    double a = 1.0, b = x;
    for (std::size_t i = 2; i <= k; ++i) {
      p = ((double)(2.0 * i - 1.0) * x * b - (double)(i - 1.0) * a) / (double)i;
      a = b;
      b = p;
    }
  }
  return p * std::sqrt((double)k + 0.5);
}

// Gauss-Jordan with partial pivoting, base/container_fixed.hpp:117-162.
std::vector<double> invert(std::vector<double> A, int n) {
  std::vector<double> inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = std::abs(A[col * n + col]);
    for (int r = col + 1; r < n; ++r)
      if (std::abs(A[r * n + col]) > best) {
        best = std::abs(A[r * n + col]);
        piv = r;
      }
    if (piv != col)
      for (int k = 0; k < n; ++k) {
        std::swap(A[col * n + k], A[piv * n + k]);
        std::swap(inv[col * n + k], inv[piv * n + k]);
      }
    const double diag = A[col * n + col];
    for (int k = 0; k < n; ++k) {
      A[col * n + k] /= diag;
      inv[col * n + k] /= diag;
    }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      const double f = A[r * n + col];
      for (int k = 0; k < n; ++k) {
        A[r * n + k] -= f * A[col * n + k];
        inv[r * n + k] -= f * inv[col * n + k];
      }
    }
  }
  return inv;
}

}  // namespace

BasisTables make_basis(int N) {
  if (N < 1 || N + 1 > kMaxNodes) throw std::invalid_argument("polynomial degree out of range");
  BasisTables b;
  b.N = N;
  b.n = N + 1;
  const int n = b.n;
  b.nodes.assign(n, 0.0);
  b.weights.assign(n, 0.0);
  b.inv_weights.assign(n, 0.0);
  gauss_lobatto(n, b.nodes.data(), b.weights.data(), b.inv_weights.data());
  // Barycentric derivative matrix, basis_detail.hpp:170-188.
  const auto wb = barycentric(b.nodes);
  b.D.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j) {
        b.D[i * n + j] = wb[j] / wb[i] * 1.0 / (b.nodes[i] - b.nodes[j]);
        b.D[i * n + i] -= b.D[i * n + j];
      }
  // D_split = 2 D - M^{-1} B, basis_detail.hpp:190-200.
  b.Dsplit.resize(n * n);
  for (int k = 0; k < n * n; ++k) b.Dsplit[k] = 2.0 * b.D[k];
  b.Dsplit[0] += 1.0 / b.weights[0];
  b.Dsplit[n * n - 1] -= 1.0 / b.weights[n - 1];
  // Dhat = -M^{-1} D^T M, basis_detail.hpp:202-215.
  b.Dhat.resize(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) b.Dhat[j * n + i] = b.D[i * n + j];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) b.Dhat[j * n + i] *= -b.weights[i] / b.weights[j];
  // Inverse Legendre Vandermonde, basis_detail.hpp:250-262.
  std::vector<double> V(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i * n + j] = legendre_normalised((std::size_t)j, b.nodes[i]);
  b.Vinv = invert(V, n);
  b.lhat_left = boundary_lift(-1.0, b.nodes, b.weights);
  b.lhat_right = boundary_lift(1.0, b.nodes, b.weights);
  return b;
}

Geometry make_geometry(const BasisTables& b, int dim, const int* cells,
                       const double* left, const double* right,
                       const int* periodic) {
  if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
  Geometry g;
  g.dim = dim;
  g.N = b.N;
  g.n = b.n;
  const int n = b.n, d = dim;
  g.nn = d == 2 ? n * n : n * n * n;
  const int nn = g.nn;
  for (int k = 0; k < d; ++k) g.cells[k] = cells[k];
  g.E = (int64_t)g.cells[0] * g.cells[1] * g.cells[2];
  const int64_t E = g.E;
  g.coords.assign(E * nn * d, 0.0);
  g.jacobian.assign(E * nn * d * d, 0.0);
  g.contravariant.assign(E * nn * d * d, 0.0);
  g.inverse_jacobian.assign(E * nn, 0.0);
  g.left_neighbors.assign(E * d * d, 0);

  auto nidx = [&](int dof, int k) {  // node index along dimension k
    return k == 0 ? dof % n : (k == 1 ? (dof / n) % n : dof / (n * n));
  };
  auto dof_of = [&](const int* ix) {
    return d == 2 ? ix[1] * n + ix[0] : (ix[2] * n + ix[1]) * n + ix[0];
  };
  auto X = [&](int64_t e, int dof, int c) -> double& { return g.coords[(e * nn + dof) * d + c]; };
  auto JM = [&](int64_t e, int dof, int r, int c) -> double& {
    return g.jacobian[((e * nn + dof) * d + r) * d + c];
  };
  auto JA = [&](int64_t e, int dof, int r, int c) -> double& {
    return g.contravariant[((e * nn + dof) * d + r) * d + c];
  };

  // Node coordinates through LinearMapping, data_container.hpp:193-221/371-414.
  double delta[3];
  for (int k = 0; k < d; ++k) delta[k] = 2.0 / (double)cells[k];
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < g.cells[0]; ++ie)
    for (int je = 0; je < g.cells[1]; ++je)
      for (int ke = 0; ke < g.cells[2]; ++ke) {
        const int64_t e = ((int64_t)ie * g.cells[1] + je) * g.cells[2] + ke;
        const int eid[3] = {ie, je, ke};
        for (int dof = 0; dof < nn; ++dof)
          for (int k = 0; k < d; ++k) {
            const double off = -1.0 + (double)eid[k] * delta[k] + 0.5 * delta[k];
            const double r = off + 0.5 * delta[k] * b.nodes[nidx(dof, k)];
            X(e, dof, k) = 0.5 * ((1.0 - r) * left[k] + (1.0 + r) * right[k]);
          }
      }
  // Covariant Jacobian by the derivative matrix, :223-258/416-482.
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e)
    for (int dof = 0; dof < nn; ++dof)
      for (int r = 0; r < d; ++r) {
        double acc[3] = {0.0, 0.0, 0.0};
        int ix[3] = {nidx(dof, 0), nidx(dof, 1), d == 3 ? nidx(dof, 2) : 0};
        const int self = ix[r];
        for (int m = 0; m < n; ++m) {
          ix[r] = m;
          const int dm = dof_of(ix);
          for (int c = 0; c < d; ++c) acc[c] += b.D[self * n + m] * X(e, dm, c);
        }
        for (int c = 0; c < d; ++c) JM(e, dof, r, c) = acc[c];
      }
  // Contravariant vectors: 2D adjugate (:260-291), 3D curl form (:484-659).
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    if (d == 2) {
      for (int dof = 0; dof < nn; ++dof) {
        JA(e, dof, 0, 0) = JM(e, dof, 1, 1);
        JA(e, dof, 0, 1) = -JM(e, dof, 1, 0);
        JA(e, dof, 1, 0) = -JM(e, dof, 0, 1);
        JA(e, dof, 1, 1) = JM(e, dof, 0, 0);
      }
      continue;
    }
    for (int c = 0; c < 3; ++c) {
      const int m = (c + 1) % 3, l = (c + 2) % 3;
      for (int dof = 0; dof < nn; ++dof) {
        // The reference accumulates the second sum into the same running
        // value (result -= ...), so reproduce that association exactly.
        auto curl = [&](int tA, int rowA, int tB, int rowB) {
          int ix[3] = {nidx(dof, 0), nidx(dof, 1), nidx(dof, 2)};
          double r = 0.0;
          const int sA = ix[tA];
          for (int ii = 0; ii < n; ++ii) {
            ix[tA] = ii;
            const int q = dof_of(ix);
            r += 0.5 * b.D[sA * n + ii] * (X(e, q, m) * JM(e, q, rowA, l) - X(e, q, l) * JM(e, q, rowA, m));
          }
          ix[tA] = sA;
          const int sB = ix[tB];
          for (int ii = 0; ii < n; ++ii) {
            ix[tB] = ii;
            const int q = dof_of(ix);
            r -= 0.5 * b.D[sB * n + ii] * (X(e, q, m) * JM(e, q, rowB, l) - X(e, q, l) * JM(e, q, rowB, m));
          }
          return r;
        };
        JA(e, dof, 0, c) = curl(1, 2, 2, 1);  // (A)_eta with row 2, (B)_zeta with row 1
        JA(e, dof, 1, c) = curl(2, 0, 0, 2);  // (A)_zeta with row 0, (B)_xi with row 2
        JA(e, dof, 2, c) = curl(0, 1, 1, 0);  // (A)_xi with row 1, (B)_eta with row 0
      }
    }
  }
  // Inverse Jacobian determinant, :293-312/661-693.
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e)
    for (int dof = 0; dof < nn; ++dof) {
      double det;
      if (d == 2) {
        det = JM(e, dof, 0, 0) * JM(e, dof, 1, 1) - JM(e, dof, 1, 0) * JM(e, dof, 0, 1);
      } else {
        const double j11 = JM(e, dof, 0, 0), j12 = JM(e, dof, 0, 1), j13 = JM(e, dof, 0, 2);
        const double j21 = JM(e, dof, 1, 0), j22 = JM(e, dof, 1, 1), j23 = JM(e, dof, 1, 2);
        const double j31 = JM(e, dof, 2, 0), j32 = JM(e, dof, 2, 1), j33 = JM(e, dof, 2, 2);
        det = j11 * (j22 * j33 - j23 * j32) - j12 * (j21 * j33 - j23 * j31) + j13 * (j21 * j32 - j22 * j31);
      }
      g.inverse_jacobian[e * nn + dof] = 1.0 / det;
    }
  // Left neighbours with periodic wrap, :314-331/695-724.
  for (int ie = 0; ie < g.cells[0]; ++ie)
    for (int je = 0; je < g.cells[1]; ++je)
      for (int ke = 0; ke < g.cells[2]; ++ke) {
        const int64_t e = ((int64_t)ie * g.cells[1] + je) * g.cells[2] + ke;
        const int eid[3] = {ie, je, ke};
        for (int r = 0; r < d; ++r)
          for (int c = 0; c < d; ++c) {
            int64_t v = eid[c];
            if (c == r) v = eid[r] > 0 ? eid[r] - 1 : (periodic[r] ? cells[r] - 1 : -1);
            g.left_neighbors[(e * d + r) * d + c] = v;
          }
      }
  // Subcell normals by the recurrence n_0 = Ja(0) + w_0 sum_m D(0,m) Ja(m),
  // n_f = n_{f-1} + w_f sum_m D(f,m) Ja(m), global_def.hpp:175-251/299-439.
  const int FN = d == 2 ? n : n * n;
  g.subcell_normals.assign(E * d * FN * (n - 1) * d, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e)
    for (int r = 0; r < d; ++r)
      for (int line = 0; line < FN; ++line) {
        int ix[3] = {0, 0, 0};
        int o = 0;
        for (int k = 0; k < d; ++k)
          if (k != r) ix[k] = (o++ == 0) ? line % n : line / n;
        double acc[3];
        for (int f = 0; f < n - 1; ++f) {
          if (f == 0) {
            ix[r] = 0;
            const int d0 = dof_of(ix);
            for (int c = 0; c < d; ++c) acc[c] = JA(e, d0, r, c);
          }
          for (int m = 0; m < n; ++m) {
            ix[r] = m;
            const int dm = dof_of(ix);
            const double wD = b.weights[f] * b.D[f * n + m];
            for (int c = 0; c < d; ++c) acc[c] += wD * JA(e, dm, r, c);
          }
          double* out = &g.subcell_normals[((((e * d + r) * FN) + line) * (n - 1) + f) * d];
          for (int c = 0; c < d; ++c) out[c] = acc[c];
        }
      }
  return g;
}

bool box_metric_table(const Geometry& g, std::vector<double>& tab, int64_t off[3]) {
  const int d = g.dim, nn = g.nn;
  const int64_t E = g.E;
  for (int64_t i = 0; i < E * nn; ++i)
    if (!(g.inverse_jacobian[i] > 0.0)) return false;
  auto pairs = [&](int c) -> int64_t {
    if (d == 2) return g.cells[1 - c];
    return (int64_t)g.cells[(c + 1) % 3] * g.cells[(c + 2) % 3];
  };
  int64_t total = 0;
  for (int c = 0; c < 3; ++c) {
    off[c] = total;
    if (c < d) total += pairs(c) * d * nn;
  }
  tab.assign(total, 0.0);
  auto ids = [&](int64_t e, int* eid) {
    eid[0] = (int)(e / ((int64_t)g.cells[1] * g.cells[2]));
    eid[1] = (int)((e / g.cells[2]) % g.cells[1]);
    eid[2] = (int)(e % g.cells[2]);
  };
  auto pair_of = [&](const int* eid, int c) -> int64_t {
    if (d == 2) return eid[1 - c];
    return (int64_t)eid[(c + 1) % 3] * g.cells[(c + 2) % 3] + eid[(c + 2) % 3];
  };
  auto ja = [&](int64_t e, int dof, int r, int c) {
    return g.contravariant[((e * nn + dof) * d + r) * d + c];
  };
  // the rows of the elements with i_c = 0
  for (int64_t e = 0; e < E; ++e) {
    int eid[3];
    ids(e, eid);
    for (int c = 0; c < d; ++c) {
      if (eid[c] != 0) continue;
      const int64_t pr = pair_of(eid, c);
      for (int r = 0; r < d; ++r)
        for (int dof = 0; dof < nn; ++dof) tab[off[c] + (pr * d + r) * nn + dof] = ja(e, dof, r, c);
    }
  }
  // every element against its rows, bit for bit
  int ok = 1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int64_t e = 0; e < E; ++e) {
    int eid[3];
    ids(e, eid);
    for (int c = 0; c < d; ++c) {
      const int64_t pr = pair_of(eid, c);
      for (int r = 0; r < d; ++r)
        for (int dof = 0; dof < nn; ++dof) {
          const double a = ja(e, dof, r, c), b = tab[off[c] + (pr * d + r) * nn + dof];
          uint64_t ua, ub;
          std::memcpy(&ua, &a, 8);
          std::memcpy(&ub, &b, 8);
          ok = ok && ua == ub;
        }
    }
  }
  return ok != 0;
}

}  // namespace dgsem
