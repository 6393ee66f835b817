// Host-side construction of the constant operators and the mesh geometry.
//
// These are computed once on the host (like the reference, which builds them
// on the host and deep_copies them to the device) and uploaded.  The file is
// compiled by g++ with -ffp-contract=off so the values are the same bits the
// reference's host build produces.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dgsem {

constexpr int kMaxNodes = 8;  // N <= 7

struct BasisTables {
  int N = 0, n = 0;
  std::vector<double> nodes, weights, inv_weights, lhat_left, lhat_right;
  std::vector<double> D, Dsplit, Dhat, Vinv;  // row-major n x n
};

// LobattoLegendreBasis<T,N>::initialize, polynomial_basis/basis.hpp:69-153.
BasisTables make_basis(int N);

// Geometry of a StructuredMesh<T,dim> mapped by LinearMapping, with the
// reference's array layouts (data_container.hpp:12-60):
//   coords[E][nn][dim], jacobian/contravariant[E][nn][dim][dim],
//   inverse_jacobian[E][nn], left_neighbors[E][dim][dim] (-1 = none).
// Element index e = (ie*ny + je)*nz + ke (LayoutRight over (ie,je,ke)).
struct Geometry {
  int dim = 0, N = 0, n = 0, nn = 0;
  int cells[3] = {1, 1, 1};
  int64_t E = 0;
  std::vector<double> coords, jacobian, contravariant, inverse_jacobian;
  std::vector<int64_t> left_neighbors;
  // Subcell normals (global_def.hpp:153-457) re-ordered for the kernels as
  // [E][dir][line][iface][comp], line = position on the face of direction
  // dir (other indices in increasing dimension order, first fastest).
  std::vector<double> subcell_normals;
};

Geometry make_geometry(const BasisTables& b, int dim, const int* cells,
                       const double* left, const double* right,
                       const int* periodic);

// The box-mesh table of the contravariant vectors (the kernels' ja_box):
// Ja^r_c of an element depends only on its indices in the directions other
// than c when every coordinate x_c depends on i_c and the node's c-index
// alone, as the linear mapping of a box gives.  Fills
// tab[off[c] + (pair * dim + r) * nn + dof] (pair = i_m * cells[l] + i_l in
// 3D with m = c+1, l = c+2 mod 3, pair = i_{1-c} in 2D) and returns true if
// every element's terms equal their table rows bit for bit and 1/J > 0 at
// every node; false otherwise (the kernels then read the per-node arrays).
bool box_metric_table(const Geometry& g, std::vector<double>& tab, int64_t off[3]);

}  // namespace dgsem
