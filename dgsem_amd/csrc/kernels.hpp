// Kernel parameter block and launch entry points shared by the ABI layer
// and the kernel translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dgsem {

constexpr int kMaxN1 = 8;  // max nodes per direction (N <= 7)

#ifndef DGSEM_NODE_KERNEL_MIN_N
#define DGSEM_NODE_KERNEL_MIN_N 7
#endif

// Split-form elements of degree N >= 7 on periodic meshes run the
// node-parallel kernel, in 2D and 3D.  With Navier-Stokes terms that kernel
// takes the viscous fluxes Ja^d . (-F_v) per node [E][dim][var][dof]
// (instead of their divergence [E][var][dof]) and adds their terms itself.
__host__ __device__ constexpr bool node_kernel_for(int N, bool shock_capturing, bool boundaries) {
  return !shock_capturing && !boundaries && N >= DGSEM_NODE_KERNEL_MIN_N;
}

enum StageMode : int {
  kModeRhs = 0,     // out = L(u)
  kModeStage2 = 1,  // out = c1 u + c2 L(u)                (parallel_ma2)
  kModeStage3 = 2,  // out = c1 u_base + c2 u + c3 L(u)    (parallel_ma3)
};

// Boundary condition kinds per face (values of DGSEM_BC_* in the C header).
enum BoundaryKind : int {
  kBcPeriodic = 0,
  kBcSlipWall = 1,
  kBcOutflow = 2,
  kBcDirichlet = 3,
  kBcNsNoSlipWall = 4,  // Navier-Stokes adiabatic walls
  kBcNsSlipWall = 5,
  kBcMixedDirichletSlipWall = 6,  // per-node choice, states [node][NV + 1]
};

struct KParams {
  const double* u;       // stage input, [E][NV][NN]
  const double* u_base;  // base state for kModeStage3
  double* out;           // L(u) or the stage result
  double* alpha;         // [E] blending factor (shock capturing), written
  const double* alpha_in;  // [E] frozen blending factor to use instead of the
                           // indicator's (shock capturing), or null
  // Box meshes (the linear mapping of a Cartesian box, nodal geometry): the
  // metric terms Ja^r_c of a node depend on its element only through the
  // element's indices in the directions other than c, so each distinct row
  // is stored once: ja_box[ja_box_off[c] + (pair * DIM + r) * NN + dof] with
  // pair = i_m * cells[l] + i_l (3D; m = c+1, l = c+2 mod 3) or i_m (2D;
  // m = 1-c).  Null unless the mesh is such a box, whose 1/J is then
  // positive at every node.
  const double* ja_box;
  int64_t ja_box_off[3];
  const double* Ja;      // nodal: [E][dim][dim][NN]; affine: [E][dim][dim]
  const double* invJ;    // nodal: [E][NN]; affine: [E]
  const double* sn;      // nodal subcell normals [E][dim][FN][N][dim]
  const int* nbr;        // [E][2*dim] (lower, upper) neighbour per direction
  const double* Vinv;    // [n][n] inverse Legendre Vandermonde
  const double* grad;    // Navier-Stokes: gradient [E][dim][var][dof]
  double* grad_out;      // written by the gradient kernel
  const double* visc;    // Navier-Stokes: viscous part of the RHS before -1/J
  double* visc_out;      // written by the viscous kernel
  int64_t E;
  int mode;
  int surface_flux, fv_flux;  // FluxKind
  int volume_flux;            // VolumeFluxKind (read by the run-time flux kernels only)
  double c1, c2, c3;
  double gamma, inv_gm1;
  double alpha_max, alpha_min, atol, threshold, sfac;
  double floor_value;
  double lhat0, lhatN;
  double mu, kappa;  // viscosity and gamma / ((gamma - 1) Pr)
  double Ds[kMaxN1][kMaxN1];
  double Dhat[kMaxN1][kMaxN1];
  double inv_w[kMaxN1];
  // boundary conditions (2D): element grid, kind per face (-x, +x, -y, +y;
  // 0 = periodic, then nbr holds -1 at that face), constant Dirichlet states
  // and optional per-node Dirichlet states
  int cells[3];
  int bc[6];
  double bc_state[6][5];
  const double* bc_values[6];
};

// Per-(dim, N) launchers, explicitly specialised in rhs_inst.hip.
// boundaries: some face is not periodic (2D only)
template <int DIM, int N>
hipError_t launch_rhs_dn(bool shock_capturing, bool nodal, bool boundaries, const KParams& p,
                         hipStream_t stream);
template <int DIM, int N>
hipError_t launch_parabolic_dn(bool shock_capturing, bool nodal, bool boundaries, const KParams& p,
                               hipStream_t stream);

// Gradient and viscous kernels of the Navier-Stokes terms: fill p.grad_out
// (with p.grad == p.grad_out) and p.visc_out from p.u; p.visc_out takes the
// per-node viscous fluxes when node_kernel_for(N, shock_capturing,
// boundaries), else the viscous part of the RHS.
hipError_t launch_parabolic(int dim, int N, bool shock_capturing, bool nodal, bool boundaries,
                            const KParams& p, hipStream_t stream);

// Launches the fused RHS / stage kernel for the configuration; returns
// hipErrorInvalidValue when (dim, N, shock capturing, nodal) is not built.
hipError_t launch_rhs(int dim, int N, bool shock_capturing, bool nodal, bool boundaries,
                      const KParams& p, hipStream_t stream);

// Layout conversion between the reference layout [e][dof][v] and the
// internal layout [e][v][dof]; to_internal selects the direction.
hipError_t launch_pack(const double* src, double* dst, int64_t E, int nn, int nv,
                       bool to_internal, hipStream_t stream);

// Same between a column-major (LayoutLeft) reference view u(ie, je[, ke],
// dof, var) over cells[0] x cells[1] x cells[2] elements and the internal
// layout.
hipError_t launch_pack_left(const double* src, double* dst, const int cells[3], int nn, int nv,
                            bool to_internal, hipStream_t stream);

// Tangent (JVP) and adjoint (VJP) of the right-hand side at the state k.u
// (adjoint_impl.hpp).  Op is J = dL/du (tangent) or J^T (adjoint) and the
// kernels write
//     out = a1 x1 + a2 x2 + b v + c Op(v)
// (x1, x2 may be null), which covers the tangent and adjoint of every SSPRK3
// stage.  Navier-Stokes needs k.grad = the forward gradient of k.u and the
// scratch arrays dgrad [E][dim][NV][NN], dvisc and qbar [E][NV][NN].
struct AdParams {
  KParams k;
  const double* v;  // tangent direction or cotangent
  const double* x1;
  const double* x2;
  double a1, a2, b, c;
  double* out;
  double* dgrad;
  double* dvisc;
  double* qbar;
};

template <int DIM, int N>
hipError_t launch_jvp_dn(bool nodal, const AdParams& a, hipStream_t stream);
template <int DIM, int N>
hipError_t launch_vjp_dn(bool nodal, const AdParams& a, hipStream_t stream);
// Launch the tangent or adjoint kernels for (dim, N); the Navier-Stokes
// forward gradient must already be in a.k.grad.
hipError_t launch_jvp(int dim, int N, bool nodal, bool navier_stokes, const AdParams& a,
                      hipStream_t stream);
hipError_t launch_vjp(int dim, int N, bool nodal, bool navier_stokes, const AdParams& a,
                      hipStream_t stream);
// Forward BR1 gradient alone (a.k.grad_out from a.k.u).
template <int DIM, int N>
hipError_t launch_gradient_dn(bool nodal, const KParams& p, hipStream_t stream);
hipError_t launch_gradient(int dim, int N, bool nodal, const KParams& p, hipStream_t stream);

// Kinetic energy sum_e sum_i w_i / |1/J_i| |m_i|^2 / (2 rho_i) into *result
// (device), and its gradient: grad += scale * dKE/du.  partial: scratch of
// ke_partials(E) doubles.
int64_t ke_partials(int64_t E);
hipError_t launch_kinetic_energy(int dim, int nn, const double* w_nodes, bool nodal,
                                 const KParams& p, double* partial, double* result,
                                 hipStream_t stream);
hipError_t launch_kinetic_energy_grad(int dim, int nn, const double* w_nodes, bool nodal,
                                      const KParams& p, double scale, double* grad,
                                      hipStream_t stream);

// Batched ensembles of 2D cases (batch_impl.hpp).  A tile is up to
// batch_slots(block, N) consecutive elements of one case.
struct BatchTile {
  int32_t case_id;
  int16_t polydeg;
  int16_t nvalid;  // elements in this tile
  int64_t e0;      // first element of the tile within its case
};

// The tile list is sorted by degree, highest first: the tiles of degree N
// are [begin[N], begin[N - 1]), and begin[0] is the number of tiles.
struct BatchRanges {
  int begin[8];
};

struct BatchLaunch {
  const BatchTile* tiles;
  int ntiles;
  BatchRanges ranges;
  const KParams* params;  // one per case
  int grid;               // persistent workgroups
  size_t lds_bytes;       // dynamic LDS: the largest tile footprint in the batch
  bool shock_capturing, nodal;
  int surface_flux, fv_flux;
  int* probe;  // non-null: store the resident workgroups per CU instead of launching
};

// Workgroup sizes the batch kernel is built for.
constexpr int kBatchBlocks[] = {128, 256};

template <int BLOCK>
hipError_t launch_batch_bs(const BatchLaunch& b, hipStream_t stream);
hipError_t launch_batch(int block, const BatchLaunch& b, hipStream_t stream);
// Element slots of one tile and its LDS bytes, for a 2D tile of degree N;
// 0 when (block, N) is not built.
int batch_slots(int block, int polydeg);
size_t batch_lds_bytes(int block, int polydeg);

}  // namespace dgsem
