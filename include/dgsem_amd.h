/*
 * dgsem_amd.h -- C ABI of the MI355X DGSEM right-hand-side library.
 *
 * The library replaces the reference's StructuredSolver::calc_rhs and the
 * SSPRK3 stage kernels (RetrO598/DGSEM src/solver/structured_solver.hpp,
 * src/time_integrator/SSPRK3.hpp) with hand-written HIP kernels for gfx950.
 * All state pointers are DEVICE pointers to double arrays in the library's
 * internal layout u[element][var][dof] (element index (ie*ny + je)*nz + ke,
 * dof = (k*(N+1) + j)*(N+1) + i).  dgsem_pack_state / dgsem_unpack_state
 * convert from / to the reference's LayoutRight u(ielem,jelem[,kelem],dof,var).
 *
 * Every function returns 0 on success and a negative DGSEM_E_* code on
 * failure; dgsem_last_error() returns the message.  No C++ exception crosses
 * the ABI.  All work is enqueued on the handle's stream; calls are
 * asynchronous with respect to the host.  One handle must not be used from
 * two host threads at once.
 */
#ifndef DGSEM_AMD_H
#define DGSEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGSEM_ABI_VERSION 2

/* error codes */
#define DGSEM_OK 0
#define DGSEM_E_INVALID (-1)     /* bad argument or configuration */
#define DGSEM_E_UNSUPPORTED (-2) /* configuration not built */
#define DGSEM_E_HIP (-3)         /* HIP runtime error */
#define DGSEM_E_ALLOC (-4)       /* device allocation failed */

/* equations (src/equations/compressible_euler{2D,3D}.hpp,
 * compressible_navier_stokes{2D,3D}.hpp) */
#define DGSEM_EULER 0
#define DGSEM_NAVIER_STOKES 1

/* volume integral policies (src/space_integral/volume_flux.hpp:686-879) */
#define DGSEM_VOLUME_SPLIT_FORM 0       /* VolumeIntegralSplitForm */
#define DGSEM_VOLUME_SHOCK_CAPTURING 1  /* VolumeIntegralShockCapturingHG */

/* two-point / numerical fluxes (src/base/numerical_flux.hpp) */
#define DGSEM_FLUX_CHANDRASHEKAR 0 /* volume flux: ChandrashekarFlux */
#define DGSEM_FLUX_LLF 0           /* surface / FV flux: LaxFriedrichsFlux */
#define DGSEM_FLUX_HLLC 1          /* surface / FV flux: HLLC with a normal */
/* volume, surface or FV flux: CentralFlux, 0.5 (F(u_L) + F(u_R)).n
 * (numerical_flux.hpp:230-258, which defines the normal form for 2D only);
 * 2D meshes, N <= 5; not in batches or the derivatives */
#define DGSEM_FLUX_CENTRAL 2

/* geometry representation on the device */
#define DGSEM_GEOMETRY_NODAL 0  /* per-node metric terms exactly as the
                                   reference computes them (curl form) */
#define DGSEM_GEOMETRY_AFFINE 1 /* exact per-element metric terms of the
                                   linear mapping; 10 doubles per element */

/* boundary conditions (src/boundary_condition); all but periodic are built
 * for 2D meshes (the reference specialises them for 1D and 2D) */
#define DGSEM_BC_PERIODIC 0  /* PeriodicBC, periodic_boundary.hpp:11-21 */
#define DGSEM_BC_SLIP_WALL 1 /* SlipWallBC, slip_wall_boundary.hpp:10-66 */
#define DGSEM_BC_OUTFLOW 2   /* OutflowBC, outflow_boundary.hpp:61-105 */
#define DGSEM_BC_DIRICHLET 3 /* DirichletBC, dirichlet_boundary.hpp:70-118 */
/* Navier-Stokes adiabatic walls: NSNoSlipAdiabaticWallBC
 * (navier_stokes_wall_boundary.hpp:39-146) and NSSlipAdiabaticWallBC
 * (:148-217).  Under Navier-Stokes the other kinds have no gradient or
 * viscous face values, which stay zero as in the reference. */
#define DGSEM_BC_NS_NOSLIP_WALL 4
#define DGSEM_BC_NS_SLIP_WALL 5
/* MixedDirichletSlipWallBC (dirichlet_boundary.hpp:198-275): per boundary
 * node either a Dirichlet state or the slip-wall reflection.  The node's
 * state and choice come from dgsem_set_boundary_values (required). */
#define DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL 6

/* Solver configuration: the reference's template arguments
 * StructuredSolver<Equations, Basis, VolumeFlux, SurfaceFlux, Mesh,
 * BoundarySet> (structured_solver.hpp:19-58) plus the runtime parameters of
 * the equation, the mesh and set_indicator_parameters (:118-123). */
typedef struct dgsem_config {
  int dim;                 /* 2 or 3 */
  int polydeg;             /* N, 1..7 (LobattoLegendreBasis<T, N>) */
  int cells[3];            /* StructuredMesh n_cells */
  double domain_left[3];   /* StructuredMesh domain_left */
  double domain_right[3];  /* StructuredMesh domain_right */
  int boundary[3];         /* DGSEM_BC_* per direction (both faces) */
  int equation;            /* DGSEM_EULER or DGSEM_NAVIER_STOKES */
  double gamma;            /* default 1.4 */
  double mu;               /* dynamic viscosity (Navier-Stokes) */
  double prandtl;          /* Prandtl number (Navier-Stokes) */
  int volume_integral;     /* DGSEM_VOLUME_* */
  int volume_flux;         /* DGSEM_FLUX_CHANDRASHEKAR or DGSEM_FLUX_CENTRAL */
  int surface_flux;        /* DGSEM_FLUX_LLF, DGSEM_FLUX_HLLC or DGSEM_FLUX_CENTRAL */
  int fv_flux;             /* subcell FV flux for shock capturing (the same kinds) */
  double alpha_max;        /* set_indicator_parameters: default 0.5 */
  double alpha_min;        /* default 0.001 */
  int alpha_smooth;        /* stored, unused (as in the reference) */
  double positivity_floor; /* <= 0 disables; applied after stage updates */
  int geometry;            /* DGSEM_GEOMETRY_* */
  /* Per-face boundary conditions, faces -x, +x, -y, +y, -z, +z (the
   * reference's BoundarySet order, boundary_set.hpp:88-111).  All zero: every
   * face takes boundary[dim].  A direction is periodic on both faces or on
   * neither. */
  int face_boundary[6];
  /* DGSEM_BC_DIRICHLET outer state per face (conservative variables), used
   * unless per-node states are set with dgsem_set_boundary_values. */
  double face_state[6][5];
} dgsem_config;

typedef struct dgsem_handle dgsem_handle;

/* Fills cfg with the reference's defaults for dim and N (periodic box
 * [-1,1]^dim, 1 cell, Euler, gamma 1.4, split form, LLF, nodal geometry). */
void dgsem_config_default(dgsem_config* cfg, int dim, int polydeg);

/* Builds basis, geometry and device buffers; replaces the StructuredSolver
 * constructor (structured_solver.hpp:45-58) with StructuredElementInitializer
 * (data_container.hpp:735-768) and LobattoLegendreBasis::initialize
 * (basis.hpp:69-153).  stream: a hipStream_t, or NULL for the null stream. */
int dgsem_create(const dgsem_config* cfg, void* stream, dgsem_handle** out);
int dgsem_destroy(dgsem_handle* h);
const char* dgsem_last_error(const dgsem_handle* h);
int dgsem_abi_version(void);

/* Number of elements, nodes per element ((N+1)^dim) and conserved variables;
 * StructuredSolver::get_ndofs (structured_solver.hpp:125). */
int dgsem_sizes(const dgsem_handle* h, int64_t* n_elements, int* ndofs, int* nvars);

/* set_indicator_parameters(alpha_max, alpha_min, alpha_smooth),
 * structured_solver.hpp:118-123. */
int dgsem_set_indicator_parameters(dgsem_handle* h, double alpha_max, double alpha_min,
                                   int alpha_smooth);

/* Host copies of the geometry in the reference's layouts
 * (StructuredElementContainer, data_container.hpp:12-60):
 * coordinates [E][nn][dim], contravariant vectors [E][nn][dim][dim],
 * inverse Jacobian [E][nn]. */
int dgsem_node_coordinates(const dgsem_handle* h, double* host_out);

/* Curvilinear meshes (the reference's StructuredElementInitializer with a
 * user Mapping, data_container.hpp:729-768).  dgsem_reference_coordinates
 * writes the reference-cube position of every node ([E][nn][dim], host),
 * the argument calc_node_coordinates hands to Mapping::eval
 * (data_container.hpp:193-221, :371-414).  dgsem_set_node_coordinates takes
 * the mapped node coordinates ([E][nn][dim], device) and recomputes the
 * Jacobian, the contravariant vectors (2D adjugate, 3D curl form), 1/J and
 * the subcell normals on the GPU, bit-identical to the host construction
 * (data_container.hpp:223-693, global_def.hpp:175-439); the host getters
 * follow.  Nodal geometry mode only.  Synchronises the handle's stream. */
int dgsem_reference_coordinates(const dgsem_handle* h, double* host_out);
int dgsem_set_node_coordinates(dgsem_handle* h, const double* device_coords);
int dgsem_contravariant_vectors(const dgsem_handle* h, double* host_out);
int dgsem_inverse_jacobian(const dgsem_handle* h, double* host_out);

/* Layout conversion on the device between the reference's LayoutRight
 * u(e, dof, var) and the internal u[e][var][dof]. */
int dgsem_pack_state(dgsem_handle* h, const double* ref_layout, double* internal);
int dgsem_unpack_state(dgsem_handle* h, const double* internal, double* ref_layout);

/* The reference's solution views are Kokkos::View<T****[*], Device>
 * (src/base/global_def.hpp:9-37), whose layout is the default execution
 * space's: LayoutRight on the OpenMP / Serial backends, LayoutLeft on the
 * HIP / CUDA backends.  These two take the layout explicitly; with
 * DGSEM_LAYOUT_RIGHT they are dgsem_pack_state / dgsem_unpack_state.  In
 * LayoutLeft, u(ie, je[, ke], dof, var) has ie fastest and var slowest. */
#define DGSEM_LAYOUT_RIGHT 0
#define DGSEM_LAYOUT_LEFT 1
int dgsem_pack_state_layout(dgsem_handle* h, int layout, const double* ref_layout,
                            double* internal);
int dgsem_unpack_state_layout(dgsem_handle* h, int layout, const double* internal,
                              double* ref_layout);

/* Per-node outer states of a DGSEM_BC_DIRICHLET face, for Dirichlet data
 * that vary along the face or in time (DirichletBC with a function of
 * (x, t), dirichlet_boundary.hpp:80-90): a device array of
 * [(elements along the face) * (N+1)][nvars] doubles, node (element index
 * along the face) * (N+1) + (node index along the face), read by every later
 * RHS or stage (the caller may rewrite it between stages).  NULL reverts to
 * the face's constant state.  For a DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL face
 * each node holds nvars + 1 doubles: the state, then a flag that is nonzero
 * where the node takes the state (the reference's use_dirichlet(x, t)) and
 * zero where it is a slip wall. */
int dgsem_set_boundary_values(dgsem_handle* h, int face, const double* device_values);

/* du = L(u): StructuredSolver::calc_rhs(sol, t), structured_solver.hpp:463-480
 * (volume integral, interface flux, boundary conditions, parabolic terms,
 * surface integral, Jacobian).  u and du are device arrays of
 * n_elements*nvars*ndofs doubles, internal layout.  t is accepted for
 * interface parity; periodic boundaries do not use it. */
int dgsem_rhs(dgsem_handle* h, const double* u, double* du, double t);

/* Blending factor per element from the last shock-capturing RHS or stage
 * (VolumeIntegralShockCapturingHG::calc_alpha, volume_flux.hpp:785-788);
 * device array of n_elements doubles. */
int dgsem_get_alpha(const dgsem_handle* h, double* alpha_out);

/* The blending factor each kind of evaluation recorded last, so that a
 * forward SSPRK3 step leaves the alpha of all three of its stages (the
 * reference recomputes alpha in every calc_rhs, structured_solver.hpp:
 * 175-198): stage 1, 2 or 3 -- the last dgsem_ssprk3_stage of that stage
 * number; stage 0 -- the last dgsem_rhs.  Device array of n_elements. */
int dgsem_get_stage_alpha(const dgsem_handle* h, int stage, double* alpha_out);

/* Freezes the blending factor (no counterpart in the reference, whose alpha
 * always comes from the indicator): while `alpha` (device array of
 * n_elements doubles, read at every launch) is set, dgsem_rhs,
 * dgsem_ssprk3_stage / _steps blend with it instead of the indicator's
 * value, and the tangent and adjoint entry points below take it as the
 * frozen alpha.  NULL returns to the indicator.  Shock capturing only. */
int dgsem_freeze_alpha(dgsem_handle* h, const double* alpha);

/* One fused SSPRK3 stage: RHS of u_k followed by the stage combination of
 * parallel_ma2 / parallel_ma3 (SSPRK3.hpp:12-143) as SSPRK3::step uses them
 * (SSPRK3.hpp:159-181):
 *   stage 1: u_out = u_k + dt L(u_k)                      (u_base unused)
 *   stage 2: u_out = 3/4 u_base + 1/4 u_k + (dt/4) L(u_k)
 *   stage 3: u_out = 1/3 u_base + 2/3 u_k + (2 dt/3) L(u_k)
 * u_out may alias u_base (stage 3 writes u^{n+1} over u^n) but never u_k. */
int dgsem_ssprk3_stage(dgsem_handle* h, int stage, double dt, const double* u_base,
                       const double* u_k, double* u_out);

/* nsteps SSPRK3 steps of fixed dt on u (SSPRK3::step, SSPRK3.hpp:159-181,
 * inside TimeIntegrator::solve's loop, integrator_base.hpp:21-40).  tmp1 and
 * tmp2 are caller-provided device arrays of the state's size. */
int dgsem_ssprk3_steps(dgsem_handle* h, double* u, double dt, int nsteps, double* tmp1,
                       double* tmp2);

/* Waits for all work on the handle's stream. */
int dgsem_synchronize(dgsem_handle* h);

/* ---------------------------------------------------------------------
 * Tangent and discrete adjoint.  The reference has no derivatives; these
 * differentiate the right-hand side this library computes exactly, up to
 * rounding, with J = dL/du at the state u: split form or shock capturing
 * (the blend (1 - alpha) split form + alpha FV with its subcell flux,
 * volume_flux.hpp:449-630, :821-869), LLF or HLLC surface and subcell
 * fluxes, Euler or Navier-Stokes (BR1); periodic meshes, and bounded 2D
 * meshes (the outer states' and the wall gradients' dependence on the inner
 * ones included).  The
 * blending factor is not differentiated: it is held at the value the
 * forward evaluation at the same state computes (the indicator at u; at each
 * stage input for the step functions), or at the array set with
 * dgsem_freeze_alpha.  With the positivity floor the step functions pass
 * through each stage's clamp (its derivative at the stage's unclamped
 * result).
 * --------------------------------------------------------------------- */
/* out = J(u) du */
int dgsem_rhs_jvp(dgsem_handle* h, const double* u, const double* du, double* out);
/* out = J(u)^T w */
int dgsem_rhs_vjp(dgsem_handle* h, const double* u, const double* w, double* out);
/* Tangent of one SSPRK3 step (the map u0 -> u3 of dgsem_ssprk3_stage 1..3)
 * given its stage states u0, u1, u2 (from dgsem_ssprk3_stage): du_out =
 * (du3/du0) du0.  tmp1, tmp2: scratch arrays of the state's size. */
int dgsem_ssprk3_step_jvp(dgsem_handle* h, double dt, const double* u0, const double* u1,
                          const double* u2, const double* du0, double* du_out, double* tmp1,
                          double* tmp2);
/* Adjoint of one SSPRK3 step: ubar_out = (du3/du0)^T ubar.  ubar_out may
 * alias ubar. */
int dgsem_ssprk3_step_vjp(dgsem_handle* h, double dt, const double* u0, const double* u1,
                          const double* u2, const double* ubar, double* ubar_out, double* tmp1,
                          double* tmp2);
/* Kinetic energy sum over nodes of w_i w_j (w_k) / |1/J| |rho v|^2 / (2 rho),
 * the quadrature of VolumeAverageEuler with the volume weights of
 * VolumeWeightedAnalyzerFunctor (volume_average_euler.hpp:30-53,
 * volume_weighted_analyzer_functor.hpp:108-134), reduced in a fixed order
 * into the device double *result. */
int dgsem_kinetic_energy(dgsem_handle* h, const double* u, double* result);
/* grad += scale * d(kinetic energy)/du */
int dgsem_kinetic_energy_grad(dgsem_handle* h, const double* u, double scale, double* grad);

/* Analysis of a state in one pass (the reference's analyzers:
 * VolumeAverageEuler, volume_average_euler.hpp:30-53, with the volume weights
 * w_i w_j (w_k) / |1/J| of VolumeWeightedAnalyzerFunctor,
 * volume_weighted_analyzer_functor.hpp:108-134; DivergenceChecker,
 * divergence_checker.hpp:6-27).  Writes 5 doubles to the device array out:
 *   [0] total volume, [1] kinetic-energy integral, [2] entropy integral
 *   (-rho (ln p - gamma ln rho) / (gamma - 1)), [3] number of non-finite
 *   nodes (skipped in the sums), [4] max over nodes of max_d |v_d| + c.
 * Reduced in a fixed order. */
int dgsem_analyze(dgsem_handle* h, const double* u, double* out);

/* dgsem_analyze without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The same kernels and the same results, bit
 * for bit; the per-block partial sums go to a static device array of
 * 1024 x 5 doubles (40 KiB) that each GPU holds until the process ends,
 * instead of a buffer dgsem_analyze allocates on its first call.  Every size
 * fits it (the reduction is grid-stride with at most 1024 blocks); the
 * handle's device must be the current device.  Calls on one GPU share the
 * array: calls that overlap on one GPU (several streams or host threads)
 * race, so serialise them. */
int dgsem_analyze_noalloc(dgsem_handle* h, const double* u, double* out);

/* Writes the device state u (internal layout) to a VTK XML unstructured-grid
 * file of Lagrange cells, as VTUOutputObserver::write_vtu does
 * (vtu_output_observer.hpp:62-214, vtu_output_observer_impl.cpp:15-117):
 * cells x-fastest, points in VTK's Lagrange order (file version 2.2), point
 * data rho, momenta, rhoE and pressure (gamma 1.4 as in the reference),
 * appended raw binary.  Synchronises the handle's stream. */
int dgsem_write_vtu(dgsem_handle* h, const double* u, const char* path);

/* ---------------------------------------------------------------------
 * Batched ensembles.  The reference has no ensemble API: its drivers run
 * one StructuredSolver per process, one case after another.  A batch
 * advances many independent 2D Euler cases (different element grids and
 * polynomial degrees; one volume integral, flux pair and geometry mode for
 * all) with one persistent kernel launch per SSPRK3 stage, the three stages
 * of a step recorded once into a HIP graph.  Each case gives bit-identical
 * results to its own single-case handle (dgsem_ssprk3_steps).
 *
 * A batch's state lives in caller-provided device "arenas" of `total`
 * doubles: case i occupies [offset_i, offset_i + count_i) in the internal
 * layout of its own handle.
 * --------------------------------------------------------------------- */
typedef struct dgsem_batch dgsem_batch;

/* Creates the per-case handles (dgsem_create of each cases[i], all on
 * `stream`).  On failure *out is still set so dgsem_batch_last_error can be
 * read; release it with dgsem_batch_destroy. */
int dgsem_batch_create(const dgsem_config* cases, int ncases, void* stream, dgsem_batch** out);
int dgsem_batch_destroy(dgsem_batch* b);
const char* dgsem_batch_last_error(const dgsem_batch* b);
int dgsem_batch_sizes(const dgsem_batch* b, int* ncases, int64_t* total);
/* Arena slice of case i and its single-case handle (owned by the batch). */
int dgsem_batch_case(dgsem_batch* b, int i, int64_t* offset, int64_t* count, dgsem_handle** h);
/* Per-case time step (host array of ncases doubles). */
int dgsem_batch_set_dt(dgsem_batch* b, const double* dt);
/* du = L(u) for every case (one launch). */
int dgsem_batch_rhs(dgsem_batch* b, const double* u, double* du);
/* nsteps SSPRK3 steps of every case with its own dt; u, tmp1, tmp2 are
 * distinct arenas.  Replays the captured graph of the three stage launches;
 * the graph is recaptured when the arenas, dt or tuning change. */
int dgsem_batch_ssprk3_steps(dgsem_batch* b, double* u, int nsteps, double* tmp1, double* tmp2);
/* Tuning: workgroup size (128 or 256), elements per tile for each degree
 * (tile_elems[N], N = 0..7; 0 = as many as one workgroup holds) and the
 * number of persistent workgroups (0 = all that fit the device at once). */
int dgsem_batch_get_tuning(const dgsem_batch* b, int* block, int* tile_elems, int* grid);
int dgsem_batch_set_tuning(dgsem_batch* b, int block, const int* tile_elems, int grid);
/* Autotuner: times the batch RHS of u (written to `scratch`, an arena) over
 * `reps` launches for each candidate workgroup size, then each degree's tile
 * size, then the grid size, and keeps the fastest. */
int dgsem_batch_autotune(dgsem_batch* b, const double* u, double* scratch, int reps);

#ifdef __cplusplus
}
#endif
#endif
