/*
 * dgsem_oracle.h -- CPU restatement of the reference DGSEM right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dgsem_amd/, include/)
 * links, loads or calls this code.  It is used by tests/, by the smoke check
 * and by bench.py's cpu_baseline leg, always as the checker, never as the
 * thing measured for the GPU result.
 *
 * The algorithm follows RetrO598/DGSEM file by file; every function in
 * dgsem_oracle.c cites the reference file:line it restates.  The code is
 * plain C99 compiled with -O2 -ffp-contract=off so that every floating point
 * operation happens in the same order and with the same rounding as the
 * reference's own host build.  Arrays use the reference's LayoutRight
 * layouts: u(ielem, jelem[, kelem], dof, var) with var fastest.
 */
#ifndef DGSEM_ORACLE_H
#define DGSEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EULER = 0, ORC_NAVIER_STOKES = 1 };
enum { ORC_SPLIT_FORM = 0, ORC_SHOCK_CAPTURING = 1 };
enum { ORC_FLUX_LLF = 0, ORC_FLUX_HLLC = 1, ORC_FLUX_CENTRAL = 2 };
/* volume fluxes of the split form */
enum { ORC_VOLUME_CHANDRASHEKAR = 0, ORC_VOLUME_CENTRAL = 2 };

typedef struct orc_case orc_case;

/* Basis tables for polynomial degree N (N+1 nodes), reference
 * src/polynomial_basis/basis.hpp:69-153.  Matrices are row-major (N+1)^2. */
void orc_basis(int N, double* nodes, double* weights, double* inv_weights,
               double* D, double* Dsplit, double* Dhat, double* Vinv,
               double* lhat_left, double* lhat_right);

/* Mesh + geometry for a Cartesian box mapped linearly, reference
 * StructuredSolver ctor (structured_solver.hpp:45-58) and
 * StructuredElementInitializer::init_elements (data_container.hpp:735-768). */
orc_case* orc_create(int dim, int N, int nx, int ny, int nz,
                     const double* domain_left, const double* domain_right,
                     const int* periodic);
void orc_destroy(orc_case* c);

/* Physics and policy selection (the reference's template arguments). */
void orc_set_equation(orc_case* c, int eq, double gamma, double mu,
                      double prandtl);
void orc_set_volume(orc_case* c, int volume, int fv_flux, double alpha_max,
                    double alpha_min);
void orc_set_surface_flux(orc_case* c, int surface_flux);
/* the split form's two-point flux: ORC_VOLUME_* (default Chandrashekar) */
void orc_set_volume_flux(orc_case* c, int volume_flux);
void orc_set_positivity_floor(orc_case* c, double floor_value);
void orc_set_threads(orc_case* c, int nthreads);

/* Array sizes (in doubles / int64) */
int64_t orc_num_elements(const orc_case* c);
int orc_nvars(const orc_case* c);
int orc_ndofs(const orc_case* c);

/* Geometry accessors (copy out, reference layouts). */
void orc_get_coordinates(const orc_case* c, double* out);   /* [E][nn][dim] */
void orc_get_jacobian(const orc_case* c, double* out);      /* [E][nn][dim][dim] */
void orc_get_contravariant(const orc_case* c, double* out); /* [E][nn][dim][dim] */
void orc_get_inverse_jacobian(const orc_case* c, double* out); /* [E][nn] */
void orc_get_left_neighbors(const orc_case* c, int64_t* out);  /* [E][dim][dim] */
/* Overwrite the metric terms (for metric-sensitivity studies and for running
 * the restatement on exactly the metrics another implementation uses). */
/* Replaces the node coordinates [E][nn][dim] (any mapping) and recomputes
 * the Jacobian, contravariant vectors, 1/J and subcell normals. */
void orc_set_coordinates(orc_case* c, const double* coords);
void orc_set_metrics(orc_case* c, const double* contravariant,
                     const double* inverse_jacobian);

/* Initial conditions evaluated at the node coordinates (reference examples). */
void orc_init_khi(const orc_case* c, double* u);
void orc_init_tgv(const orc_case* c, double* u);
void orc_init_sedov(const orc_case* c, double* u, double E0, double sigma,
                    double p_background);

/* StructuredSolver::calc_rhs (structured_solver.hpp:463-480).  alpha (may be
 * NULL) receives the blending factor per element for shock capturing. */
void orc_rhs(orc_case* c, const double* u, double* du, double* alpha);
/* Individual pieces, for unit tests (du is accumulated, not cleared). */
void orc_volume_integral(orc_case* c, const double* u, double* du,
                         double* alpha);
void orc_calc_alpha(orc_case* c, const double* u, double* alpha);
/* Fills the face-flux storage (E, 2*dim*(N+1)^(dim-1), nvars). */
void orc_interface_flux(orc_case* c, const double* u, double* surface_flux);
void orc_surface_integral(orc_case* c, const double* surface_flux, double* du);
void orc_apply_jacobian(orc_case* c, double* du);
/* BR1 gradient (E, nn, nvars, dim) and the full parabolic contribution. */
void orc_gradients(orc_case* c, const double* u, double* gradient);

/* Boundary condition of one face of a 2D mesh (face = 2*dim + upper); the
 * mesh must be created non-periodic in that direction.  state: constant
 * Dirichlet state (conservative); values: per-node Dirichlet states
 * [face nodes][nvars] (kept by pointer), or NULL. */
#define ORC_BC_PERIODIC 0
#define ORC_BC_SLIP_WALL 1
#define ORC_BC_OUTFLOW 2
#define ORC_BC_DIRICHLET 3
/* Navier-Stokes adiabatic walls (navier_stokes_wall_boundary.hpp:39-219):
 * the inviscid flux of a slip wall, plus gradient and viscous face values */
#define ORC_BC_NS_NOSLIP_WALL 4
#define ORC_BC_NS_SLIP_WALL 5
/* MixedDirichletSlipWallBC (dirichlet_boundary.hpp:198-275): values holds
 * [face nodes][nvars + 1], the Dirichlet state and a flag, nonzero where the
 * node takes the Dirichlet state and zero where it is a slip wall (the
 * reference's use_dirichlet(coord, t)) */
#define ORC_BC_MIXED_DIRICHLET_SLIP_WALL 6
void orc_set_boundary(orc_case* c, int face, int kind, const double* state,
                      const double* values);

/* test hook: blend with a fixed alpha [E] (kept by reference) instead of the
 * indicator's; NULL restores the indicator */
void orc_set_alpha_override(orc_case* c, const double* alpha);
/* the positivity floor of a stage update, applied to u in place */
void orc_apply_floor(const orc_case* c, double* u);
/* SSPRK3::step (SSPRK3.hpp:159-181), nsteps times with fixed dt. */
void orc_ssprk3(orc_case* c, double* u, double dt, int nsteps);

/* Pointwise kernels exposed for unit tests. */
double orc_ln_mean(double x, double y);
void orc_flux_chandrashekar(int dim, double gamma, const double* ul,
                            const double* ur, const double* normal,
                            double* f);
void orc_flux_llf(int dim, double gamma, const double* ul, const double* ur,
                  const double* normal, double* f);
void orc_flux_hllc(int dim, double gamma, const double* ul, const double* ur,
                   const double* normal, double* f);
void orc_flux_central(int dim, double gamma, const double* ul, const double* ur,
                      const double* normal, double* f);
void orc_physical_flux(int dim, double gamma, const double* u,
                       const double* normal, double* f);

/* gradient_variables and viscous_flux (compressible_navier_stokes{2D,3D}.hpp)
 * with the case's gamma, mu and kappa; grad is [dim][nvars] */
void orc_gradient_variables(const orc_case* c, const double* u, double* q);
void orc_viscous_flux(const orc_case* c, const double* q, const double* grad, int dim, double* f);

/* Volume-weighted kinetic energy sum (volume_average_euler.hpp:30-53 with the
 * weights of volume_weighted_analyzer_functor.hpp:108-134). */
double orc_kinetic_energy(const orc_case* c, const double* u);
/* [volume, kinetic energy, entropy, non-finite nodes, max |v_d| + c] */
void orc_analyze(const orc_case* c, const double* u, double* out);

#ifdef __cplusplus
}
#endif
#endif
