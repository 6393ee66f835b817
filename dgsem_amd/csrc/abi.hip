// C ABI of the MI355X DGSEM right-hand side (include/dgsem_amd.h).
//
// The handle owns the device copies of the geometry, the neighbour table,
// the basis tables and the per-element blending factor; state arrays belong
// to the caller.  Everything is enqueued on the handle's stream, with no
// allocation, host synchronisation or host<->device traffic inside dgsem_rhs
// and dgsem_ssprk3_stage, so both can be captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "../../include/dgsem_amd.h"
#include "handle.hpp"
#include "host_setup.hpp"
#include "kernels.hpp"

namespace {

int fail(dgsem_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

int hip_fail(dgsem_handle* h, hipError_t e, const char* where) {
  return fail(h, DGSEM_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(h, expr)                                \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return hip_fail(h, _e, #expr); \
  } while (0)

template <class T>
int upload(dgsem_handle* h, T** dst, const std::vector<T>& src) {
  const size_t bytes = src.size() * sizeof(T);
  hipError_t e = hipMalloc((void**)dst, bytes ? bytes : sizeof(T));
  if (e != hipSuccess) return fail(h, DGSEM_E_ALLOC, "hipMalloc failed");
  if (bytes) {
    e = hipMemcpy(*dst, src.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(h, e, "hipMemcpy");
  }
  return DGSEM_OK;
}

// A zero-filled device array of `count` doubles.
int alloc_zero(dgsem_handle* h, double** dst, size_t count) {
  const size_t bytes = (count ? count : 1) * sizeof(double);
  if (hipMalloc((void**)dst, bytes) != hipSuccess) return fail(h, DGSEM_E_ALLOC, "hipMalloc failed");
  const hipError_t e = hipMemset(*dst, 0, bytes);
  return e == hipSuccess ? DGSEM_OK : hip_fail(h, e, "hipMemset");
}

void release(dgsem_handle* h) {
  hipFree(h->d_Ja);
  hipFree(h->d_invJ);
  hipFree(h->d_sn);
  hipFree(h->d_nbr);
  hipFree(h->d_Vinv);
  hipFree(h->d_ja_box);
  hipFree(h->d_alpha);
  hipFree(h->d_grad);
  hipFree(h->d_visc);
  hipFree(h->d_ad_dgrad);
  hipFree(h->d_ad_dvisc);
  hipFree(h->d_ad_qbar);
  hipFree(h->d_wq);
  hipFree(h->d_ke_partial);
  hipFree(h->d_an_partial);
  hipFree(h->d_ad_alpha);
  hipFree(h->d_ad_out);
  hipFree(h->d_fl_r);
  hipFree(h->d_fl_w);
  h->d_ad_dgrad = h->d_ad_dvisc = h->d_ad_qbar = h->d_wq = h->d_ke_partial = nullptr;
  h->d_an_partial = h->d_ad_alpha = h->d_ad_out = h->d_fl_r = h->d_fl_w = nullptr;
  h->d_Ja = h->d_invJ = h->d_sn = h->d_Vinv = h->d_alpha = h->d_grad = h->d_visc = nullptr;
  h->d_nbr = nullptr;
  h->d_ja_box = nullptr;
}

// Geometry in the kernels' layouts.  Nodal: Ja[E][dir][comp][nn] and
// invJ[E][nn] copied from the reference-layout arrays.  Affine: the exact
// metric terms of the linear map of each element.
void device_geometry(const dgsem_handle* h, std::vector<double>& Ja, std::vector<double>& invJ) {
  const auto& g = h->geo;
  const int d = g.dim, nn = g.nn;
  if (h->cfg.geometry == DGSEM_GEOMETRY_NODAL) {
    Ja.assign(g.E * d * d * nn, 0.0);
    for (int64_t e = 0; e < g.E; ++e)
      for (int dof = 0; dof < nn; ++dof)
        for (int r = 0; r < d; ++r)
          for (int c = 0; c < d; ++c)
            Ja[((e * d + r) * d + c) * nn + dof] = g.contravariant[((e * nn + dof) * d + r) * d + c];
    invJ = g.inverse_jacobian;
  } else {
    // x = L + (xi_cell + 1)/2 * h: dx/dxi = h/2.  Ja^r_c = delta_rc * prod_{k != r} h_k/2,
    // 1/J = prod_k 2/h_k.
    double half[3] = {1.0, 1.0, 1.0};
    for (int k = 0; k < d; ++k)
      half[k] = 0.5 * (h->cfg.domain_right[k] - h->cfg.domain_left[k]) / (double)h->cfg.cells[k];
    double ja[3][3] = {{0}};
    double det = 1.0;
    for (int k = 0; k < d; ++k) det *= half[k];
    for (int r = 0; r < d; ++r) {
      double prod = 1.0;
      for (int k = 0; k < d; ++k)
        if (k != r) prod *= half[k];
      ja[r][r] = prod;
    }
    Ja.assign(g.E * d * d, 0.0);
    invJ.assign(g.E, 1.0 / det);
    for (int64_t e = 0; e < g.E; ++e)
      for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) Ja[(e * d + r) * d + c] = ja[r][c];
  }
}

// (lower, upper) neighbour per direction: periodic wrap, or -1 at a
// boundary face of a non-periodic direction.
std::vector<int> neighbour_table(const dgsem::Geometry& g, const int* periodic) {
  const int d = g.dim;
  std::vector<int> nb(g.E * 2 * d);
  for (int ie = 0; ie < g.cells[0]; ++ie)
    for (int je = 0; je < g.cells[1]; ++je)
      for (int ke = 0; ke < g.cells[2]; ++ke) {
        const int64_t e = ((int64_t)ie * g.cells[1] + je) * g.cells[2] + ke;
        const int id[3] = {ie, je, ke};
        for (int r = 0; r < d; ++r)
          for (int side = 0; side < 2; ++side) {
            int o[3] = {id[0], id[1], id[2]};
            o[r] += side ? 1 : -1;
            if (!periodic[r] && (o[r] < 0 || o[r] >= g.cells[r])) {
              nb[e * 2 * d + 2 * r + side] = -1;
              continue;
            }
            o[r] = (o[r] + g.cells[r]) % g.cells[r];
            nb[e * 2 * d + 2 * r + side] = (int)(((int64_t)o[0] * g.cells[1] + o[1]) * g.cells[2] + o[2]);
          }
      }
  return nb;
}

void fill_params(dgsem_handle* h) {
  dgsem::KParams& p = h->base;
  const auto& c = h->cfg;
  const auto& b = h->basis;
  std::memset(&p, 0, sizeof(p));
  p.Ja = h->d_Ja;
  p.invJ = h->d_invJ;
  p.sn = h->d_sn;
  p.nbr = h->d_nbr;
  p.Vinv = h->d_Vinv;
  p.alpha = h->d_alpha;
  p.alpha_in = nullptr;
  p.ja_box = h->d_ja_box;
  for (int c = 0; c < 3; ++c) p.ja_box_off[c] = h->ja_box_off[c];
  p.E = h->E;
  p.surface_flux = c.surface_flux;
  p.fv_flux = c.fv_flux;
  p.volume_flux = c.volume_flux;
  p.gamma = c.gamma;
  p.inv_gm1 = 1.0 / (c.gamma - 1.0);
  p.alpha_max = c.alpha_max;
  p.alpha_min = c.alpha_min;
  // atol = max(100 eps, eps^0.75), volume_flux.hpp:772-773
  const double eps = std::numeric_limits<double>::epsilon();
  p.atol = std::max(100.0 * eps, std::pow(eps, 0.75));
  // threshold and sharpness, indicator.hpp:443-452
  p.threshold = 0.5 * std::pow(10.0, -1.8 * std::pow((double)b.n, 0.25));
  p.sfac = std::log((1.0 - 0.0001) / 0.0001);
  p.floor_value = c.positivity_floor;
  p.lhat0 = b.lhat_left[0];
  p.lhatN = b.lhat_right[b.n - 1];
  for (int i = 0; i < b.n; ++i) {
    p.inv_w[i] = b.inv_weights[i];
    for (int j = 0; j < b.n; ++j) {
      p.Ds[i][j] = b.Dsplit[i * b.n + j];
      p.Dhat[i][j] = b.Dhat[i * b.n + j];
    }
  }
  // Navier-Stokes: kappa = gamma / ((gamma - 1) Pr), compressible_navier_stokes3D.hpp:19-22
  p.mu = c.mu;
  p.kappa = c.gamma / ((c.gamma - 1.0) * c.prandtl);
  p.grad = h->d_grad;
  p.grad_out = h->d_grad;
  p.visc_out = h->d_visc;
  for (int k = 0; k < 3; ++k) p.cells[k] = c.cells[k];
  for (int f = 0; f < 6; ++f) {
    p.bc[f] = f < 2 * c.dim ? c.face_boundary[f] : dgsem::kBcPeriodic;
    for (int v = 0; v < 5; ++v) p.bc_state[f][v] = c.face_state[f][v];
    p.bc_values[f] = nullptr;
  }
}

// Per-face boundary kinds: a face left at 0 takes boundary[its direction].
void resolve_faces(dgsem_config& c) {
  for (int f = 0; f < 6; ++f) {
    if (f >= 2 * c.dim) {
      c.face_boundary[f] = DGSEM_BC_PERIODIC;
      continue;
    }
    if (c.face_boundary[f] == DGSEM_BC_PERIODIC) c.face_boundary[f] = c.boundary[f / 2];
  }
}

bool has_boundaries(const dgsem_config& c) {
  for (int f = 0; f < 2 * c.dim; ++f)
    if (c.face_boundary[f] != DGSEM_BC_PERIODIC) return true;
  return false;
}

// Navier-Stokes viscous buffer: the node-parallel RHS kernel takes the
// per-node fluxes of every direction, the line kernels the viscous RHS part.
int64_t visc_width(const dgsem_config& c) {
  const bool sc = c.volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING;
  return dgsem::node_kernel_for(c.polydeg, sc, has_boundaries(c)) ? c.dim : 1;
}

int check_config(dgsem_handle* h, const dgsem_config& c) {
  if (c.dim != 2 && c.dim != 3) return fail(h, DGSEM_E_INVALID, "dim must be 2 or 3");
  if (c.polydeg < 1 || c.polydeg > 7) return fail(h, DGSEM_E_INVALID, "polydeg must be 1..7");
  for (int k = 0; k < c.dim; ++k) {
    if (c.cells[k] < 1) return fail(h, DGSEM_E_INVALID, "cells must be positive");
    if (!(c.domain_right[k] > c.domain_left[k]))
      return fail(h, DGSEM_E_INVALID, "domain_right must exceed domain_left");
  }
  // faces resolved by resolve_faces
  for (int k = 0; k < c.dim; ++k) {
    const int lo = c.face_boundary[2 * k], hi = c.face_boundary[2 * k + 1];
    for (int kind : {lo, hi})
      if (kind < DGSEM_BC_PERIODIC || kind > DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL)
        return fail(h, DGSEM_E_INVALID, "unknown boundary condition");
      else if ((kind == DGSEM_BC_NS_NOSLIP_WALL || kind == DGSEM_BC_NS_SLIP_WALL) &&
               c.equation != DGSEM_NAVIER_STOKES)
        return fail(h, DGSEM_E_INVALID, "the adiabatic walls are Navier-Stokes boundary conditions");
    if ((lo == DGSEM_BC_PERIODIC) != (hi == DGSEM_BC_PERIODIC))
      return fail(h, DGSEM_E_INVALID, "a direction is periodic on both faces or on neither");
  }
  if (has_boundaries(c) && c.dim != 2)
    return fail(h, DGSEM_E_UNSUPPORTED, "non-periodic boundaries are built for 2D meshes only");
  const int64_t E = (int64_t)c.cells[0] * c.cells[1] * (c.dim == 3 ? c.cells[2] : 1);
  if (E > std::numeric_limits<int>::max() / 2) return fail(h, DGSEM_E_INVALID, "mesh too large");
  if (c.equation != DGSEM_EULER && c.equation != DGSEM_NAVIER_STOKES)
    return fail(h, DGSEM_E_INVALID, "unknown equation");
  if (c.equation == DGSEM_NAVIER_STOKES && !(c.mu >= 0.0 && c.prandtl > 0.0))
    return fail(h, DGSEM_E_INVALID, "Navier-Stokes needs mu >= 0 and Pr > 0");
  if (c.volume_integral != DGSEM_VOLUME_SPLIT_FORM && c.volume_integral != DGSEM_VOLUME_SHOCK_CAPTURING)
    return fail(h, DGSEM_E_INVALID, "unknown volume integral");
  if (c.volume_flux != DGSEM_FLUX_CHANDRASHEKAR && c.volume_flux != DGSEM_FLUX_CENTRAL)
    return fail(h, DGSEM_E_INVALID, "unknown volume flux");
  for (int kind : {c.surface_flux, c.fv_flux})
    if (kind != DGSEM_FLUX_LLF && kind != DGSEM_FLUX_HLLC && kind != DGSEM_FLUX_CENTRAL)
      return fail(h, DGSEM_E_INVALID, "unknown surface or FV flux");
  // CentralFlux takes a normal in 2D only (numerical_flux.hpp:247-257); the
  // kernels that read flux kinds at run time are the line kernels (N <= 5)
  if (dgsem::uses_central_flux(c)) {
    if (c.dim != 2)
      return fail(h, DGSEM_E_UNSUPPORTED, "CentralFlux is built for 2D meshes (as in the reference)");
    if (c.polydeg > 5) return fail(h, DGSEM_E_UNSUPPORTED, "CentralFlux is built for N <= 5");
  }
  if (c.geometry != DGSEM_GEOMETRY_NODAL && c.geometry != DGSEM_GEOMETRY_AFFINE)
    return fail(h, DGSEM_E_INVALID, "unknown geometry mode");
  if (!(c.gamma > 1.0)) return fail(h, DGSEM_E_INVALID, "gamma must exceed 1");
  if (c.dim == 3 && c.polydeg == 1) return fail(h, DGSEM_E_UNSUPPORTED, "3D N=1 is not built");
  if (c.polydeg == 6) return fail(h, DGSEM_E_UNSUPPORTED, "N=6 is not built");
  return DGSEM_OK;
}

hipStream_t stream_of(const dgsem_handle* h) { return h->stream; }

int launch(dgsem_handle* h, dgsem::KParams p) {
  for (int f = 0; f < 2 * h->cfg.dim; ++f)
    if (p.bc[f] == DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL && !p.bc_values[f])
      return fail(h, DGSEM_E_INVALID, "a mixed Dirichlet/slip-wall face needs its per-node states");
  const bool sc = h->cfg.volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING;
  const bool nodal = h->cfg.geometry == DGSEM_GEOMETRY_NODAL;
  if (h->cfg.equation == DGSEM_NAVIER_STOKES) {
    // BR1 gradient and viscous part of the stage input, then the inviscid
    // kernel adds the viscous part before -1/J
    hipError_t e = dgsem::launch_parabolic(h->cfg.dim, h->cfg.polydeg, sc, nodal, has_boundaries(h->cfg),
                                           p, stream_of(h));
    if (e != hipSuccess) return hip_fail(h, e, "parabolic kernel launch");
    p.visc = h->d_visc;
  } else {
    p.visc = nullptr;
  }
  hipError_t e = dgsem::launch_rhs(h->cfg.dim, h->cfg.polydeg, sc, nodal, has_boundaries(h->cfg), p,
                                   stream_of(h));
  if (e != hipSuccess) return hip_fail(h, e, "rhs kernel launch");
  return DGSEM_OK;
}

}  // namespace

namespace dgsem {

int stage_params(dgsem_handle* h, int stage, double dt, const double* u_base, const double* u_k,
                 double* u_out, KParams* out) {
  if (!u_k || !u_out || u_out == u_k) return fail(h, DGSEM_E_INVALID, "u_out must not alias u_k");
  KParams p = h->base;
  p.u = u_k;
  p.out = u_out;
  // each stage records its blending factor in its own slot (0: plain RHS)
  if (stage >= 0 && stage <= 3) p.alpha = h->d_alpha + stage * h->E;
  // coefficients exactly as SSPRK3::step builds them, SSPRK3.hpp:162-180
  if (stage == 0) {
    p.mode = kModeRhs;
  } else if (stage == 1) {
    p.mode = kModeStage2;
    p.c1 = 1.0;
    p.c2 = dt;
  } else if (stage == 2 || stage == 3) {
    if (!u_base) return fail(h, DGSEM_E_INVALID, "u_base required for stages 2 and 3");
    p.mode = kModeStage3;
    p.u_base = u_base;
    if (stage == 2) {
      p.c1 = 3.0 / 4.0;
      p.c2 = 1.0 / 4.0;
      p.c3 = 1.0 / 4.0 * dt;
    } else {
      p.c1 = 1.0 / 3.0;
      p.c2 = 2.0 / 3.0;
      p.c3 = 2.0 / 3.0 * dt;
    }
  } else {
    return fail(h, DGSEM_E_INVALID, "stage must be 0, 1, 2 or 3");
  }
  *out = p;
  return DGSEM_OK;
}

int launch_params(dgsem_handle* h, const KParams& p) { return launch(h, p); }

}  // namespace dgsem

extern "C" {

int dgsem_abi_version(void) { return DGSEM_ABI_VERSION; }

void dgsem_config_default(dgsem_config* c, int dim, int polydeg) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->dim = dim;
  c->polydeg = polydeg;
  for (int k = 0; k < 3; ++k) {
    c->cells[k] = 1;
    c->domain_left[k] = -1.0;
    c->domain_right[k] = 1.0;
    c->boundary[k] = DGSEM_BC_PERIODIC;
  }
  c->equation = DGSEM_EULER;
  c->gamma = 1.4;
  c->mu = dim == 3 ? 6.25e-4 : 5.0e-3;  // compressible_navier_stokes{3D,2D}.hpp defaults
  c->prandtl = dim == 3 ? 0.72 : 0.73;
  c->volume_integral = DGSEM_VOLUME_SPLIT_FORM;
  c->volume_flux = DGSEM_FLUX_CHANDRASHEKAR;
  c->surface_flux = DGSEM_FLUX_LLF;
  c->fv_flux = DGSEM_FLUX_LLF;
  c->alpha_max = 0.5;  // structured_solver.hpp:140-142
  c->alpha_min = 0.001;
  c->alpha_smooth = 0;
  c->positivity_floor = 0.0;
  c->geometry = DGSEM_GEOMETRY_NODAL;
}

int dgsem_create(const dgsem_config* cfg, void* stream, dgsem_handle** out) {
  if (!cfg || !out) return DGSEM_E_INVALID;
  *out = nullptr;
  dgsem_handle* h = new (std::nothrow) dgsem_handle();
  if (!h) return DGSEM_E_ALLOC;
  h->cfg = *cfg;
  if (h->cfg.dim == 2) h->cfg.cells[2] = 1;
  resolve_faces(h->cfg);
  int rc = check_config(h, h->cfg);
  if (rc != DGSEM_OK) {
    // keep the handle so the caller can read the message
    *out = h;
    return rc;
  }
  h->stream = (hipStream_t)stream;
  if (hipGetDevice(&h->device) != hipSuccess) {
    *out = h;
    return fail(h, DGSEM_E_HIP, "no HIP device");
  }
  int periodic[3] = {1, 1, 1};
  for (int k = 0; k < h->cfg.dim; ++k) periodic[k] = h->cfg.face_boundary[2 * k] == DGSEM_BC_PERIODIC;
  try {
    h->basis = dgsem::make_basis(h->cfg.polydeg);
    h->geo = dgsem::make_geometry(h->basis, h->cfg.dim, h->cfg.cells, h->cfg.domain_left,
                                  h->cfg.domain_right, periodic);
  } catch (const std::exception& ex) {
    *out = h;
    return fail(h, DGSEM_E_INVALID, ex.what());
  }
  h->nvars = h->cfg.dim + 2;
  h->nn = h->geo.nn;
  h->E = h->geo.E;
  std::vector<double> Ja, invJ;
  device_geometry(h, Ja, invJ);
  std::vector<double> sn;
  if (h->cfg.geometry == DGSEM_GEOMETRY_NODAL && h->cfg.volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING)
    sn = h->geo.subcell_normals;
  if ((rc = upload(h, &h->d_Ja, Ja)) || (rc = upload(h, &h->d_invJ, invJ)) ||
      (rc = upload(h, &h->d_sn, sn)) || (rc = upload(h, &h->d_nbr, neighbour_table(h->geo, periodic))) ||
      (rc = upload(h, &h->d_Vinv, h->basis.Vinv)) ||
      (rc = upload(h, &h->d_alpha, std::vector<double>(4 * h->E, 0.0))) ||
      // Navier-Stokes: the per-node gradient only where the line kernels
      // read it (the node-parallel path keeps the viscous fluxes alone; the
      // tangent and adjoint allocate the gradient when they first need it)
      (h->cfg.equation == DGSEM_NAVIER_STOKES &&
       ((visc_width(h->cfg) == 1 &&
         (rc = alloc_zero(h, &h->d_grad, (size_t)h->E * h->cfg.dim * h->nvars * h->nn))) ||
        (rc = alloc_zero(h, &h->d_visc, (size_t)h->E * visc_width(h->cfg) * h->nvars * h->nn))))) {
    release(h);
    *out = h;
    return rc;
  }
  // periodic nodal box meshes: the metric terms' table for the line kernels
  if (h->cfg.geometry == DGSEM_GEOMETRY_NODAL && !has_boundaries(h->cfg)) {
    std::vector<double> tab;
    if (dgsem::box_metric_table(h->geo, tab, h->ja_box_off) && (rc = upload(h, &h->d_ja_box, tab))) {
      release(h);
      *out = h;
      return rc;
    }
  }
  fill_params(h);
  *out = h;
  return DGSEM_OK;
}

int dgsem_destroy(dgsem_handle* h) {
  if (!h) return DGSEM_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  else hipDeviceSynchronize();
  release(h);
  delete h;
  return DGSEM_OK;
}

const char* dgsem_last_error(const dgsem_handle* h) { return h ? h->err.c_str() : "null handle"; }

int dgsem_sizes(const dgsem_handle* h, int64_t* E, int* ndofs, int* nvars) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (E) *E = h->E;
  if (ndofs) *ndofs = h->nn;
  if (nvars) *nvars = h->nvars;
  return DGSEM_OK;
}

int dgsem_set_boundary_values(dgsem_handle* h, int face, const double* device_values) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (face < 0 || face >= 2 * h->cfg.dim) return fail(h, DGSEM_E_INVALID, "face out of range");
  const int kind = h->cfg.face_boundary[face];
  if (kind != DGSEM_BC_DIRICHLET && kind != DGSEM_BC_MIXED_DIRICHLET_SLIP_WALL)
    return fail(h, DGSEM_E_INVALID, "per-node states need a Dirichlet or mixed face");
  h->base.bc_values[face] = device_values;
  return DGSEM_OK;
}

int dgsem_set_indicator_parameters(dgsem_handle* h, double amax, double amin, int smooth) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  h->cfg.alpha_max = amax;
  h->cfg.alpha_min = amin;
  h->cfg.alpha_smooth = smooth;
  h->base.alpha_max = amax;
  h->base.alpha_min = amin;
  return DGSEM_OK;
}

int dgsem_node_coordinates(const dgsem_handle* h, double* out) {
  if (!h || !h->d_Ja || !out) return DGSEM_E_INVALID;
  std::memcpy(out, h->geo.coords.data(), h->geo.coords.size() * sizeof(double));
  return DGSEM_OK;
}

int dgsem_contravariant_vectors(const dgsem_handle* h, double* out) {
  if (!h || !h->d_Ja || !out) return DGSEM_E_INVALID;
  std::memcpy(out, h->geo.contravariant.data(), h->geo.contravariant.size() * sizeof(double));
  return DGSEM_OK;
}

int dgsem_inverse_jacobian(const dgsem_handle* h, double* out) {
  if (!h || !h->d_Ja || !out) return DGSEM_E_INVALID;
  std::memcpy(out, h->geo.inverse_jacobian.data(), h->geo.inverse_jacobian.size() * sizeof(double));
  return DGSEM_OK;
}

int dgsem_pack_state(dgsem_handle* h, const double* ref, double* internal) {
  if (!h || !h->d_Ja || !ref || !internal) return DGSEM_E_INVALID;
  hipError_t e = dgsem::launch_pack(ref, internal, h->E, h->nn, h->nvars, true, h->stream);
  return e == hipSuccess ? DGSEM_OK : hip_fail(h, e, "pack");
}

int dgsem_unpack_state(dgsem_handle* h, const double* internal, double* ref) {
  if (!h || !h->d_Ja || !ref || !internal) return DGSEM_E_INVALID;
  hipError_t e = dgsem::launch_pack(internal, ref, h->E, h->nn, h->nvars, false, h->stream);
  return e == hipSuccess ? DGSEM_OK : hip_fail(h, e, "unpack");
}

int dgsem_pack_state_layout(dgsem_handle* h, int layout, const double* ref, double* internal) {
  if (!h || !h->d_Ja || !ref || !internal) return DGSEM_E_INVALID;
  if (layout == DGSEM_LAYOUT_RIGHT) return dgsem_pack_state(h, ref, internal);
  if (layout != DGSEM_LAYOUT_LEFT) return fail(h, DGSEM_E_INVALID, "unknown layout");
  hipError_t e = dgsem::launch_pack_left(ref, internal, h->cfg.cells, h->nn, h->nvars, true,
                                         h->stream);
  return e == hipSuccess ? DGSEM_OK : hip_fail(h, e, "pack");
}

int dgsem_unpack_state_layout(dgsem_handle* h, int layout, const double* internal, double* ref) {
  if (!h || !h->d_Ja || !ref || !internal) return DGSEM_E_INVALID;
  if (layout == DGSEM_LAYOUT_RIGHT) return dgsem_unpack_state(h, internal, ref);
  if (layout != DGSEM_LAYOUT_LEFT) return fail(h, DGSEM_E_INVALID, "unknown layout");
  hipError_t e = dgsem::launch_pack_left(internal, ref, h->cfg.cells, h->nn, h->nvars, false,
                                         h->stream);
  return e == hipSuccess ? DGSEM_OK : hip_fail(h, e, "unpack");
}

int dgsem_rhs(dgsem_handle* h, const double* u, double* du, double /*t*/) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (!u || !du || u == du) return fail(h, DGSEM_E_INVALID, "u and du must be distinct device arrays");
  dgsem::KParams p = h->base;
  p.u = u;
  p.out = du;
  p.mode = dgsem::kModeRhs;
  const int rc = launch(h, p);
  if (rc == DGSEM_OK) h->alpha_last = 0;
  return rc;
}

int dgsem_get_alpha(const dgsem_handle* h, double* alpha) {
  if (!h || !h->d_Ja || !alpha) return DGSEM_E_INVALID;
  hipError_t e = hipMemcpyAsync(alpha, h->d_alpha + h->alpha_last * h->E, h->E * sizeof(double),
                                hipMemcpyDeviceToDevice, h->stream);
  return e == hipSuccess ? DGSEM_OK : DGSEM_E_HIP;
}

int dgsem_get_stage_alpha(const dgsem_handle* h, int stage, double* alpha) {
  if (!h || !h->d_Ja || !alpha || stage < 0 || stage > 3) return DGSEM_E_INVALID;
  hipError_t e = hipMemcpyAsync(alpha, h->d_alpha + stage * h->E, h->E * sizeof(double),
                                hipMemcpyDeviceToDevice, h->stream);
  return e == hipSuccess ? DGSEM_OK : DGSEM_E_HIP;
}

int dgsem_freeze_alpha(dgsem_handle* h, const double* alpha) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (alpha && h->cfg.volume_integral != DGSEM_VOLUME_SHOCK_CAPTURING)
    return fail(h, DGSEM_E_INVALID, "a frozen blending factor needs the shock-capturing volume integral");
  h->base.alpha_in = alpha;
  return DGSEM_OK;
}

int dgsem_ssprk3_stage(dgsem_handle* h, int stage, double dt, const double* u_base,
                       const double* u_k, double* u_out) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (stage < 1 || stage > 3) return fail(h, DGSEM_E_INVALID, "stage must be 1, 2 or 3");
  dgsem::KParams p;
  int rc = dgsem::stage_params(h, stage, dt, u_base, u_k, u_out, &p);
  if (rc == DGSEM_OK) rc = launch(h, p);
  if (rc == DGSEM_OK) h->alpha_last = stage;
  return rc;
}

int dgsem_ssprk3_steps(dgsem_handle* h, double* u, double dt, int nsteps, double* tmp1, double* tmp2) {
  if (!h || !h->d_Ja) return DGSEM_E_INVALID;
  if (!u || !tmp1 || !tmp2 || u == tmp1 || u == tmp2 || tmp1 == tmp2)
    return fail(h, DGSEM_E_INVALID, "u, tmp1, tmp2 must be distinct device arrays");
  for (int s = 0; s < nsteps; ++s) {
    int rc;
    if ((rc = dgsem_ssprk3_stage(h, 1, dt, nullptr, u, tmp1))) return rc;
    if ((rc = dgsem_ssprk3_stage(h, 2, dt, u, tmp1, tmp2))) return rc;
    if ((rc = dgsem_ssprk3_stage(h, 3, dt, u, tmp2, u))) return rc;
  }
  return DGSEM_OK;
}

int dgsem_synchronize(dgsem_handle* h) {
  if (!h) return DGSEM_E_INVALID;
  hipError_t e = h->stream ? hipStreamSynchronize(h->stream) : hipDeviceSynchronize();
  return e == hipSuccess ? DGSEM_OK : hip_fail(h, e, "synchronize");
}

}  // extern "C"
