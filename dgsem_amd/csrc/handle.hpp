// Internal definition of the library handle, shared by the C ABI
// (abi.hip) and the batched-ensemble API (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/dgsem_amd.h"
#include "host_setup.hpp"
#include "kernels.hpp"

struct dgsem_handle {
  dgsem_config cfg{};
  dgsem::BasisTables basis;
  dgsem::Geometry geo;
  hipStream_t stream = nullptr;
  int device = 0;
  int nvars = 0, nn = 0;
  int64_t E = 0;
  // device buffers
  double* d_Ja = nullptr;
  double* d_invJ = nullptr;
  double* d_sn = nullptr;
  int* d_nbr = nullptr;
  double* d_Vinv = nullptr;
  double* d_ja_box = nullptr;  // box-mesh table of the metric terms (KParams::ja_box)
  int64_t ja_box_off[3] = {0, 0, 0};
  double* d_alpha = nullptr;  // blending factor per element, [4][E]: slot 0 the last
                              // dgsem_rhs, slots 1-3 the last SSPRK3 stage of each kind
  int alpha_last = 0;         // slot of the last launch that wrote one
  double* d_grad = nullptr;  // Navier-Stokes gradient [E][dim][var][dof]
  double* d_visc = nullptr;  // Navier-Stokes viscous RHS part [E][var][dof] or
                             // per-node fluxes [E][dim][var][dof] (node_kernel_for)
  dgsem::KParams base{};
  // tangent / adjoint scratch, allocated on first use: d_ad_dgrad
  // [E][dim][NV][NN], d_ad_dvisc and d_ad_qbar [E][NV][NN] (Navier-Stokes),
  // d_wq the tensor-product quadrature weight of each node, d_ke_partial the
  // per-block sums of the kinetic-energy reduction
  double* d_ad_dgrad = nullptr;
  double* d_ad_dvisc = nullptr;
  double* d_ad_qbar = nullptr;
  double* d_wq = nullptr;
  double* d_ke_partial = nullptr;
  double* d_an_partial = nullptr;  // per-block partials of dgsem_analyze
  // shock capturing without a frozen alpha: the alpha of the linearisation
  // state [E] and the forward RHS that computes it (discarded)
  double* d_ad_alpha = nullptr;
  double* d_ad_out = nullptr;
  // step derivatives with the positivity floor: a stage input's RHS and the
  // cotangent of the step's result after the last clamp [E][NV][NN]
  double* d_fl_r = nullptr;
  double* d_fl_w = nullptr;
  std::string err;
};

namespace dgsem {

// The kernel parameters of SSPRK3 stage 1, 2 or 3 (coefficients as
// SSPRK3::step builds them), or of the plain right-hand side for stage 0
// (u_k -> u_out = L(u_k)).  Returns DGSEM_OK or an error code with the
// handle's message set.
int stage_params(dgsem_handle* h, int stage, double dt, const double* u_base, const double* u_k,
                 double* u_out, KParams* out);

// Enqueues the kernels for one evaluation with parameters p.
int launch_params(dgsem_handle* h, const KParams& p);

// Uploads the tensor-product quadrature weight of each node (h->d_wq) and
// the reduction scratch on first use.
int prepare_weights(dgsem_handle* h);

// CentralFlux anywhere in the configuration: as the volume flux, the
// surface flux or, with shock capturing, the subcell flux.
inline bool uses_central_flux(const dgsem_config& c) {
  return c.volume_flux == DGSEM_FLUX_CENTRAL || c.surface_flux == DGSEM_FLUX_CENTRAL ||
         (c.volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING && c.fv_flux == DGSEM_FLUX_CENTRAL);
}

}  // namespace dgsem
