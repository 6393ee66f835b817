// Fused DGSEM right-hand side / SSPRK3 stage kernel for gfx950.
//
// One workgroup owns EPW elements.  Work is split three ways:
//   * node threads load the element state, form per-node primitives and,
//     for shock capturing, the modal indicator (LDS tensor sweeps);
//   * line threads -- one per (element, direction, line of N+1 nodes) --
//     evaluate all two-point fluxes on their line (flux differencing with
//     D_split), the first-order subcell fluxes of troubled elements, and the
//     numerical fluxes on the two element faces the line ends on (face
//     prolongation is the identity on Gauss-Lobatto nodes), lifting them with
//     l_hat.  Every accumulation stays in registers;
//   * node threads finally sum the per-direction contributions, apply
//     -1/J and either store L(u) or the SSPRK3 stage combination.
// The state is read from HBM once (plus the neighbours' face traces) and the
// result written once; there are no atomics, so the result does not depend
// on scheduling.
//
// Reference kernels replaced (RetrO598/DGSEM): volume_integral
// (volume_flux.hpp:221-410 split form, :449-630 subcell FV, :743-879 policy),
// HGIndicator{2D,3D} (indicator.hpp:313-585), interface_flux
// (surface_flux.hpp:45-271), surface_integral (surface_integral.hpp:35-154),
// jacobian_proj (jacobian.hpp:22-105), parallel_ma2/ma3 (SSPRK3.hpp:12-143).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "physics.hpp"

namespace dgsem {
namespace detail {

template <int DIM, int N>
struct Shape {
  static constexpr int n = N + 1;
  static constexpr int NV = DIM + 2;
  static constexpr int NN = DIM == 2 ? n * n : n * n * n;
  static constexpr int FN = DIM == 2 ? n : n * n;  // lines per direction
  static constexpr int LINES = DIM * FN;            // line threads per element
  static constexpr int NP = DIM + 5;                // rho, v[DIM], beta, kin, p, c
  // LDS planes hold one padding slot per row of n nodes, so that line
  // threads walking rows (node stride n) hit distinct banks.
  static constexpr int PN = NN + NN / n;
  // Element slot strides of the LDS state (NV planes) and primitive (NP
  // planes) arrays, padded to 8 doubles modulo 32 so that the slots of the
  // elements sharing a wave start in different banks.
  static constexpr int slot_pad(int x) { return (8 - x % 32 + 32) % 32; }
  static constexpr int SU = NV * PN + slot_pad(NV * PN);
  static constexpr int SP = NP * PN + slot_pad(NP * PN);
};

// Padded LDS position of a node (see Shape::PN).
template <int N>
__device__ __forceinline__ int lds_pos(int node) {
  return node + node / (N + 1);
}

// Metric term Ja^r_c at node `dof` of the element with indices eid on a box
// mesh, from the table described at KParams::ja_box.
template <int DIM, int N>
__device__ __forceinline__ double box_ja(const KParams& p, const int* eid, int c, int r, int dof) {
  constexpr int NN = Shape<DIM, N>::NN;
  int64_t pair;
  if constexpr (DIM == 3) {
    const int m = c == 2 ? 0 : c + 1, l = c == 0 ? 2 : c - 1;  // (c + 1) % 3, (c + 2) % 3
    pair = (int64_t)eid[m] * p.cells[l] + eid[l];
  } else {
    pair = eid[1 - c];
  }
  return p.ja_box[p.ja_box_off[c] + (pair * DIM + r) * NN + dof];
}

// Elements per workgroup: keeps EPW * LINES near 192 threads.
template <int DIM, int N>
constexpr int elems_per_block() {
  constexpr int L = Shape<DIM, N>::LINES;
  constexpr int e = 192 / L;
  return e < 1 ? 1 : e;
}

// dof of the k-th node on line `line` of direction `dir`; the line index
// enumerates the remaining node indices in increasing dimension order.
template <int DIM, int N>
__device__ __forceinline__ int line_dof(int dir, int line, int k) {
  constexpr int n = N + 1;
  if constexpr (DIM == 2) {
    return dir == 0 ? line * n + k : k * n + line;
  } else {
    const int a = line % n, b = line / n;
    if (dir == 0) return (b * n + a) * n + k;
    if (dir == 1) return (b * n + k) * n + a;
    return (k * n + b) * n + a;
  }
}

// Maps hardware block ids (dispatched round-robin over the 8 XCDs) to
// contiguous element ranges per XCD, so face neighbours share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb / 8, r = nb % 8;
  const int x = b % 8, i = b / 8;
  return x * q + (x < r ? x : r) + i;
}

// Hennemann-Gassner modal indicator for the EPW elements of the block,
// indicator.hpp:337-417 (2D) / :454-574 (3D), including the 3D total-energy
// sum that skips the modes (i = N, j < N, k < N) and gamma fixed at 1.4.
//
// The modal coefficients come from one sweep of V^-1 per direction.  A
// sweep gives each thread one line of n nodes (EPW n^(DIM-1) lines), so every
// index into V^-1 is a compile-time constant and its entries are scalar
// loads; the first sweep forms the indicator variable of its line's nodes
// itself.  Each output is the sum over the line in increasing order, with
// the same fused multiply-adds as a node-parallel sweep.  A thread reads its
// line before it writes it, so every sweep works in place in one scratch
// plane per element, at the padded node positions of the state planes
// (lds_pos): the line reads and writes hit distinct banks.  The energy sums
// follow the reference's order, one thread per element.
//
// OVERLAP: the sweeps and the energy sums keep only EPW n^(DIM-1) and EPW
// threads busy; the others meanwhile form the per-node primitives, prim(i)
// for i < EPW NN, a slice per phase.  s_w (indicator_scratch() doubles) then
// lies apart from the primitives' array; without OVERLAP it may be that
// array, and prim goes unused.
template <int DIM, int N, int EPW>
constexpr int indicator_scratch() {
  return EPW * Shape<DIM, N>::PN;
}

// Whether a shock-capturing kernel overlaps the primitives with the
// indicator: only where its separate scratch leaves the LDS-bound number of
// blocks per CU (160 KiB of LDS) unchanged -- in 2D at N = 3 it would cut
// three blocks to two, and the step measured 23% slower.
template <int DIM, int N, int EPW, bool SC>
constexpr bool overlap_prims() {
  using S = Shape<DIM, N>;
  constexpr int lds_cu = 160 * 1024;
  constexpr int base = 8 * (EPW * (S::SU + S::SP) + EPW);
  constexpr int with = base + 8 * indicator_scratch<DIM, N, EPW>();
  return SC && lds_cu / with == lds_cu / base;
}

template <int DIM, int N, int EPW, int BLOCK, bool OVERLAP, class Prim>
__device__ void block_alpha(const KParams& p, const double* s_u, double* s_w, double* s_alpha,
                            int nvalid, Prim&& prim) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NN = S::NN, NV = S::NV, PN = S::PN;
  constexpr int LN = NN / n;    // lines per direction and element
  constexpr int T0 = EPW * LN;  // threads of a sweep
  static_assert(!OVERLAP || T0 < BLOCK, "threads left for the primitives");
  static_assert(OVERLAP || indicator_scratch<DIM, N, EPW>() <= EPW * S::SP,
                "indicator scratch in the primitives' LDS");
  constexpr int PHASES = DIM + 1;  // the sweeps and the energy sums
  double* A = s_w;
  const double* V = p.Vinv;
  const int tid = threadIdx.x;
  // the primitives' slice of phase f, on the threads a sweep leaves idle
  auto prims_slice = [&](int f) {
    if (OVERLAP && tid >= T0) {
      const int lo = f * (EPW * NN) / PHASES, hi = (f + 1) * (EPW * NN) / PHASES;
      for (int i = lo + (tid - T0); i < hi; i += BLOCK - T0) prim(i);
    }
  };
  // sweep along xi, lines of fixed (j, k), from the indicator variable
  prims_slice(0);
  for (int t = tid; t < EPW * LN; t += BLOCK) {
    const int slot = t / LN, row = (t % LN) * n;  // the line's node i = 0
    double x[n];
#pragma unroll
    for (int ii = 0; ii < n; ++ii) {
      double s[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) s[v] = s_u[slot * S::SU + v * PN + lds_pos<N>(row + ii)];
      x[ii] = slot < nvalid ? indicator_variable<DIM>(s) : 0.0;
    }
#pragma unroll
    for (int i = 0; i < n; ++i) {
      double r = 0.0;
#pragma unroll
      for (int ii = 0; ii < n; ++ii) r += V[i * n + ii] * x[ii];
      A[slot * PN + lds_pos<N>(row + i)] = r;
    }
  }
  __syncthreads();
  // sweep along eta, lines of fixed (i, k)
  prims_slice(1);
  for (int t = tid; t < EPW * LN; t += BLOCK) {
    const int slot = t / LN, l = t % LN;
    const int base = (l / n) * n * n + l % n;
    double x[n];
#pragma unroll
    for (int jj = 0; jj < n; ++jj) x[jj] = A[slot * PN + lds_pos<N>(base + jj * n)];
#pragma unroll
    for (int j = 0; j < n; ++j) {
      double r = 0.0;
#pragma unroll
      for (int jj = 0; jj < n; ++jj) r += V[j * n + jj] * x[jj];
      A[slot * PN + lds_pos<N>(base + j * n)] = r;
    }
  }
  __syncthreads();
  const double* modal = A;
  if constexpr (DIM == 3) {
    // sweep along zeta, lines of fixed (i, j)
    prims_slice(2);
    for (int t = tid; t < EPW * LN; t += BLOCK) {
      const int slot = t / LN, base = t % LN;
      double x[n];
#pragma unroll
      for (int kk = 0; kk < n; ++kk) x[kk] = A[slot * PN + lds_pos<N>(base + kk * n * n)];
#pragma unroll
      for (int k = 0; k < n; ++k) {
        double r = 0.0;
#pragma unroll
        for (int kk = 0; kk < n; ++kk) r += V[k * n + kk] * x[kk];
        A[slot * PN + lds_pos<N>(base + k * n * n)] = r;
      }
    }
    __syncthreads();
  }
  prims_slice(DIM);
  if (tid < nvalid) {
    const double* m = modal + tid * PN;
    auto M = [&](int k, int j, int i) { return m[lds_pos<N>((k * n + j) * n + i)]; };
    double total = 0.0, clip1 = 0.0, clip2 = 0.0;
    if constexpr (DIM == 2) {
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) total += M(0, j, i) * M(0, j, i);
      for (int j = 0; j + 1 < n; ++j)
        for (int i = 0; i + 1 < n; ++i) clip1 += M(0, j, i) * M(0, j, i);
      for (int j = 0; j + 2 < n; ++j)
        for (int i = 0; i + 2 < n; ++i) clip2 += M(0, j, i) * M(0, j, i);
    } else {
      for (int k = 0; k < n - 2; ++k)
        for (int j = 0; j < n - 2; ++j)
          for (int i = 0; i < n - 2; ++i) clip2 += M(k, j, i) * M(k, j, i);
      clip1 = clip2;
      for (int j = 0; j < n - 1; ++j)
        for (int i = 0; i < n - 1; ++i) clip1 += M(n - 2, j, i) * M(n - 2, j, i);
      for (int k = 0; k < n - 2; ++k)
        for (int i = 0; i < n - 1; ++i) clip1 += M(k, n - 2, i) * M(k, n - 2, i);
      for (int k = 0; k < n - 2; ++k)
        for (int j = 0; j < n - 2; ++j) clip1 += M(k, j, n - 2) * M(k, j, n - 2);
      total = clip1;
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) total += M(n - 1, j, i) * M(n - 1, j, i);
      for (int k = 0; k < n - 1; ++k)
        for (int i = 0; i < n; ++i) total += M(k, n - 1, i) * M(k, n - 1, i);
    }
    const double ef1 = total != 0.0 ? (total - clip1) / total : 0.0;
    const double ef2 = clip1 != 0.0 ? (clip1 - clip2) / clip1 : 0.0;
    const double energy = dmax(ef1, ef2);
    double a = 1.0 / (1.0 + exp(-p.sfac / p.threshold * (energy - p.threshold)));
    if (a < p.alpha_min) a = 0.0;
    if (a > 1.0 - p.alpha_min) a = 1.0;
    s_alpha[tid] = dmin(p.alpha_max, a);
  }
  __syncthreads();
}

// Minimum waves per SIMD the register allocator must allow (1 = no limit).
#ifndef DGSEM_WAVES_3D
#define DGSEM_WAVES_3D 1
#endif
#ifndef DGSEM_WAVES_2D
#define DGSEM_WAVES_2D 1
#endif
// Nodal-metric 3D kernels sit just above the three-waves register budget
// (170 VGPRs); asking for three waves costs a few spilled registers and
// measures faster than running two waves.
template <int DIM, int N, bool NODAL>
constexpr int min_waves_per_simd() {
  if constexpr (DIM == 3 && NODAL && N <= 3 && DGSEM_WAVES_3D == 1) return 3;
  return DIM == 3 ? DGSEM_WAVES_3D : DGSEM_WAVES_2D;
}

#ifndef DGSEM_PAIR_FENCE_MIN_N
#define DGSEM_PAIR_FENCE_MIN_N 99
#endif
template <int DIM, int N>
constexpr bool pair_fence() {
  return N >= DGSEM_PAIR_FENCE_MIN_N;
}

// Lower and upper face neighbours of element e along direction dir on the
// structured grid of cells[] (element index (ie * ny + je) * nz + ke): the
// periodic wrap, or -1 at a non-periodic boundary face (BCS: kinds from
// p.bc) -- what the neighbour table holds, computed rather than loaded so
// that the neighbour traces do not wait on a dependent load.  The grid
// sizes come in as values: selecting them per thread from the parameter
// block would index it dynamically and turn the reads into vector loads.
template <int DIM, bool BCS>
__device__ __forceinline__ void face_neighbours(const KParams& p, int nx, int ny, int nz, int64_t e,
                                                int dir, int64_t& lo, int64_t& hi) {
  const int ei = (int)e;
  const int stride = dir == 0 ? ny * nz : (dir == 1 ? nz : 1);
  const int count = dir == 0 ? nx : (dir == 1 ? ny : nz);
  const int idx = dir == 0 ? ei / stride : (dir == 1 ? (ei / nz) % ny : ei % nz);
  bool wrap_lo = true, wrap_hi = true;
  if constexpr (BCS) {
    const int blo = dir == 0 ? p.bc[0] : (dir == 1 ? p.bc[2] : p.bc[4]);
    const int bhi = dir == 0 ? p.bc[1] : (dir == 1 ? p.bc[3] : p.bc[5]);
    wrap_lo = blo == kBcPeriodic;
    wrap_hi = bhi == kBcPeriodic;
  }
  lo = idx > 0 ? e - stride : (wrap_lo ? e + (int64_t)(count - 1) * stride : -1);
  hi = idx < count - 1 ? e + stride : (wrap_hi ? e - (int64_t)(count - 1) * stride : -1);
}

// Outer state of a non-periodic face of a 2D compressible-Euler mesh, seen
// from the line of the element that owns the face (face = 2 * dir + upper):
//   slip wall  slip_wall_boundary.hpp:12-62: the inner state with the normal
//              velocity reflected about the (unnormalised) normal nrm; the
//              Navier-Stokes adiabatic walls use the same reflection for
//              their inviscid flux (navier_stokes_wall_boundary.hpp:12-27,
//              :41-77, :150-161);
//   outflow    outflow_boundary.hpp:61-105: the mean of the inner element's
//              nodes on the line normal to the face (this thread's line, su
//              holds the element's padded LDS state);
//   Dirichlet  dirichlet_boundary.hpp:70-118: the face's constant state or,
//              if set, the per-node state of boundary node `bnode`;
//   mixed      dirichlet_boundary.hpp:198-275: per node either the given
//              state or the slip-wall reflection, chosen by the node's flag
//              (the last of its NV + 1 values).
template <int DIM, int N>
__device__ __forceinline__ void boundary_state(const KParams& p, int face, int64_t bnode,
                                               const double* su, const int* pdofs,
                                               const double* uin, const double* nrm,
                                               double* uout) {
  constexpr int n = N + 1, NV = DIM + 2, PN = Shape<DIM, N>::PN;
  static_assert(DIM == 2, "non-periodic boundaries are built for 2D meshes");
  const int kind = p.bc[face];
  const double* vals = p.bc_values[face];
  const bool mixed_dirichlet = kind == kBcMixedDirichletSlipWall && vals[bnode * (NV + 1) + NV] != 0.0;
  if (mixed_dirichlet) {
#pragma unroll
    for (int v = 0; v < NV; ++v) uout[v] = vals[bnode * (NV + 1) + v];
  } else if (kind == kBcSlipWall || kind == kBcNsNoSlipWall || kind == kBcNsSlipWall ||
             kind == kBcMixedDirichletSlipWall) {
    const double rho = uin[0], vx = uin[1] / rho, vy = uin[2] / rho;
    const double nn = nrm[0] * nrm[0] + nrm[1] * nrm[1];
    const double vn = (vx * nrm[0] + vy * nrm[1]) / nn;
    uout[0] = rho;
    uout[1] = rho * (vx - 2.0 * vn * nrm[0]);
    uout[2] = rho * (vy - 2.0 * vn * nrm[1]);
    uout[3] = uin[3];
  } else if (kind == kBcOutflow) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < n; ++k) acc += su[v * PN + pdofs[k]];
      uout[v] = acc / (double)n;
    }
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) uout[v] = vals ? vals[bnode * NV + v] : p.bc_state[face][v];
  }
}

// SF / FF: surface and subcell-FV flux kinds (kFluxLLF, kFluxHLLC) as
// template parameters, so a kernel only carries the register footprint of
// the fluxes it uses; SF = FF = kFluxAny is the 2D instantiation that reads
// the volume, surface and subcell flux kinds from p (CentralFlux).
// BCS: some face of the (2D) mesh has a non-periodic boundary condition.
template <int DIM, int N, int EPW, bool SC, bool NODAL, int SF, int FF, bool BCS, bool BOX>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES))
    __attribute__((amdgpu_waves_per_eu(min_waves_per_simd<DIM, N, NODAL>())))
    rhs_kernel(const KParams p) {
  static_assert(!BOX || (NODAL && !BCS), "box-mesh tables serve periodic nodal meshes");
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES, NP = S::NP;
  constexpr int BLOCK = EPW * L;
  // s_p holds the node primitives during the flux phase; before it, the
  // indicator's modal sweeps use it as scratch and after it, once every line
  // thread is done reading primitives, the per-direction sums (s_du).
  static_assert(NP >= NV && NP >= 2, "s_p must be able to hold s_du and scratch");
  constexpr int PN = S::PN;
  __shared__ double s_u[EPW * S::SU];
  __shared__ double s_p[EPW * S::SP];
  __shared__ double s_alpha[EPW];
  constexpr bool OVERLAP = overlap_prims<DIM, N, EPW, SC>();
  __shared__ double s_ind[OVERLAP ? indicator_scratch<DIM, N, EPW>() : 1];
  double* const s_du = s_p;

  const int tid = threadIdx.x;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e0 = (int64_t)blk * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

#include "element_body.inc"
}

}  // namespace detail
}  // namespace dgsem

#include "node_kernel_impl.hpp"

namespace dgsem {
namespace detail {

template <int DIM, int N, bool SC, bool NODAL, int SF, int FF, bool BCS>
hipError_t launch_one(const KParams& p, hipStream_t stream) {
  static_assert(SF != kFluxAny || (DIM == 2 && N <= 5), "run-time fluxes: 2D line kernels");
  if constexpr (SF != kFluxAny && use_node_kernel<DIM, N, SC, BCS>())
    return launch_node_kernel<DIM, N, NODAL, SF>(p, stream);
  constexpr int EPW = elems_per_block<DIM, N>();
  constexpr int BLOCK = EPW * Shape<DIM, N>::LINES;
  const int64_t nb = (p.E + EPW - 1) / EPW;
  if (nb <= 0) return hipSuccess;
  if (nb > INT32_MAX) return hipErrorInvalidValue;
  // nodal metric terms of a periodic box mesh come from the box table
  if constexpr (NODAL && !BCS) {
    if (p.ja_box) {
      hipLaunchKernelGGL((rhs_kernel<DIM, N, EPW, SC, NODAL, SF, FF, BCS, true>), dim3((unsigned)nb),
                         dim3(BLOCK), 0, stream, p);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((rhs_kernel<DIM, N, EPW, SC, NODAL, SF, FF, BCS, false>), dim3((unsigned)nb),
                     dim3(BLOCK), 0, stream, p);
  return hipGetLastError();
}

template <int DIM, int N, bool SC, bool NODAL, bool BCS>
hipError_t launch_flux(const KParams& p, hipStream_t s) {
  // CentralFlux anywhere (2D, N <= 5; the ABI refuses other cases): the
  // instantiation that selects every flux at run time
  if constexpr (DIM == 2 && N <= 5) {
    if (p.volume_flux == kVolumeCentral || p.surface_flux == kFluxCentral ||
        (SC && p.fv_flux == kFluxCentral))
      return launch_one<DIM, N, SC, NODAL, kFluxAny, kFluxAny, BCS>(p, s);
  } else {
    if (p.volume_flux == kVolumeCentral || p.surface_flux == kFluxCentral ||
        (SC && p.fv_flux == kFluxCentral))
      return hipErrorInvalidValue;
  }
  const bool shllc = p.surface_flux == kFluxHLLC;
  if constexpr (SC) {
    const bool fhllc = p.fv_flux == kFluxHLLC;
    if (shllc)
      return fhllc ? launch_one<DIM, N, SC, NODAL, kFluxHLLC, kFluxHLLC, BCS>(p, s)
                   : launch_one<DIM, N, SC, NODAL, kFluxHLLC, kFluxLLF, BCS>(p, s);
    return fhllc ? launch_one<DIM, N, SC, NODAL, kFluxLLF, kFluxHLLC, BCS>(p, s)
                 : launch_one<DIM, N, SC, NODAL, kFluxLLF, kFluxLLF, BCS>(p, s);
  } else {
    return shllc ? launch_one<DIM, N, SC, NODAL, kFluxHLLC, kFluxLLF, BCS>(p, s)
                 : launch_one<DIM, N, SC, NODAL, kFluxLLF, kFluxLLF, BCS>(p, s);
  }
}

template <int DIM, int N, bool BCS>
hipError_t launch_geo(bool sc, bool nodal, const KParams& p, hipStream_t s) {
  if (sc)
    return nodal ? launch_flux<DIM, N, true, true, BCS>(p, s)
                 : launch_flux<DIM, N, true, false, BCS>(p, s);
  return nodal ? launch_flux<DIM, N, false, true, BCS>(p, s)
               : launch_flux<DIM, N, false, false, BCS>(p, s);
}

// Non-periodic boundaries are built for 2D meshes only.
template <int DIM, int N>
hipError_t launch_dn(bool sc, bool nodal, bool bcs, const KParams& p, hipStream_t s) {
  if constexpr (DIM == 2) {
    if (bcs) return launch_geo<DIM, N, true>(sc, nodal, p, s);
  } else {
    if (bcs) return hipErrorInvalidValue;
  }
  return launch_geo<DIM, N, false>(sc, nodal, p, s);
}

}  // namespace detail
}  // namespace dgsem
