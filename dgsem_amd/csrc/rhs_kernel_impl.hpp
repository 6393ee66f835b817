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
};

// Padded LDS position of a node (see Shape::PN).
template <int N>
__device__ __forceinline__ int lds_pos(int node) {
  return node + node / (N + 1);
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
template <int DIM, int N, int EPW, int BLOCK>
__device__ void block_alpha(const KParams& p, const double* s_u, double* s_w, double* s_alpha,
                            int nvalid) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NN = S::NN, NV = S::NV;
  double* A = s_w;
  double* B = s_w + EPW * NN;
  const int tid = threadIdx.x;
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int slot = i / NN, node = i % NN;
    double s[NV];
    for (int v = 0; v < NV; ++v) s[v] = s_u[(slot * NV + v) * S::PN + lds_pos<N>(node)];
    A[i] = slot < nvalid ? indicator_variable<DIM>(s) : 0.0;
  }
  __syncthreads();
  // sweep along xi
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int node = i % NN, ix = node % n, row = i - ix;
    double r = 0.0;
    for (int ii = 0; ii < n; ++ii) r += p.Vinv[ix * n + ii] * A[row + ii];
    B[i] = r;
  }
  __syncthreads();
  // sweep along eta
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int node = i % NN, jx = (node / n) % n, base = i - jx * n;
    double r = 0.0;
    for (int jj = 0; jj < n; ++jj) r += p.Vinv[jx * n + jj] * B[base + jj * n];
    A[i] = r;
  }
  __syncthreads();
  const double* modal = A;
  if constexpr (DIM == 3) {
    for (int i = tid; i < EPW * NN; i += BLOCK) {
      const int node = i % NN, kx = node / (n * n), base = i - kx * n * n;
      double r = 0.0;
      for (int kk = 0; kk < n; ++kk) r += p.Vinv[kx * n + kk] * A[base + kk * n * n];
      B[i] = r;
    }
    __syncthreads();
    modal = B;
  }
  if (tid < nvalid) {
    const double* m = modal + tid * NN;
    double total = 0.0, clip1 = 0.0, clip2 = 0.0;
    if constexpr (DIM == 2) {
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) total += m[j * n + i] * m[j * n + i];
      for (int j = 0; j + 1 < n; ++j)
        for (int i = 0; i + 1 < n; ++i) clip1 += m[j * n + i] * m[j * n + i];
      for (int j = 0; j + 2 < n; ++j)
        for (int i = 0; i + 2 < n; ++i) clip2 += m[j * n + i] * m[j * n + i];
    } else {
      auto M = [&](int k, int j, int i) { return m[(k * n + j) * n + i]; };
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

// SF / FF: surface and subcell-FV flux kinds (kFluxLLF, kFluxHLLC) as
// template parameters, so a kernel only carries the register footprint of
// the fluxes it uses.
template <int DIM, int N, int EPW, bool SC, bool NODAL, int SF, int FF>
__global__ void __launch_bounds__((EPW * Shape<DIM, N>::LINES))
    __attribute__((amdgpu_waves_per_eu(min_waves_per_simd<DIM, N, NODAL>())))
    rhs_kernel(const KParams p) {
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES, NP = S::NP;
  constexpr int BLOCK = EPW * L;
  // s_p holds the node primitives during the flux phase; before it, the
  // indicator's modal sweeps use it as scratch and after it, once every line
  // thread is done reading primitives, the per-direction sums (s_du).
  static_assert(NP >= NV && NP >= 2, "s_p must be able to hold s_du and scratch");
  constexpr int PN = S::PN;
  __shared__ double s_u[EPW * NV * PN];
  __shared__ double s_p[EPW * NP * PN];
  __shared__ double s_alpha[EPW];
  double* const s_du = s_p;

  const int tid = threadIdx.x;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e0 = (int64_t)blk * EPW;
  const int nvalid = (int)((p.E - e0) < EPW ? (p.E - e0) : EPW);

  // Line-thread identity, and the global loads the line phase needs (its
  // metric row, the neighbours' traces at both line ends and the upper
  // neighbour's metric), issued first so they overlap the LDS staging.
  const int slot = tid / L, rem = tid % L, dir = rem / FN, line = rem % FN;
  const bool valid = slot < nvalid;
  const int64_t e = e0 + slot;
  int dofs[n], pdofs[n];  // global-memory and padded LDS positions
#pragma unroll
  for (int k = 0; k < n; ++k) {
    dofs[k] = line_dof<DIM, N>(dir, line, k);
    pdofs[k] = lds_pos<N>(dofs[k]);
  }
  double ja[NODAL ? n : 1][DIM];
  double nb_lo_u[NV], nb_hi_u[NV], nb_ja[DIM];
  double own_sg = 1.0, nb_sg = 1.0;
  if (valid) {
    if constexpr (NODAL) {
#pragma unroll
      for (int k = 0; k < n; ++k)
#pragma unroll
        for (int c = 0; c < DIM; ++c) ja[k][c] = p.Ja[((e * DIM + dir) * DIM + c) * NN + dofs[k]];
    } else {
#pragma unroll
      for (int c = 0; c < DIM; ++c) ja[0][c] = p.Ja[(e * DIM + dir) * DIM + c];
    }
    const int64_t nb_lo = p.nbr[e * 2 * DIM + 2 * dir], nb_hi = p.nbr[e * 2 * DIM + 2 * dir + 1];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      nb_lo_u[v] = p.u[(nb_lo * NV + v) * NN + dofs[n - 1]];
      nb_hi_u[v] = p.u[(nb_hi * NV + v) * NN + dofs[0]];
    }
    if constexpr (NODAL) {
      // the upper neighbour's metric row is loaded where it is used (below),
      // which keeps the register count at three waves per SIMD
      own_sg = p.invJ[e * NN + dofs[0]] >= 0.0 ? 1.0 : -1.0;
    } else {
#pragma unroll
      for (int c = 0; c < DIM; ++c) nb_ja[c] = p.Ja[(nb_hi * DIM + dir) * DIM + c];
      own_sg = p.invJ[e] >= 0.0 ? 1.0 : -1.0;
      nb_sg = p.invJ[nb_hi] >= 0.0 ? 1.0 : -1.0;
    }
  }

  // 1. stage input of the block's elements: one contiguous chunk
  {
    const double* src = p.u + e0 * NV * NN;
    const int cnt = nvalid * NV * NN;
    for (int i = tid; i < EPW * NV * NN; i += BLOCK) {
      const int row = i / NN, node = i - row * NN;  // row = slot * NV + var
      s_u[row * PN + lds_pos<N>(node)] = i < cnt ? src[i] : 1.0;
    }
  }
  if (tid < EPW) s_alpha[tid] = 0.0;
  __syncthreads();

  // 2. blending factor (indicator sweeps use s_p as scratch), then per-node
  // primitives
  if constexpr (SC) block_alpha<DIM, N, EPW, BLOCK>(p, s_u, s_p, s_alpha, nvalid);
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int slot = i / NN, node = i % NN;
    if (slot >= nvalid) continue;
    double s[NV];
    const int pn = lds_pos<N>(node);
    for (int v = 0; v < NV; ++v) s[v] = s_u[(slot * NV + v) * PN + pn];
    const NodePrims<DIM> q = node_prims<DIM>(p.gamma, s);
    double* d = s_p + slot * NP * PN + pn;
    d[0] = q.rho;
    for (int k = 0; k < DIM; ++k) d[(1 + k) * PN] = q.v[k];
    d[(DIM + 1) * PN] = q.beta;
    d[(DIM + 2) * PN] = q.kin;
    d[(DIM + 3) * PN] = q.p;
    d[(DIM + 4) * PN] = q.c;
  }
  __syncthreads();

  // 3. line threads
  double acc[n][NV];
#pragma unroll
  for (int k = 0; k < n; ++k)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[k][v] = 0.0;

  if (valid) {
    double alpha = 0.0, coef = 1.0;
    bool troubled = false;
    if constexpr (SC) {
      alpha = s_alpha[slot];
      troubled = fabs(alpha - 0.0) > p.atol;
      if (troubled) coef = 1.0 - alpha;
    }

    // flux differencing along the line: D_split is zero on the diagonal
    const double* pr = s_p + slot * NP * PN;
    auto load_prims = [&](int pdof) {
      NodePrims<DIM> q;
      q.rho = pr[pdof];
#pragma unroll
      for (int c = 0; c < DIM; ++c) q.v[c] = pr[(1 + c) * PN + pdof];
      q.beta = pr[(DIM + 1) * PN + pdof];
      q.kin = pr[(DIM + 2) * PN + pdof];
      q.p = 0.0;
      q.c = 0.0;
      return q;
    };
#pragma unroll
    for (int a = 0; a < n - 1; ++a) {
      const NodePrims<DIM> qa = load_prims(pdofs[a]);
#pragma unroll
      for (int b = a + 1; b < n; ++b) {
        // A scheduling barrier per pair keeps the unrolled pair loop from
        // interleaving pairs, which bounds the live registers.
        if constexpr (pair_fence<DIM, N>()) __builtin_amdgcn_sched_barrier(0);
        const NodePrims<DIM> qb = load_prims(pdofs[b]);
        double nrm[DIM], f[NV];
        if constexpr (NODAL) {
#pragma unroll
          for (int c = 0; c < DIM; ++c) nrm[c] = 0.5 * (ja[a][c] + ja[b][c]);
        } else {
#pragma unroll
          for (int c = 0; c < DIM; ++c) nrm[c] = ja[0][c];
        }
        flux_chandrashekar<DIM>(p.inv_gm1, qa, qb, nrm, f);
        const double wa = coef * p.Ds[a][b], wb = coef * p.Ds[b][a];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          acc[a][v] += wa * f[v];
          acc[b][v] += wb * f[v];
        }
      }
    }

    const double* su = s_u + slot * NV * PN;
    if constexpr (SC) {
      if (troubled) {
        // first-order finite volumes on the Gauss-Lobatto subcells
#pragma unroll
        for (int fi = 1; fi < n; ++fi) {
          double ul[NV], ur[NV], nrm[DIM], f[NV];
          for (int v = 0; v < NV; ++v) {
            ul[v] = su[v * PN + pdofs[fi - 1]];
            ur[v] = su[v * PN + pdofs[fi]];
          }
          if constexpr (NODAL) {
            const double* snp = p.sn + ((((e * DIM + dir) * FN) + line) * (n - 1) + (fi - 1)) * DIM;
            for (int c = 0; c < DIM; ++c) nrm[c] = snp[c];
          } else {
            for (int c = 0; c < DIM; ++c) nrm[c] = ja[0][c];
          }
          if constexpr (FF == kFluxHLLC)
            flux_hllc<DIM>(p.gamma, ul, ur, nrm, f);
          else
            flux_llf<DIM>(p.gamma, ul, ur, nrm, f);
          const double wl = alpha * p.inv_w[fi - 1], wr = alpha * p.inv_w[fi];
          for (int v = 0; v < NV; ++v) {
            acc[fi - 1][v] += wl * f[v];
            acc[fi][v] -= wr * f[v];
          }
        }
      }
    }

    // surface fluxes on the two faces the line ends on.  Each interface flux
    // is evaluated with the upper element's metric at its first node, as the
    // reference's per-element interface kernel does, so both neighbours
    // compute bit-identical values.  One code path serves both ends.
    const int dof_up = pdofs[0], dof_lo = pdofs[n - 1];  // padded LDS positions
    if constexpr (NODAL) {
      const int64_t nb_hi = p.nbr[e * 2 * DIM + 2 * dir + 1];
#pragma unroll
      for (int c = 0; c < DIM; ++c) nb_ja[c] = p.Ja[((nb_hi * DIM + dir) * DIM + c) * NN + dofs[0]];
      nb_sg = p.invJ[nb_hi * NN + dofs[0]] >= 0.0 ? 1.0 : -1.0;
    }
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
      const bool lower = side == 0;
      double ul[NV], ur[NV], nrm[DIM], f[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        ul[v] = lower ? nb_lo_u[v] : su[v * PN + dof_lo];
        ur[v] = lower ? su[v * PN + dof_up] : nb_hi_u[v];
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) nrm[c] = lower ? own_sg * ja[0][c] : nb_sg * nb_ja[c];
      if constexpr (SF == kFluxLLF) {
        // own side from the hoisted node quantities, the neighbour's trace
        // with one reciprocal and one square root
        const int dof_own = lower ? dof_up : dof_lo;
        NodePrims<DIM> qo;
        qo.rho = pr[dof_own];
#pragma unroll
        for (int c = 0; c < DIM; ++c) qo.v[c] = pr[(1 + c) * PN + dof_own];
        qo.kin = pr[(DIM + 2) * PN + dof_own];
        qo.p = pr[(DIM + 3) * PN + dof_own];
        qo.c = pr[(DIM + 4) * PN + dof_own];
        qo.beta = 0.0;
        const NodePrims<DIM> qn = face_prims<DIM>(p.gamma, lower ? ul : ur);
        flux_llf_prims<DIM>(ul, lower ? qn : qo, ur, lower ? qo : qn, nrm, f);
      } else {
        flux_hllc<DIM>(p.gamma, ul, ur, nrm, f);
      }
      if (lower) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[0][v] -= f[v] * p.lhat0;
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[n - 1][v] += f[v] * p.lhatN;
      }
    }
  }

  // 4. per-direction contributions into LDS, one direction at a time; s_du
  // aliases the primitives, so every thread must be done reading them first
  __syncthreads();
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    if (valid && dir == d) {
#pragma unroll
      for (int k = 0; k < n; ++k)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const int idx = (slot * NV + v) * PN + pdofs[k];
          s_du[idx] = d == 0 ? acc[k][v] : s_du[idx] + acc[k][v];
        }
    }
    __syncthreads();
  }

  // 5. -1/J and the stage combination
  for (int i = tid; i < EPW * NN; i += BLOCK) {
    const int sl = i / NN, node = i % NN;
    if (sl >= nvalid) continue;
    const int pn = lds_pos<N>(node);
    const int64_t ee = e0 + sl;
    const double fac = NODAL ? -p.invJ[ee * NN + node] : -p.invJ[ee];
    double r[NV];
    const int64_t g = ee * NV * NN + node;
    if (p.visc) {
      for (int v = 0; v < NV; ++v) r[v] = (s_du[(sl * NV + v) * PN + pn] + p.visc[g + v * NN]) * fac;
    } else {
      for (int v = 0; v < NV; ++v) r[v] = s_du[(sl * NV + v) * PN + pn] * fac;
    }
    if (p.mode == kModeRhs) {
      for (int v = 0; v < NV; ++v) p.out[g + v * NN] = r[v];
    } else {
      double s[NV];
      for (int v = 0; v < NV; ++v) {
        const double uk = s_u[(sl * NV + v) * PN + pn];
        s[v] = p.mode == kModeStage2 ? p.c1 * uk + p.c2 * r[v]
                                     : p.c1 * p.u_base[g + v * NN] + p.c2 * uk + p.c3 * r[v];
      }
      if (p.floor_value > 0.0) {
        if (s[0] < p.floor_value) s[0] = p.floor_value;
        double m2 = 0.0;
        for (int k = 0; k < DIM; ++k) m2 += s[k + 1] * s[k + 1];
        const double kin = 0.5 * m2 / s[0];
        if ((p.gamma - 1.0) * (s[NV - 1] - kin) < p.floor_value)
          s[NV - 1] = p.floor_value / (p.gamma - 1.0) + kin;
      }
      for (int v = 0; v < NV; ++v) p.out[g + v * NN] = s[v];
    }
  }
  if constexpr (SC) {
    if (tid < nvalid) p.alpha[e0 + tid] = s_alpha[tid];
  }
}

template <int DIM, int N, bool SC, bool NODAL, int SF, int FF>
hipError_t launch_one(const KParams& p, hipStream_t stream) {
  constexpr int EPW = elems_per_block<DIM, N>();
  constexpr int BLOCK = EPW * Shape<DIM, N>::LINES;
  const int64_t nb = (p.E + EPW - 1) / EPW;
  if (nb <= 0) return hipSuccess;
  if (nb > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((rhs_kernel<DIM, N, EPW, SC, NODAL, SF, FF>), dim3((unsigned)nb), dim3(BLOCK),
                     0, stream, p);
  return hipGetLastError();
}

template <int DIM, int N, bool SC, bool NODAL>
hipError_t launch_flux(const KParams& p, hipStream_t s) {
  const bool shllc = p.surface_flux == kFluxHLLC;
  if constexpr (SC) {
    const bool fhllc = p.fv_flux == kFluxHLLC;
    if (shllc)
      return fhllc ? launch_one<DIM, N, SC, NODAL, kFluxHLLC, kFluxHLLC>(p, s)
                   : launch_one<DIM, N, SC, NODAL, kFluxHLLC, kFluxLLF>(p, s);
    return fhllc ? launch_one<DIM, N, SC, NODAL, kFluxLLF, kFluxHLLC>(p, s)
                 : launch_one<DIM, N, SC, NODAL, kFluxLLF, kFluxLLF>(p, s);
  } else {
    return shllc ? launch_one<DIM, N, SC, NODAL, kFluxHLLC, kFluxLLF>(p, s)
                 : launch_one<DIM, N, SC, NODAL, kFluxLLF, kFluxLLF>(p, s);
  }
}

template <int DIM, int N>
hipError_t launch_dn(bool sc, bool nodal, const KParams& p, hipStream_t s) {
  if (sc) return nodal ? launch_flux<DIM, N, true, true>(p, s) : launch_flux<DIM, N, true, false>(p, s);
  return nodal ? launch_flux<DIM, N, false, true>(p, s) : launch_flux<DIM, N, false, false>(p, s);
}

}  // namespace detail
}  // namespace dgsem
