// Persistent kernel that advances an ensemble of independent 2D cases of
// different element grids and polynomial degrees in one launch per RK stage.
//
// The host cuts every case into tiles of at most EPW(N) consecutive
// elements (EPW(N) = BLOCK / LINES(N) element slots fit one workgroup) and
// sorts the tiles by degree, highest first, and by estimated cost within a
// degree.  Each workgroup walks the tile list with a grid-size stride; a
// tile names its case, whose KParams (state, geometry, basis and stage
// coefficients) sit in a device array, and the degree's loop it falls in
// selects the instantiation of the element body.  The element body is the
// same source as the single-case kernel (element_body.inc), so every case
// gets bit-identical results to its own single-case run.  Degrees 1..5 are
// built: one kernel holds every degree's body, so its register budget is
// that of the largest, and the N = 7 body alone would cut the whole kernel
// to one wave per SIMD.
//
// batch_inst.hip is compiled with machine-level loop-invariant code motion
// off (-mllvm -disable-machine-licm): with it, the FP64 literals of every
// body (series and polynomial coefficients) are hoisted out of the tile
// loops and held in registers across the whole kernel, which made the
// kernel spill.  Without it they are materialised where they are used.
#pragma once

#include <hip/hip_runtime.h>

#include "rhs_kernel_impl.hpp"

namespace dgsem {

namespace detail {

// LDS doubles one tile of degree N needs with BLOCK threads.
template <int N, int BLOCK>
constexpr int batch_lds_doubles() {
  using S = Shape<2, N>;
  constexpr int EPW = BLOCK / S::LINES;
  return EPW * (S::SU + S::SP) + EPW;
}

template <int N, int BLOCK, bool SC, bool NODAL, int SF, int FF, bool BCS = false>
__device__ __forceinline__ void batch_tile(const KParams& p, const int64_t e0, const int nvalid,
                                           double* smem) {
  constexpr int DIM = 2;
  using S = Shape<DIM, N>;
  constexpr int n = S::n, NV = S::NV, NN = S::NN, FN = S::FN, L = S::LINES, NP = S::NP;
  constexpr int PN = S::PN;
  constexpr int EPW = BLOCK / L;
  static_assert(EPW >= 1, "workgroup too small for one element");
  static_assert(NP >= NV && NP >= 2, "s_p must be able to hold s_du and scratch");
  double* const s_u = smem;
  double* const s_p = smem + EPW * S::SU;
  double* const s_alpha = s_p + EPW * S::SP;
  // the indicator works in s_p, before the primitives: the LDS is sized for
  // the largest degree of the ensemble, and 2D kernels are LDS-bound
  constexpr bool OVERLAP = false;
  double* const s_ind = nullptr;
  double* const s_du = s_p;
  // The thread's roles (line, direction, node positions) do not change from
  // tile to tile; computed from a plain threadIdx.x the compiler would hoist
  // them out of the tile loop and keep every degree's copy live in registers
  // across the whole kernel.  The empty asm makes tid opaque per tile.
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  constexpr bool BOX = false;  // ensemble cases read the full nodal metrics
#include "element_body.inc"
}

// The tiles of one degree: those of [begin, end) whose index is congruent
// to blockIdx.x modulo the grid, so that the loops of all degrees together
// visit the degree-sorted list with one grid-size stride.
template <int N, int BLOCK, bool SC, bool NODAL, int SF, int FF>
__device__ __forceinline__ void batch_range(const BatchTile* __restrict__ tiles, int begin, int end,
                                            const KParams* __restrict__ params, double* smem) {
  const int g = (int)gridDim.x;
  for (int t = begin + (((int)blockIdx.x - begin) % g + g) % g; t < end; t += g) {
    const BatchTile tile = tiles[t];
    batch_tile<N, BLOCK, SC, NODAL, SF, FF>(params[tile.case_id], tile.e0, tile.nvalid, smem);
    __syncthreads();  // the next tile reuses the LDS
  }
}

// One loop per degree, each with its own body, rather than one loop that
// switches on the tile's degree: the register allocator then sees one
// degree's live values at a time.
template <int BLOCK, bool SC, bool NODAL, int SF, int FF>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
    batch_kernel(const BatchTile* __restrict__ tiles, BatchRanges r,
                 const KParams* __restrict__ params) {
  extern __shared__ double smem[];
  batch_range<5, BLOCK, SC, NODAL, SF, FF>(tiles, r.begin[5], r.begin[4], params, smem);
  batch_range<4, BLOCK, SC, NODAL, SF, FF>(tiles, r.begin[4], r.begin[3], params, smem);
  batch_range<3, BLOCK, SC, NODAL, SF, FF>(tiles, r.begin[3], r.begin[2], params, smem);
  batch_range<2, BLOCK, SC, NODAL, SF, FF>(tiles, r.begin[2], r.begin[1], params, smem);
  batch_range<1, BLOCK, SC, NODAL, SF, FF>(tiles, r.begin[1], r.begin[0], params, smem);
}

template <int BLOCK, bool SC, bool NODAL, int SF, int FF>
hipError_t launch_batch_t(const BatchLaunch& b, hipStream_t stream) {
  auto kern = batch_kernel<BLOCK, SC, NODAL, SF, FF>;
  if (b.probe) {
    // occupancy query for the host's grid sizing; no launch.  It also raises
    // the kernel's dynamic-LDS limit to b.lds_bytes, which every later launch
    // (including ones recorded into a graph) relies on.
    int nb = 0;
    hipError_t e = hipFuncSetAttribute((const void*)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)b.lds_bytes);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, BLOCK, b.lds_bytes);
    if (e != hipSuccess) return e;
    *b.probe = nb;
    return hipSuccess;
  }
  if (b.ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern, dim3((unsigned)b.grid), dim3(BLOCK), b.lds_bytes, stream, b.tiles,
                     b.ranges, b.params);
  return hipGetLastError();
}

template <int BLOCK, bool SC, bool NODAL>
hipError_t launch_batch_flux(const BatchLaunch& b, hipStream_t s) {
  const bool shllc = b.surface_flux == kFluxHLLC;
  if constexpr (SC) {
    const bool fhllc = b.fv_flux == kFluxHLLC;
    if (shllc)
      return fhllc ? launch_batch_t<BLOCK, SC, NODAL, kFluxHLLC, kFluxHLLC>(b, s)
                   : launch_batch_t<BLOCK, SC, NODAL, kFluxHLLC, kFluxLLF>(b, s);
    return fhllc ? launch_batch_t<BLOCK, SC, NODAL, kFluxLLF, kFluxHLLC>(b, s)
                 : launch_batch_t<BLOCK, SC, NODAL, kFluxLLF, kFluxLLF>(b, s);
  } else {
    return shllc ? launch_batch_t<BLOCK, SC, NODAL, kFluxHLLC, kFluxLLF>(b, s)
                 : launch_batch_t<BLOCK, SC, NODAL, kFluxLLF, kFluxLLF>(b, s);
  }
}

template <int BLOCK>
hipError_t launch_batch_block(const BatchLaunch& b, hipStream_t s) {
  if (b.shock_capturing)
    return b.nodal ? launch_batch_flux<BLOCK, true, true>(b, s)
                   : launch_batch_flux<BLOCK, true, false>(b, s);
  return b.nodal ? launch_batch_flux<BLOCK, false, true>(b, s)
                 : launch_batch_flux<BLOCK, false, false>(b, s);
}

}  // namespace detail
}  // namespace dgsem
