// Batched ensembles of independent 2D Euler cases (include/dgsem_amd.h,
// dgsem_batch_*).  Every case keeps its own single-case handle (geometry,
// neighbour table, basis tables, blending factor); the batch adds one state
// arena layout, a device array of per-case kernel parameters per RK stage, a
// cost-sorted tile list and a HIP graph of the three stage launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/dgsem_amd.h"
#include "handle.hpp"
#include "kernels.hpp"

struct dgsem_batch {
  std::vector<dgsem_handle*> cases;
  std::vector<int64_t> offset, count;  // doubles of each case in a state arena
  int64_t total = 0;
  hipStream_t stream = nullptr;      // caller's stream: all work is ordered on it
  hipStream_t cap_stream = nullptr;  // private stream used only to capture the graph
  int num_cu = 0;
  bool sc = false, nodal = false;
  int surface_flux = 0, fv_flux = 0;
  std::vector<double> dt;
  bool dt_set = false;
  // tuning: workgroup size, elements per tile for each N, persistent
  // workgroups (0 = as many as fit the device at once)
  int block = 256;
  int tile[8] = {0};
  int grid = 0;
  // derived launch state
  dgsem::BatchTile* d_tiles = nullptr;
  int ntiles = 0;
  dgsem::BatchRanges ranges{};  // degree ranges of the sorted tile list
  size_t lds = 0;
  int grid_eff = 0;
  bool tiles_valid = false;
  // kernel parameters: [stage 0..3][case]; stage 0 is the plain RHS
  dgsem::KParams* d_params = nullptr;
  const void* bound[4][3] = {};  // (u_k, u_out, u_base) each stage's block was built for
  bool bound_valid[4] = {false, false, false, false};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* graph_ptrs[3] = {};
  std::string err;
};

namespace {

int bfail(dgsem_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

int bhip(dgsem_batch* b, hipError_t e, const char* where) {
  return bfail(b, DGSEM_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

void drop_graph(dgsem_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Rough relative cost of a tile: the two-point fluxes dominate, N (N+1)^2
// per 2D element, plus per-node work.
double tile_cost(int N, int nvalid) {
  const double n = N + 1;
  return nvalid * (N * n * n + 4.0 * n * n);
}

dgsem::BatchLaunch base_launch(const dgsem_batch* b) {
  dgsem::BatchLaunch L{};
  L.tiles = b->d_tiles;
  L.ntiles = b->ntiles;
  L.ranges = b->ranges;
  L.grid = b->grid_eff;
  L.lds_bytes = b->lds;
  L.shock_capturing = b->sc;
  L.nodal = b->nodal;
  L.surface_flux = b->surface_flux;
  L.fv_flux = b->fv_flux;
  return L;
}

// Builds the tile list for the current tuning, uploads it and sizes the
// persistent grid.  Synchronises the stream first: kernels in flight may
// still read the old list.
int build_tiles(dgsem_batch* b) {
  std::vector<dgsem::BatchTile> tiles;
  size_t lds = 0;
  for (size_t i = 0; i < b->cases.size(); ++i) {
    const int N = b->cases[i]->cfg.polydeg;
    const int slots = dgsem::batch_slots(b->block, N);
    if (slots <= 0) return bfail(b, DGSEM_E_UNSUPPORTED, "polynomial degree not built for batches");
    const int t = std::max(1, std::min(b->tile[N] > 0 ? b->tile[N] : slots, slots));
    lds = std::max(lds, dgsem::batch_lds_bytes(b->block, N));
    for (int64_t e0 = 0; e0 < b->cases[i]->E; e0 += t) {
      dgsem::BatchTile tl;
      tl.case_id = (int32_t)i;
      tl.polydeg = (int16_t)N;
      tl.nvalid = (int16_t)std::min<int64_t>(t, b->cases[i]->E - e0);
      tl.e0 = e0;
      tiles.push_back(tl);
    }
  }
  // by degree, highest first (the kernel runs one loop per degree), then by
  // cost, so that each workgroup starts with its largest tiles
  std::stable_sort(tiles.begin(), tiles.end(), [](const dgsem::BatchTile& x, const dgsem::BatchTile& y) {
    if (x.polydeg != y.polydeg) return x.polydeg > y.polydeg;
    return tile_cost(x.polydeg, x.nvalid) > tile_cost(y.polydeg, y.nvalid);
  });
  if (tiles.size() > (size_t)INT32_MAX) return bfail(b, DGSEM_E_INVALID, "too many tiles");
  // begin[N]: the first tile of degree N or lower
  for (int N = 7; N >= 0; --N) {
    size_t k = 0;
    while (k < tiles.size() && tiles[k].polydeg > N) ++k;
    b->ranges.begin[N] = (int)k;
  }
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bhip(b, e, "hipStreamSynchronize");
  hipFree(b->d_tiles);
  b->d_tiles = nullptr;
  e = hipMalloc((void**)&b->d_tiles, std::max<size_t>(1, tiles.size()) * sizeof(dgsem::BatchTile));
  if (e != hipSuccess) return bfail(b, DGSEM_E_ALLOC, "hipMalloc failed");
  e = hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(dgsem::BatchTile),
                hipMemcpyHostToDevice);
  if (e != hipSuccess) return bhip(b, e, "hipMemcpy");
  b->ntiles = (int)tiles.size();
  b->lds = lds;
  // resident workgroups per CU for this (block, LDS); also sets the kernel's
  // dynamic-LDS limit, which the graph capture relies on
  dgsem::BatchLaunch L = base_launch(b);
  int per_cu = 0;
  L.probe = &per_cu;
  e = dgsem::launch_batch(b->block, L, b->stream);
  if (e != hipSuccess) return bhip(b, e, "batch occupancy query");
  if (per_cu < 1) return bfail(b, DGSEM_E_UNSUPPORTED, "batch kernel does not fit a CU");
  const int64_t full = (int64_t)per_cu * b->num_cu;
  int64_t g = b->grid > 0 ? b->grid : full;
  g = std::max<int64_t>(1, std::min<int64_t>(g, b->ntiles));
  b->grid_eff = (int)g;
  b->tiles_valid = true;
  drop_graph(b);
  return DGSEM_OK;
}

// Kernel parameters of `stage` for every case, for the arena pointers given.
int bind_stage(dgsem_batch* b, int stage, const double* u_k, double* u_out, const double* u_base) {
  if (b->bound_valid[stage] && b->bound[stage][0] == u_k && b->bound[stage][1] == u_out &&
      b->bound[stage][2] == u_base)
    return DGSEM_OK;
  const size_t nc = b->cases.size();
  std::vector<dgsem::KParams> ps(nc);
  for (size_t i = 0; i < nc; ++i) {
    const int64_t o = b->offset[i];
    dgsem_handle* h = b->cases[i];
    const int rc = dgsem::stage_params(h, stage, stage ? b->dt[i] : 0.0, u_base ? u_base + o : nullptr,
                                       u_k + o, u_out + o, &ps[i]);
    if (rc != DGSEM_OK) return bfail(b, rc, "case " + std::to_string(i) + ": " + h->err);
  }
  hipError_t e = hipStreamSynchronize(b->stream);  // kernels in flight may read the old block
  if (e != hipSuccess) return bhip(b, e, "hipStreamSynchronize");
  e = hipMemcpy(b->d_params + stage * nc, ps.data(), nc * sizeof(dgsem::KParams),
                hipMemcpyHostToDevice);
  if (e != hipSuccess) return bhip(b, e, "hipMemcpy");
  b->bound[stage][0] = u_k;
  b->bound[stage][1] = u_out;
  b->bound[stage][2] = u_base;
  b->bound_valid[stage] = true;
  return DGSEM_OK;
}

int launch_stage(dgsem_batch* b, int stage, hipStream_t s) {
  dgsem::BatchLaunch L = base_launch(b);
  L.params = b->d_params + stage * b->cases.size();
  hipError_t e = dgsem::launch_batch(b->block, L, s);
  return e == hipSuccess ? DGSEM_OK : bhip(b, e, "batch kernel launch");
}

// The batch kernel is built for the Chandrashekar volume flux with LLF or
// HLLC surface and subcell fluxes.
int check_case(const dgsem_config& c0, const dgsem_config& c) {
  return c.dim == 2 && c.equation == DGSEM_EULER && c.volume_integral == c0.volume_integral &&
         c.surface_flux == c0.surface_flux && c.fv_flux == c0.fv_flux && c.geometry == c0.geometry &&
         c.volume_flux == DGSEM_FLUX_CHANDRASHEKAR && !dgsem::uses_central_flux(c);
}

}  // namespace

extern "C" {

int dgsem_batch_create(const dgsem_config* cfgs, int ncases, void* stream, dgsem_batch** out) {
  if (!cfgs || !out || ncases < 1) return DGSEM_E_INVALID;
  *out = nullptr;
  dgsem_batch* b = new (std::nothrow) dgsem_batch();
  if (!b) return DGSEM_E_ALLOC;
  *out = b;
  b->stream = (hipStream_t)stream;
  for (int i = 0; i < ncases; ++i) {
    if (!check_case(cfgs[0], cfgs[i]))
      return bfail(b, DGSEM_E_UNSUPPORTED,
                   "batch cases must be 2D Euler with one volume integral, flux pair and geometry "
                   "mode, the Chandrashekar volume flux and LLF or HLLC fluxes (case " +
                       std::to_string(i) + ")");
    if (dgsem::batch_slots(256, cfgs[i].polydeg) <= 0)
      return bfail(b, DGSEM_E_UNSUPPORTED, "polynomial degree not built for batches");
  }
  b->sc = cfgs[0].volume_integral == DGSEM_VOLUME_SHOCK_CAPTURING;
  b->nodal = cfgs[0].geometry == DGSEM_GEOMETRY_NODAL;
  b->surface_flux = cfgs[0].surface_flux;
  b->fv_flux = cfgs[0].fv_flux;
  for (int i = 0; i < ncases; ++i) {
    dgsem_handle* h = nullptr;
    const int rc = dgsem_create(&cfgs[i], stream, &h);
    if (rc != DGSEM_OK) {
      std::string msg = "case " + std::to_string(i) + ": " + (h ? h->err : std::string("create"));
      dgsem_destroy(h);
      return bfail(b, rc, msg);
    }
    b->cases.push_back(h);
    // the batch kernel reads every face neighbour: periodic meshes only
    for (int f = 0; f < 2 * h->cfg.dim; ++f)
      if (h->cfg.face_boundary[f] != DGSEM_BC_PERIODIC)
        return bfail(b, DGSEM_E_UNSUPPORTED,
                     "batch cases must be periodic (case " + std::to_string(i) + ")");
    const int64_t cnt = h->E * h->nvars * h->nn;
    b->offset.push_back(b->total);
    b->count.push_back(cnt);
    b->total += cnt;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return bhip(b, e, "device query");
  e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bhip(b, e, "hipStreamCreateWithFlags");
  e = hipMalloc((void**)&b->d_params, 4 * (size_t)ncases * sizeof(dgsem::KParams));
  if (e != hipSuccess) return bfail(b, DGSEM_E_ALLOC, "hipMalloc failed");
  b->dt.assign(ncases, 0.0);
  return build_tiles(b);
}

int dgsem_batch_destroy(dgsem_batch* b) {
  if (!b) return DGSEM_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  else hipDeviceSynchronize();
  drop_graph(b);
  if (b->cap_stream) hipStreamDestroy(b->cap_stream);
  hipFree(b->d_tiles);
  hipFree(b->d_params);
  for (dgsem_handle* h : b->cases) dgsem_destroy(h);
  delete b;
  return DGSEM_OK;
}

const char* dgsem_batch_last_error(const dgsem_batch* b) { return b ? b->err.c_str() : "null batch"; }

int dgsem_batch_sizes(const dgsem_batch* b, int* ncases, int64_t* total) {
  if (!b || b->cases.empty()) return DGSEM_E_INVALID;
  if (ncases) *ncases = (int)b->cases.size();
  if (total) *total = b->total;
  return DGSEM_OK;
}

int dgsem_batch_case(dgsem_batch* b, int i, int64_t* offset, int64_t* count, dgsem_handle** h) {
  if (!b || i < 0 || i >= (int)b->cases.size()) return DGSEM_E_INVALID;
  if (offset) *offset = b->offset[i];
  if (count) *count = b->count[i];
  if (h) *h = b->cases[i];
  return DGSEM_OK;
}

int dgsem_batch_set_dt(dgsem_batch* b, const double* dt) {
  if (!b || !dt) return DGSEM_E_INVALID;
  for (size_t i = 0; i < b->cases.size(); ++i)
    if (!(dt[i] >= 0.0)) return bfail(b, DGSEM_E_INVALID, "dt must be non-negative");
  b->dt.assign(dt, dt + b->cases.size());
  b->dt_set = true;
  for (int s = 1; s < 4; ++s) b->bound_valid[s] = false;
  drop_graph(b);
  return DGSEM_OK;
}

int dgsem_batch_rhs(dgsem_batch* b, const double* u, double* du) {
  if (!b || b->cases.empty()) return DGSEM_E_INVALID;
  if (!u || !du || u == du) return bfail(b, DGSEM_E_INVALID, "u and du must be distinct device arrays");
  int rc;
  if (!b->tiles_valid && (rc = build_tiles(b))) return rc;
  if ((rc = bind_stage(b, 0, u, du, nullptr))) return rc;
  return launch_stage(b, 0, b->stream);
}

int dgsem_batch_ssprk3_steps(dgsem_batch* b, double* u, int nsteps, double* tmp1, double* tmp2) {
  if (!b || b->cases.empty()) return DGSEM_E_INVALID;
  if (!u || !tmp1 || !tmp2 || u == tmp1 || u == tmp2 || tmp1 == tmp2)
    return bfail(b, DGSEM_E_INVALID, "u, tmp1, tmp2 must be distinct device arrays");
  if (!b->dt_set) return bfail(b, DGSEM_E_INVALID, "call dgsem_batch_set_dt first");
  if (nsteps < 0) return bfail(b, DGSEM_E_INVALID, "nsteps must be non-negative");
  int rc;
  if (!b->tiles_valid && (rc = build_tiles(b))) return rc;
  // the stage sequence of SSPRK3::step, SSPRK3.hpp:159-181
  if ((rc = bind_stage(b, 1, u, tmp1, nullptr)) || (rc = bind_stage(b, 2, tmp1, tmp2, u)) ||
      (rc = bind_stage(b, 3, tmp2, u, u)))
    return rc;
  if (b->exec && (b->graph_ptrs[0] != u || b->graph_ptrs[1] != tmp1 || b->graph_ptrs[2] != tmp2))
    drop_graph(b);
  if (!b->exec) {
    hipError_t e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bhip(b, e, "hipStreamBeginCapture");
    for (int s = 1; s <= 3 && rc == DGSEM_OK; ++s) rc = launch_stage(b, s, b->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc != DGSEM_OK) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return bhip(b, e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      return bhip(b, e, "hipGraphInstantiate");
    }
    b->graph_ptrs[0] = u;
    b->graph_ptrs[1] = tmp1;
    b->graph_ptrs[2] = tmp2;
  }
  for (int s = 0; s < nsteps; ++s) {
    hipError_t e = hipGraphLaunch(b->exec, b->stream);
    if (e != hipSuccess) return bhip(b, e, "hipGraphLaunch");
  }
  return DGSEM_OK;
}

int dgsem_batch_get_tuning(const dgsem_batch* b, int* block, int* tile_elems, int* grid) {
  if (!b) return DGSEM_E_INVALID;
  if (block) *block = b->block;
  if (tile_elems)
    for (int N = 0; N < 8; ++N) {
      const int slots = dgsem::batch_slots(b->block, N);
      tile_elems[N] = slots > 0 ? std::max(1, std::min(b->tile[N] > 0 ? b->tile[N] : slots, slots)) : 0;
    }
  if (grid) *grid = b->grid_eff;
  return DGSEM_OK;
}

int dgsem_batch_set_tuning(dgsem_batch* b, int block, const int* tile_elems, int grid) {
  if (!b) return DGSEM_E_INVALID;
  if (dgsem::batch_slots(block, 1) <= 0) return bfail(b, DGSEM_E_INVALID, "workgroup size not built");
  if (grid < 0) return bfail(b, DGSEM_E_INVALID, "grid must be >= 0");
  b->block = block;
  for (int N = 0; N < 8; ++N) b->tile[N] = tile_elems ? std::max(0, tile_elems[N]) : 0;
  b->grid = grid;
  return build_tiles(b);
}

// Coordinate search over the workgroup size, then each degree's tile size,
// then the persistent grid, timing the batch right-hand side u -> scratch.
int dgsem_batch_autotune(dgsem_batch* b, const double* u, double* scratch, int reps) {
  if (!b || !u || !scratch || u == scratch || reps < 1) return DGSEM_E_INVALID;
  hipEvent_t t0, t1;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) return bhip(b, e, "hipEventCreate");
  auto measure = [&](float* ms) -> int {
    int rc = dgsem_batch_rhs(b, u, scratch);  // warm-up, also builds tiles / params
    if (rc) return rc;
    hipError_t er = hipEventRecord(t0, b->stream);
    for (int r = 0; r < reps && er == hipSuccess; ++r) {
      rc = launch_stage(b, 0, b->stream);
      if (rc) return rc;
    }
    if (er == hipSuccess) er = hipEventRecord(t1, b->stream);
    if (er == hipSuccess) er = hipEventSynchronize(t1);
    if (er == hipSuccess) er = hipEventElapsedTime(ms, t0, t1);
    return er == hipSuccess ? DGSEM_OK : bhip(b, er, "timing");
  };
  int rc = DGSEM_OK;
  int best_block = b->block, best_tile[8];
  std::memcpy(best_tile, b->tile, sizeof(best_tile));
  float best = 1e30f;
  for (int blk : dgsem::kBatchBlocks) {
    int zero[8] = {0};
    if ((rc = dgsem_batch_set_tuning(b, blk, zero, 0))) break;
    float ms = 0;
    if ((rc = measure(&ms))) break;
    if (ms < best) {
      best = ms;
      best_block = blk;
      std::memcpy(best_tile, zero, sizeof(best_tile));
    }
  }
  std::vector<int> degrees;
  for (dgsem_handle* h : b->cases)
    if (std::find(degrees.begin(), degrees.end(), h->cfg.polydeg) == degrees.end())
      degrees.push_back(h->cfg.polydeg);
  for (size_t di = 0; di < degrees.size() && rc == DGSEM_OK; ++di) {
    const int N = degrees[di];
    const int slots = dgsem::batch_slots(best_block, N);
    for (int div : {2, 4}) {
      const int t = slots / div;
      if (t < 1) break;
      int trial[8];
      std::memcpy(trial, best_tile, sizeof(trial));
      trial[N] = t;
      if ((rc = dgsem_batch_set_tuning(b, best_block, trial, 0))) break;
      float ms = 0;
      if ((rc = measure(&ms))) break;
      if (ms < best) {
        best = ms;
        std::memcpy(best_tile, trial, sizeof(best_tile));
      }
    }
  }
  int best_grid = 0;
  if (rc == DGSEM_OK) {
    if ((rc = dgsem_batch_set_tuning(b, best_block, best_tile, 0)) == DGSEM_OK) {
      const int full = b->grid_eff;
      for (int g : {full / 2, full * 2}) {
        if (g < 1) continue;
        if ((rc = dgsem_batch_set_tuning(b, best_block, best_tile, g))) break;
        float ms = 0;
        if ((rc = measure(&ms))) break;
        if (ms < best) {
          best = ms;
          best_grid = g;
        }
      }
    }
  }
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (rc != DGSEM_OK) return rc;
  return dgsem_batch_set_tuning(b, best_block, best_tile, best_grid);
}

}  // extern "C"
