// VTK XML unstructured-grid output of a state as Lagrange cells: the file
// VTUOutputObserver writes (vtu_output_observer.hpp:62-214) through VTK's
// XML writer (vtu_output_observer_impl.cpp:15-117), produced here without
// the VTK library:
//   * cells in the reference's output order, x fastest (je * nx + ie in 2D,
//     (ke * ny + je) * nx + ie in 3D, vtu_output_observer.hpp:103-106,
//     :180-184), each with its (N+1)^d points in VTK's Lagrange order;
//   * point data rho, momenta, rhoE and the pressure from cons_to_prim with
//     gamma fixed at 1.4, as the reference computes it
//     (vtu_output_observer.hpp:125-129, :205-209; state_conversion.hpp:31-47);
//   * appended raw binary blocks, each preceded by its UInt64 byte count
//     (the writer's SetDataModeToAppended + EncodeAppendedDataOff).
// The state is reordered on the GPU (one thread per node) and copied to the
// host one variable at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dgsem_amd.h"
#include "handle.hpp"

namespace {

constexpr double kOutputGamma = 1.4;  // the reference's output pressure
constexpr uint8_t kVtkLagrangeQuadrilateral = 70;
constexpr uint8_t kVtkLagrangeHexahedron = 72;

// Position of node (i, j) of an order-N Lagrange quadrilateral in VTK's
// point order: the 4 vertices counter-clockwise from (0, 0); the edge nodes
// of the edges (0,1), (1,2), (3,2), (0,3), each running in increasing i or
// j; then the interior nodes with i fastest.
int vtk_quad_index(int i, int j, int N) {
  const bool ib = i == 0 || i == N, jb = j == 0 || j == N;
  if (ib && jb) return i ? (j ? 2 : 1) : (j ? 3 : 0);
  if (jb) return 4 + (i - 1) + (j ? 2 * (N - 1) : 0);             // edges along i
  if (ib) return 4 + (j - 1) + (i ? (N - 1) : 3 * (N - 1));       // edges along j
  return 4 + 4 * (N - 1) + (i - 1) + (N - 1) * (j - 1);
}

// Position of node (i, j, k) of an order-N Lagrange hexahedron in VTK's
// point order (XML file version 2.2 numbering): the 8 vertices (the bottom
// quad, then the top one); the 12 edges (0,1), (1,2), (3,2), (0,3),
// (4,5), (5,6), (7,6), (4,7), (0,4), (1,5), (2,6), (3,7), each in increasing
// parametric coordinate; the faces x-, x+, y-, y+, z-, z+, each with its
// lower remaining axis fastest; the interior with i fastest, then j, k.
int vtk_hex_index(int i, int j, int k, int N) {
  const int m = N - 1;
  const bool ib = i == 0 || i == N, jb = j == 0 || j == N, kb = k == 0 || k == N;
  const int nb = (int)ib + (int)jb + (int)kb;
  if (nb == 3) return (i ? (j ? 2 : 1) : (j ? 3 : 0)) + (k ? 4 : 0);
  int off = 8;
  if (nb == 2) {
    if (!ib) return off + (i - 1) + (j ? 2 * m : 0) + (k ? 4 * m : 0);
    if (!jb) return off + (j - 1) + (i ? m : 3 * m) + (k ? 4 * m : 0);
    return off + 8 * m + (k - 1) + m * (i ? (j ? 2 : 1) : (j ? 3 : 0));
  }
  off += 12 * m;
  if (nb == 1) {
    if (ib) return off + (j - 1) + m * (k - 1) + (i ? m * m : 0);
    off += 2 * m * m;
    if (jb) return off + (i - 1) + m * (k - 1) + (j ? m * m : 0);
    off += 2 * m * m;
    return off + (i - 1) + m * (j - 1) + (k ? m * m : 0);
  }
  off += 6 * m * m;
  return off + (i - 1) + m * ((j - 1) + m * (k - 1));
}

// VTK point position of every dof (dof = (k * n + j) * n + i).
std::vector<int> vtk_permutation(int dim, int N) {
  const int n = N + 1, nn = dim == 2 ? n * n : n * n * n;
  std::vector<int> perm(nn);
  for (int dof = 0; dof < nn; ++dof) {
    const int i = dof % n, j = (dof / n) % n, k = dof / (n * n);
    perm[dof] = dim == 2 ? vtk_quad_index(i, j, N) : vtk_hex_index(i, j, k, N);
  }
  return perm;
}

// Output cell of solver element e (element index (ie * ny + je) * nz + ke).
__host__ __device__ inline int64_t output_cell(int64_t e, int dim, int nx, int ny, int nz) {
  if (dim == 2) {
    const int64_t ie = e / ny, je = e % ny;
    return je * nx + ie;
  }
  const int64_t ie = e / ((int64_t)ny * nz), je = (e / nz) % ny, ke = e % nz;
  return (ke * ny + je) * nx + ie;
}

// out[v][point] for v < NV (conservative variables) and out[NV][point]
// (pressure), points in output order.
__global__ void vtu_pack_kernel(const double* __restrict__ u, double* __restrict__ out,
                                const int* __restrict__ perm, int dim, int64_t E, int nn, int nv,
                                int nx, int ny, int nz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * nn) return;
  const int64_t e = t / nn;
  const int dof = (int)(t - e * nn);
  const int64_t P = E * nn;
  const int64_t g = output_cell(e, dim, nx, ny, nz) * nn + perm[dof];
  double s[5];
  for (int v = 0; v < nv; ++v) {
    s[v] = u[(e * nv + v) * nn + dof];
    out[v * P + g] = s[v];
  }
  // cons_to_prim, state_conversion.hpp:31-47
  double vsq = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double vel = s[d + 1] / s[0];
    vsq += vel * vel;
  }
  out[nv * P + g] = (kOutputGamma - 1.0) * (s[nv - 1] - 0.5 * s[0] * vsq);
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool put(FILE* f, const void* p, size_t bytes) { return std::fwrite(p, 1, bytes, f) == bytes; }

int vfail(dgsem_handle* h, int code, const std::string& msg) {
  h->err = msg;
  return code;
}

}  // namespace

extern "C" int dgsem_write_vtu(dgsem_handle* h, const double* u, const char* path) {
  if (!h || !h->d_Ja || !u || !path) return DGSEM_E_INVALID;
  const int dim = h->cfg.dim, N = h->cfg.polydeg, nn = h->nn, nv = h->nvars;
  const int nx = h->cfg.cells[0], ny = h->cfg.cells[1], nz = dim == 3 ? h->cfg.cells[2] : 1;
  const int64_t E = h->E, P = E * nn;
  const int nout = nv + 1;
  const std::vector<int> perm = vtk_permutation(dim, N);

  // reorder on the device
  int* d_perm = nullptr;
  double* d_out = nullptr;
  hipError_t err = hipMalloc((void**)&d_perm, perm.size() * sizeof(int));
  if (err == hipSuccess) err = hipMalloc((void**)&d_out, (size_t)nout * P * sizeof(double));
  if (err != hipSuccess) {
    hipFree(d_perm);
    return vfail(h, DGSEM_E_ALLOC, "vtu: device allocation failed");
  }
  struct Free {
    int* a;
    double* b;
    ~Free() {
      hipFree(a);
      hipFree(b);
    }
  } release{d_perm, d_out};
  err = hipMemcpyAsync(d_perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, h->stream);
  if (err != hipSuccess) return vfail(h, DGSEM_E_HIP, "vtu: permutation upload failed");
  const int block = 256;
  const int64_t grid = (P + block - 1) / block;
  if (grid > INT32_MAX) return vfail(h, DGSEM_E_INVALID, "vtu: mesh too large");
  hipLaunchKernelGGL(vtu_pack_kernel, dim3((unsigned)grid), dim3(block), 0, h->stream, u, d_out,
                     d_perm, dim, E, nn, nv, nx, ny, nz);
  err = hipGetLastError();
  if (err != hipSuccess) return vfail(h, DGSEM_E_HIP, "vtu: pack kernel launch failed");

  // header
  static const char* names2[] = {"rho", "rhou", "rhov", "rhoE", "pressure"};
  static const char* names3[] = {"rho", "rhou", "rhov", "rhow", "rhoE", "pressure"};
  const char** names = dim == 2 ? names2 : names3;
  const int64_t C = E;
  const uint64_t bytes_var = (uint64_t)P * sizeof(double);
  const uint64_t bytes_pts = 3 * bytes_var;
  const uint64_t bytes_conn = (uint64_t)P * sizeof(int64_t);
  const uint64_t bytes_offs = (uint64_t)C * sizeof(int64_t);
  const uint64_t bytes_type = (uint64_t)C;
  std::string xml =
      "<?xml version=\"1.0\"?>\n"
      "<VTKFile type=\"UnstructuredGrid\" version=\"2.2\" byte_order=\"LittleEndian\" "
      "header_type=\"UInt64\">\n  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"" +
      std::to_string(P) + "\" NumberOfCells=\"" + std::to_string(C) + "\">\n      <PointData>\n";
  uint64_t offset = 0;
  auto array = [&](const char* type, const char* name, int comps, uint64_t bytes) {
    xml += "        <DataArray type=\"" + std::string(type) + "\" Name=\"" + name + "\"";
    if (comps > 1) xml += " NumberOfComponents=\"" + std::to_string(comps) + "\"";
    xml += " format=\"appended\" offset=\"" + std::to_string(offset) + "\"/>\n";
    offset += sizeof(uint64_t) + bytes;
  };
  for (int v = 0; v < nout; ++v) array("Float64", names[v], 1, bytes_var);
  xml += "      </PointData>\n      <CellData>\n      </CellData>\n      <Points>\n";
  array("Float64", "Points", 3, bytes_pts);
  xml += "      </Points>\n      <Cells>\n";
  array("Int64", "connectivity", 1, bytes_conn);
  array("Int64", "offsets", 1, bytes_offs);
  array("UInt8", "types", 1, bytes_type);
  xml += "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";

  File file;
  file.f = std::fopen(path, "wb");
  if (!file.f) return vfail(h, DGSEM_E_INVALID, std::string("vtu: cannot open ") + path);
  bool ok = put(file.f, xml.data(), xml.size());

  // point data, one variable at a time through a host buffer
  std::vector<double> buf((size_t)P * 3);
  for (int v = 0; v < nout && ok; ++v) {
    err = hipMemcpyAsync(buf.data(), d_out + (size_t)v * P, bytes_var, hipMemcpyDeviceToHost, h->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(h->stream);
    if (err != hipSuccess) return vfail(h, DGSEM_E_HIP, "vtu: device to host copy failed");
    ok = put(file.f, &bytes_var, sizeof(bytes_var)) && put(file.f, buf.data(), bytes_var);
  }
  // points in output order (z = 0 in 2D)
  if (ok) {
    const auto& X = h->geo.coords;  // [E][nn][dim]
    for (int64_t e = 0; e < E; ++e) {
      const int64_t c = output_cell(e, dim, nx, ny, nz);
      for (int dof = 0; dof < nn; ++dof) {
        const int64_t g = c * nn + perm[dof];
        for (int d = 0; d < 3; ++d) buf[3 * g + d] = d < dim ? X[(e * nn + dof) * dim + d] : 0.0;
      }
    }
    ok = put(file.f, &bytes_pts, sizeof(bytes_pts)) && put(file.f, buf.data(), bytes_pts);
  }
  // cells: each cell owns its nn consecutive points
  if (ok) {
    std::vector<int64_t> ids((size_t)std::max<int64_t>(P, C));
    for (int64_t i = 0; i < P; ++i) ids[i] = i;
    ok = put(file.f, &bytes_conn, sizeof(bytes_conn)) && put(file.f, ids.data(), bytes_conn);
    for (int64_t c = 0; c < C && ok; ++c) ids[c] = (c + 1) * nn;
    ok = ok && put(file.f, &bytes_offs, sizeof(bytes_offs)) && put(file.f, ids.data(), bytes_offs);
    std::vector<uint8_t> types((size_t)C, dim == 2 ? kVtkLagrangeQuadrilateral : kVtkLagrangeHexahedron);
    ok = ok && put(file.f, &bytes_type, sizeof(bytes_type)) && put(file.f, types.data(), bytes_type);
  }
  static const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";
  ok = ok && put(file.f, tail, sizeof(tail) - 1);
  if (!ok) return vfail(h, DGSEM_E_INVALID, std::string("vtu: write failed: ") + path);
  return DGSEM_OK;
}
