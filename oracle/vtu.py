"""Oracle for the VTU output (test infrastructure only: used by tests/).

Restates what the reference's VTUOutputObserver writes
(src/observer/vtu_output_observer.hpp:62-214 and the VTK XML writer calls of
vtu_output_observer_impl.cpp:15-117) as numpy arrays, and reads such a file
back.  The point order inside a cell is VTK's Lagrange numbering
(vtkLagrangeQuadrilateral / vtkLagrangeHexahedron::PointIndexFromIJK, which
the reference calls at vtu_output_observer.hpp:108-117, :186-197).  VTK is
not vendored in the reference (CMakeLists.txt:80, find_package(VTK
REQUIRED), no version pinned) nor installed here; the numbering restated
below is VTK's published one for XML file version 2.2 (VTK >= 9.1): vertices,
edges, faces, interior, each group in increasing parametric order.
"""
import xml.etree.ElementTree as ET

import numpy as np

OUTPUT_GAMMA = 1.4  # vtu_output_observer.hpp:127, :207


def lagrange_quad_order(N):
    """VTK point index of every (i, j), as a dict: vertices counter-clockwise
    from (0, 0), the edges (0,1) bottom, (1,2) right, (3,2) top, (0,3) left
    in increasing i or j, then the interior with i fastest."""
    idx = {(0, 0): 0, (N, 0): 1, (N, N): 2, (0, N): 3}
    pos = 4
    inner = range(1, N)
    for i in inner:
        idx[(i, 0)] = pos; pos += 1
    for j in inner:
        idx[(N, j)] = pos; pos += 1
    for i in inner:
        idx[(i, N)] = pos; pos += 1
    for j in inner:
        idx[(0, j)] = pos; pos += 1
    for j in inner:
        for i in inner:
            idx[(i, j)] = pos; pos += 1
    assert pos == (N + 1) ** 2
    return idx


def lagrange_hex_order(N):
    """VTK point index of every (i, j, k), as a dict (file version 2.2)."""
    idx = {}
    corners = [(0, 0, 0), (N, 0, 0), (N, N, 0), (0, N, 0),
               (0, 0, N), (N, 0, N), (N, N, N), (0, N, N)]
    for p, c in enumerate(corners):
        idx[c] = p
    pos = 8
    inner = range(1, N)
    # edges (0,1), (1,2), (3,2), (0,3) at k = 0, the same at k = N
    for k in (0, N):
        for i in inner:
            idx[(i, 0, k)] = pos; pos += 1
        for j in inner:
            idx[(N, j, k)] = pos; pos += 1
        for i in inner:
            idx[(i, N, k)] = pos; pos += 1
        for j in inner:
            idx[(0, j, k)] = pos; pos += 1
    # vertical edges above vertices 0, 1, 2, 3
    for (i, j) in ((0, 0), (N, 0), (N, N), (0, N)):
        for k in inner:
            idx[(i, j, k)] = pos; pos += 1
    # faces x-, x+ (j fastest, then k); y-, y+ (i, then k); z-, z+ (i, then j)
    for i in (0, N):
        for k in inner:
            for j in inner:
                idx[(i, j, k)] = pos; pos += 1
    for j in (0, N):
        for k in inner:
            for i in inner:
                idx[(i, j, k)] = pos; pos += 1
    for k in (0, N):
        for j in inner:
            for i in inner:
                idx[(i, j, k)] = pos; pos += 1
    for k in inner:
        for j in inner:
            for i in inner:
                idx[(i, j, k)] = pos; pos += 1
    assert pos == (N + 1) ** 3
    return idx


def permutation(dim, N):
    """VTK position of every dof, dof = (k * n + j) * n + i."""
    n = N + 1
    if dim == 2:
        order = lagrange_quad_order(N)
        return np.array([order[(dof % n, dof // n)] for dof in range(n * n)])
    order = lagrange_hex_order(N)
    return np.array([order[(dof % n, (dof // n) % n, dof // (n * n))] for dof in range(n ** 3)])


def expected_arrays(dim, N, cells, coords, u):
    """Arrays of the file the reference writes for state u (reference layout
    [E][dof][var], element (ie * ny + je) * nz + ke) at node coordinates
    coords [E][dof][dim]: dict of point-data arrays, points [P][3],
    connectivity, offsets, types."""
    nn = (N + 1) ** dim
    E = u.shape[0]
    nx, ny = cells[0], cells[1]
    nz = cells[2] if dim == 3 else 1
    perm = permutation(dim, N)
    e = np.arange(E)
    if dim == 2:
        ie, je = e // ny, e % ny
        cell = je * nx + ie
    else:
        ie, je, ke = e // (ny * nz), (e // nz) % ny, e % nz
        cell = (ke * ny + je) * nx + ie
    g = (cell[:, None] * nn + perm[None, :]).ravel()
    P = E * nn
    names = ["rho", "rhou", "rhov", "rhoE"] if dim == 2 else ["rho", "rhou", "rhov", "rhow", "rhoE"]
    data = {}
    flat = u.reshape(P, dim + 2)
    for v, name in enumerate(names):
        arr = np.empty(P)
        arr[g] = flat[:, v]
        data[name] = arr
    # cons_to_prim (state_conversion.hpp:31-47) with gamma 1.4
    vsq = np.zeros(P)
    for d in range(dim):
        vel = flat[:, d + 1] / flat[:, 0]
        vsq = vsq + vel * vel
    p = np.empty(P)
    p[g] = (OUTPUT_GAMMA - 1.0) * (flat[:, dim + 1] - 0.5 * flat[:, 0] * vsq)
    data["pressure"] = p
    pts = np.zeros((P, 3))
    pts[g, :dim] = coords.reshape(P, dim)
    return {"point_data": data, "points": pts, "connectivity": np.arange(P),
            "offsets": (np.arange(E) + 1) * nn,
            "types": np.full(E, 70 if dim == 2 else 72, dtype=np.uint8)}


_DTYPES = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1", "Int32": "<i4", "UInt64": "<u8"}


def read_vtu(path):
    """Reads an appended-raw VTK XML unstructured grid (UInt64 headers)."""
    raw = open(path, "rb").read()
    start = raw.index(b"<AppendedData")
    underscore = raw.index(b"_", start) + 1
    head = raw[:start].decode() + "</VTKFile>"
    root = ET.fromstring(head)
    assert root.get("type") == "UnstructuredGrid" and root.get("header_type") == "UInt64"
    assert root.get("byte_order") == "LittleEndian"
    piece = root.find("UnstructuredGrid/Piece")

    def block(arr):
        off = underscore + int(arr.get("offset"))
        nbytes = int(np.frombuffer(raw, "<u8", 1, off)[0])
        a = np.frombuffer(raw, _DTYPES[arr.get("type")], nbytes // np.dtype(_DTYPES[arr.get("type")]).itemsize,
                          off + 8)
        comps = int(arr.get("NumberOfComponents", "1"))
        return a.reshape(-1, comps) if comps > 1 else a

    out = {"version": root.get("version"),
           "points_count": int(piece.get("NumberOfPoints")),
           "cells_count": int(piece.get("NumberOfCells")),
           "point_data": {a.get("Name"): block(a) for a in piece.find("PointData")},
           "points": block(piece.find("Points/DataArray"))}
    for a in piece.find("Cells"):
        out[a.get("Name")] = block(a)
    return out


def write_vtu(path, arrays):
    """Writes arrays as returned by expected_arrays in the same file format
    (appended raw, UInt64 headers), for round-trip checks of read_vtu."""
    P = len(arrays["connectivity"])
    C = len(arrays["types"])
    parts, offset, lines = [], 0, []

    def add(type_, name, a, comps=1):
        nonlocal offset
        b = np.ascontiguousarray(a, dtype=_DTYPES[type_]).tobytes()
        extra = f' NumberOfComponents="{comps}"' if comps > 1 else ""
        lines.append(f'<DataArray type="{type_}" Name="{name}"{extra} format="appended" offset="{offset}"/>')
        parts.append(np.uint64(len(b)).tobytes() + b)
        offset += 8 + len(b)

    for name, a in arrays["point_data"].items():
        add("Float64", name, a)
    pd, lines = lines, []
    add("Float64", "Points", arrays["points"], 3)
    pts, lines = lines, []
    add("Int64", "connectivity", arrays["connectivity"])
    add("Int64", "offsets", arrays["offsets"])
    add("UInt8", "types", arrays["types"])
    cells = lines
    xml = ('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="2.2" '
           'byte_order="LittleEndian" header_type="UInt64">\n<UnstructuredGrid>\n'
           f'<Piece NumberOfPoints="{P}" NumberOfCells="{C}">\n<PointData>\n' + "\n".join(pd) +
           '\n</PointData>\n<CellData>\n</CellData>\n<Points>\n' + "\n".join(pts) +
           '\n</Points>\n<Cells>\n' + "\n".join(cells) +
           '\n</Cells>\n</Piece>\n</UnstructuredGrid>\n<AppendedData encoding="raw">\n_')
    with open(path, "wb") as f:
        f.write(xml.encode())
        for p in parts:
            f.write(p)
        f.write(b"\n</AppendedData>\n</VTKFile>\n")
