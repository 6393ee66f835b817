"""Pins of the VTU oracle (oracle/vtu.py): VTK's Lagrange point numbering,
checked against the orders VTK documents for its fixed-order cells, and by
the numbering's published rules at every order; and a file round trip.

* order 1: the linear VTK_QUAD / VTK_HEXAHEDRON vertex order;
* order 2 quadrilateral: VTK_BIQUADRATIC_QUAD's documented order (vertices,
  mid-edge nodes of (0,1), (1,2), (2,3), (3,0), centre);
* order 2 hexahedron: VTK_QUADRATIC_HEXAHEDRON's documented order for the
  20 vertex and mid-edge nodes (edges (0,1), (1,2), (2,3), (3,0), (4,5),
  (5,6), (6,7), (7,4), (0,4), (1,5), (2,6), (3,7)) -- the file version 2.2
  numbering of Lagrange hexahedra agrees with it;
* every order: a permutation; vertices, edges, faces and interior in
  contiguous groups; nodes along each edge in increasing parameter.
"""
import numpy as np
import pytest

from oracle import vtu


def ijk(dim, N):
    n = N + 1
    nn = n ** dim
    return [(d % n, (d // n) % n, d // (n * n)) for d in range(nn)]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("N", range(1, 8))
def test_permutation_is_a_bijection(dim, N):
    p = vtu.permutation(dim, N)
    assert sorted(p) == list(range((N + 1) ** dim))


def test_order_one_is_the_linear_cell():
    assert list(vtu.permutation(2, 1)) == [0, 1, 3, 2]  # dofs (0,0),(1,0),(0,1),(1,1)
    assert list(vtu.permutation(3, 1)) == [0, 1, 3, 2, 4, 5, 7, 6]


def test_biquadratic_quad():
    order = vtu.lagrange_quad_order(2)
    expect = {(0, 0): 0, (2, 0): 1, (2, 2): 2, (0, 2): 3,
              (1, 0): 4, (2, 1): 5, (1, 2): 6, (0, 1): 7, (1, 1): 8}
    assert order == expect


def test_quadratic_hexahedron_vertices_and_edges():
    order = vtu.lagrange_hex_order(2)
    verts = [(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 2), (2, 0, 2), (2, 2, 2), (0, 2, 2)]
    edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
             (0, 4), (1, 5), (2, 6), (3, 7)]
    for p, v in enumerate(verts):
        assert order[v] == p
    for p, (a, b) in enumerate(edges):
        mid = tuple((x + y) // 2 for x, y in zip(verts[a], verts[b]))
        assert order[mid] == 8 + p, (a, b)


@pytest.mark.parametrize("N", [3, 4, 7])
def test_groups_and_edge_direction(N):
    m = N - 1
    quad = vtu.lagrange_quad_order(N)
    on = lambda c: sum(x in (0, N) for x in c)  # noqa: E731
    for c, p in quad.items():
        group = 0 if on(c) == 2 else (1 if on(c) == 1 else 2)
        lo = [0, 4, 4 + 4 * m][group]
        hi = [4, 4 + 4 * m, (N + 1) ** 2][group]
        assert lo <= p < hi
    # along every edge the index grows with the running coordinate
    for j in (0, N):
        assert [quad[(i, j)] for i in range(1, N)] == sorted(quad[(i, j)] for i in range(1, N))
    hexo = vtu.lagrange_hex_order(N)
    bounds = [(0, 8), (8, 8 + 12 * m), (8 + 12 * m, 8 + 12 * m + 6 * m * m),
              (8 + 12 * m + 6 * m * m, (N + 1) ** 3)]
    for c, p in hexo.items():
        lo, hi = bounds[3 - on(c)]
        assert lo <= p < hi
    # faces in the order x-, x+, y-, y+, z-, z+
    base = 8 + 12 * m
    face_of = []
    for c, p in hexo.items():
        if on(c) == 1:
            axis = [x in (0, N) for x in c].index(True)
            face_of.append(((p - base) // (m * m), 2 * axis + (c[axis] == N)))
    assert all(a == b for a, b in face_of)


def test_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    N, cells = 2, (3, 2)
    E, nn = 6, 9
    u = rng.random((E, nn, 4)) + 1.0
    coords = rng.random((E, nn, 2))
    arrays = vtu.expected_arrays(2, N, cells, coords, u)
    path = tmp_path / "a.vtu"
    vtu.write_vtu(path, arrays)
    back = vtu.read_vtu(path)
    assert back["version"] == "2.2"
    assert back["points_count"] == E * nn and back["cells_count"] == E
    for k, v in arrays["point_data"].items():
        assert np.array_equal(back["point_data"][k], v)
    assert np.array_equal(back["points"], arrays["points"])
    for k in ("connectivity", "offsets", "types"):
        assert np.array_equal(back[k], arrays[k])
