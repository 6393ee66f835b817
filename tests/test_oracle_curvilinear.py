"""Pins of the oracle's geometry for mapped (curvilinear) meshes
(orc_set_coordinates, the reference's StructuredElementInitializer with a
user Mapping):

* the linear mapping through set_coordinates reproduces the built-in
  geometry bit for bit;
* a polynomial mapping of low degree is represented exactly: the Jacobian
  matrix and the metric terms equal the analytic ones (2D adjugate; 3D
  curl form exact when X_m grad X_l has degree <= N);
* free-stream preservation on a warped periodic mesh: a uniform state has a
  zero right-hand side (the discrete metric identities of the curl form), in
  2D and 3D, and with every element blended with the FV subcells (which needs
  the subcell normals to telescope: a dropped quadrature weight or a
  transposed derivative index in their recurrence breaks it).
"""
import numpy as np
import pytest

from oracle import oracle as O


def warp(ref, a=0.1):
    """Periodic sine warp of [-1, 1]^d (faces stay planar and matching)."""
    s = np.prod(np.sin(np.pi * ref), axis=-1)
    out = ref.copy()
    for k in range(ref.shape[-1]):
        out[..., k] = ref[..., k] + a * s * (1.0 + 0.5 * k)
    return out


def test_linear_mapping_reproduces_builtin_geometry():
    for dim, cells in ((2, (4, 3)), (3, (2, 3, 2))):
        L, R = [-1.0, 0.0, 0.5][:dim], [1.0, 2.0, 1.5][:dim]
        a = O.Case(dim, 3, cells, L, R)
        b = O.Case(dim, 3, cells, L, R)
        r = b.reference_coordinates()
        x = 0.5 * ((1.0 - r) * np.array(L) + (1.0 + r) * np.array(R))
        assert np.array_equal(x, a.coordinates())
        b.set_coordinates(x)
        for get in ("jacobian", "contravariant", "inverse_jacobian"):
            assert np.array_equal(getattr(a, get)(), getattr(b, get)()), get


def quad_map(ref):
    x, y = ref[..., 0], ref[..., 1]
    z = ref[..., 2] if ref.shape[-1] == 3 else 0.0
    out = np.empty_like(ref)
    out[..., 0] = x + 0.1 * x * y + 0.05 * y * y
    out[..., 1] = y - 0.08 * x * x + 0.1 * x * z
    if ref.shape[-1] == 3:
        out[..., 2] = z + 0.07 * x * y + 0.03 * z * z
    return out


def quad_jacobian(ref, h):
    """d x_c / d xi_r per element (reference cube of side 2, element side h_r
    in reference units: d ref / d xi = h / 2)."""
    x, y = ref[..., 0], ref[..., 1]
    d = ref.shape[-1]
    z = ref[..., 2] if d == 3 else 0.0
    J = np.zeros(ref.shape[:-1] + (d, d))  # [.., r, c]
    # d out_c / d ref_k
    g = np.zeros(ref.shape[:-1] + (d, d))  # [.., k, c]
    g[..., 0, 0] = 1 + 0.1 * y
    g[..., 1, 0] = 0.1 * x + 0.1 * y
    g[..., 0, 1] = -0.16 * x + (0.1 * z if d == 3 else 0.0)
    g[..., 1, 1] = 1.0
    if d == 3:
        g[..., 2, 1] = 0.1 * x
        g[..., 0, 2] = 0.07 * y
        g[..., 1, 2] = 0.07 * x
        g[..., 2, 2] = 1 + 0.06 * z
    for r in range(d):
        J[..., r, :] = g[..., r, :] * (h[r] / 2)
    return J


@pytest.mark.parametrize("dim,cells", [(2, (3, 4)), (3, (2, 2, 3))])
def test_polynomial_mapping_is_exact(dim, cells):
    N = 3
    c = O.Case(dim, N, cells, [-1.0] * dim, [1.0] * dim)
    ref = c.reference_coordinates()
    c.set_coordinates(quad_map(ref))
    h = [2.0 / k for k in cells]
    J = quad_jacobian(ref, h)
    assert np.max(np.abs(c.jacobian() - J)) <= 1e-13
    # contravariant vectors = rows of the cofactor matrix: Ja^r_c = J det(J) (J^-1)_{c r}
    det = np.linalg.det(J)
    cof = np.einsum("...,...cr->...rc", det, np.linalg.inv(J))
    assert np.max(np.abs(c.contravariant() - cof)) <= 1e-13
    assert np.max(np.abs(c.inverse_jacobian() * det - 1.0)) <= 1e-12


@pytest.mark.parametrize("dim,N,cells", [(2, 3, (5, 4)), (2, 5, (3, 3)), (3, 3, (3, 3, 2)),
                                         (3, 4, (2, 2, 2))])
@pytest.mark.parametrize("volume", [O.SPLIT_FORM, O.SHOCK_CAPTURING])
def test_free_stream_preservation_on_warped_mesh(dim, N, cells, volume):
    # alpha_min = 0 blends every element with alpha = 1e-4 (the indicator of
    # a constant state), so the FV subcell fluxes take part
    c = O.Case(dim, N, cells, [-1.0] * dim, [1.0] * dim, volume=volume, alpha_min=0.0)
    c.set_coordinates(warp(c.reference_coordinates()))
    assert np.min(c.inverse_jacobian()) > 0
    u = np.zeros(c.shape)
    v = [0.3, -0.2, 0.1][:dim]
    u[..., 0] = 1.2
    u[..., 1:dim + 1] = 1.2 * np.array(v)
    u[..., dim + 1] = 1.0 / 0.4 + 0.5 * 1.2 * np.dot(v, v)
    du, alpha = c.rhs(u, with_alpha=True)
    if volume == O.SHOCK_CAPTURING:
        assert np.all(alpha > 1e-5)
    assert np.max(np.abs(du)) <= 1e-11


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 4)])
def test_warped_mesh_is_not_trivially_steady(dim, N):
    """The same test would pass for any RHS that vanishes identically: a
    non-uniform state on the warped mesh has a clearly non-zero RHS."""
    cells = (3,) * dim
    c = O.Case(dim, N, cells, [-1.0] * dim, [1.0] * dim)
    c.set_coordinates(warp(c.reference_coordinates()))
    x = c.coordinates()
    u = np.zeros(c.shape)
    u[..., 0] = 1.0
    u[..., dim + 1] = 2.5 + 0.2 * np.sin(np.pi * x[..., 0])  # a pressure wave
    assert np.max(np.abs(c.rhs(u))) > 1e-3
