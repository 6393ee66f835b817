"""Pins the restatement's basis tables (oracle/dgsem_oracle.c, following
polynomial_basis/basis_detail.hpp) against closed forms and the summation-by-
parts identities they must satisfy."""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from oracle import oracle as O

CLOSED_FORM = {
    1: ([-1.0, 1.0], [1.0, 1.0]),
    2: ([-1.0, 0.0, 1.0], [1 / 3, 4 / 3, 1 / 3]),
    3: ([-1.0, -1 / np.sqrt(5), 1 / np.sqrt(5), 1.0], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
    4: ([-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0],
        [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
}


@pytest.mark.parametrize("N", sorted(CLOSED_FORM))
def test_gll_closed_forms(N):
    b = O.basis(N)
    x, w = CLOSED_FORM[N]
    np.testing.assert_allclose(b["nodes"], x, rtol=0, atol=2e-16)
    np.testing.assert_allclose(b["weights"], w, rtol=4e-16, atol=0)


@pytest.mark.parametrize("N", range(1, 8))
def test_gll_against_legendre_roots(N):
    b = O.basis(N)
    # interior GLL nodes are the roots of P_N'
    interior = np.sort(npleg.legroots(npleg.legder([0] * N + [1])))
    np.testing.assert_allclose(b["nodes"][1:-1], interior, atol=1e-14)
    assert abs(b["weights"].sum() - 2.0) < 1e-14
    # weights integrate polynomials of degree <= 2N-1 exactly
    for k in range(2 * N):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(b["weights"], b["nodes"] ** k) - exact) < 1e-13
    np.testing.assert_array_equal(b["inv_weights"], 1.0 / b["weights"])


@pytest.mark.parametrize("N", range(1, 8))
def test_derivative_matrix_exact_on_polynomials_and_sbp(N):
    b = O.basis(N)
    x, w, D = b["nodes"], b["weights"], b["D"]
    n = N + 1
    assert np.max(np.abs(D @ np.ones(n))) < 1e-13
    for k in range(1, N + 1):
        np.testing.assert_allclose(D @ x ** k, k * x ** (k - 1), atol=2e-12)
    M, B = np.diag(w), np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    np.testing.assert_allclose(M @ D + (M @ D).T, B, atol=1e-13)


@pytest.mark.parametrize("N", range(1, 8))
def test_split_and_weak_derivative_operators(N):
    b = O.basis(N)
    x, w, D, Ds, Dh = b["nodes"], b["weights"], b["D"], b["Dsplit"], b["Dhat"]
    n = N + 1
    M = np.diag(w)
    # D_split = 2D - M^{-1}B: M D_split is skew-symmetric, diagonal ~ 0
    S = M @ Ds
    np.testing.assert_allclose(S + S.T, 0.0, atol=2e-13)
    assert np.max(np.abs(np.diag(Ds))) < 1e-13
    # weak derivative + boundary lift reproduces the strong derivative (SBP)
    rng = np.random.default_rng(N)
    q = rng.standard_normal(n)
    lift = np.zeros(n)
    lift[0], lift[-1] = -q[0] / w[0], q[-1] / w[-1]
    np.testing.assert_allclose(Dh @ q + lift, D @ q, atol=1e-12)
    # l_hat at the end points: e_0 / w_0 and e_N / w_N
    np.testing.assert_array_equal(b["lhat_left"], np.eye(n)[0] / w[0])
    np.testing.assert_array_equal(b["lhat_right"], np.eye(n)[-1] / w[-1])


@pytest.mark.parametrize("N", range(1, 8))
def test_inverse_vandermonde(N):
    b = O.basis(N)
    n = N + 1
    V = np.stack([npleg.legval(b["nodes"], np.eye(n)[j]) * np.sqrt(j + 0.5) for j in range(n)], axis=1)
    np.testing.assert_allclose(b["Vinv"] @ V, np.eye(n), atol=1e-12)
