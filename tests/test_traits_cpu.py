"""The plugin traits of the C++ mirror (dgsem_amd/include/dgsem_hip.hpp):
each equation and numerical-flux member the reference's kernels call --
flux(u, dim), flux(u, normal), get_wave_speed(u_ll, u_rr, dim | normal)
(compressible_euler3D.hpp:26-166), LaxFriedrichsFlux / ChandrashekarFlux /
HLLCFlux / CentralFlux::numerical_flux (numerical_flux.hpp:39-594),
gradient_variables and viscous_flux (compressible_navier_stokes3D.hpp:55-122)
and utils::prim_to_cons -- held against the oracle's pointwise functions or
their closed forms, on seeded random states in 2D and 3D.  Host code only
(tests/native/traits_abi.cpp, compiled here with hipcc)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GAMMA = 1.4
TOL = 1e-13


@pytest.fixture(scope="module")
def traits(tmp_path_factory):
    out = tmp_path_factory.mktemp("traits") / "libtraits.so"
    lib_dir = os.path.join(ROOT, "dgsem_amd")
    subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++20", "-O2", "-fPIC", "-shared", "-o", str(out),
                    os.path.join(ROOT, "tests", "native", "traits_abi.cpp"), f"-L{lib_dir}",
                    "-ldgsem_amd", f"-Wl,-rpath,{lib_dir}"], check=True)
    L = C.CDLL(str(out))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.tr_eval.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp, dp, C.c_int, dp]
    L.tr_wave_speed.restype = C.c_double
    L.tr_wave_speed.argtypes = [C.c_int, C.c_double, dp, dp, C.c_void_p, C.c_int]
    L.tr_gradient_variables.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
    L.tr_viscous_flux.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, C.c_int, dp]
    L.tr_prim_to_cons.argtypes = [C.c_int, C.c_double, dp, dp]
    return L


def states(dim, seed, count=40):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        pair = []
        for _ in range(2):
            rho, p = rng.uniform(0.3, 3.0), rng.uniform(0.3, 3.0)
            v = rng.normal(0.0, 0.8, dim)
            pair.append(np.concatenate([[rho], rho * v, [p / (GAMMA - 1) + 0.5 * rho * v @ v]]))
        n = rng.normal(0.0, 1.0, dim)
        out.append((pair[0], pair[1], n))
    return out


def close(a, b):
    return np.linalg.norm(a - b) <= TOL * max(1.0, np.linalg.norm(b))


def call(L, dim, kind, ul, ur, n, d=0):
    f = np.zeros(dim + 2)
    L.tr_eval(dim, kind, GAMMA, ul, ur, np.ascontiguousarray(n, dtype=np.float64), d, f)
    return f


@pytest.mark.parametrize("dim", [2, 3])
def test_equation_flux_members(traits, dim):
    for ul, ur, n in states(dim, 1 + dim):
        assert close(call(traits, dim, 1, ul, ur, n), O.physical_flux(dim, GAMMA, ul, n))
        for d in range(dim):
            e = np.eye(dim)[d]
            assert close(call(traits, dim, 0, ul, ur, n, d), O.physical_flux(dim, GAMMA, ul, e))


@pytest.mark.parametrize("dim", [2, 3])
def test_wave_speeds(traits, dim):
    def prims(u):
        rho = u[0]
        v = u[1:1 + dim] / rho
        p = (GAMMA - 1) * (u[-1] - 0.5 * rho * v @ v)
        return v, np.sqrt(GAMMA * p / rho)

    for ul, ur, n in states(dim, 7 + dim):
        (vl, cl), (vr, cr) = prims(ul), prims(ur)
        ref_n = max(abs(vl @ n), abs(vr @ n)) + max(cl, cr) * np.linalg.norm(n)
        got_n = traits.tr_wave_speed(dim, GAMMA, ul, ur, np.ascontiguousarray(n).ctypes.data, 0)
        assert abs(got_n - ref_n) <= TOL * ref_n
        for d in range(dim):
            ref_d = max(abs(vl[d]) + cl, abs(vr[d]) + cr)
            assert abs(traits.tr_wave_speed(dim, GAMMA, ul, ur, None, d) - ref_d) <= TOL * ref_d


@pytest.mark.parametrize("dim", [2, 3])
def test_numerical_fluxes_against_oracle(traits, dim):
    for ul, ur, n in states(dim, 11 + dim):
        assert close(call(traits, dim, 3, ul, ur, n), O.flux_llf(dim, GAMMA, ul, ur, n))
        assert close(call(traits, dim, 4, ul, ur, n), O.flux_chandrashekar(dim, GAMMA, ul, ur, n))
        assert close(call(traits, dim, 5, ul, ur, n), O.flux_hllc(dim, GAMMA, ul, ur, n))
        for d in range(dim):
            e = np.eye(dim)[d]
            fl, fr = O.physical_flux(dim, GAMMA, ul, e), O.physical_flux(dim, GAMMA, ur, e)
            assert close(call(traits, dim, 6, ul, ur, n, d), 0.5 * (fl + fr))
            # LLF in dimension d: the dimension-wise wave speed max(|v_d| + c)
            lam = traits.tr_wave_speed(dim, GAMMA, ul, ur, None, d)
            assert close(call(traits, dim, 2, ul, ur, n, d), 0.5 * (fl + fr) - 0.5 * lam * (ur - ul))
        if dim == 2:
            fl, fr = O.physical_flux(dim, GAMMA, ul, n), O.physical_flux(dim, GAMMA, ur, n)
            assert close(call(traits, dim, 7, ul, ur, n), 0.5 * (fl + fr))


def test_chandrashekar_is_consistent_and_symmetric(traits):
    """f(u, u) is the physical flux and f(a, b) = f(b, a)."""
    for ul, ur, n in states(3, 21):
        assert np.allclose(call(traits, 3, 4, ul, ul, n), O.physical_flux(3, GAMMA, ul, n),
                           rtol=1e-13, atol=1e-13)
        assert np.array_equal(call(traits, 3, 4, ul, ur, n), call(traits, 3, 4, ur, ul, n))


@pytest.mark.parametrize("dim", [2, 3])
def test_navier_stokes_members_against_oracle(traits, dim):
    mu, pr = 0.05, 0.72
    oc = O.Case(dim, 2, (2,) * dim, [0.0] * dim, [1.0] * dim, eq=O.NAVIER_STOKES, mu=mu,
                prandtl=pr)
    rng = np.random.default_rng(31 + dim)
    for ul, _, _ in states(dim, 41 + dim, count=20):
        q = np.zeros(dim + 2)
        traits.tr_gradient_variables(dim, GAMMA, mu, pr, ul, q)
        assert close(q, oc.gradient_variables(ul))
        grad = rng.normal(0.0, 1.0, (dim, dim + 2))
        for d in range(dim):
            f = np.zeros(dim + 2)
            traits.tr_viscous_flux(dim, GAMMA, mu, pr, q, np.ascontiguousarray(grad), d, f)
            assert close(f, oc.viscous_flux(q, grad, d))


@pytest.mark.parametrize("dim", [2, 3])
def test_prim_to_cons(traits, dim):
    rng = np.random.default_rng(5)
    for _ in range(10):
        rho, p, v = rng.uniform(0.5, 2), rng.uniform(0.5, 2), rng.normal(0, 1, dim)
        prim = np.concatenate([[rho], v, [p]])
        cons = np.zeros(dim + 2)
        traits.tr_prim_to_cons(dim, GAMMA, prim, cons)
        ref = np.concatenate([[rho], rho * v, [p / (GAMMA - 1) + 0.5 * rho * v @ v]])
        assert close(cons, ref)
