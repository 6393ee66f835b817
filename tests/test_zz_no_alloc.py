"""dgsem_analyze_noalloc, the analysis entry point that allocates nothing
(its block partials live in a static device array of the library), called
through the C ABI and compared with dgsem_analyze, which allocates its
partials on first use: two calls in a row, the second on a smaller mesh,
bit for bit (both run the same two kernels in the same order)."""
import ctypes as C
import math

import numpy as np
import pytest

import dgsem_amd as dg
from dgsem_amd import _abi, tgv

pytestmark = pytest.mark.gpu


def _lib():
    L = _abi.load()
    L.dgsem_analyze_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.dgsem_analyze_noalloc.restype = C.c_int
    return L


def _solver(dim, N, cells):
    cfg = dg.SolverConfig(dim=dim, polydeg=N, cells=cells, domain_left=(0.0,) * dim,
                          domain_right=(2 * math.pi,) * dim)
    s = dg.StructuredSolver(cfg)
    x = s.node_coordinates()
    if dim == 3:
        u = tgv.tgv_state(x)
    else:
        rho = 1.0 + 0.2 * np.sin(x[..., 0]) * np.cos(x[..., 1])
        u = np.stack([rho, rho * 0.3, -rho * 0.1, 2.5 + 0.1 * np.cos(x[..., 0])], axis=-1)
    return s, s.from_reference_layout(u)


def _both(L, s, u):
    import torch

    a = torch.full((5,), -1.0, dtype=torch.float64, device="cuda")
    b = torch.full((5,), -2.0, dtype=torch.float64, device="cuda")
    assert L.dgsem_analyze(s._h, C.c_void_p(u.data_ptr()), C.c_void_p(a.data_ptr())) == 0
    rc = L.dgsem_analyze_noalloc(s._h, C.c_void_p(u.data_ptr()), C.c_void_p(b.data_ptr()))
    assert rc == 0, L.dgsem_last_error(s._h).decode()
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_analyze_matches_allocating_entry_point(gpu):
    L = _lib()
    big, ub = _solver(3, 3, (12, 10, 8))      # 61,440 nodes: 240 blocks of partials
    small, us = _solver(2, 2, (3, 4))         # 108 nodes: one block
    for s, u, name in ((big, ub, "3D N=3 12x10x8"), (small, us, "2D N=2 3x4")):
        print(f"case analyze_noalloc {name}", flush=True)
        ref, got = _both(L, s, u)
        assert np.array_equal(ref, got), (name, ref, got)
        assert got[0] > 0 and got[3] == 0
    # again on the large mesh after the small one: stale partials beyond the
    # small call's block are never read
    ref, got = _both(L, big, ub)
    assert np.array_equal(ref, got)


def test_noalloc_analyze_rejects_null_arguments(gpu):
    L = _lib()
    s, u = _solver(2, 2, (2, 2))
    assert L.dgsem_analyze_noalloc(s._h, None, None) == -1  # DGSEM_E_INVALID
    assert L.dgsem_analyze_noalloc(None, C.c_void_p(u.data_ptr()), None) == -1
