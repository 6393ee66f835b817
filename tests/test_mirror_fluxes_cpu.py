"""The C++ mirror (dgsem_amd/include/dgsem_hip.hpp) composes solvers from the
reference's flux traits (numerical_flux.hpp:39-594) and turns each into the
DGSEM_FLUX_* value of its role: CentralFlux is a volume, surface and
subcell flux of 2D solvers; ChandrashekarFlux a volume flux only; LLF and
HLLC surface and subcell fluxes only.  A solver that names a flux in a role
the kernels do not take fails to compile with a message.  Host compilation
only (hipcc -fsyntax-only), no GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

TEMPLATE = """
#include "dgsem_hip.hpp"
using namespace dgsem_hip;
using Eq = equations::CompressibleEuler{dim}D<double>;
using B = Basis::LobattoLegendreBasis<double, 3>;
using Volume = {volume};
using Mesh = StructuredMesh<double, {dim}>;
using Boundaries = BoundarySet<{bcs}>;
using Solver = StructuredSolver<Eq, B, Volume, {surface}<Eq>, Mesh, Boundaries>;
static_assert(Solver::NDIMS == {dim});
int check() {{ return Volume::volume_flux * 10 + {surface}<Eq>::kind; }}
"""


def compile_snippet(tmp_path, dim, volume, surface):
    src = tmp_path / "snippet.cpp"
    bcs = ", ".join(["PeriodicBC"] * (2 * dim))
    src.write_text(TEMPLATE.format(dim=dim, volume=volume, surface=surface, bcs=bcs))
    return subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++20", "-fsyntax-only",
                           f"-I{os.path.join(ROOT, 'dgsem_amd', 'include')}", str(src)],
                          capture_output=True, text=True, timeout=120)


SPLIT = "VolumeIntegralSplitForm<B, Eq, {}>"
SC = "VolumeIntegralShockCapturingHG<B, Eq, {}, {}, HGIndicator<B, Eq>>"


@pytest.mark.parametrize("volume,surface", [
    (SPLIT.format("CentralFlux"), "CentralFlux"),
    (SPLIT.format("CentralFlux"), "LaxFriedrichsFlux"),
    (SPLIT.format("ChandrashekarFlux"), "CentralFlux"),
    (SC.format("CentralFlux", "CentralFlux"), "HLLCFlux"),
    (SC.format("ChandrashekarFlux", "LaxFriedrichsFlux"), "LaxFriedrichsFlux"),
])
def test_2d_solvers_take_central_flux(tmp_path, volume, surface):
    r = compile_snippet(tmp_path, 2, volume, surface)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.parametrize("dim,volume,surface,message", [
    (3, SPLIT.format("ChandrashekarFlux"), "CentralFlux", "CentralFlux takes a normal in 2D only"),
    (3, SPLIT.format("CentralFlux"), "LaxFriedrichsFlux", "CentralFlux takes a normal in 2D only"),
    (2, SPLIT.format("LaxFriedrichsFlux"), "LaxFriedrichsFlux", "Chandrashekar or central volume flux"),
    (2, SPLIT.format("ChandrashekarFlux"), "ChandrashekarFlux", "Chandrashekar or central volume flux"),
])
def test_unbuilt_flux_roles_fail_to_compile(tmp_path, dim, volume, surface, message):
    r = compile_snippet(tmp_path, dim, volume, surface)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr[-3000:]
