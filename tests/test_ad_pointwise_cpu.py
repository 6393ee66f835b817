"""The hand-written tangents and adjoints of the pointwise physics
(dgsem_amd/csrc/physics_ad.hpp), compiled for the CPU: tangent against
central finite differences, adjoint against tangent by the dot-product
identity, for the Chandrashekar flux (both branches of the logarithmic mean),
the LLF flux, the HLLC flux (tangent by dual numbers, adjoint in reverse
mode and against the dual-number Jacobian columns) and the Navier-Stokes
viscous flux (tests/native/ad_check.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pointwise_derivatives(tmp_path):
    exe = tmp_path / "ad_check"
    subprocess.run(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-o", str(exe),
                    os.path.join(ROOT, "tests", "native", "ad_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count(" ok") == 12
