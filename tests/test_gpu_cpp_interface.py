"""The reference-style C++ interface (dgsem_amd/include/dgsem_hip.hpp): a
program composed like the reference's KHI example runs SSPRK3 on the GPU and
matches the restatement of the reference after the same steps."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "dgsem_amd", "examples", "khi_shock_capturing")


@pytest.mark.gpu
def test_cpp_khi_example_matches_oracle(gpu, tmp_path):
    assert os.path.exists(EXE), "build the library first (dgsem_amd/build.py)"
    nx, ny, steps, dt, amax = 16, 12, 30, 2e-4, 0.002
    out = tmp_path / "u.bin"
    print(f"case cpp khi {nx}x{ny} steps={steps}", flush=True)
    subprocess.run([EXE, str(nx), str(ny), str(steps), str(dt), str(amax), str(out)],
                   check=True, timeout=60)
    c = O.Case(2, 3, (nx, ny), [-1, -1], [1, 1], volume=O.SHOCK_CAPTURING, alpha_max=amax,
               alpha_min=0.001)
    ref = c.ssprk3(c.init_khi(), dt, steps)
    got = np.fromfile(out, dtype=np.float64).reshape(ref.shape)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-10
