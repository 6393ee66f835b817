"""Builds libdgsem_amd.so in-tree: host setup with g++ (-ffp-contract=off, so
basis and geometry are bitwise the reference's host values) and the HIP
kernels + C ABI with hipcc for gfx950."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgsem_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=on: multiply-adds are fused only where the source writes them
# as one expression, so the fusion (and with it every rounding) is fixed by
# the source text and the same element arithmetic gives the same bits in
# every kernel that includes element_body.inc.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=on",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-std=c++20", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-Wall",
              "-fopenmp"]

# Per-source compiler options.  The batch kernel loops over tiles; with
# machine-level loop-invariant code motion the FP64 literals of every
# degree's body are hoisted out of those loops and kept in registers across
# the kernel, which made it spill.
EXTRA_FLAGS = {"batch_inst.hip": ["-mllvm", "-disable-machine-licm"]}

# (source, compiler, object name, extra defines)
INSTANCES = [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (2, 7), (3, 2), (3, 3), (3, 4), (3, 5), (3, 7)]
SOURCES = [
    ("host_setup.cpp", "g++", "host_setup", ()),
    ("rhs_dispatch.hip", "hipcc", "rhs_dispatch", ()),
    ("abi.hip", "hipcc", "abi", ()),
    ("batch.hip", "hipcc", "batch", ()),
    ("adjoint.hip", "hipcc", "adjoint", ()),
    ("analysis.hip", "hipcc", "analysis", ()),
    ("vtu.hip", "hipcc", "vtu", ()),
    ("geometry.hip", "hipcc", "geometry", ()),
] + [("batch_inst.hip", "hipcc", f"batch_inst_b{b}", (f"DGSEM_BATCH_BLOCK={b}",))
     for b in (128, 256)] + [(f"{kind}_inst.hip", "hipcc", f"{kind}_inst_{d}d_n{n}",
       (f"DGSEM_INST_DIM={d}", f"DGSEM_INST_N={n}"))
     for kind in ("rhs", "adjoint") for d, n in INSTANCES]


def _deps_newer(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    for f in os.listdir(CSRC):
        if os.path.getmtime(os.path.join(CSRC, f)) > t:
            return True
    inc = os.path.join(HERE, "..", "include", "dgsem_amd.h")
    return os.path.getmtime(inc) > t


def build(force: bool = True, verbose: bool = False, defines=(), out: str | None = None,
          build_dir: str | None = None) -> str:
    """Compiles and links the library.  defines: extra -D flags for the HIP
    sources (tuning knobs such as DGSEM_WAVES_3D); out/build_dir: where the
    variant goes (defaults: the in-tree library)."""
    out = out or LIB
    bdir = build_dir or BUILD
    os.makedirs(bdir, exist_ok=True)

    def one(src):
        name, tool, oname, extra = src
        obj = os.path.join(bdir, oname + ".o")
        if not force and not _deps_newer(name, obj):
            return obj
        path = os.path.join(CSRC, name)
        if tool == "g++":
            cmd = ["g++", *HOST_FLAGS, "-c", path, "-o", obj]
        else:
            cmd = [HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(name, []), *[f"-D{d}" for d in (*defines, *extra)],
                   "-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr}")
        return obj

    workers = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(one, SOURCES))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", *objs, "-lgomp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    if out == LIB:
        build_examples()
    if verbose:
        print("built", out)
    return out


EXAMPLES = ["khi_shock_capturing", "double_mach_reflection", "viscous_shock_tube",
            "taylor_green_observers", "astro_jet", "riemann2d"]


def build_examples() -> None:
    """C++ programs written against include/dgsem_hip.hpp (the reference-style
    template interface), linked to the in-tree library."""
    for name in EXAMPLES:
        src = os.path.join(HERE, "examples", name + ".cpp")
        exe = os.path.join(HERE, "examples", name)
        cmd = [HIPCC, "-std=c++20", "-O2", "-o", exe, src, f"-L{HERE}", "-ldgsem_amd",
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed: {' '.join(cmd)}\n{r.stderr}")


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
