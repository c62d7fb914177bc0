"""Build recipe of the native library: hipcc for gfx950, in-tree output.

    python -m proxpde_amd.build          # compile csrc/ -> proxpde_amd/libproxpde_amd.so

Every translation unit (HIP kernels and the host C++ around them) is compiled
with hipcc; objects go to proxpde_amd/_build/, the shared library sits next to
this file so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libproxpde_amd.so")
ARCH = os.environ.get("PROXPDE_AMD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-function",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".cpp", ".hip")):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


TORCH_SRC = os.path.join(HERE, "torch_ext", "ops.cpp")
TORCH_LIB = os.path.join(HERE, "libproxpde_torch.so")


def build_torch_ext(verbose: bool = False) -> str:
    """LibTorch autograd operators (host C++ only: g++ against the torch headers,
    linked to libproxpde_amd.so)."""
    import torch.utils.cpp_extension as ce
    import torch

    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    inc = sum([["-I", p] for p in ce.include_paths()], [])
    libdirs = ce.library_paths()
    tmp = TORCH_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}", *inc,
           TORCH_SRC, "-o", tmp, *[f"-L{d}" for d in libdirs], "-ltorch", "-ltorch_cpu", "-lc10",
           f"-L{HERE}", "-lproxpde_amd", "-Wl,-rpath,$ORIGIN", *[f"-Wl,-rpath,{d}" for d in libdirs]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"torch extension build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, TORCH_LIB)
    if verbose:
        print(f"built {TORCH_LIB}")
    return TORCH_LIB


if __name__ == "__main__":
    build(verbose=True)
    if "--torch" in sys.argv:
        build_torch_ext(verbose=True)
    sys.exit(0)
