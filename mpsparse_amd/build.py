"""In-tree build of the native parts.

1. ``libmpsparse_amd.so`` -- every HIP kernel plus the C ABI, one hipcc
   invocation for gfx950 (``include/mpsparse_amd.h``).
2. ``_C*.so`` -- the PyTorch-ROCm extension (pybind surface of the reference),
   plain C++ against torch's headers, linked to (1) with an ``$ORIGIN`` rpath.

Run ``python -m mpsparse_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("MSP_OFFLOAD_ARCH", "gfx950")

LIB = os.path.join(HERE, "libmpsparse_amd.so")
HIP_SOURCES = [os.path.join(CSRC, "capi.hip")]
HIP_DEPS = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".hip"))] + [
    os.path.join(ROOT, "include", "mpsparse_amd.h")]

HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
    # never let the compiler form packed FP32 ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_C" + suffix)


def _run(cmd, quiet=False):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip() and not quiet:
        print(out)


def build_lib(verbose=False):
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, "-o", LIB + ".tmp", *HIP_SOURCES]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_ext(verbose=False):
    import torch
    from torch.utils import cpp_extension

    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    incs = cpp_extension.include_paths() + [sysconfig.get_paths()["include"], os.path.join(ROCM, "include")]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    out = ext_path()
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-fvisibility=hidden",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=_C", "-DTORCH_API_INCLUDE_EXTENSION_H",
           "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
           *[f"-I{p}" for p in incs],
           os.path.join(CSRC, "torch_ext.cpp"),
           "-L" + HERE, "-lmpsparse_amd", "-Wl,-rpath,$ORIGIN",
           "-L" + torch_lib, "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python",
           "-Wl,-rpath," + torch_lib,
           "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd, quiet=True)
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def build_all(verbose=False):
    build_lib(verbose)
    build_ext(verbose)
    build_oracle()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB, ext_path())
