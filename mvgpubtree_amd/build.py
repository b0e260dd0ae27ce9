"""Build recipe for the HIP library (gfx950) and the C++ API check program.

``build()`` compiles every time it is called:
  * ``mvgpubtree_amd/libmvbt.so``   - kernels + C ABI (include/mvbt.h)
  * ``mvgpubtree_amd/capi_example`` - C++ program using include/gpu_btree.h
  * ``mvgpubtree_amd/device_api_example`` - user kernels on include/mvbt_device.hpp
  * ``mvgpubtree_amd/libmvbt_driver.so`` - the plain per-batch / per-request
    caller loops over the C ABI (baselines of the batched and served forms)
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmvbt.so")
EXAMPLE = os.path.join(PKG, "capi_example")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MVBT_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_lib(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "mvbt_capi.hip")]
    _run(cmd, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_example(verbose: bool = False) -> str:
    """C++ host program over include/gpu_btree.h, linked against libmvbt.so."""
    src = os.path.join(CSRC, "capi_example.cpp")
    cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-command-line-argument", f"-I{os.path.join(ROOT, 'include')}", "-o",
           EXAMPLE, src, f"-L{PKG}", "-lmvbt", f"-Wl,-rpath,{PKG}"]
    _run(cmd, verbose)
    return EXAMPLE


def build_device_example(verbose: bool = False) -> str:
    """User-kernel program over include/mvbt_device.hpp (device code: gfx950)."""
    src = os.path.join(CSRC, "device_api_example.hip")
    out = os.path.join(PKG, "device_api_example")
    cmd = [HIPCC, *[f for f in HIP_FLAGS if f != "-fPIC"], f"-I{os.path.join(ROOT, 'include')}",
           "-o", out, src, f"-L{PKG}", "-lmvbt", f"-Wl,-rpath,{PKG}"]
    _run(cmd, verbose)
    return out


def build_driver(verbose: bool = False) -> str:
    """The plain per-batch / per-request caller loops over the C ABI
    (libmvbt_driver.so, host code), the baselines of the batched and served forms."""
    out = os.path.join(PKG, "libmvbt_driver.so")
    tmp = out + ".tmp"
    src = os.path.join(CSRC, "per_batch_driver.cpp")
    cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-o", tmp, src, f"-L{PKG}", "-lmvbt",
           f"-Wl,-rpath,{PKG}"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_example(verbose)
    build_device_example(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
