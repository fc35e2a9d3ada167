"""Build recipe for the native libraries (gfx950 HIP + host C++).

    python -m latticeboltzmanncuda_amd.build

produces, in-tree under latticeboltzmanncuda_amd/lib/:
  liblbm_hip.so   the C-ABI of include/lbm.h (HIP kernels for gfx950)

Every .hip file is compiled to an object with hipcc (parallel), then
linked into one shared library.  Compilation always runs (no staleness
check), so a call after a source change always rebuilds.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "lib", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

HIP_SOURCES = ["lbm_capi.hip", "k_standard.hip", "k_reference.hip",
               "k_common.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    # surface real warnings, drop the host pass's note about the GPU-only
    # target feature
    lines = [l for l in r.stdout.splitlines()
             if "packed-fp32-ops" not in l and l.strip()]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def _compile_hip(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip(jobs: int | None = None) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(HIP_SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    out = os.path.join(LIBDIR, "liblbm_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o",
          tmp])
    os.replace(tmp, out)
    return out


def build_all() -> list[str]:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    return [build_hip()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
