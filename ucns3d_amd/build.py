"""Build the native libraries in-tree.

libucnsprep.so  host FP64 preprocessor (g++ -fopenmp)
libucnsgpu.so   HIP residual/RK3 library for gfx950 (hipcc)
ucns_fortran_host  the Fortran binding module (fortran/ucns_gpu.f90) compiled
                with amdflang into a test host that drives the library the
                way UCNS3D's TIME_MARCHING would (tests/fortran/host_march.f90)

The built .so files live in ucns3d_amd/lib/ so they travel with the source
tree; nothing is installed into site-packages.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")

# packed-FP32 instructions are disabled for every HIP translation unit
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_prep() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libucnsprep.so")
    src = os.path.join(CSRC, "prep", "ucns_prep.cpp")
    _run([CXX, "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-I" + INCLUDE, src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_gpu() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libucnsgpu.so")
    src = os.path.join(CSRC, "ucns_gpu.hip")
    _run([HIPCC, *HIP_FLAGS, "-I" + INCLUDE, src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_fortran() -> str | None:
    """the Fortran binding module and the Fortran test host linked against
    both libraries (None when amdflang is not installed)"""
    if not os.path.exists(FC):
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    mod_dir = os.path.join(LIBDIR, "fortran_mod")
    os.makedirs(mod_dir, exist_ok=True)
    out = os.path.join(LIBDIR, "ucns_fortran_host")
    srcs = [os.path.join(HERE, "fortran", "ucns_gpu.f90"), os.path.join(REPO, "tests", "fortran", "host_march.f90")]
    _run([FC, "-O2", "-J", mod_dir, *srcs, "-o", out + ".tmp", "-L" + LIBDIR, "-lucnsgpu", "-lucnsprep",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> list[str]:
    return [p for p in (build_prep(), build_gpu(), build_fortran()) if p]


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        if w == "prep":
            print(build_prep())
        elif w == "gpu":
            print(build_gpu())
        elif w == "fortran":
            print(build_fortran())
        else:
            for p in build_all():
                print(p)
