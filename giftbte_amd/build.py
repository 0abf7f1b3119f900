"""Build libgiftbte_hip.so (gfx950) and the CPU oracle, in-tree.

    python -m giftbte_amd.build          # both
    python -m giftbte_amd.build --lib    # only the HIP library

hipcc compiles every HIP translation unit for gfx950 with packed FP32
instructions disabled; the host-only sources are compiled by the same driver.
The oracle (oracle/oracle.cpp) is built with g++ into oracle/liboracle.so; it
is test infrastructure and never linked into the product library.  When the
reference sources are present, oracle/_ref is built from them by
oracle/ref.mk.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgiftbte_hip.so")
ORACLE_SRC = os.path.join(ROOT, "oracle", "oracle.cpp")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REFERENCE = "/root/reference"

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=ROOT):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


# sources with no HIP code, compiled by the host compiler alone
HOST_ONLY = {"ndchol.cpp"}


def host_cxx() -> str:
    """The clang++ that hipcc drives for host code."""
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++") or "", shutil.which("g++") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("no host C++ compiler")


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_lib() -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    srcs = ["problem.cpp", "bands.cpp", "outputs.cpp", "reorder.cpp", "ndchol.cpp", "tiles.cpp", "capi_problem.cpp", "capi_solver.cpp",
            "solver.hip", "solver3d.hip", "keff_device.hip",
            "adjoint.hip"]
    for s in srcs:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s + ".o")
        if s in HOST_ONLY:
            # plain host C++ (no device pass): its dense kernels are built for
            # two host instruction sets and picked at load time
            cmd = [host_cxx(), "-O3", "-std=c++17", "-fPIC", "-I", CSRC, "-c", src, "-o", obj]
        else:
            cmd = [hipcc()] + HIP_FLAGS + ["-I", INC, "-I", CSRC, "-c", src, "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> str:
    tmp = ORACLE_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INC, ORACLE_SRC, "-o", tmp])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_ref() -> bool:
    """Compile the reference into oracle/_ref when its sources are here."""
    if not os.path.isdir(os.path.join(REFERENCE, "src")):
        return False
    jobs = str(min(8, os.cpu_count() or 1))
    _run(["make", "-f", os.path.join("oracle", "ref.mk"), "-j", jobs])
    return True


def main(argv=None) -> None:
    argv = sys.argv[1:] if argv is None else argv
    if "--lib" in argv:
        build_lib()
        return
    build_lib()
    build_oracle()
    build_ref()


if __name__ == "__main__":
    main()
