"""Builds librelccsd_hip.so (gfx950) in-tree with hipcc.

    python -m molfdir_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librelccsd_hip.so")
SOURCES = ["rcc_zgemm.hip", "rcc_kernels.hip", "rcc_split.hip", "rcc_split_i8.hip",
           "rcc_split_i8w.hip", "rcc_triples.hip", "rcc_frontend.hip", "rcc_props.hip",
           "rcc_blocked.hip",
           "rcc_solver.hip"]
# sources compiled without the VGPR form of the MFMA accumulators (none now)
NO_VGPR_FORM = set()
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RCC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: no AGPR <-> VGPR copies per K-tile
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


FORTRAN = os.path.join(HERE, "fortran")
FC = os.environ.get("FC_RCC", "flang")


def build_fortran(verbose: bool = False) -> str:
    """The ISO_C_BINDING module and the CCENER/CCEQN-sequence driver."""
    exe = os.path.join(FORTRAN, "ccdrv_hip.x")
    cmd = [FC, "-O2", "-J", FORTRAN, os.path.join(FORTRAN, "relccsd_hip.f90"),
           os.path.join(FORTRAN, "ccdrv_hip.f90"), "-L" + HERE, "-lrelccsd_hip",
           "-Wl,-rpath,$ORIGIN/..", "-o", exe]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return exe


def build(verbose: bool = False) -> str:
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(CSRC, s.replace(".hip", ".o"))
        flags = FLAGS
        if s in NO_VGPR_FORM:
            flags = [f for f in FLAGS if f not in ("-mllvm", "-amdgpu-mfma-vgpr-form")]
        cmd = [HIPCC, *flags, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
