"""Build the native parts in-tree.

    python -m softflow_amd.build          # HIP library + softflow_core module
    python -m softflow_amd.build --oracle # also the CPU oracle (test infrastructure)

Outputs:
    softflow_amd/lib/libsoftflow_hip.so                  hipcc, gfx950 device code + C-ABI
    softflow_amd/softflow_core<EXT_SUFFIX>               g++, pybind11 engine over the C-ABI
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
HIP_LIB = os.path.join(LIBDIR, "libsoftflow_hip.so")
HIP_SRCS = ["sf_api.hip", "sf_lbm.hip", "sf_ibm.hip", "sf_membrane.hip", "sf_mlp.hip", "sf_batch.hip"]
ARCH = os.environ.get("SOFTFLOW_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
    # IEEE double without FMA contraction: the FP64 path reproduces the
    # reference's operation order bit for bit
    "-ffp-contract=off",
    "-Wno-unused-value", "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in HIP_SRCS:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        _run(["hipcc", *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run(["hipcc", "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", HIP_LIB, *objs])
    for o in objs:
        os.remove(o)
    return HIP_LIB


def core_path() -> str:
    return os.path.join(PKG, "softflow_core" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_core() -> str:
    import pybind11

    out = core_path()
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "engine.cpp"), os.path.join(CSRC, "output.cpp"),
           os.path.join(CSRC, "bindings.cpp"), "-pthread",
           "-L", LIBDIR, "-lsoftflow_hip", "-Wl,-rpath,$ORIGIN/lib",
           "-o", out]
    _run(cmd)
    return out


def build_oracle(ref: bool = False) -> None:
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    if ref and os.path.isdir("/root/reference/src"):
        _run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "ref", "ref-release"])


def build_all(oracle: bool = True) -> None:
    build_hip()
    build_core()
    if oracle:
        build_oracle(ref=os.path.isdir("/root/reference/src"))


if __name__ == "__main__":
    build_hip()
    build_core()
    if "--oracle" in sys.argv:
        build_oracle(ref=True)
