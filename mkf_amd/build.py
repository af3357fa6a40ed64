"""Build libmkf_amd.so (gfx950 HIP kernels + C ABI) and the oracle library.

Plain compiler invocations, no build system: hipcc for the .hip sources,
gcc for the host-only C sources, hipcc to link. Outputs stay in the tree
(mkf_amd/libmkf_amd.so, oracle/liboracle.so) so they travel with the
repository snapshot to the GPU machine.
"""

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mkf_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(ROOT, "mkf_amd", "libmkf_amd.so")
ORACLE_SRC = os.path.join(ROOT, "oracle", "mkf_oracle.c")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-Wall", f"-I{INCLUDE}"]

HIP_SOURCES = ["mkf_kernels.hip", "mkf_api.hip", "mkf_batch.hip", "mkf_serve.hip", "mkf_stream.hip"]
C_SOURCES = ["binmap_util.c", "measures.c"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose=False):
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O3", "-march=x86-64-v2", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-pthread",
          ORACLE_SRC, "-o", tmp, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    if verbose:
        print(f"built {ORACLE_LIB}")
    return ORACLE_LIB


def build_all(verbose=False):
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    if "--clean" in sys.argv:
        shutil.rmtree(BUILD, ignore_errors=True)
