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
from concurrent.futures import ThreadPoolExecutor

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

HIP_SOURCES = ["mkf_api.hip", "mkf_batch.hip", "mkf_serve.hip", "mkf_stream.hip"]
# mkf_kernels.hip is compiled once per part (-DMKF_KPART=i), each part holding
# a disjoint set of launchers and their kernel instantiations.
KERNEL_SOURCE = "mkf_kernels.hip"
KERNEL_PARTS = 9
C_SOURCES = ["binmap_util.c", "measures.c"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False, jobs=None):
    """Compiles every source (in parallel, one compiler process per source)
    and links the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for part in range(KERNEL_PARTS):
        obj = os.path.join(BUILD, f"{KERNEL_SOURCE}.p{part}.o")
        cmds.append([HIPCC, *HIP_FLAGS, f"-DMKF_KPART={part}", "-c", os.path.join(CSRC, KERNEL_SOURCE), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    n = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose=False):
    tmp = ORACLE_LIB + ".tmp"
    # x86-64-v3 (AVX2, BMI2, FMA) rather than -march=native: the library is
    # built here and run on the GPU machine's host, whose CPU may differ.
    _run(["gcc", "-O3", "-march=x86-64-v3", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-pthread",
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
