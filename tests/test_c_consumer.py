"""A C program written against the reference's API (tests/c/consumer.c)
links libmkf_amd.so in place of the reference library: it compiles and links
with gcc here, and on a GPU its histograms and measures equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import mkf_amd as M

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_consumer(out_dir):
    exe = os.path.join(str(out_dir), "consumer")
    libdir = os.path.dirname(M.LIB_PATH)
    cmd = ["gcc", "-O2", "-std=gnu11", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{ROOT}/include",
           os.path.join(HERE, "c", "consumer.c"), "-o", exe, f"-L{libdir}", "-lmkf_amd", f"-Wl,-rpath,{libdir}",
           f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_c_consumer_builds(tmp_path):
    assert os.path.exists(build_consumer(tmp_path))


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(97, 61, 45), (40, 33, 20), (1025, 5, 4)])
def test_c_consumer_matches_oracle(tmp_path, orc, dims):
    exe = build_consumer(tmp_path)
    X, Y, Z = dims
    fb, ht = str(tmp_path / "field.bin"), str(tmp_path / "hist.txt")
    print(f"C consumer {dims}", flush=True)
    subprocess.run([exe, str(X), str(Y), str(Z), fb, ht], check=True, timeout=120)
    f = np.fromfile(fb, dtype=np.float32).reshape(Z, Y, X)
    lines = open(ht).read().splitlines()
    bm = M.bin_map(">=", 0.1)
    ref = orc.bpfd_field(f, bm.t[0], bm.m)[0]
    for line in lines[:3]:
        np.testing.assert_array_equal(np.array([int(v) for v in line.split()], dtype=np.uint64), ref)
    kv, vv, nm = lines[3].split()
    m, _ = orc.measures(ref, float(np.float32(3.0) / np.float32(X + Y + Z)))  # the consumer's float division
    assert int(nm) == 2
    assert np.float32(kv) == m[0] and np.float32(vv) == m[3]
