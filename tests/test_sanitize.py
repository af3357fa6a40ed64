"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, built
from the sources with gcc / g++ (CPU only; the GPU kernels are not built
here):
* the library's host C (binmap_util.c, measures.c) and the C oracle, driven
  by tests/c/sanitize_host.c: every host helper and the oracle's histogram
  paths run with no sanitizer report;
* the serving loop's first-fit allocator (mkf_amd/csrc/mkf_arena.hpp),
  driven by tests/c/arena_test.cpp against a brute-force granule model."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    srcs = [os.path.join(HERE, "c", "sanitize_host.c"), os.path.join(ROOT, "mkf_amd", "csrc", "binmap_util.c"),
            os.path.join(ROOT, "mkf_amd", "csrc", "measures.c"), os.path.join(ROOT, "oracle", "mkf_oracle.c")]
    cmd = ["gcc", "-O1", "-g", "-std=gnu11", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", f"-I{ROOT}/include", *srcs, "-o", exe, "-lm", "-pthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout and "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_serving_arena_against_a_model(tmp_path):
    exe = str(tmp_path / "arena_test")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", f"-I{ROOT}/mkf_amd/csrc", os.path.join(HERE, "c", "arena_test.cpp"),
           "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ)
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "arena: ok" in r.stdout
