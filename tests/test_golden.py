"""Known-answer fixtures (tests/golden/known_answers.json, written by
tests/golden/make_fixtures.py with plain loops from the definitions): the
oracle on the CPU and the library's kernels on the GPU reproduce every packed
binary map and histogram bit for bit."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import mkf_amd as M

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")
CASES = json.load(open(PATH))["cases"]


def field_of(case):
    X, Y, Z = case["dims"]
    vals = [float(v) for v in case["field"]]  # "nan", "inf", "-inf" parse as floats
    return np.array(vals, dtype=np.float32 if case["dtype"] == "f32" else np.float64).reshape(Z, Y, X)


def ids(case):
    return f"{case['dims']}-{case['dtype']}-{case['relop'] or 'none'}"


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_oracle_reproduces_known_answers(orc, case):
    f = field_of(case)
    bm = M.bin_map(case["relop"], case["t"])
    assert bm.m == case["mask"]
    X = case["dims"][0]
    words = orc.pbm_c(f, bm.t[0], bm.m)
    np.testing.assert_array_equal(words.ravel(), np.array(case["pbm"], dtype=np.uint32))
    np.testing.assert_array_equal(orc.pbm_numpy(f, bm.t[0], bm.m).ravel(), np.array(case["pbm"], dtype=np.uint32))
    hist = orc.bpfd_chunk(words, X) if X >= 32 else orc.bpfd_direct_c(words, X)
    np.testing.assert_array_equal(hist, np.array(case["hist"], dtype=np.uint64))
    np.testing.assert_array_equal(orc.bpfd_numpy(orc.threshold_bits(f, bm.t[0], bm.m)),
                                  np.array(case["hist"], dtype=np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=ids)
def test_gpu_reproduces_known_answers(lib, cuda, case):
    import torch

    from tests.gpu_util import dev_bpfd_from_pbm, dev_fused, dev_pbm

    print(f"known answer {ids(case)}", flush=True)
    field = torch.from_numpy(field_of(case)).to(cuda)
    want_w = np.array(case["pbm"], dtype=np.uint32)
    want_h = np.array(case["hist"], dtype=np.uint64)
    words, dw, org, _ = dev_pbm(lib, field, case["relop"], case["t"])
    np.testing.assert_array_equal(words.ravel(), want_w)
    h, _ = dev_bpfd_from_pbm(lib, dw, org)
    np.testing.assert_array_equal(h, want_h)
    h, _ = dev_fused(lib, field, case["relop"], case["t"])
    np.testing.assert_array_equal(h, want_h)
