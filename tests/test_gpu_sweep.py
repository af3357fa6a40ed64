"""GPU parity sweep: seeded random shapes, relations, thresholds and tilings.
Every kernel variant of the fused histogram (1, 4, 8 or 16 voxels per lane
for floats; 1 or 2 for doubles) and of the map histogram (1, 4 or 16) must equal
the oracle bit for bit, and the map must equal the oracle's words.

Row lengths cluster around the 64-, 256- and 1024-voxel segment edges (one
voxel short, exact, one or a few over), where the kernels switch between
whole and partial segments and read the voxel right of a segment."""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M
from tests.gpu_util import dev_fused, dev_pbm, host_np

pytestmark = pytest.mark.gpu

RELOPS = ["<", "=", "<=", ">", "!=", ">=", "", "!"]
EDGES = [64, 128, 256, 512, 1024, 2048, 3072]


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    if rng.random() < 0.7:
        X = int(rng.choice(EDGES)) + int(rng.choice([-1, 0, 1, 2, 3, 5]))
    else:
        X = int(rng.integers(2, 3100))
    X = max(2, X)
    Y = int(rng.integers(2, 10))
    Z = int(rng.integers(2, 10))
    f32 = bool(rng.random() < 0.7)
    relop = RELOPS[int(rng.integers(0, len(RELOPS)))]
    thr = float(rng.choice([0.0, 0.3, -0.4, 1.5]))
    vs = [1, 4, 8, 16] if f32 else [1, 2]
    V = int(rng.choice(vs))
    Ty = int(rng.choice([2, 4] if V >= 8 else [2, 4, 8]))
    TP = int(rng.integers(1, 13))
    waves = int(rng.integers(1, 17))
    vb = int(rng.choice([1, 4, 16]))
    Tyb = int(rng.choice([2, 4] if vb == 16 else [2, 4, 8]))
    return (X, Y, Z), f32, relop, thr, (TP, waves, Ty, V), (int(rng.integers(1, 13)), waves, Tyb, vb), seed


CASES = [_case(s) for s in range(160)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"s{c[-1]}")
def test_random_case(lib, orc, cuda, case):
    (X, Y, Z), f32, relop, thr, tiling, btiling, seed = case
    print(f"sweep seed {seed}: dims {(X, Y, Z)} {'f32' if f32 else 'f64'} {relop!r} {thr} "
          f"fused {tiling} bpfd {btiling}", flush=True)
    g = torch.Generator(device=cuda).manual_seed(seed)
    f = torch.randn((Z, Y, X), generator=g, device=cuda, dtype=torch.float64)
    # smooth along x so that uniform runs and mixed cells both occur
    f = (f + torch.roll(f, 1, 2) + torch.roll(f, 2, 2)) / 3.0
    if f32:
        f = f.to(torch.float32)
    f_np = host_np(f)
    bm = M.bin_map(relop, thr)
    words_ref = orc.pbm_c(f_np, bm.t[0], bm.m)
    h_ref = orc.bpfd_chunk(words_ref, X) if X >= 32 else orc.bpfd_direct_c(words_ref, X)
    assert int(h_ref.sum()) == (X - 1) * (Y - 1) * (Z - 1)

    h, _ = dev_fused(lib, f, relop, thr, tiling=tiling)
    np.testing.assert_array_equal(h, h_ref)

    words, dw, org, _ = dev_pbm(lib, f, relop, thr)
    np.testing.assert_array_equal(words, words_ref)
    hb = torch.zeros(256, dtype=torch.int64, device=cuda)
    t = M.MkfTiling(*btiling)
    M.check(lib.mkfBPFDAsync(hb.data_ptr(), C.byref(org), dw.data_ptr(), C.byref(t), None), "mkfBPFDAsync")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hb.cpu().numpy().view(np.uint64), h_ref)
