"""The allocation-free fused histogram + moments entry point
(mkfFusedBPFDMomentsNoAlloc) against the allocating one
(mkfFusedBPFDMoments): two calls in a row, the second on a smaller field, give
bit-identical histograms and moments (the same kernels and fixed-order
sums), the histograms equal the oracle's, and the device's free memory does
not move across the calls. A launch that would need more partials than the
static scratch holds is refused before any work is queued, and the next call
still works."""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M
from mkf_amd import synth

pytestmark = pytest.mark.gpu


def run(lib, fn, field, bm, tiling=None):
    Z, Y, X = field.shape
    h = torch.zeros(256, dtype=torch.int64, device=field.device)
    mom = torch.zeros(M.MKF_MOM_N, dtype=torch.float64, device=field.device)
    fi = M.FieldInfo(field.data_ptr(), (X, Y, Z), M.ENC_F32 if field.dtype == torch.float32 else M.ENC_F64)
    t = M.MkfTiling(*tiling) if tiling is not None else None
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(field.device)[0]
    rc = fn(h.data_ptr(), mom.data_ptr(), None, fi.ref, C.byref(bm), C.byref(t) if t else None, None)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(field.device)[0]
    return rc, h.cpu().numpy().view(np.uint64).copy(), mom.cpu().numpy().copy(), free0 - free1


def test_no_alloc_matches_allocating_entry_point(lib, orc, cuda):
    # the larger first; the persistent kernel's moments form at 16 and 8
    # voxels per lane (its partials are per resident workgroup, not per tile)
    cases = [((96, 130, 700), torch.float32, (">=", 0.1), None),
             ((30, 41, 257), torch.float64, ("<", -0.2), None),
             ((40, 60, 1025), torch.float32, (">=", 0.0), (4, 16, 4, 16)),
             ((20, 30, 600), torch.float32, ("<=", 0.3), (2, 16, 4, 8)),
             ((1, 50, 70), torch.float32, (">", 0.0), None)]  # no cells: moments from the map pass
    for k, (shape, dt, lv, tiling) in enumerate(cases):
        print(f"no-alloc case {k}: {shape} {dt} {lv} {tiling}", flush=True)
        f = synth.grf_field(shape, 2.0, seed=90 + k, device=cuda, dtype=dt)
        bm = M.bin_map(*lv)
        rc, h_ref, m_ref, _ = run(lib, lib.mkfFusedBPFDMoments, f, bm, tiling)
        M.check(rc, "mkfFusedBPFDMoments")
        for rep in range(2):
            rc, h, m, dfree = run(lib, lib.mkfFusedBPFDMomentsNoAlloc, f, bm, tiling)
            M.check(rc, "mkfFusedBPFDMomentsNoAlloc")
            np.testing.assert_array_equal(h, h_ref, err_msg=f"case {k} call {rep}")
            assert np.array_equal(m.view(np.uint64), m_ref.view(np.uint64)), f"moments differ: case {k} call {rep}"
            if k > 0 or rep > 0:
                assert dfree == 0, f"device free memory moved by {dfree} bytes"
        ref, _ = orc.bpfd_field(f.cpu().numpy(), bm.t[0], bm.m)
        if min(shape) >= 2:
            np.testing.assert_array_equal(h, ref)
        assert m[4] + m[5] == f.numel()


def test_no_alloc_refuses_launches_beyond_its_scratch(lib, cuda):
    f = synth.grf_field((700, 512, 256), 2.0, seed=95, device=cuda, dtype=torch.float32)
    bm = M.bin_map(">=", 0.0)
    # 1 plane pair per tile: 699 x 128 tiles of partials, more than fit
    rc, h, m, _ = run(lib, lib.mkfFusedBPFDMomentsNoAlloc, f, bm, (1, 16, 4, 4))
    assert rc == M.MKF_ERR_BUDGET and "scratch" in M.last_error()
    assert not h.any() and not m.any()
    rc, h, m, _ = run(lib, lib.mkfFusedBPFDMomentsNoAlloc, f, bm)
    M.check(rc, "mkfFusedBPFDMomentsNoAlloc after a refusal")
    rc, h_ref, m_ref, _ = run(lib, lib.mkfFusedBPFDMoments, f, bm)
    M.check(rc)
    np.testing.assert_array_equal(h, h_ref)
    assert np.array_equal(m.view(np.uint64), m_ref.view(np.uint64))
