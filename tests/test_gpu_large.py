"""A field with more than 2^32 cells (2050 x 2048 x 1100 float32, 18.5 GB of
HBM): every count the kernels keep must be 64-bit where it sums over the
grid. At this size the oracle checks a slab, and the whole field is checked
by properties that hold at any size:
* the histogram's total is the exact cell count (X-1)(Y-1)(Z-1) > 2^32;
* '>=' and '<' at one threshold give bin-reversed histograms (complement);
* the drop-in pair (PBM, then its histogram) equals the fused kernel;
* the phase-moment counts add up to the voxel count.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M

pytestmark = pytest.mark.gpu

X, Y, Z = 2050, 2048, 1100


@pytest.fixture(scope="module")
def big(cuda):
    free, _ = torch.cuda.mem_get_info(cuda)
    need = X * Y * Z * 4 + 2 * (2 ** 30)
    if free < need:
        pytest.skip(f"needs {need / 2**30:.1f} GiB of free device memory, {free / 2**30:.1f} available")
    g = torch.Generator(device=cuda).manual_seed(7)
    f = torch.rand((Z, Y, X), generator=g, device=cuda, dtype=torch.float32)
    yield f
    del f
    torch.cuda.empty_cache()


def fused(lib, f, relop, thr, dims=(X, Y, Z)):
    h = torch.zeros(256, dtype=torch.int64, device=f.device)
    fi = M.FieldInfo(f.data_ptr(), dims, M.ENC_F32)
    bm = M.bin_map(relop, thr)
    M.check(lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bm), None, None), "mkfFusedBPFD")
    torch.cuda.synchronize()
    return h.cpu().numpy().view(np.uint64).copy()


def test_cell_count_beyond_32_bits_and_complement(lib, big):
    print("large field: fused >= and <", flush=True)
    cells = (X - 1) * (Y - 1) * (Z - 1)
    assert cells > 2 ** 32
    ge = fused(lib, big, ">=", 0.5)
    lt = fused(lib, big, "<", 0.5)
    assert int(ge.sum()) == cells and int(lt.sum()) == cells
    np.testing.assert_array_equal(lt, ge[::-1])
    # i.i.d. uniform voxels: every pattern has probability 1/256
    assert np.all(np.abs(ge.astype(np.float64) / cells - 1 / 256) < 1e-4)


def check_moments(m, big, inside):
    """The library's moments against torch's float64 sums of the same field,
    slab by slab (the field is uniform in [0, 1), so no sum cancels)."""
    want = np.zeros(6)
    for z0 in range(0, Z, 100):
        sl = big[z0:z0 + 100].double()
        sel1 = inside(sl)
        for ph, sel in ((1, sel1), (0, ~sel1)):
            v = torch.where(sel, sl, torch.zeros((), dtype=torch.float64, device=sl.device))
            want[2 * ph] += float(v.sum())
            want[2 * ph + 1] += float((v * v).sum())
            want[4 + ph] += int(sel.sum())
            del v
        del sl, sel1
    assert m[4] == want[4] and m[5] == want[5]
    np.testing.assert_allclose(m[:4], want[:4], rtol=1e-10)


def test_dropin_pair_and_moments_match_fused(lib, big, cuda):
    print("large field: drop-in pair and moments", flush=True)
    ref = fused(lib, big, ">", 0.25)
    fi = M.FieldInfo(big.data_ptr(), (X, Y, Z), M.ENC_F32)
    bm = M.bin_map(">", 0.25)
    org = M.BMOrg()
    nw = lib.setBMO(C.byref(org), M.dims((X, Y, Z)), b"\x00")
    words = torch.empty(nw, dtype=torch.int32, device=cuda)
    mom = torch.zeros(M.MKF_MOM_N, dtype=torch.float64, device=cuda)
    M.check(lib.mkfPBMMomentsAsync(words.data_ptr(), mom.data_ptr(), C.byref(org), fi.ref, C.byref(bm), None),
            "mkfPBMMomentsAsync")
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    M.check(lib.mkfBPFDAsync(h.data_ptr(), C.byref(org), words.data_ptr(), None, None), "mkfBPFDAsync")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), ref)
    m = mom.cpu().numpy()
    assert m[4] + m[5] == X * Y * Z
    check_moments(m, big, lambda v: v > 0.25)
    assert abs(m[5] / (X * Y * Z) - 0.75) < 1e-4  # P(u > 0.25)
    del words


def test_slab_against_oracle(lib, orc, big):
    """The first 6 planes, addressed in the same allocation, against the oracle."""
    print("large field: slab vs oracle", flush=True)
    zs = 6
    h = fused(lib, big, ">=", 0.5, dims=(X, Y, zs))
    f = big[:zs].cpu().numpy()
    bm = M.bin_map(">=", 0.5)
    np.testing.assert_array_equal(h, orc.bpfd_field(f, bm.t[0], bm.m)[0])


def test_fused_moments_and_batch_on_the_large_field(lib, big, cuda):
    """The one-pass histogram + moments, and the batched kernel (whose
    per-workgroup cell counter is flushed before it passes 2^31) with the
    large field next to a small one, give the fused kernel's histograms."""
    from mkf_amd.batch import Batch

    print("large field: fused moments and batch", flush=True)
    cells = (X - 1) * (Y - 1) * (Z - 1)
    ge = fused(lib, big, ">=", 0.5)
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    mom = torch.zeros(M.MKF_MOM_N, dtype=torch.float64, device=cuda)
    fi = M.FieldInfo(big.data_ptr(), (X, Y, Z), M.ENC_F32)
    bm = M.bin_map(">=", 0.5)
    M.check(lib.mkfFusedBPFDMoments(h.data_ptr(), mom.data_ptr(), None, fi.ref, C.byref(bm), None, None),
            "mkfFusedBPFDMoments")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), ge)
    m = mom.cpu().numpy()
    assert m[4] + m[5] == X * Y * Z
    check_moments(m, big, lambda v: v >= 0.5)
    small = torch.rand((40, 50, 60), generator=torch.Generator(device=cuda).manual_seed(3), device=cuda)
    # two levels: the per-level form; three of one family: the Gray-plane form
    ge25 = fused(lib, big, ">=", 0.25)
    b = Batch([big, big, small], [[(">=", 0.5), ("<", 0.5)], [("<", 0.5), (">=", 0.25), (">=", 0.5)],
                                  [(">=", 0.5)]])
    try:
        b.run()
        torch.cuda.synchronize()
        hb = b.hist(0).cpu().numpy().view(np.uint64)
        np.testing.assert_array_equal(hb[0], ge)
        np.testing.assert_array_equal(hb[1], ge[::-1])
        assert int(hb[0].sum()) == cells
        hg = b.hist(1).cpu().numpy().view(np.uint64)
        np.testing.assert_array_equal(hg[0], ge[::-1])
        np.testing.assert_array_equal(hg[1], ge25)
        np.testing.assert_array_equal(hg[2], ge)
        np.testing.assert_array_equal(b.hist(2).cpu().numpy().view(np.uint64)[0],
                                      fused(lib, small, ">=", 0.5, dims=(60, 50, 40)))
    finally:
        b.close()
