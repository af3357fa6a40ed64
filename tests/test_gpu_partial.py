"""Partial PBM organisations (profID & 0x0F != 0; src/binMapUtil.c:60-87):
setBMO describes a buffer of min(2, Y) rows of min(2, Z) planes, and
binMapCUDA maps exactly those rows and planes of the field at the table's
pointer. Stepping the pointer over every row pair (y, y+1) of every plane pair
(z, z+1) and adding each window's histogram (mkfCUDAGetBPFD) gives the whole
field's histogram; each window's map words equal the oracle's map of those
four rows. The reference's own kernels write the whole field into this
buffer, so there is no reference result to compare against beyond the
organisation itself."""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,dtype,prof", [((9, 11, 70), torch.float32, 0x01), ((7, 6, 33), torch.float64, 0x02),
                                              ((5, 8, 1030), torch.float32, 0x21), ((6, 5, 40), torch.float64, 0x31)])
def test_partial_windows_sum_to_the_field(lib, orc, cuda, shape, dtype, prof):
    Z, Y, X = shape
    g = torch.Generator(device=cuda).manual_seed(X + Y)
    f = torch.randn(shape, generator=g, device=cuda, dtype=dtype)
    bm = M.bin_map(">=", 0.1)
    host = f.cpu().numpy()
    full = orc.bpfd_field(host, bm.t[0], bm.m)[0]
    ref_words = orc.pbm_c(host, bm.t[0], bm.m)
    org = M.BMOrg()
    nw = lib.setBMO(C.byref(org), M.dims((X, Y, Z)), bytes([prof]))
    rowWS = (X + 31) // 32
    assert nw == 4 * rowWS and (org.rowPairs, org.planePairs, org.rowWS, org.planeWS) == (1, 1, rowWS, 2 * rowWS)
    words = torch.full((nw + 64,), -1, dtype=torch.int32, device=cuda)  # guard words after the buffer
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    eb = f.element_size()
    enc = M.ENC_F32 if dtype == torch.float32 else M.ENC_F64
    print(f"partial windows {shape} {dtype} profID 0x{prof:02x}", flush=True)
    for z in range(Z - 1):
        for y in range(Y - 1):
            fi = M.FieldInfo(f.data_ptr() + (z * Y * X + y * X) * eb, (X, Y, Z), enc, prof)
            o = M.BMOrg()
            assert lib.binMapCUDA(None, words.data_ptr(), C.byref(o), fi.ref, C.byref(bm)) == words.data_ptr(), \
                M.last_error()
            assert (o.rowPairs, o.planePairs) == (1, 1)
            assert lib.mkfCUDAGetBPFD(None, h.data_ptr(), C.byref(o), words.data_ptr(), 2 + (y & 1)) == h.data_ptr()
            if (y + z) % 5 == 0:
                got = words.cpu().numpy().view(np.uint32)
                win = ref_words[z:z + 2, y:y + 2, :].reshape(-1)
                np.testing.assert_array_equal(got[:nw], win, err_msg=f"window y={y} z={z}")
                assert np.all(got[nw:] == 0xFFFFFFFF), "words written past the partial buffer"
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), full)


def test_partial_library_buffer_and_flat_fields(lib, orc, cuda):
    """The library-owned buffer (pW NULL) takes the partial organisation too,
    and a field of one plane has no plane pair: its window has no cells."""
    X, Y, Z = 64, 3, 4
    f = torch.rand((Z, Y, X), device=cuda, dtype=torch.float32)
    bm = M.bin_map("<", 0.5)
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32, 0x01)
    org = M.BMOrg()
    w = lib.binMapCUDA(None, None, C.byref(org), fi.ref, C.byref(bm))
    assert w, M.last_error()
    h = np.zeros(256, dtype=np.uint64)
    assert lib.mkfCUDAGetBPFDH(None, h.ctypes.data, C.byref(org), w, M.MKFCU_PROFILE_FAST)
    host = f.cpu().numpy()[:2, :2, :]
    np.testing.assert_array_equal(h, orc.bpfd_field(np.ascontiguousarray(host), bm.t[0], bm.m)[0])
    assert int(h.sum()) == X - 1
    flat = torch.rand((1, 5, 50), device=cuda, dtype=torch.float32)
    fi = M.FieldInfo(flat.data_ptr(), (50, 5, 1), M.ENC_F32, 0x01)
    org = M.BMOrg()
    w = lib.binMapCUDA(None, None, C.byref(org), fi.ref, C.byref(bm))
    assert w and org.planePairs == 0
    h[:] = 0
    assert lib.mkfCUDAGetBPFDH(None, h.ctypes.data, C.byref(org), w, M.MKFCU_PROFILE_FAST)
    assert int(h.sum()) == 0
    lib.binMapCUDACleanup()
